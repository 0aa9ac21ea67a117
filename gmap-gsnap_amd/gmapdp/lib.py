"""ctypes binding of libgmapdp.so (declarations in include/gmapdp.h)."""
import ctypes
import os

import numpy as np

GDP_FULL, GDP_UPPER, GDP_LOWER, GDP_STD = 0, 1, 2, 3
STD_UPPER, STD_LOWER = 1, 2
HIGHQ, MEDQ, LOWQ, ENDQ = 0, 1, 2, 3

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path():
    return os.path.join(os.path.dirname(_HERE), "lib", "libgmapdp.so")


class Problem(ctypes.Structure):
    _fields_ = [("qoff", ctypes.c_int64), ("goff", ctypes.c_int64), ("galtoff", ctypes.c_int64),
                ("rl", ctypes.c_int32), ("gl", ctypes.c_int32), ("lband", ctypes.c_int16),
                ("uband", ctypes.c_int16), ("kind", ctypes.c_int8), ("width", ctypes.c_int8),
                ("mismatchtype", ctypes.c_int8), ("jump_late", ctypes.c_int8), ("open", ctypes.c_int8),
                ("extend", ctypes.c_int8), ("keep_scores", ctypes.c_int8), ("std_flags", ctypes.c_int8),
                ("saturation", ctypes.c_int32)]


class TbRequest(ctypes.Structure):
    _fields_ = [("fill", ctypes.c_int32), ("r", ctypes.c_int32), ("c", ctypes.c_int32),
                ("queryoffset", ctypes.c_int32), ("genomeoffset", ctypes.c_int32),
                ("chroffset", ctypes.c_uint32), ("chrhigh", ctypes.c_uint32), ("watsonp", ctypes.c_int8),
                ("revp", ctypes.c_int8), ("genestrand", ctypes.c_int8), ("flags", ctypes.c_int8),
                ("pair_off", ctypes.c_int64), ("max_pairs", ctypes.c_int32), ("endc", ctypes.c_int32)]


class Pair(ctypes.Structure):
    _fields_ = [("querypos", ctypes.c_int32), ("genomepos", ctypes.c_int32), ("cdna", ctypes.c_int8),
                ("comp", ctypes.c_int8), ("genome", ctypes.c_int8), ("genomealt", ctypes.c_int8),
                ("gapp", ctypes.c_int16), ("genomejump", ctypes.c_int16)]


PROBLEM_DTYPE = np.dtype([("qoff", "<i8"), ("goff", "<i8"), ("galtoff", "<i8"), ("rl", "<i4"), ("gl", "<i4"),
                          ("lband", "<i2"), ("uband", "<i2"), ("kind", "i1"), ("width", "i1"),
                          ("mismatchtype", "i1"), ("jump_late", "i1"), ("open", "i1"), ("extend", "i1"),
                          ("keep_scores", "i1"), ("std_flags", "i1"), ("saturation", "<i4")])
TBREQ_DTYPE = np.dtype([("fill", "<i4"), ("r", "<i4"), ("c", "<i4"), ("queryoffset", "<i4"),
                        ("genomeoffset", "<i4"), ("chroffset", "<u4"), ("chrhigh", "<u4"), ("watsonp", "i1"),
                        ("revp", "i1"), ("genestrand", "i1"), ("flags", "i1"), ("pair_off", "<i8"),
                        ("max_pairs", "<i4"), ("endc", "<i4")])
PAIR_DTYPE = np.dtype([("querypos", "<i4"), ("genomepos", "<i4"), ("cdna", "i1"), ("comp", "i1"),
                       ("genome", "i1"), ("genomealt", "i1"), ("gapp", "<i2"), ("genomejump", "<i2")])
# Pair records of the splice-junction end callers: a known-site gap holder
# (gapp 2, Pairpool_push_gapholder with knownp, pairpool.c:374-452) spans an
# intron, so genomejump is 32 bits here (Pair.genomejump is an int, pairdef.h).
SJ_PAIR_DTYPE = np.dtype([("querypos", "<i4"), ("genomepos", "<i4"), ("cdna", "i1"), ("comp", "i1"),
                          ("genome", "i1"), ("genomealt", "i1"), ("gapp", "<i2"), ("pad", "<i2"),
                          ("genomejump", "<i4")])
SJ_END_DTYPE = np.dtype([("qoff", "<i8"), ("goff", "<i8"), ("rl", "<i4"), ("gl", "<i4"), ("roffset", "<i4"),
                         ("anchor", "<i4"), ("far", "<i4"), ("contlength", "<i4"), ("extraband", "<i2"),
                         ("genestrand", "i1"), ("jump_late", "i1"), ("pad", "<i4"), ("defect_rate", "<f8")])
SJ_RESULT_DTYPE = np.dtype([("pair_off", "<i8"), ("npairs", "<i4"), ("counters", "<i4", (6,)), ("pad", "<i4")])
SG_CALL_DTYPE = np.dtype([("qoff", "<i8"), ("rl", "<i4"), ("gl", "<i4"), ("roffset", "<i4"), ("goffset", "<i4"),
                          ("chroffset", "<u4"), ("chrhigh", "<u4"), ("extraband", "<i4"), ("watsonp", "i1"),
                          ("genestrand", "i1"), ("jump_late", "i1"), ("widebandp", "i1"), ("defect_rate", "<f8")])
SG_RESULT_DTYPE = np.dtype([("pair_off", "<i8"), ("npairs", "<i4"), ("counters", "<i4", (5,)), ("simple", "<i4"),
                            ("pad", "<i4")])
EG_CALL_DTYPE = np.dtype([("qoff", "<i8"), ("rl", "<i4"), ("gl", "<i4"), ("roffset", "<i4"), ("goffset", "<i4"),
                          ("chroffset", "<u4"), ("chrhigh", "<u4"), ("extraband", "<i4"), ("watsonp", "i1"),
                          ("genestrand", "i1"), ("jump_late", "i1"), ("endalign", "i1"), ("defect_rate", "<f8")])
EG_RESULT_DTYPE = np.dtype([("pair_off", "<i8"), ("npairs", "<i4"), ("counters", "<i4", (5,)), ("finalscore", "<i4"),
                            ("advanced", "<i4")])
CD_CALL_DTYPE = np.dtype([("qoffL", "<i8"), ("qoffR", "<i8"), ("rlengthL", "<i4"), ("rlengthR", "<i4"),
                          ("glength", "<i4"), ("roffsetL", "<i4"), ("rev_roffsetR", "<i4"), ("goffset", "<i4"),
                          ("chroffset", "<u4"), ("chrhigh", "<u4"), ("extraband", "<i4"), ("watsonp", "i1"),
                          ("genestrand", "i1"), ("jump_late", "i1"), ("pad", "i1"), ("defect_rate", "<f8")])
CD_RESULT_DTYPE = np.dtype([("pair_off", "<i8"), ("npairs", "<i4"), ("traceback_score", "<i4"), ("incomplete", "<i4"),
                            ("advanced", "<i4"), ("finalscore", "<i4"), ("pad", "<i4")])
assert SG_CALL_DTYPE.itemsize == 48 and SG_RESULT_DTYPE.itemsize == 40
assert EG_CALL_DTYPE.itemsize == 48 and EG_RESULT_DTYPE.itemsize == 40
assert CD_CALL_DTYPE.itemsize == 64 and CD_RESULT_DTYPE.itemsize == 32
SEGMENT_DTYPE = np.dtype([("coord", "<u4"), ("bound", "<u4"), ("length", "<i4"), ("left", "i1"), ("revcomp", "i1"),
                          ("reverse", "<i2"), ("out_off", "<i8")])
ALIGN_REQ_DTYPE = np.dtype([("tag", "<u8"), ("read", "<u8"), ("rl", "<i4"), ("gl", "<i4"), ("roffset", "<i4"),
                            ("goffset", "<i4"), ("chroffset", "<u4"), ("chrhigh", "<u4"), ("extraband", "<i4"),
                            ("watsonp", "i1"), ("widebandp", "i1"), ("jump_late", "i1"), ("genestrand", "i1"),
                            ("mismatchtype", "i1"), ("open", "i1"), ("extend", "i1"), ("scoring", "i1"),
                            ("pad", "<i4"), ("defect_rate", "<f8")])
ALIGN_RES_DTYPE = np.dtype([("tag", "<u8"), ("status", "<i4"), ("finalscore", "<i4"), ("counts", "<i4", (6,)),
                            ("width", "<i4"), ("pad", "<i4"), ("pairs", "<u8"), ("t_submit", "<f8"),
                            ("t_done", "<f8")])
INTRON_GAP_DTYPE = np.dtype([("fills", "<i4", (4,)), ("span", "<i4"), ("cdna_direction", "i1"), ("finalp", "i1"),
                             ("halfp", "i1"), ("pad", "i1"), ("prob_off", "<i8")])
PLAN_PARAMS_DTYPE = np.dtype([("g16_max_span", "<i4"), ("g64_min_rows", "<i4"), ("mix_span", "<i4"), ("mix_order", "<i4")])
SERVER_CFG_DTYPE = np.dtype([("dir_budget_bytes", "<i8"), ("pair_capacity", "<i8"), ("max_batch", "<i4"),
                             ("slots", "<i4"), ("max_rl", "<i4"), ("max_gl", "<i4"), ("result_buffers", "<i4"),
                             ("long_rows", "<i4"), ("plan", PLAN_PARAMS_DTYPE)])
SERVER_STATS_DTYPE = np.dtype([("submitted", "<i8"), ("completed", "<i8"), ("batches", "<i8"),
                               ("dir_bytes_peak", "<i8"), ("inflight_peak", "<i4"), ("pad", "<i4"),
                               ("prep_seconds", "<f8"), ("retire_seconds", "<f8"), ("batch_seconds", "<f8"),
                               ("device_seconds", "<f8")])
assert PROBLEM_DTYPE.itemsize == 48 and TBREQ_DTYPE.itemsize == 48 and PAIR_DTYPE.itemsize == 16
assert SEGMENT_DTYPE.itemsize == 24 and ALIGN_REQ_DTYPE.itemsize == 64 and ALIGN_RES_DTYPE.itemsize == 72
GENOME_ASCII, GENOME_COMP = 0, 1

_lib = None


def runtime_first():
    """torch bundles its own libamdhip64.so.7 under the same soname as
    /opt/rocm's; whichever is loaded first serves the whole process, and torch
    cannot run on the other one.  Load torch's before any library of this
    package so that torch and the kernels share one HIP runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path=None):
    """Loads libgmapdp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path()
    if not os.path.exists(p):
        raise RuntimeError("libgmapdp.so not built (%s); run __graft_entry__.build()" % p)
    runtime_first()
    L = ctypes.CDLL(p)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sig = {
        "gdp_init": (ctypes.c_int, [ctypes.c_int]),
        "gdp_last_error": (ctypes.c_char_p, []),
        "gdp_descriptor_bytes": (i32, []),
        "gdp_plan": (ctypes.c_int, [vp, i32, vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "gdp_plan_ex": (ctypes.c_int, [vp, i32, i32, vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "gdp_plan_bytes": (i64, [i32]),
        "gdp_plan_tuned": (ctypes.c_int, [vp, i32, i32, vp, vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "gdp_graph_create": (vp, [vp, i32, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, i32, vp, i32, ctypes.c_uint32,
                                  vp, vp]),
        "gdp_graph_launch": (ctypes.c_int, [vp, vp]),
        "gdp_graph_begin": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "gdp_graph_end": (ctypes.c_int, [vp]),
        "gdp_best_endpoint_ex": (ctypes.c_int, [vp, vp, i32, i32, vp, vp, vp, vp, vp]),
        "gdp_graph_destroy": (None, [vp]),
        "gdp_plan_kernels": (ctypes.c_uint32, [vp, i32]),
        "gdp_fill_ex": (ctypes.c_int, [vp, i32, ctypes.c_uint32, vp, vp, vp, vp, vp, vp]),
        "gdp_align_each": (ctypes.c_int, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, ctypes.c_uint32, vp, vp,
                                          vp]),
        "gdp_cells": (i64, [vp]),
        "gdp_fill": (ctypes.c_int, [vp, i32, vp, vp, vp, vp, vp, vp]),
        "gdp_best_endpoint": (ctypes.c_int, [vp, vp, i32, i32, vp, vp, vp]),
        "gdp_bridge_cdna": (ctypes.c_int, [vp, vp, i32, vp, vp, vp, vp, vp]),
        "gdp_bridge_intron": (ctypes.c_int, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gdp_sj_create": (vp, []),
        "gdp_sj_destroy": (None, [vp]),
        "gdp_sj_run": (ctypes.c_int, [vp, i32, vp, i32, vp, i64, vp, i64, i32, i32, vp]),
        "gdp_sj_npairs": (i64, [vp]),
        "gdp_sj_results": (vp, [vp]),
        "gdp_sj_pairs": (vp, [vp]),
        "gdp_sg_create": (vp, []),
        "gdp_sg_destroy": (None, [vp]),
        "gdp_sg_run": (ctypes.c_int, [vp, vp, i32, vp, i64, vp, i32, ctypes.c_uint32, i32, i32, i32, i32, i32, vp]),
        "gdp_sg_npairs": (i64, [vp]),
        "gdp_sg_results": (vp, [vp]),
        "gdp_sg_pairs": (vp, [vp]),
        "gdp_eg_create": (vp, []),
        "gdp_eg_destroy": (None, [vp]),
        "gdp_eg_run": (ctypes.c_int, [vp, i32, vp, i32, vp, i64, vp, i32, ctypes.c_uint32, i32, i32, i32, i32, i32, i32,
                                      vp]),
        "gdp_cd_create": (vp, []),
        "gdp_cd_destroy": (None, [vp]),
        "gdp_cd_run": (ctypes.c_int, [vp, vp, i32, vp, i64, vp, i32, ctypes.c_uint32, i32, i32, vp]),
        "gdp_cd_npairs": (i64, [vp]),
        "gdp_cd_results": (vp, [vp]),
        "gdp_cd_pairs": (vp, [vp]),
        "gdp_eg_npairs": (i64, [vp]),
        "gdp_eg_results": (vp, [vp]),
        "gdp_eg_pairs": (vp, [vp]),
        "gdp_traceback": (ctypes.c_int, [vp, vp, i32, vp, vp, vp, vp, ctypes.c_uint32, vp, vp, vp]),
        "gdp_traceback_ex": (ctypes.c_int, [vp, vp, i32, vp, vp, vp, vp, i32, ctypes.c_uint32, vp, vp, vp]),
        "gdp_traceback_ends": (ctypes.c_int, [vp, vp, i32, vp, vp, vp, vp, i32, ctypes.c_uint32, vp, vp, vp, vp]),
        "gdp_genome_segments": (ctypes.c_int, [vp, i32, ctypes.c_uint32, vp, i32, vp, vp]),
        "gdp_download_results": (ctypes.c_int, [vp, i32, vp, vp, vp, vp, vp]),
        "gdp_genomecomp_words": (i64, [ctypes.c_uint64]),
        "gdp_genomecomp_pack": (ctypes.c_int, [vp, ctypes.c_uint64, vp]),
        "gdp_server_create": (vp, [vp, vp, i32, ctypes.c_uint32]),
        "gdp_server_submit": (ctypes.c_int, [vp, vp, i32]),
        "gdp_server_poll": (i32, [vp, vp, i32, i32]),
        "gdp_server_get_stats": (ctypes.c_int, [vp, vp]),
        "gdp_server_destroy": (None, [vp]),
        "gdp_server_now": (ctypes.c_double, []),
        "gdp_export_dense": (ctypes.c_int, [vp, i32, vp, vp, vp, vp, vp]),
        "gdp_pairdistance": (ctypes.POINTER(ctypes.c_int16), [ctypes.c_int]),
        "gdp_nt_to_int": (ctypes.POINTER(ctypes.c_int32), []),
        "gdp_use8p_size": (i32, [ctypes.c_int]),
        "gdp_consistent": (i32, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "gdp_compute_bands": (None, [ctypes.POINTER(i32), ctypes.POINTER(i32), i32, i32, i32, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L
