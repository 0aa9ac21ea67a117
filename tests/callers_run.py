#!/usr/bin/env python3
"""Runs calls of the reference's own DP callers (Dynprog_single_gap,
Dynprog_end5_gap, Dynprog_end3_gap, through oracle/ref_callers_shim.c) from
one of their two builds, over the genome-skip test genome, and writes the
returned pair lists and counters.

    python3 tests/callers_run.py {ref|gpu|ref_large|gpu_large} CASES.npz OUT.npz [genome]

With `genome` the cases are Dynprog_genome_gap calls (save_genome_cases),
with `cdna` Dynprog_cdna_gap calls (save_cdna_cases).

ref: oracle/_ref/libgmapref_callers.so (the reference's SIMD fills);
gpu: oracle/_ref/libgmapref_callers_gpu.so (the same callers over the
drop-in libgmapdp_dynprog.so, whose fills and walks run on the GPU);
ref_large, gpu_large: the same from oracle/_ref_large, GMAP's large-genome
configuration (64-bit Univcoord_T), the latter over
libgmapdp_dynprog_large.so.
Run as a process of its own: the two builds define the same symbols, and the
drop-in must resolve its host symbols against the callers' build alone.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gmap-gsnap_amd")]

MAXPAIRS = 8192


def callers_lib(which):
    d = "_ref_large" if which.endswith("_large") else "_ref"
    return os.path.join(ROOT, "oracle", d,
                        "libgmapref_callers.so" if which.startswith("ref") else "libgmapref_callers_gpu.so")


def run(which, cases_path, out_path):
    import golden_io
    lib = callers_lib(which)
    L = ctypes.CDLL(lib, mode=os.RTLD_LAZY)
    _p, _i, _u, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_double
    L.refc_init.argtypes = [_i, _p, _u]
    L.refc_new.argtypes = [_i] * 5
    L.refc_new.restype = _p
    L.refc_single_gap.argtypes = [_p, _p, _p, _i, _i, _i, _i, _u, _u, _i, _i, _i, _i, _i, _d, _p, _p, _i]
    L.refc_end_gap.argtypes = [_i, _p, _p, _p, _i, _i, _i, _i, _u, _u, _i, _i, _i, _i, _d, _i, _p, _p, _i]
    seq, gc = golden_io.skip_genome_image()
    L.refc_init(0, gc.ctypes.data, len(seq))
    # GMAP's default stage-3 workspaces: max_rlength 500, max_glength 2000
    ws1 = L.refc_new(60, 50, 20, 0, 0)
    ws2 = L.refc_new(60, 50, 20, 0, 1)
    z = np.load(cases_path, allow_pickle=False)
    c = {k: z[k] for k in z.files}
    n = len(c["rl"])
    npairs = np.zeros(n, np.int64)
    counters = np.zeros((n, 6), np.int32)
    pairs = []
    out = np.zeros((MAXPAIRS, 8), np.int32)
    for k in range(n):
        q = bytes(c["rbuf"][c["roff"][k]:c["roff"][k] + c["rl"][k]]) + b"\0"
        quc = q.upper()
        qb, qub = ctypes.create_string_buffer(q, len(q)), ctypes.create_string_buffer(quc, len(quc))
        cnt = np.zeros(6, np.int32)
        args = (c["rl"][k], c["gl"][k], c["roffset"][k], c["goffset"][k], c["chroffset"][k], c["chrhigh"][k],
                c["watsonp"][k], c["genestrand"][k], c["jump_late"][k], c["extraband"][k])
        args = [int(a) for a in args]
        if c["single"][k]:
            m = L.refc_single_gap(ws1, qb, qub, *args, int(c["widebandp"][k]), float(c["defect_rate"][k]),
                                  cnt.ctypes.data, out.ctypes.data, MAXPAIRS)
        else:
            m = L.refc_end_gap(int(c["five"][k]), ws2, qb, qub, *args, float(c["defect_rate"][k]),
                               int(c["endalign"][k]), cnt.ctypes.data, out.ctypes.data, MAXPAIRS)
        npairs[k] = m
        counters[k] = cnt
        if m > 0:
            pairs.append(out[:m].copy())
        print_every = max(1, n // 4)
        if k % print_every == 0:
            print("callers_run %s: %d of %d" % (which, k, n), flush=True)
    np.savez_compressed(out_path, npairs=npairs, counters=counters,
                        pairs=np.concatenate(pairs) if pairs else np.zeros((0, 8), np.int32))


GENOME_KEYS = ("rl", "glL", "glR", "roffset", "goffsetL", "rev_goffsetR", "chroffset", "chrhigh", "cdna_direction",
               "watsonp", "genestrand", "jump_late", "extraband", "maxpeelback", "halfp", "finalp")


def run_genome(which, cases_path, out_path):
    """Dynprog_genome_gap calls (helpers.genome_gap_calls) through refc_genome_gap."""
    import golden_io
    L = ctypes.CDLL(callers_lib(which), mode=os.RTLD_LAZY)
    _p, _i, _u, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_double
    L.refc_init.argtypes = [_i, _p, _u]
    L.refc_new.argtypes = [_i] * 5
    L.refc_new.restype = _p
    L.refc_genome_gap.argtypes = [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _u, _u, _i, _i, _i, _i, _i, _d, _i, _i, _i,
                                  _p, _p, _p, _i]
    seq, gc = golden_io.skip_genome_image()
    L.refc_init(0, gc.ctypes.data, len(seq))
    wsL = L.refc_new(60, 50, 20, 0, 1)
    wsR = L.refc_new(60, 50, 20, 0, 1)
    z = np.load(cases_path, allow_pickle=False)
    c = {k: z[k] for k in z.files}
    n = len(c["rl"])
    npairs = np.zeros(n, np.int64)
    counters = np.zeros((n, 10), np.int32)
    probs = np.zeros((n, 2), np.float64)
    pairs = []
    out = np.zeros((MAXPAIRS, 8), np.int32)
    for k in range(n):
        q = bytes(c["rbuf"][c["roff"][k]:c["roff"][k] + c["rl"][k]]) + b"\0"
        qb, qub = ctypes.create_string_buffer(q, len(q)), ctypes.create_string_buffer(q.upper(), len(q))
        cnt = np.zeros(10, np.int32)
        pr = np.zeros(2, np.float64)
        a = [int(c[key][k]) for key in GENOME_KEYS]
        # GENOME_KEYS in the call's order, defect_rate between extraband and maxpeelback
        m = L.refc_genome_gap(wsL, wsR, qb, qub, *a[:13], float(c["defect_rate"][k]), *a[13:], cnt.ctypes.data,
                              pr.ctypes.data, out.ctypes.data, MAXPAIRS)
        npairs[k] = m
        counters[k] = cnt
        probs[k] = pr
        if m > 0:
            pairs.append(out[:m].copy())
        if k % max(1, n // 4) == 0:
            print("callers_run %s genome: %d of %d" % (which, k, n), flush=True)
    np.savez_compressed(out_path, npairs=npairs, counters=counters, probs=probs,
                        pairs=np.concatenate(pairs) if pairs else np.zeros((0, 8), np.int32))


CDNA_KEYS = ("qdp5", "qdp3", "rlengthL", "rlengthR", "glength", "goffset", "chroffset", "chrhigh", "watsonp",
             "genestrand", "jump_late", "extraband")


def run_cdna(which, cases_path, out_path):
    """Dynprog_cdna_gap calls (helpers.cdna_gap_calls) through refc_cdna_gap."""
    import golden_io
    L = ctypes.CDLL(callers_lib(which), mode=os.RTLD_LAZY)
    _p, _i, _u, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_double
    L.refc_init.argtypes = [_i, _p, _u]
    L.refc_new.argtypes = [_i] * 5
    L.refc_new.restype = _p
    L.refc_cdna_gap.argtypes = [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _u, _u, _i, _i, _i, _i, _d, _p, _p, _i]
    seq, gc = golden_io.skip_genome_image()
    L.refc_init(0, gc.ctypes.data, len(seq))
    wsL = L.refc_new(60, 50, 20, 0, 1)
    wsR = L.refc_new(60, 50, 20, 0, 1)
    z = np.load(cases_path, allow_pickle=False)
    c = {k: z[k] for k in z.files}
    n = len(c["qlen"])
    npairs = np.zeros(n, np.int64)
    counters = np.zeros((n, 3), np.int32)
    pairs = []
    out = np.zeros((MAXPAIRS, 8), np.int32)
    for k in range(n):
        q = bytes(c["qbuf"][c["qoff"][k]:c["qoff"][k] + c["qlen"][k]]) + b"\0"
        qb, qub = ctypes.create_string_buffer(q, len(q)), ctypes.create_string_buffer(q.upper(), len(q))
        cnt = np.zeros(3, np.int32)
        a = [int(c[key][k]) for key in CDNA_KEYS]
        m = L.refc_cdna_gap(wsL, wsR, qb, qub, *a, float(c["defect_rate"][k]), cnt.ctypes.data, out.ctypes.data,
                            MAXPAIRS)
        npairs[k] = m
        counters[k] = cnt
        if m > 0:
            pairs.append(out[:m].copy())
        if k % max(1, n // 4) == 0:
            print("callers_run %s cdna: %d of %d" % (which, k, n), flush=True)
    np.savez_compressed(out_path, npairs=npairs, counters=counters,
                        pairs=np.concatenate(pairs) if pairs else np.zeros((0, 8), np.int32))


def save_cdna_cases(path, cases):
    """The cases of helpers.cdna_gap_calls as arrays."""
    arr = {k: np.array([int(c[k]) for c in cases], np.int64) for k in CDNA_KEYS}
    arr["defect_rate"] = np.array([c["defect_rate"] for c in cases], np.float64)
    arr["qlen"] = np.array([len(c["query"]) for c in cases], np.int64)
    arr["qoff"] = np.concatenate([[0], np.cumsum(arr["qlen"])[:-1]]).astype(np.int64)
    arr["qbuf"] = np.frombuffer(b"".join(c["query"] for c in cases), np.uint8)
    np.savez_compressed(path, **arr)


def save_genome_cases(path, cases):
    """The cases of helpers.genome_gap_calls as arrays."""
    arr = {k: np.array([int(c[k]) for c in cases], np.int64) for k in GENOME_KEYS}
    arr["defect_rate"] = np.array([c["defect_rate"] for c in cases], np.float64)
    arr["roff"] = np.concatenate([[0], np.cumsum(arr["rl"])[:-1]]).astype(np.int64)
    arr["rbuf"] = np.frombuffer(b"".join(c["rseq"] for c in cases), np.uint8)
    np.savez_compressed(path, **arr)


def save_cases(path, cases):
    """The cases of helpers.caller_cases as arrays."""
    keys = ("single", "five", "rl", "gl", "roffset", "goffset", "chroffset", "chrhigh", "watsonp", "genestrand",
            "jump_late", "extraband", "widebandp", "endalign")
    arr = {k: np.array([int(c[k]) for c in cases], np.int64) for k in keys}
    arr["defect_rate"] = np.array([c["defect_rate"] for c in cases], np.float64)
    arr["roff"] = np.concatenate([[0], np.cumsum(arr["rl"])[:-1]]).astype(np.int64)
    arr["rbuf"] = np.frombuffer(b"".join(c["rseq"] for c in cases), np.uint8)
    np.savez_compressed(path, **arr)


if __name__ == "__main__":
    mode = sys.argv[4] if len(sys.argv) > 4 else ""
    {"genome": run_genome, "cdna": run_cdna}.get(mode, run)(sys.argv[1], sys.argv[2], sys.argv[3])
