"""GMAP's single-gap and end-gap callers for many gaps at once:
Dynprog_single_gap (dynprog_single.c:428-675) run by the library
(gdp_sg_run), and Dynprog_end5_gap / Dynprog_end3_gap (dynprog_end.c:
1293-1650, 1923-2243) run likewise (gdp_eg_run).

Per call the reference picks the mismatch type and gap penalties by defect
rate, cuts the genome window (reverse complement on the minus strand), takes
single_gap_simple's ungapped list when the window is as long as the read and
that list has at most one mismatch, and otherwise fills the band of
Dynprog_compute_bands with the int8 or int16 kernel and walks back from the
corner.  Here the windows are cut from a device genome, both lists are made
on the device and a join kernel keeps the one the reference returns.

The end callers fill both triangles of the read end against its window
(ENDQ scores, end penalties by defect rate), take the best end cell of the
mode asked (or the nogaps diagonal), walk back from it, drop the result when
it has too many mismatches, and trim indels from the far end; all of it on
the device.

The cDNA gap caller (Dynprog_cdna_gap, dynprog_cdna.c:786-1180, run by
gdp_cd_run) fills both ends' triangles with every score kept, bridges them
on the device, walks both ends and joins the walks around the insertion
pairs or a gap holder, as the reference assembles its list.
"""
import ctypes

import numpy as np

from .batch import _check, init
from .lib import (CD_CALL_DTYPE, CD_RESULT_DTYPE, EG_CALL_DTYPE, EG_RESULT_DTYPE, PAIR_DTYPE, SG_CALL_DTYPE,
                  SG_RESULT_DTYPE)

_batch = None


def _view(ptr, n, dtype, copy=True):
    """n records of dtype at a library pointer, copied or as a view."""
    if n == 0:
        return np.zeros(0, dtype)
    v = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8)), (n * dtype.itemsize,)).view(dtype)
    return v.copy() if copy else v


class SingleGapBatch:
    """A gdp_sg_batch: device and host buffers reused across runs."""

    def __init__(self):
        self.L = init()
        self.h = self.L.gdp_sg_create()

    def run(self, calls, qbuf, genome, max_rlength=500, max_glength=2000, user_penalties=None, stream=None,
            copy=True):
        """calls: SG_CALL_DTYPE rows; qbuf: the reads; genome: a
        serve.DeviceGenome.  user_penalties: (open, extend) for GMAP's
        --indel-open / --indel-extend, None for the defect-rate penalties.
        Returns (results SG_RESULT_DTYPE, pairs PAIR_DTYPE); with
        copy=False the pairs are a view of the batch's pinned buffer, valid
        until its next run."""
        L = self.L
        calls = np.ascontiguousarray(calls, SG_CALL_DTYPE)
        q = qbuf if isinstance(qbuf, np.ndarray) else np.frombuffer(bytes(qbuf), np.uint8)
        q = np.ascontiguousarray(q).view(np.uint8)
        uo, ue = user_penalties if user_penalties is not None else (0, 0)
        _check(L, L.gdp_sg_run(self.h, calls.ctypes.data, len(calls), q.ctypes.data, len(q), genome.ptr, genome.fmt,
                               genome.nbases, max_rlength, max_glength, 1 if user_penalties is not None else 0,
                               uo, ue, stream), "gdp_sg_run")
        res = _view(L.gdp_sg_results(self.h), len(calls), SG_RESULT_DTYPE)
        pairs = _view(L.gdp_sg_pairs(self.h), int(L.gdp_sg_npairs(self.h)), PAIR_DTYPE, copy)
        return res, pairs

    def close(self):
        if self.h:
            self.L.gdp_sg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def single_gap(calls, qbuf, genome, max_rlength=500, max_glength=2000, user_penalties=None):
    """Dynprog_single_gap for every call.  Returns per call None (the
    reference's NULL list, with traceback_score NEG_INFINITY_32 when the
    lengths do not fit the workspace) or (the returned list's pairs from its
    head as PAIR_DTYPE records, {traceback_score, nmatches, nmismatches,
    nopens, nindels}, simple)."""
    global _batch
    if _batch is None:
        _batch = SingleGapBatch()
    res, pairs = _batch.run(calls, qbuf, genome, max_rlength, max_glength, user_penalties)
    out = [None] * len(res)
    for k in np.nonzero(res["npairs"] >= 0)[0]:
        o, m = int(res["pair_off"][k]), int(res["npairs"][k])
        out[k] = (pairs[o:o + m], res["counters"][k].astype(np.int32), bool(res["simple"][k]))
    return out, res


_ebatch = None


class EndGapBatch:
    """A gdp_eg_batch: device and host buffers reused across runs."""

    def __init__(self):
        self.L = init()
        self.h = self.L.gdp_eg_create()

    def run(self, five, calls, qbuf, genome, max_rlength=500, max_glength=2000, user_penalties=None,
            require_pos_score=False, stream=None, copy=True):
        """calls: EG_CALL_DTYPE rows; returns (results EG_RESULT_DTYPE,
        pairs PAIR_DTYPE); copy as in SingleGapBatch.run."""
        L = self.L
        calls = np.ascontiguousarray(calls, EG_CALL_DTYPE)
        q = qbuf if isinstance(qbuf, np.ndarray) else np.frombuffer(bytes(qbuf), np.uint8)
        q = np.ascontiguousarray(q).view(np.uint8)
        uo, ue = user_penalties if user_penalties is not None else (0, 0)
        _check(L, L.gdp_eg_run(self.h, 1 if five else 0, calls.ctypes.data, len(calls), q.ctypes.data, len(q),
                               genome.ptr, genome.fmt, genome.nbases, max_rlength, max_glength,
                               1 if user_penalties is not None else 0, uo, ue, 1 if require_pos_score else 0,
                               stream), "gdp_eg_run")
        res = _view(L.gdp_eg_results(self.h), len(calls), EG_RESULT_DTYPE)
        pairs = _view(L.gdp_eg_pairs(self.h), int(L.gdp_eg_npairs(self.h)), PAIR_DTYPE, copy)
        return res, pairs

    def close(self):
        if self.h:
            self.L.gdp_eg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def end_gap(five, calls, qbuf, genome, max_rlength=500, max_glength=2000, user_penalties=None,
            require_pos_score=False):
    """Dynprog_end5_gap (five) or Dynprog_end3_gap for every call.  Returns
    (per call None or (pairs from the returned list's head, {traceback_score,
    nmatches, nmismatches, nopens, nindels}), the EG_RESULT_DTYPE rows)."""
    global _ebatch
    if _ebatch is None:
        _ebatch = EndGapBatch()
    res, pairs = _ebatch.run(five, calls, qbuf, genome, max_rlength, max_glength, user_penalties, require_pos_score)
    out = [None] * len(res)
    for k in np.nonzero(res["npairs"] >= 0)[0]:
        o, m = int(res["pair_off"][k]), int(res["npairs"][k])
        out[k] = (pairs[o:o + m], res["counters"][k].astype(np.int32))
    return out, res


class CdnaGapBatch:
    """A gdp_cd_batch: device and host buffers reused across runs."""

    def __init__(self):
        self.L = init()
        self.h = self.L.gdp_cd_create()

    def run(self, calls, qbuf, genome, max_rlength=500, max_glength=2000, stream=None, copy=True):
        """calls: CD_CALL_DTYPE rows; returns (results CD_RESULT_DTYPE, pairs
        PAIR_DTYPE); copy as in SingleGapBatch.run."""
        L = self.L
        calls = np.ascontiguousarray(calls, CD_CALL_DTYPE)
        q = qbuf if isinstance(qbuf, np.ndarray) else np.frombuffer(bytes(qbuf), np.uint8)
        q = np.ascontiguousarray(q).view(np.uint8)
        _check(L, L.gdp_cd_run(self.h, calls.ctypes.data, len(calls), q.ctypes.data, len(q), genome.ptr, genome.fmt,
                               genome.nbases, max_rlength, max_glength, stream), "gdp_cd_run")
        res = _view(L.gdp_cd_results(self.h), len(calls), CD_RESULT_DTYPE)
        pairs = _view(L.gdp_cd_pairs(self.h), int(L.gdp_cd_npairs(self.h)), PAIR_DTYPE, copy)
        return res, pairs

    def close(self):
        if self.h:
            self.L.gdp_cd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


_cbatch = None


def cdna_gap(calls, qbuf, genome, max_rlength=500, max_glength=2000):
    """Dynprog_cdna_gap for every call.  Returns (per call None or (pairs
    from the returned list's head, traceback_score, incompletep), the
    CD_RESULT_DTYPE rows)."""
    global _cbatch
    if _cbatch is None:
        _cbatch = CdnaGapBatch()
    res, pairs = _cbatch.run(calls, qbuf, genome, max_rlength, max_glength)
    out = [None] * len(res)
    for k in np.nonzero(res["npairs"] >= 0)[0]:
        o, m = int(res["pair_off"][k]), int(res["npairs"][k])
        out[k] = (pairs[o:o + m], int(res["traceback_score"][k]), bool(res["incomplete"][k]))
    return out, res
