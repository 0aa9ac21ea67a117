"""Splice-junction end alignment composed from the batched kernels:
Dynprog_end5_splicejunction (dynprog_end.c:1652-1922) and
Dynprog_end3_splicejunction (2248-2745) for many ends at once.

Per end the reference fills the upper and lower triangles of the reversed
read against the reversed genome (ENDQ scores, end penalties by defect
rate, bands widened to the far corner, int8 when the read or the genome is
shorter than use8p_size, the opposite tie rule), takes the best cell of the
last row (find_best_endpoint_to_queryend_indels), and when that score is
not negative walks back twice: a local walk to the continuation column,
a known gap holder, and a local walk from where the first stopped to the
anchor.  It then trims indel pairs from the read's end.  The 3' form is
the same on the forward strings with the caller's own tie rule, its gap
holder spanning far - anchor, and its list returned in the trimmed order.
Here the whole caller runs in the library (gdp_sj_run): the fills keep only
their best last-row cell (GDP_KEEP_BEST_QUERYEND), one launch merges the
triangles, each set of walks is one traceback launch (GDP_TB_LOCAL), and a
join kernel writes each end's list around its gap holder.
"""
import numpy as np

from .lib import SJ_END_DTYPE, SJ_PAIR_DTYPE, SJ_RESULT_DTYPE



def end_penalties(defect_rate):
    """END_OPEN/END_EXTEND by defect rate (dynprog_end.c:88-94, 1688-1700;
    thresholds dynprog.h:58-59)."""
    if defect_rate < 0.003:
        return -10, -2
    if defect_rate < 0.014:
        return -8, -2
    return -6, -2


def wide_bands(rl, gl, extraband):
    """Dynprog_compute_bands with widebandp (dynprog.c:1245-1261): (lband, uband)."""
    return (extraband, gl - rl + extraband) if gl >= rl else (rl - gl + extraband, extraband)


def end5_splicejunction(ends, max_rlength=1000, max_glength=2000):
    """ends: sequence of dicts with rev_rseq, rev_gseq, rev_galt (bytes in
    the caller's orientation, read backwards from the last base), rev_roffset,
    anchor, far (rev_goffset_anchor / rev_goffset_far), genestrand,
    jump_late, extraband, defect_rate, contlength.  Returns per end None (the
    reference's NULL list) or (the returned list's pairs from its last
    element to its head -- the push order -- as gdp_pair records,
    {traceback_score, missscore, nmatches, nmismatches, nopens, nindels}).
    The records are SJ_PAIR_DTYPE: the known-site gap holder has gapp 2 and
    its 32-bit genomejump is the intron's length."""
    return _splicejunction(ends, True, max_rlength, max_glength)


def end3_splicejunction(ends, max_rlength=1000, max_glength=2000):
    """As end5_splicejunction for Dynprog_end3_splicejunction: the same keys
    with the strings forward (rsequence, gsequence) and rev_roffset, anchor,
    far standing for roffset, goffset_anchor, goffset_far.  The pairs come
    from the returned list's last element to its head: the push order
    reversed, since that caller returns its trimmed list unreversed."""
    return _splicejunction(ends, False, max_rlength, max_glength)


def pack_ends(ends, five):
    """The ends as gdp_sj_end rows and the query / genome byte buffers in
    sweep order (the 5' form's reversed strings forward)."""
    n = len(ends)
    arr = np.zeros(n, SJ_END_DTYPE)
    rl = np.array([len(e["rev_rseq"]) for e in ends], np.int64)
    gl = np.array([len(e["rev_gseq"]) for e in ends], np.int64)
    if five:
        qbuf = b"".join(e["rev_rseq"][::-1] for e in ends)
        gbuf = b"".join(e["rev_gseq"][::-1] + e["rev_galt"][::-1] for e in ends)
    else:
        qbuf = b"".join(e["rev_rseq"] for e in ends)
        gbuf = b"".join(e["rev_gseq"] + e["rev_galt"] for e in ends)
    arr["qoff"] = np.concatenate([[0], np.cumsum(rl)[:-1]]) if n else 0
    arr["goff"] = np.concatenate([[0], np.cumsum(2 * gl)[:-1]]) if n else 0
    arr["rl"], arr["gl"] = rl, gl
    for f, key in (("roffset", "rev_roffset"), ("anchor", "anchor"), ("far", "far"), ("contlength", "contlength"),
                   ("extraband", "extraband"), ("genestrand", "genestrand"), ("defect_rate", "defect_rate")):
        arr[f] = [e[key] for e in ends]
    arr["jump_late"] = [bool(e["jump_late"]) for e in ends]
    return arr, qbuf, gbuf


class SpliceBatch:
    """gdp_sj_run through one gdp_sj_batch (device buffers reused across
    runs): packed ends in, per-end results and the flat pair array out."""

    def __init__(self):
        from .batch import init
        self.L = init()
        self.h = self.L.gdp_sj_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.gdp_sj_destroy(self.h)
            self.h = None

    def run(self, five, arr, qbuf, gbuf, max_rlength=1000, max_glength=2000, stream=None, copy=True):
        """(per-end SJ_RESULT_DTYPE rows, SJ_PAIR_DTYPE pairs); with
        copy=False the pairs are a view of the batch's pinned buffer, valid
        until its next run."""
        import ctypes
        arr = np.ascontiguousarray(arr, SJ_END_DTYPE)
        q = np.frombuffer(bytes(qbuf) or b"\0", np.uint8)
        g = np.frombuffer(bytes(gbuf) or b"\0", np.uint8)
        rc = self.L.gdp_sj_run(self.h, int(five), arr.ctypes.data, len(arr), q.ctypes.data, len(qbuf), g.ctypes.data,
                               len(gbuf), max_rlength, max_glength, stream)
        if rc != 0:
            raise RuntimeError("gdp_sj_run failed (%d): %s" % (rc, self.L.gdp_last_error().decode()))
        n, npairs = len(arr), int(self.L.gdp_sj_npairs(self.h))
        res = np.ctypeslib.as_array(ctypes.cast(self.L.gdp_sj_results(self.h), ctypes.POINTER(ctypes.c_uint8)),
                                    (max(n, 1) * SJ_RESULT_DTYPE.itemsize,))[: n * SJ_RESULT_DTYPE.itemsize]
        res = res.view(SJ_RESULT_DTYPE).copy() if n else np.zeros(0, SJ_RESULT_DTYPE)
        if npairs:
            pp = np.ctypeslib.as_array(ctypes.cast(self.L.gdp_sj_pairs(self.h), ctypes.POINTER(ctypes.c_uint8)),
                                       (npairs * SJ_PAIR_DTYPE.itemsize,))
            pairs = pp.view(SJ_PAIR_DTYPE)
            if copy:
                pairs = pairs.copy()
        else:
            pairs = np.zeros(0, SJ_PAIR_DTYPE)
        return res, pairs


_batch = None


def _splicejunction(ends, five, max_rlength, max_glength):
    global _batch
    if _batch is None:
        _batch = SpliceBatch()
    arr, qbuf, gbuf = pack_ends(ends, five)
    res, pairs = _batch.run(five, arr, qbuf, gbuf, max_rlength, max_glength)
    out = [None] * len(ends)
    for k in np.nonzero(res["npairs"] >= 0)[0]:
        o, m = int(res["pair_off"][k]), int(res["npairs"][k])
        out[k] = (pairs[o:o + m], res["counters"][k].astype(np.int32))
    return out
