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
holder spanning far - anchor, and its list returned in the trimmed order.  Here the fills keep
only their best last-row cell (GDP_KEEP_BEST_QUERYEND), one launch merges
the triangles, and each set of walks is one traceback launch
(GDP_TB_LOCAL); the host joins the two walks.
"""
import numpy as np

from .batch import DeviceBatch
from .lib import PROBLEM_DTYPE, SJ_PAIR_DTYPE, TBREQ_DTYPE, load

GDP_KIND_UPPER, GDP_KIND_LOWER = 1, 2
GDP_KEEP_BEST_QUERYEND = 3
GDP_TB_LOCAL = 4
ENDQ = 3
FULLMATCH = 3  # dynprog.h:42
INDEL_COMP = ord("-")  # comp.h:9


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


def _splicejunction(ends, five, max_rlength, max_glength):
    L = load()
    use8 = int(L.gdp_use8p_size(ENDQ))
    out = [None] * len(ends)
    rl_all = np.array([len(e["rev_rseq"]) for e in ends], np.int64)
    gl_all = np.array([len(e["rev_gseq"]) for e in ends], np.int64)
    # dynprog_end.c:1702-1717: empty or oversized ends give NULL
    live = np.nonzero((rl_all > 0) & (rl_all <= max_rlength) & (gl_all > 0) & (gl_all <= max_glength))[0]
    if len(live) == 0:
        return out
    E = [ends[k] for k in live]
    rl, gl = rl_all[live], gl_all[live]
    defect = np.array([e["defect_rate"] for e in E])
    extra = np.array([e["extraband"] for e in E], np.int64)
    late = np.array([bool(e["jump_late"]) for e in E])
    roff = np.array([e["rev_roffset"] for e in E], np.int64)
    anchor = np.array([e["anchor"] for e in E], np.int64)
    far = np.array([e["far"] for e in E], np.int64)
    endc1 = np.array([e["contlength"] for e in E], np.int64)
    gs = np.array([e["genestrand"] for e in E], np.int64)
    # sweep order: the 5' form's reversed strings forward
    if five:
        qbuf = b"".join(e["rev_rseq"][::-1] for e in E)
        gbuf = b"".join(e["rev_gseq"][::-1] + e["rev_galt"][::-1] for e in E)
    else:
        qbuf = b"".join(e["rev_rseq"] for e in E)
        gbuf = b"".join(e["rev_gseq"] + e["rev_galt"] for e in E)
    qoff = np.concatenate([[0], np.cumsum(rl)[:-1]])
    goff = np.concatenate([[0], np.cumsum(2 * gl)[:-1]])
    # end penalties by defect rate (dynprog_end.c:88-94), bands widened to the far corner
    open_ = np.where(defect < 0.003, -10, np.where(defect < 0.014, -8, -6))
    lb = np.where(gl >= rl, extra, rl - gl + extra)
    ub = np.where(gl >= rl, gl - rl + extra, extra)
    probs = np.zeros(2 * len(live), PROBLEM_DTYPE)
    for j, kind in enumerate((GDP_KIND_UPPER, GDP_KIND_LOWER)):
        p = probs[j::2]
        p["qoff"], p["goff"], p["galtoff"] = qoff, goff, goff + gl
        p["rl"], p["gl"], p["lband"], p["uband"] = rl, gl, lb, ub
        p["kind"], p["mismatchtype"] = kind, ENDQ
        p["width"] = np.where((rl < use8) | (gl < use8), 8, 16)
        p["jump_late"] = late != five  # the 5' fills take the opposite tie rule
        p["open"], p["extend"], p["keep_scores"] = open_, -2, GDP_KEEP_BEST_QUERYEND
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    ul = np.stack([np.arange(0, 2 * len(live), 2), np.arange(1, 2 * len(live), 2)], 1).astype(np.int32)
    best = b.best_endpoint(ul, queryend_indels=True)
    sel = np.nonzero(best[:, 0] >= 0)[0]  # finalscore < 0: no splice (1797-1800)
    if len(sel) == 0:
        return out

    def requests(r, c, endc, goffs):
        req = np.zeros(len(sel), TBREQ_DTYPE)
        req["fill"] = 2 * sel + (c < r)  # the triangle holding the cell
        req["r"], req["c"] = r, c
        req["queryoffset"], req["genomeoffset"] = roff[sel], goffs
        req["revp"], req["genestrand"], req["watsonp"] = int(five), gs[sel], 1
        req["flags"], req["endc"] = GDP_TB_LOCAL, endc
        req["max_pairs"] = rl[sel] + gl[sel] + 4
        return req

    p1, c1, stop = b.traceback(requests(best[sel, 1], best[sel, 2], endc1[sel], far[sel]), ends=True)
    p2, c2 = b.traceback(requests(stop[:, 0], stop[:, 1], 0, anchor[sel]))
    cnt = c1[:, 1:] + c2[:, 1:]
    jump = (anchor - far) if five else (far - anchor)

    def widen(p):
        w = np.zeros(len(p), SJ_PAIR_DTYPE)
        for f in ("querypos", "genomepos", "cdna", "comp", "genome", "genomealt", "gapp", "genomejump"):
            w[f] = p[f]
        return w

    for j, i in enumerate(sel):
        gap = np.zeros(1, SJ_PAIR_DTYPE)  # Pairpool_push_gapholder, knownp (pairpool.c:374-452)
        gap["querypos"] = gap["genomepos"] = -1
        gap["cdna"] = gap["comp"] = gap["genome"] = gap["genomealt"] = ord(" ")
        gap["gapp"], gap["genomejump"] = 2, jump[i]
        pairs = np.concatenate([widen(p1[j]), gap, widen(p2[j])])
        # the reference reverses the list and drops indel pairs from the read's
        # end, which are the first pushed (1905-1910, 2485-2490); the 3' form
        # returns that trimmed list as it is
        k = 0
        while k < len(pairs) and pairs["comp"][k] == INDEL_COMP:
            k += 1
        out[live[i]] = (pairs[k:] if five else pairs[k:][::-1].copy(),
                        np.array([cnt[j, 0], cnt[j, 0] - rl[i] * FULLMATCH, cnt[j, 1], cnt[j, 2], cnt[j, 3],
                                  cnt[j, 4]], np.int32))
    return out


def _random_ends(n, seed=1007):
    """Splice-junction ends of the end-extension mix: read ends of 10-150
    bases against their genome window plus 0-20 bases (3 % substitutions,
    1 % indels), extraband 6, continuation columns up to the read length."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    ends = []
    for k in range(n):
        rl = int(rng.integers(10, 151))
        gl = rl + int(rng.integers(0, 21))
        g = acgt[rng.integers(0, 4, gl)]
        q = g[:rl].copy()
        sub = rng.random(rl) < 0.03
        q[sub] = acgt[rng.integers(0, 4, int(sub.sum()))]
        anchor = int(rng.integers(1000, 1 << 30))
        ends.append({"rev_rseq": q.tobytes(), "rev_gseq": g.tobytes(), "rev_galt": g.tobytes(),
                     "rev_roffset": rl + 100, "anchor": anchor, "far": anchor - int(rng.integers(50, 5000)),
                     "genestrand": 0, "jump_late": bool(k % 2), "extraband": 6, "defect_rate": 0.001,
                     "contlength": int(rng.integers(0, rl // 2 + 1))})
    return ends


if __name__ == "__main__":
    # python3 -m gmapdp.ends [N]: times end5_splicejunction over N ends of the
    # end-extension mix (after one untimed call of the same size).
    import sys
    import time

    import torch

    n = int(sys.argv[1]) if len(sys.argv) > 1 else 200_000
    ends = _random_ends(n)
    end5_splicejunction(ends)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = end5_splicejunction(ends)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    walked = sum(r is not None for r in res)
    print("end5_splicejunction: %d ends (%d walked) in %.3f s = %.0f ends/s" % (n, walked, dt, n / dt), flush=True)
