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
from .lib import PAIR_DTYPE, PROBLEM_DTYPE, TBREQ_DTYPE, load

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
    {traceback_score, missscore, nmatches, nmismatches, nopens, nindels})."""
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
    live = []
    probs, qparts, gparts = [], [], []
    qo = go = 0
    for k, e in enumerate(ends):
        rl, gl = len(e["rev_rseq"]), len(e["rev_gseq"])
        if rl <= 0 or rl > max_rlength or gl <= 0 or gl > max_glength:
            continue  # dynprog_end.c:1702-1717
        open_, ext = end_penalties(e["defect_rate"])
        lb, ub = wide_bands(rl, gl, e["extraband"])
        width = 8 if (rl < use8 or gl < use8) else 16
        # sweep order: the 5' form's reversed strings forward
        q, g, ga = e["rev_rseq"], e["rev_gseq"], e["rev_galt"]
        if five:
            q, g, ga = q[::-1], g[::-1], ga[::-1]
        for kind in (GDP_KIND_UPPER, GDP_KIND_LOWER):
            p = np.zeros(1, PROBLEM_DTYPE)
            p["qoff"], p["goff"], p["galtoff"] = qo, go, go + gl
            p["rl"], p["gl"], p["lband"], p["uband"] = rl, gl, lb, ub
            p["kind"], p["width"], p["mismatchtype"] = kind, width, ENDQ
            p["jump_late"] = int(bool(e["jump_late"]) != five)  # the 5' fills take the opposite tie rule
            p["open"], p["extend"], p["keep_scores"] = open_, ext, GDP_KEEP_BEST_QUERYEND
            probs.append(p)
        qparts.append(q)
        gparts.append(g + ga)
        qo += rl
        go += 2 * gl
        live.append(k)
    if not live:
        return out
    b = DeviceBatch(np.concatenate(probs), b"".join(qparts), b"".join(gparts))
    b.fill()
    ul = np.stack([np.arange(0, 2 * len(live), 2), np.arange(1, 2 * len(live), 2)], 1).astype(np.int32)
    best = b.best_endpoint(ul, queryend_indels=True)

    def requests(sel, cells, endc_of, goff_of):
        req = np.zeros(len(sel), TBREQ_DTYPE)
        for j, (i, (r, c)) in enumerate(zip(sel, cells)):
            e = ends[live[i]]
            req[j]["fill"] = 2 * i + (0 if c >= r else 1)  # the triangle holding the cell
            req[j]["r"], req[j]["c"] = r, c
            req[j]["queryoffset"], req[j]["genomeoffset"] = e["rev_roffset"], goff_of(e)
            req[j]["revp"], req[j]["genestrand"], req[j]["watsonp"] = int(five), e["genestrand"], 1
            req[j]["flags"], req[j]["endc"] = GDP_TB_LOCAL, endc_of(e)
            req[j]["max_pairs"] = len(e["rev_rseq"]) + len(e["rev_gseq"]) + 4
        return req

    sel = [i for i in range(len(live)) if best[i, 0] >= 0]  # finalscore < 0: no splice (1797-1800)
    if not sel:
        return out
    p1, c1, stop = b.traceback(requests(sel, [(best[i, 1], best[i, 2]) for i in sel], lambda e: e["contlength"],
                                        lambda e: e["far"]), ends=True)
    p2, c2 = b.traceback(requests(sel, [tuple(x) for x in stop], lambda e: 0, lambda e: e["anchor"]))
    for j, i in enumerate(sel):
        e = ends[live[i]]
        gap = np.zeros(1, PAIR_DTYPE)  # Pairpool_push_gapholder, knownp (pairpool.c:374-452)
        gap["querypos"] = gap["genomepos"] = -1
        gap["cdna"] = gap["comp"] = gap["genome"] = gap["genomealt"] = ord(" ")
        gap["gapp"], gap["genomejump"] = 1, (e["anchor"] - e["far"]) if five else (e["far"] - e["anchor"])
        pairs = np.concatenate([p1[j], gap, p2[j]])
        cnt = c1[j, 1:] + c2[j, 1:]
        # the reference reverses the list and drops indel pairs from the read's
        # end, which are the first pushed (1905-1910, 2485-2490); the 3' form
        # returns that trimmed list as it is
        k = 0
        while k < len(pairs) and pairs["comp"][k] == INDEL_COMP:
            k += 1
        out[live[i]] = (pairs[k:] if five else pairs[k:][::-1].copy(), np.array([cnt[0], cnt[0] - len(e["rev_rseq"]) * FULLMATCH, cnt[1], cnt[2],
                                             cnt[3], cnt[4]], np.int32))
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
