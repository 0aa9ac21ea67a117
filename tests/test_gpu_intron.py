"""Intron gap bridge on the GPU (gdp_bridge_intron, the batched form of
bridge_intron_gap_{8,16}_ud, dynprog_genome.c:866-2363): four one-sided fills
per gap kept with every score, the bridge, and the two ends' tracebacks from
the requests it writes, against the oracle.  The splice-site probabilities
are synthetic per-column values (coarse steps, many exact ties, sums at and
around the rule's 2 * 0.85, all zero), handed over as the caller's site
model would."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

SMALL = 1024  # GDP_INTRON_SITES_SMALL (gdp_kernels.hip)


def _probs(rng, n, mode):
    if mode == 0:
        return rng.integers(0, 21, n) / 20.0  # steps of 0.05: sums land on 1.7 exactly
    if mode == 1:
        return np.zeros(n)
    if mode == 2:
        return rng.random(n)
    return rng.integers(0, 5, n) / 4.0


def _gap_row(k, c, prob_off):
    from gmapdp.lib import INTRON_GAP_DTYPE
    g = np.zeros(1, INTRON_GAP_DTYPE)
    g["fills"] = [4 * k, 4 * k + 1, 4 * k + 2, 4 * k + 3]
    g["span"] = c["rightoffset"] - c["leftoffset"]
    g["cdna_direction"], g["finalp"], g["halfp"] = c["cdna_direction"], int(c["finalp"]), int(c["halfp"])
    g["prob_off"] = prob_off
    return g[0]


def test_intron_bridge_and_end_tracebacks(oracle):
    from gmapdp.batch import DeviceBatch
    seq, _ = H.intron_genome()
    rng = np.random.default_rng(31)
    cases = H.intron_gap_cases(rng, 180, seq) + H.intron_gap_cases(np.random.default_rng(32), 3, seq,
                                                                    long_reads=True)
    assert sum(c["glengthL"] >= SMALL for c in cases) == 3  # both launches take gaps
    # reads unrelated to their segments: joins below 0
    for c in [c for c in cases if c["rlength"] >= 60][:12]:
        poor = dict(c)
        poor["q"] = H.rand_seq(rng, c["rlength"])
        cases.append(poor)
    # a gap whose segments are no longer than the read: not bridged
    short = dict(cases[1])
    short["gL"], short["gLalt"] = short["gL"][:short["rlength"]], short["gLalt"][:short["rlength"]]
    short["glengthL"] = short["rlength"]
    cases.append(short)
    specs = [s for c in cases for s in H.intron_gap_specs(c)]
    probs, qbuf, gbuf = H.pack_batch(specs, keep_scores=True)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    site, rows, pl = [], [], 0
    for k, c in enumerate(cases):
        pL, pR = _probs(rng, c["glengthL"], k % 4), _probs(rng, c["glengthR"], (k // 4) % 4)
        site += [pL, pR]
        rows.append(_gap_row(k, c, pl))
        pl += c["glengthL"] + c["glengthR"]
    # fills in the wrong order: not bridged
    bad = _gap_row(0, cases[0], 0)
    bad["fills"] = [1, 0, 2, 3]
    rows.append(bad)
    gaps = np.array(rows)
    site = np.concatenate(site)
    # host-side bounds of what the kernel reads
    assert (gaps["fills"] < len(specs)).all() and (gaps["fills"] >= 0).all()
    assert all(r["prob_off"] + c["glengthL"] + c["glengthR"] <= len(site) for r, c in zip(rows, cases))
    tmpl = []
    for k in range(len(gaps)):
        kk = min(k, len(cases) - 1)
        sp = H.intron_gap_specs(cases[kk])
        tmpl.append(H.tb_request(4 * kk + 2, sp[2], 0, 0, qoff=11, goff=13))
        tmpl.append(H.tb_request(4 * kk, sp[0], 0, 0, qoff=3, goff=5))
        for q in tmpl[-2:]:
            q["max_pairs"] = 2 * (cases[kk]["rlength"] + cases[kk]["glengthL"]) + 8
    got, gotp, reqs = b.bridge_intron(gaps, site, np.array(tmpl))
    g2, gp2 = b.bridge_intron(gaps, site)
    assert (g2 == got).all() and (gp2 == gotp).all()
    assert tuple(got[-1]) == (np.iinfo(np.int32).min, -1, -1, -1, -1)
    assert tuple(got[-2]) == (np.iinfo(np.int32).min, -1, -1, -1, -1)
    found = neg = 0
    walks = []
    for k, c in enumerate(cases[:-1]):
        sp = H.intron_gap_specs(c)
        Hs = [H.oracle_fill(oracle, s)[0] for s in sp]
        r = rows[k]
        pL = site[r["prob_off"]: r["prob_off"] + c["glengthL"]]
        pR = site[r["prob_off"] + c["glengthL"]: r["prob_off"] + c["glengthL"] + c["glengthR"]]
        want, want_p = oracle.bridge_intron(c["width"], *Hs, c["gL"], c["gLalt"], c["gRseg"][::-1],
                                            c["gRsegalt"][::-1], pL, pR, c["rlength"], c["lbandL"], c["ubandL"],
                                            c["lbandR"], c["ubandR"], c["leftoffset"], c["rightoffset"],
                                            c["cdna_direction"], c["finalp"], c["halfp"])
        assert tuple(got[k]) == want and tuple(gotp[k]) == want_p, ("bridge", k, tuple(got[k]), want)
        fin, rL, rR, cL, cR = want
        if fin < 0:
            neg += 1
            assert all(reqs[2 * k + j]["r"] == 0 and reqs[2 * k + j]["c"] == 0 for j in (0, 1))
            continue
        found += 1
        for j, (rr, cc, ui) in enumerate(((rR, cR, 2), (rL, cL, 0))):
            q = reqs[2 * k + j]
            kind = 1 if cc >= rr else 2
            assert (q["fill"], q["r"], q["c"]) == (4 * k + ui + (kind - 1), rr, cc)
            s = sp[ui + kind - 1]
            Dir = np.where(H.computed_mask(s, b.layout[4 * k + ui + kind - 1]), H.oracle_fill(oracle, s)[1],
                           0).astype(np.uint8)
            qo, go = int(q["queryoffset"]), int(q["genomeoffset"])
            walks.append((q, oracle.traceback(kind, Dir, rr, cc, s.q, s.q.upper(), s.g, s.galt, qo, go, s.revp)))
    assert found > 120 and neg >= 4
    pairs, counts = b.traceback(np.array([w[0] for w in walks]))
    for j, (_, (po, co)) in enumerate(walks):
        assert (H.pairs_to_records(pairs[j]) == po).all() and (counts[j, 1:] == co).all(), ("end walk", j)
