"""The oracle's intron bridge (orc_bridge_intron, a restatement of
bridge_intron_gap_{8,16}_ud on its site-level path, dynprog_genome.c:866-2363)
against the reference: the committed vectors made by the reference's own
function with its own splice-site models, and, where the reference build is
present, the function itself on fresh intron gaps."""
import numpy as np
import pytest

import golden_io
import helpers as H


def _oracle(O, c, probL, probR):
    Hs = [H.oracle_fill(O, s)[0] for s in H.intron_gap_specs(c)]
    gRf = c["gRseg"][::-1]
    return O.bridge_intron(c["width"], *Hs, c["gL"], c["gLalt"], gRf, c["gRsegalt"][::-1], probL, probR, c["rlength"], c["lbandL"],
                           c["ubandL"], c["lbandR"], c["ubandR"], c["leftoffset"], c["rightoffset"],
                           c["cdna_direction"], c["finalp"], c["halfp"])


def test_intron_bridge_matches_reference_vectors(oracle):
    cases = golden_io.load_intron_bridge()
    assert len(cases) >= 600
    # both widths, all three cdna directions, both strands, halfp, joins
    # below 0, and joins moved to a canonical dinucleotide by the rule
    assert {c["width"] for c, _, _ in cases} == {8, 16}
    assert {c["cdna_direction"] for c, _, _ in cases} == {-1, 0, 1}
    assert {c["watsonp"] for c, _, _ in cases} == {False, True}
    assert sum(w[0] < 0 for _, w, _ in cases) >= 3 and sum(c["halfp"] for c, _, _ in cases) > 100
    for k, (c, want, want_p) in enumerate(cases):
        got, got_p = _oracle(oracle, c, c["probL"], c["probR"])
        assert got == want and got_p == want_p, ("intron vector", k, got, want, got_p, want_p)


def test_intron_bridge_matches_reference_function(oracle, reference):
    seq, gc = H.intron_genome()
    reference.genome_setup(gc, len(seq))
    cases = H.intron_gap_cases(np.random.default_rng(5), 3000, seq)
    for t, c in enumerate(cases):
        Hs = [H.oracle_fill(oracle, s)[0] for s in H.intron_gap_specs(c)]
        pL, pR = reference.site_probs(c["glengthL"], c["glengthR"], c["leftoffset"], c["rightoffset"],
                                      c["chroffset"], c["chrhigh"], c["cdna_direction"], c["watsonp"])
        want = reference.bridge_intron(c["width"], *Hs, c["gL"], c["gLalt"], c["gRseg"], c["gRsegalt"], c["rlength"],
                                       c["lbandL"], c["ubandL"], c["lbandR"], c["ubandR"], c["leftoffset"],
                                       c["rightoffset"], c["chroffset"], c["chrhigh"], c["cdna_direction"],
                                       c["watsonp"], c["finalp"], c["halfp"], c["late"])
        assert _oracle(oracle, c, pL, pR) == want, ("intron", t)
        if want[0][0] >= 0:
            # the probabilities returned are the models' at the chosen
            # sites: the per-column arrays follow the reference's positions
            assert want[1] == (pL[want[0][3]], pR[want[0][4]]), ("site positions", t)


def test_intron_bridge_scan_order(oracle):
    """Closed forms on flat tables over a genome with no canonical site: every
    candidate scores 0, so the result is the first candidate of the scan with
    the highest probability sum -- the diagonal pair of row 1 when all
    probabilities are equal, the column of the single best left site
    otherwise."""
    rl, gl = 9, 14
    z = [np.zeros((gl + 1) * (rl + 1), np.int16) for _ in range(4)]
    g = b"A" * gl
    flat = np.full(gl, 0.25)
    res, pr = oracle.bridge_intron(16, *z, g, g, g, g, flat, flat, rl, 2, 7, 2, 7, 0, 1000, 1, False, False)
    assert res == (0, 1, rl - 1, 1, rl - 1) and pr == (0.25, 0.25)
    # one left site better than the rest, at column 6: rows 4 and 5 reach it
    # past their diagonal (cL < rL + uband), row 6 on it, rows 7 and 8 below
    # it; the scan meets row 4's first
    pL = flat.copy()
    pL[6] = 0.75
    res, pr = oracle.bridge_intron(16, *z, g, g, g, g, pL, flat, rl, 2, 3, 2, 3, 0, 1000, 1, False, False)
    assert res == (0, 4, rl - 4, 6, rl - 4) and pr == (0.75, 0.25)
