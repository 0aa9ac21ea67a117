"""The oracle's cDNA gap bridge (orc_bridge_cdna, a restatement of
bridge_cdna_gap_{8,16}_ud, dynprog_cdna.c:123-644) against the reference:
the committed vectors made by the reference's own function, and, where the
reference build is present, the function itself on fresh random tables."""
import numpy as np
import pytest

import golden_io


def test_bridge_matches_reference_vectors(oracle):
    cases = golden_io.load_bridge()
    assert len(cases) >= 100
    # the vectors cover both widths, both tie rules and joins that find nothing
    assert {c[0][0] for c in cases} == {8, 16} and {c[0][-1] for c in cases} == {0, 1}
    assert any(c[1][1] == -1 for c in cases) and sum(c[1][1] > 0 for c in cases) > 100
    for k, (args, want) in enumerate(cases):
        assert oracle.bridge_cdna(*args) == want, ("bridge vector", k)


def _random_case(rng, t):
    gl = int(rng.integers(1, 30))
    rlL = int(rng.integers(max(1, gl - 1), 40))
    rlR = int(rng.integers(max(1, gl - 1), 40))
    w = int(rng.choice([8, 16]))
    lo, hi = (-128, 128) if w == 8 else (-300, 60)
    sh = [(gl + 1) * (rlL + 1), (rlL + 1) * (gl + 1), (gl + 1) * (rlR + 1), (rlR + 1) * (gl + 1)]
    A = [rng.integers(lo, hi, z) for z in sh]
    if t % 3 == 0:
        A = [a // 40 for a in A]
    if t % 5 == 0:
        A = [np.zeros_like(a) for a in A]
    bands = [int(x) for x in rng.integers(0, 12, 4)]
    left = int(rng.integers(0, 500))
    return (w, *A, gl, rlL, rlR, *bands, int(rng.integers(-12, -5)), int(rng.integers(-3, 0)), left,
            left + int(rng.integers(0, 60)), int(rng.integers(0, 2)))


def test_bridge_matches_reference_function(oracle, reference):
    rng = np.random.default_rng(9)
    for t in range(1500):
        args = _random_case(rng, t)
        assert oracle.bridge_cdna(*args) == reference.bridge_cdna(*args), ("bridge", t)


def test_bridge_penalty_and_offsets(oracle):
    """Closed forms on a flat gap: with every score 0 the best join is the
    last (jump late) or first (jump early) candidate of the scan among those
    without a genomic insertion, and a right offset right at the left one
    leaves no rows for the right end."""
    gl, rl = 6, 8
    z = [np.zeros((gl + 1) * (rl + 1), np.int16) for _ in range(4)]
    late = oracle.bridge_cdna(16, *z, gl, rl, rl, 2, 2, 2, 2, -10, -1, 0, 100, 1)
    early = oracle.bridge_cdna(16, *z, gl, rl, rl, 2, 2, 2, 2, -10, -1, 0, 100, 0)
    # score 0 needs cR = gl - cL; late keeps cL = gl-1 with the last rows of
    # both bands, early keeps cL = 1, cR = gl-1 with the first rows
    assert late == (0, gl - 1, 1, min(rl - 1, gl - 1 + 2), min(rl - 1, 1 + 2))
    assert early == (0, 1, gl - 1, 1, gl - 1 - 2)
    none = oracle.bridge_cdna(16, *z, gl, rl, rl, 2, 2, 2, 2, -10, -1, 50, 50, 1)
    assert none == (-32768, -1, -1, -1, -1)
