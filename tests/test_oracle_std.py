"""The oracle's Dynprog_standard restatement against the reference's scalar
build (oracle/_ref_std: dynprog.c compiled without SIMD).

Skipped where the reference tree was not available to build oracle/_ref_std.
The scalar build clears its matrices on every fill (dynprog.c:206-224,
487-500), so every cell of the dense outputs is compared, in band or not.
"""
import numpy as np
import pytest

import helpers as H
import pyoracle as po

pytestmark = pytest.mark.ref


def _std_cases(seed, n):
    return [(s, s.lband, s.uband, s.saturation, s.upperp, s.lowerp) for s in H.random_std_specs(seed, n)]


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_fill_std_matches_reference(oracle, reference_std, seed):
    for s, lb, ub, sat, up, lo in _std_cases(seed, 250):
        args = (s.q, s.g, s.galt, s.mt, s.open, s.ext, lb, ub, s.late, s.revp, sat, up, lo)
        Ho, Do = oracle.fill_std(*args)
        Hr, Dr = reference_std.fill_std(*args)
        assert (Ho == Hr).all() and (Do == Dr).all(), (s.rl, s.gl, lb, ub, sat, up, lo, s.late, s.revp)


@pytest.mark.parametrize("seed", [21, 22])
def test_traceback_std_matches_reference(oracle, reference_std, seed):
    rng = np.random.default_rng(seed)
    for s, lb, ub, sat, up, lo in _std_cases(seed, 200):
        args = (s.q, s.g, s.galt, s.mt, s.open, s.ext, lb, ub, s.late, s.revp, sat, up, lo)
        Ho, Do = oracle.fill_std(*args)
        reference_std.fill_std(*args)
        B = po.band_mask_full(s.rl, s.gl, lb, ub)
        if B[s.gl, s.rl]:
            r, c = s.rl, s.gl
        else:
            cells = np.argwhere(B)
            if not len(cells):
                continue
            c, r = (int(x) for x in cells[rng.integers(0, len(cells))])
        ta = reference_std.traceback(r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp, genestrand=0)
        to = oracle.traceback(po.KIND_FULL, Do, r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
        assert to[0].shape == ta[0].shape and (to[0] == ta[0]).all() and (to[1] == ta[1]).all(), (r, c, lb, ub)


def test_fill_std_long_saturating(oracle, reference_std):
    """Long dissimilar sequences drive H far below zero; with a floor of -40
    every stored score clamps while the unclamped chain keeps running."""
    rng = np.random.default_rng(5)
    for rl, gl in ((900, 1000), (1000, 700)):
        q = H.rand_seq(rng, rl)
        g = H.rand_seq(rng, gl)
        for late in (False, True):
            for sat in (-32768, -40):
                args = (q, g, g, 1, -8, -2, rl + 20, gl + 20, late, False, sat, True, True)
                Ho, Do = oracle.fill_std(*args)
                Hr, Dr = reference_std.fill_std(*args)
                assert (Ho == Hr).all() and (Do == Dr).all()
