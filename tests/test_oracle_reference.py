"""The CPU oracle against the reference's own AVX2 DP objects (oracle/_ref).

Skipped where the reference tree was not available to build oracle/_ref.
Every fill runs in the reference twice, with 0x00 and 0xff left in its
workspace; cells and tracebacks that agree between the two runs are the
reference's deterministic output, and the oracle must reproduce all of it.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.ref


def _ref_fill(R, s, garbage):
    if s.kind == 0:
        return R.fill_full(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp,
                           garbage=garbage)
    if s.kind == 1:
        return R.fill_upper(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.uband, s.late, s.revp, garbage=garbage)
    return R.fill_lower(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.late, s.revp, garbage=garbage)


def test_tables_match_reference(oracle, reference):
    for mt in range(4):
        assert (oracle.pairdistance(mt) == reference.pairdistance(mt)).all()
        assert oracle.use8p(mt) == reference.use8p(mt)
    for c in range(128):
        for g in range(0, 128, 3):
            assert oracle.consistent(0, c, g) == reference.consistent(0, c, g)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fills_and_tracebacks_match_reference(oracle, reference, seed):
    rng = np.random.default_rng(seed)
    ndet = 0
    for s in H.random_specs(100 + seed, 700):
        Ho, Do = H.oracle_fill(oracle, s)
        B = H.band_mask(s)
        start = None
        if s.kind == 0 and s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl:
            start = (s.rl, s.gl)
        elif s.kind != 0:
            cells = np.argwhere(B)
            if len(cells):
                a, b = cells[rng.integers(0, len(cells))]
                start = (int(b), int(a)) if s.kind == 1 else (int(a), int(b))
        Ha, Da = _ref_fill(reference, s, 0)
        ta = reference.traceback(s.kind, None, *start, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp) if start else None
        Hb, Db = _ref_fill(reference, s, 255)
        tb = reference.traceback(s.kind, None, *start, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp) if start else None
        det = (Ha == Hb) & (Da == Db)
        assert det[B].all()
        assert (Ho[det] == Ha[det]).all() and (Do[det] == Da[det]).all(), (s.kind, s.width, s.rl, s.gl)
        if start and ta[0].shape == tb[0].shape and (ta[0] == tb[0]).all() and (ta[1] == tb[1]).all():
            to = oracle.traceback(s.kind, Do, *start, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
            assert to[0].shape == ta[0].shape and (to[0] == ta[0]).all() and (to[1] == ta[1]).all()
            ndet += 1
    assert ndet > 600
