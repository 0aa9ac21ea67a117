"""GPU parity of the scalar standard fill (problem kind 3, Dynprog_standard,
dynprog.c:1267-1785) and its traceback (Dynprog_traceback_std, 1795-1947)
against the oracle, bit for bit, and against the reference's own vectors
(tests/golden/std_golden.npz).

The oracle's restatement is pinned to the reference's scalar build by
tests/test_oracle_std.py and to the committed vectors by
tests/test_oracle_golden.py.
"""
import numpy as np
import pytest

import golden_io
import helpers as H

pytestmark = pytest.mark.gpu


def _batch(specs, keep=True):
    from gmapdp.batch import DeviceBatch
    probs, qbuf, gbuf = H.pack_batch(specs, keep_scores=keep)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    return b


def _start(s, rng):
    B = H.band_mask(s)
    if B[s.gl, s.rl]:
        return s.rl, s.gl
    cells = np.argwhere(B)
    if not len(cells):
        return None
    c, r = (int(x) for x in cells[rng.integers(0, len(cells))])
    return r, c


def _check(oracle, specs, b, seed, keep=True):
    fin = b.final_scores()
    rng = np.random.default_rng(seed)
    reqs, starts = [], []
    for i, s in enumerate(specs):
        Ho, Do = H.oracle_fill(oracle, s)
        Dg, Sg = b.export(i)
        assert (Dg == Do).all(), ("directions differ", i, s.rl, s.gl, s.lband, s.uband, s.saturation)
        if keep:
            assert (Sg[1:, 1:].astype(np.int32) == Ho[1:, 1:]).all(), ("scores differ", i, s.rl, s.gl)
        in_band = s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl
        want = int(Ho[s.gl, s.rl]) if in_band else np.iinfo(np.int32).min
        assert fin[i] == want, ("final score", i, fin[i], want)
        st = _start(s, rng)
        if st:
            reqs.append(H.tb_request(i, s, *st))
            starts.append((i,) + st)
    pairs, counts = b.traceback(np.array(reqs))
    for k, (i, r, c) in enumerate(starts):
        s = specs[i]
        _, Do = H.oracle_fill(oracle, s)
        po, co = oracle.traceback(0, Do, r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
        got = H.pairs_to_records(pairs[k])
        assert counts[k, 0] == len(po) and (got == po).all(), ("pairs differ", i)
        assert (counts[k, 1:] == co).all(), ("counters differ", i)


@pytest.mark.parametrize("seed,keep", [(31, True), (32, False)])
def test_std_random(oracle, seed, keep):
    specs = H.random_std_specs(seed, 300)
    b = _batch(specs, keep)
    _check(oracle, specs, b, seed, keep)


def test_std_golden_vectors(oracle):
    """The reference's own outputs: dense directions, scores and pairs."""
    g, specs = golden_io.load_std()
    b = _batch(specs)
    reqs, idx = [], []
    for i, s in enumerate(specs):
        Dg, Sg = b.export(i)
        Hs = np.zeros(Sg.shape, np.int32)
        Hs[1:, 1:] = Sg[1:, 1:]
        golden_io.check_std(g, i, Dg, Hs)
        exp = golden_io.std_traceback_expect(g, i)
        if exp is not None:
            reqs.append(H.tb_request(i, s, exp[0], exp[1]))
            idx.append(i)
    pairs, counts = b.traceback(np.array(reqs))
    for k, i in enumerate(idx):
        _, _, want, cnt = golden_io.std_traceback_expect(g, i)
        got = H.pairs_to_records(pairs[k])
        assert got.shape == want.shape and (got == want).all(), ("golden pairs", i)
        assert (counts[k, 1:] == cnt).all(), ("golden counters", i)


def test_std_long_and_wide(oracle):
    """Long queries, the widest band the standard fill takes (span 526), and
    a floor that clamps most stored scores."""
    rng = np.random.default_rng(9)
    specs = []
    for k in range(12):
        rl = int(rng.integers(400, 1200))
        gl = rl + int(rng.integers(0, 200))
        g = H.rand_seq(rng, gl)
        q = H.mutate(rng, g, 0.2 if k % 2 else 0.6, 0.02)[:rl]
        q = q + H.rand_seq(rng, rl - len(q))
        lb, ub = (263, 263) if k < 4 else (8, gl - rl + 8)
        specs.append(H.Spec(3, 32, q, g, g, k % 4, -8, -2, lb, ub, bool(k % 3), bool(k % 2),
                            saturation=-40 if k % 4 == 1 else -32768))
    b = _batch(specs)
    _check(oracle, specs, b, 9)


def test_std_mixed_with_simd_kinds(oracle):
    """Standard fills planned in one batch with the SIMD kinds: the single
    sweep's ring is sized for the two-word entries, paired sweeps unchanged."""
    specs = H.random_std_specs(41, 60) + H.random_specs(42, 120)
    rng = np.random.default_rng(3)
    rng.shuffle(specs)
    b = _batch(specs, keep=False)
    std = [i for i, s in enumerate(specs) if s.kind == 3]
    sub = [specs[i] for i in std]
    fin = b.final_scores()
    for j, i in enumerate(std):
        s = sub[j]
        Ho, Do = H.oracle_fill(oracle, s)
        Dg, _ = b.export(i)
        assert (Dg == Do).all(), ("directions differ", i)
        in_band = s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl
        assert fin[i] == (int(Ho[s.gl, s.rl]) if in_band else np.iinfo(np.int32).min)
    for i, s in enumerate(specs):
        if s.kind == 3:
            continue
        _, Do = H.oracle_fill(oracle, s)
        Dg, _ = b.export(i)
        B = H.band_mask(s)
        assert (Dg[B] == Do[B]).all(), ("SIMD-kind directions differ", i, s.kind)
