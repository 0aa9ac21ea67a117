"""The paired sweep (gdp_fill2_kernel): full-band int16 problems that share
band and tie rule are swept two per lane, padded to the largest problem of
their wave.  Every fill and traceback must still equal the oracle's."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _class_specs(seed, n, lb, ub, late, rl_range, alph_q=b"ACGT", genome_n=0.0, alt=0.0):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        rl = int(rng.integers(*rl_range))
        gl = max(1, rl + int(rng.integers(-lb, ub + 1)))
        g = H.rand_seq(rng, gl, genome_n)
        q = H.mutate(rng, g, float(rng.choice([0.02, 0.1, 0.3])), 0.03)[:rl]
        q = q + H.rand_seq(rng, rl - len(q))
        if alph_q != b"ACGT":
            q = bytes(alph_q[rng.integers(0, len(alph_q))] if rng.random() < 0.1 else ch for ch in q)
        galt = bytes(ch if rng.random() >= alt else H.ALPH[rng.integers(0, 4)] for ch in g) if alt else g
        mt = int(rng.integers(0, 4))
        op = int(rng.choice([-8, -7, -6, -10, -20]))
        ex = int(rng.choice([-3, -2, -1, -5]))
        specs.append(H.Spec(0, 16, q, g, galt, mt, op, ex, lb, ub, late, bool(rng.integers(0, 2))))
    return specs


def _check(oracle, specs, seed):
    from test_gpu_parity import _check_fills, _check_tracebacks, _run
    b = _run(oracle, specs, keep_scores=False)
    _check_fills(oracle, specs, b)
    _check_tracebacks(oracle, specs, b, seed)


@pytest.mark.parametrize("late", [False, True])
def test_paired_same_band_padded(oracle, late):
    specs = []
    for k, (lb, ub) in enumerate([(5, 30), (0, 0), (1, 2), (12, 12), (20, 50), (3, 200)]):
        specs += _class_specs(100 + k, 37, lb, ub, late, (1, 190))
    _check(oracle, specs, 5)


def test_paired_wide_bands_64_lane_groups(oracle):
    specs = _class_specs(7, 10, 40, 300, True, (250, 600)) + _class_specs(8, 9, 250, 30, False, (300, 500))
    _check(oracle, specs, 6)


def test_paired_general_lookup(oracle):
    # genome N / IUPAC and alternate alleles take the general score lookup
    specs = _class_specs(9, 30, 8, 25, True, (20, 150), alph_q=b"ACGTNRYacgt", genome_n=0.05, alt=0.1)
    specs += _class_specs(10, 30, 8, 25, False, (20, 150), genome_n=0.0, alt=0.05)
    _check(oracle, specs, 7)


def test_paired_mixed_with_single(oracle):
    # paired and single sweeps in one plan, in interleaved order
    specs = _class_specs(11, 40, 10, 40, True, (100, 160)) + H.random_specs(12, 120)
    order = np.random.default_rng(3).permutation(len(specs))
    _check(oracle, [specs[k] for k in order], 8)
