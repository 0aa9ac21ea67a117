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


def _check(oracle, specs, seed, params=None):
    from test_gpu_parity import _check_fills, _check_tracebacks, _run
    b = _run(oracle, specs, keep_scores=False, params=params)
    _check_fills(oracle, specs, b)
    _check_tracebacks(oracle, specs, b, seed)
    return b


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


@pytest.mark.parametrize("params", [(0, 8001), (240, 64), (48, 200)])
def test_paired_lane_group_choice(oracle, params):
    """64-lane groups on narrow bands (the autotuner's and the serving
    loop's latency choice) and 16-lane groups up to a smaller span give the
    same bits as every other sweep."""
    specs = []
    for k, (lb, ub) in enumerate([(0, 0), (2, 5), (12, 12), (5, 60), (30, 90), (3, 200)]):
        specs += _class_specs(300 + k, 21, lb, ub, bool(k % 2), (1, 400))
    b = _check(oracle, specs, 8, params)
    d = b.d_descs.cpu().numpy()[: b.n * 96].reshape(b.n, 96)
    layout, glog = d[:, 58], d[:, 59]  # gdp_dev_problem.layout / .glog
    rl = np.array([s.rl for s in specs])
    # the span the planner sweeps: bands narrowed to one past the matrix
    span = np.array([min(s.lband, s.rl + 1) + min(s.uband, s.gl + 1) for s in specs])
    want64 = (span > params[0]) | (rl >= params[1])
    assert (layout == 1).all()
    assert ((glog == 6) == want64).all()


def test_graph_replays_equal_eager(oracle):
    """fill + traceback captured in one HIP graph: every replay, with eager
    work and synchronisations between replays, gives the eager results."""
    import torch
    from gmapdp.batch import DeviceBatch
    specs = _class_specs(40, 60, 6, 40, True, (20, 300)) + H.random_specs(41, 100)
    probs, qbuf, gbuf = H.pack_batch(specs, keep_scores=False)
    b = DeviceBatch(probs, qbuf, gbuf)
    reqs = np.array([H.tb_request(i, s, s.rl, s.gl) for i, s in enumerate(specs)
                     if s.kind == 0 and s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl])
    off = np.zeros(len(reqs), np.int64)
    off[1:] = np.cumsum(reqs["max_pairs"].astype(np.int64))[:-1]
    reqs["pair_off"] = off
    h = b.prepare_traceback(reqs)
    b.fill()
    b.traceback_device(h)
    torch.cuda.synchronize()
    # direction words up to the sink that empty slots of partial waves share
    # (its contents are whatever slot wrote last)
    hdr = b.d_descs.cpu().numpy()[b.n * 96: b.n * 96 + 256]
    sink = int(hdr[176:184].view(np.int64)[0])
    assert 0 < sink <= b.dir_words
    want = (b.d_dirs[:sink].clone(), b.d_final.clone(), h["d_pairs"].clone(), h["d_counts"].clone())
    g = b.graph(h)
    try:
        for k in range(4):
            for t in (b.d_dirs, b.d_final, h["d_pairs"], h["d_counts"]):
                t.zero_()
            torch.cuda.synchronize()
            b.graph_launch(g)
            torch.cuda.synchronize()
            got = (b.d_dirs[:sink], b.d_final, h["d_pairs"], h["d_counts"])
            for x, y in zip(got, want):
                assert torch.equal(x, y), k
    finally:
        b.graph_destroy(g)


def _mixed_specs(seed, n, late, rl_range, lb_range, ub_extra, alph_q=b"ACGT", genome_n=0.0, alt=0.0):
    """Problems each with a band of its own: lower bands from lb_range and
    upper bands gl - rl + U(ub_extra), as Dynprog_compute_bands widens them,
    so that few share a band and most land in mixed waves."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        lb = int(rng.integers(*lb_range))
        s = _class_specs(seed * 1000 + k, 1, lb, lb + int(rng.integers(*ub_extra)), late, rl_range, alph_q, genome_n,
                         alt)[0]
        specs.append(s)
    return specs


@pytest.mark.parametrize("params", [(240, 8001, 16), (112, 256, 24), (0, 8001, 40), (240, 1, 8), (240, 8001, 16, 1),
                                    (112, 256, 8, 1), (240, 1, 24, 1), (0, 8001, 32, 1)])
def test_mixed_band_waves(oracle, params):
    """Waves whose problems have different bands (gdp_plan_params.mix_span):
    the sweep is laid out for the wave's largest bands and each half applies
    its own problem's row starts, block prefix, bottom edge and top edge.
    Both tie rules, narrow bands (bottom edges near column 0), 16- and
    64-lane groups, the general lookup, and both ways of forming the mixed
    waves (mix_order 1 bins every problem by band, then by rows, and bins
    what is left again at twice the spread)."""
    specs = []
    specs += _mixed_specs(51, 48, False, (1, 120), (0, 40), (0, 60))
    specs += _mixed_specs(52, 48, True, (20, 300), (3, 64), (0, 120))
    specs += _mixed_specs(53, 24, True, (1, 40), (0, 6), (0, 8))
    specs += _mixed_specs(54, 20, False, (30, 160), (5, 30), (0, 40), alph_q=b"ACGTNRY", genome_n=0.05, alt=0.1)
    b = _check(oracle, specs, 9, params)
    d = b.d_descs.cpu().numpy()[: b.n * 96].reshape(b.n, 96)
    layout = d[:, 58]
    slb = d[:, 82].astype(np.int16) | (d[:, 83].astype(np.int16) << 8)
    lband = np.array([min(s.lband, s.rl + 1) for s in specs])
    paired = layout == 1
    assert paired.sum() > 100
    assert (slb[paired] > lband[paired]).sum() > 20, "no wave mixed lower bands"
    assert (slb[paired] >= lband[paired]).all()


def test_mixed_band_waves_end_on_uniform_ones(oracle):
    """With mix_span the problems that fill waves of their own band keep
    them; the rest go to mixed waves; the fills stay exact either way."""
    specs = _class_specs(61, 20, 9, 30, False, (50, 150)) + _mixed_specs(62, 30, False, (40, 150), (5, 15), (0, 40))
    _check(oracle, specs, 10, (240, 8001, 12))

