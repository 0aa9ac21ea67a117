"""GPU parity: the HIP fill/traceback against the CPU oracle, bit for bit.

The oracle (oracle/gdp_oracle.c) is pinned to the reference's AVX2 build by
tests/test_oracle_reference.py (where the reference tree exists) and by the
committed golden vectors in tests/golden (everywhere).
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _run(oracle, specs, keep_scores=True, params=None):
    from gmapdp.batch import DeviceBatch
    probs, qbuf, gbuf = H.pack_batch(specs, keep_scores=keep_scores)
    b = DeviceBatch(probs, qbuf, gbuf, params=params)
    b.fill()
    return b


def _check_fills(oracle, specs, b):
    fin = b.final_scores()
    for i, s in enumerate(specs):
        Ho, Do = H.oracle_fill(oracle, s)
        Dg, Sg = b.export(i)
        B = H.band_mask(s)
        M = H.computed_mask(s, b.layout[i])
        assert (Dg[B] == Do[B]).all(), ("in-band directions differ", i, s.kind, s.width, s.rl, s.gl)
        assert (Dg[M] == Do[M]).all(), ("out-of-band block directions differ", i, s.kind, s.width)
        if b.problems["keep_scores"][i]:
            assert (Sg[B] == Ho[B]).all(), ("in-band scores differ", i, s.kind, s.width, s.rl, s.gl)
            assert (Sg[M] == Ho[M]).all(), ("out-of-band block scores differ", i, s.kind, s.width)
        if s.kind == 0:
            in_band = s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl
            want = int(Ho[s.gl, s.rl]) if in_band else np.iinfo(np.int32).min
            assert fin[i] == want, ("final score", i, fin[i], want)


def _check_tracebacks(oracle, specs, b, seed):
    rng = np.random.default_rng(seed)
    reqs, starts = [], []
    for i, s in enumerate(specs):
        B = H.band_mask(s)
        if s.kind == 0:
            if not (s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl):
                continue
            r, c = s.rl, s.gl
        else:
            cells = np.argwhere(B)
            if len(cells) == 0:
                continue
            a, bb = cells[rng.integers(0, len(cells))]
            r, c = (int(bb), int(a)) if s.kind == 1 else (int(a), int(bb))
        reqs.append(H.tb_request(i, s, r, c))
        starts.append((i, r, c))
    pairs, counts = b.traceback(np.array(reqs))
    for k, (i, r, c) in enumerate(starts):
        s = specs[i]
        Ho, Do = H.oracle_fill(oracle, s)
        Do = np.where(H.computed_mask(s, b.layout[i]), Do, 0).astype(np.uint8)
        po, co = oracle.traceback(s.kind, Do, r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
        got = H.pairs_to_records(pairs[k])
        assert counts[k, 0] == len(po), ("pair count", i, counts[k, 0], len(po))
        assert (got == po).all(), ("pairs differ", i, s.kind)
        assert (counts[k, 1:] == co).all(), ("counters differ", i, counts[k, 1:], co)


@pytest.mark.parametrize("seed,keep", [(11, True), (12, True), (13, False), (14, False)])
def test_random_all_kinds(oracle, seed, keep):
    # keep=False lets full-band int16 problems take the paired sweep
    specs = H.random_specs(seed, 400)
    b = _run(oracle, specs, keep_scores=keep)
    _check_fills(oracle, specs, b)
    _check_tracebacks(oracle, specs, b, seed)


def test_wide_bands_and_long_queries(oracle):
    rng = np.random.default_rng(7)
    specs = []
    for _ in range(40):
        rl = int(rng.integers(300, 1000))
        gl = int(rl * rng.uniform(1.0, 1.5))
        g = H.rand_seq(rng, gl)
        q = H.mutate(rng, g, 0.05, 0.01)[:rl]
        q = q + H.rand_seq(rng, rl - len(q))
        eb = int(rng.integers(3, 65))
        specs.append(H.Spec(0, 16, q, g, g, 0, -8, -3, eb, gl - rl + eb, bool(rng.integers(0, 2)), False))
    b = _run(oracle, specs)
    _check_fills(oracle, specs, b)
    _check_tracebacks(oracle, specs, b, 7)


@pytest.mark.parametrize("keep", [True, False])
def test_edge_shapes(oracle, keep):
    # keep=False sends the one-sided fills with bands under 32 columns to the
    # narrow sweep (one problem per lane) and full-band int16 to the paired one
    specs = []
    for rl, gl, lb, ub in [(1, 1, 0, 0), (1, 5, 0, 4), (5, 1, 4, 0), (2, 40, 0, 38), (40, 2, 38, 0),
                           (33, 33, 0, 0), (64, 64, 1, 1), (31, 95, 3, 67), (16, 16, 15, 15),
                           # one-sided bands at the narrow sweep's row widths (8, 16, 32 cells)
                           (40, 40, 7, 7), (40, 40, 8, 8), (40, 48, 15, 15), (40, 56, 16, 16), (50, 81, 31, 31),
                           (50, 82, 32, 32)]:
        for width in (8, 16):
            for kind in (0, 1, 2):
                rng = np.random.default_rng(rl * 1000 + gl)
                g = H.rand_seq(rng, gl)
                q = H.rand_seq(rng, rl)
                for late in (False, True):
                    specs.append(H.Spec(kind, width, q, g, g, 3, -10, -2, lb, ub, late, late))
    b = _run(oracle, specs, keep_scores=keep)
    if not keep:
        assert (b.layout == 2).any() and (b.layout == 1).any()
    _check_fills(oracle, specs, b)
    _check_tracebacks(oracle, specs, b, 3)


def test_end_extension_pipeline(oracle):
    """Config-4 style problems: upper and lower fills, best end cell over both
    triangles, traceback through the triangle holding it (dynprog_end.c:1576-1607)."""
    from gmapdp.batch import DeviceBatch
    from gmapdp.workload import Workload
    w = Workload(4, 300, genome_len=1 << 20)
    p, m = w.problems, w.meta
    b = DeviceBatch(p, w.qbuf.tobytes(), w.gbuf.tobytes())
    b.fill()
    ul = np.stack([np.arange(0, len(p), 2), np.arange(1, len(p), 2)], 1).astype(np.int32)
    for qi in (False, True):
        best = b.best_endpoint(ul, queryend_indels=qi)
        starts = []
        for k in range(len(ul)):
            pu = p[ul[k, 0]]
            revp = bool(m["revp"][ul[k, 0]])
            q = bytes(w.qbuf[pu["qoff"]: pu["qoff"] + pu["rl"]])
            g = bytes(w.gbuf[pu["goff"]: pu["goff"] + pu["gl"]])
            if revp:
                q, g = q[::-1], g[::-1]
            su = H.Spec(1, int(pu["width"]), q, g, g, 3, int(pu["open"]), int(pu["extend"]), int(pu["lband"]),
                        int(pu["uband"]), bool(pu["jump_late"]), revp)
            sl = H.Spec(2, su.width, q, g, g, 3, su.open, su.ext, su.lband, su.uband, su.late, revp)
            Hu, Du = H.oracle_fill(oracle, su)
            Hl, Dl = H.oracle_fill(oracle, sl)
            want = oracle.best_endpoint(qi, Hu, Hl, su.width, su.rl, su.gl, su.lband, su.uband, su.late)
            assert tuple(best[k]) == want, ("best end cell", k, tuple(best[k]), want)
            fs, br, bc = want
            if br == 0 and bc == 0:
                continue
            kind = 1 if bc >= br else 2
            s = su if kind == 1 else sl
            D = np.where(H.computed_mask(s, b.layout[ul[k, 0 if kind == 1 else 1]]), Du if kind == 1 else Dl,
                         0).astype(np.uint8)
            qo, go = int(m["queryoffset"][ul[k, 0]]), int(m["genomeoffset"][ul[k, 0]])
            po, co = oracle.traceback(kind, D, br, bc, s.q, s.q.upper(), s.g, s.galt, qo, go, revp)
            req = H.tb_request(int(ul[k, 0 if kind == 1 else 1]), s, br, bc, qoff=qo, goff=go)
            starts.append((req, po, co))
        pairs, counts = b.traceback(np.array([x[0] for x in starts]))
        for j, (_, po, co) in enumerate(starts):
            assert (H.pairs_to_records(pairs[j]) == po).all() and (counts[j, 1:] == co).all(), ("end traceback", j)


@pytest.mark.parametrize("qi", [False, True])
def test_end_best_cell_kept_during_the_fill(oracle, qi):
    """keep_scores 2 / 3: each one-sided fill keeps its triangle's best end
    cell; the merged pair equals the scan of fully kept scores and the
    oracle's find_best_endpoint, and the fills' directions are unchanged."""
    from gmapdp.batch import DeviceBatch
    from gmapdp.workload import Workload
    w = Workload(4, 400, genome_len=1 << 20)
    ul = np.stack([np.arange(0, len(w.problems), 2), np.arange(1, len(w.problems), 2)], 1).astype(np.int32)
    full = DeviceBatch(w.problems, w.qbuf.tobytes(), w.gbuf.tobytes())
    full.fill()
    want = full.best_endpoint(ul, queryend_indels=qi)
    p = w.problems.copy()
    p["keep_scores"] = 3 if qi else 2
    b = DeviceBatch(p, w.qbuf.tobytes(), w.gbuf.tobytes())
    b.fill()
    got = b.best_endpoint(ul, queryend_indels=qi)
    assert (got == want).all(), np.argwhere((got != want).any(1))[:5]
    other = b.best_endpoint(ul, queryend_indels=not qi)  # kept for the other scan
    assert (other[:, 0] == np.iinfo(np.int32).min).all()
    for k in range(0, len(ul), 37):
        pu = p[ul[k, 0]]
        revp = bool(w.meta["revp"][ul[k, 0]])
        q = bytes(w.qbuf[pu["qoff"]: pu["qoff"] + pu["rl"]])
        g = bytes(w.gbuf[pu["goff"]: pu["goff"] + pu["gl"]])
        if revp:
            q, g = q[::-1], g[::-1]
        su = H.Spec(1, int(pu["width"]), q, g, g, 3, int(pu["open"]), int(pu["extend"]), int(pu["lband"]),
                    int(pu["uband"]), bool(pu["jump_late"]), revp)
        sl = H.Spec(2, su.width, q, g, g, 3, su.open, su.ext, su.lband, su.uband, su.late, revp)
        Hu, Du = H.oracle_fill(oracle, su)
        Hl, Dl = H.oracle_fill(oracle, sl)
        assert tuple(got[k]) == oracle.best_endpoint(qi, Hu, Hl, su.width, su.rl, su.gl, su.lband, su.uband,
                                                     su.late), ("best end cell", k)
        for i, s_, D in ((ul[k, 0], su, Du), (ul[k, 1], sl, Dl)):
            Dg, _ = b.export(int(i))
            B = H.band_mask(s_)
            assert (Dg[B] == D[B]).all(), ("directions", k)


def test_traceback_nogaps(oracle):
    # traceback_nogaps (dynprog_end.c:648-722) from min(rl, gl) on both axes
    specs = [s for s in H.random_specs(21, 200, kinds=(0,))]
    b = _run(oracle, specs, keep_scores=False)
    reqs = []
    for i, s in enumerate(specs):
        m = min(s.rl, s.gl)
        rq = H.tb_request(i, s, m, m, genestrand=i % 3)
        rq["flags"] = 2  # GDP_TB_NOGAPS
        reqs.append(rq)
    pairs, counts = b.traceback(np.array(reqs))
    for i, s in enumerate(specs):
        m = min(s.rl, s.gl)
        want, wc = oracle.traceback_nogaps(m, m, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp, i % 3)
        got = H.pairs_to_records(pairs[i])
        assert counts[i, 0] == len(want) and (got == want).all(), ("nogaps pairs", i)
        assert (counts[i, 1:4] == wc[:3]).all() and counts[i, 4] == 0 and counts[i, 5] == 0, ("nogaps counters", i)


def test_end_mix_device_requests(oracle):
    """Config-4 mix: best end cells turned into traceback requests on the
    device (gdp_best_endpoint_ex) give the requests and pairs of the host
    path, for the single-gap half and the end half in one traceback launch,
    also when the whole step is replayed from a captured graph."""
    import torch
    from gmapdp.batch import DeviceBatch
    from gmapdp.lib import PAIR_DTYPE, TBREQ_DTYPE
    from gmapdp.workload import EndMix
    w = EndMix(150, 100, genome_len=1 << 20)
    p = w.problems
    b = DeviceBatch(p, w.qbuf.tobytes(), w.gbuf.tobytes())
    reqs = np.concatenate([w.single_requests(), w.end_templates()])
    reqs["pair_off"][1:] = np.cumsum(reqs["max_pairs"].astype(np.int64))[:-1]
    h = b.prepare_traceback(reqs, genome=w.genome())
    d_ul = torch.from_numpy(w.ul.reshape(-1).copy()).to(b.dev)
    d_tmpl = torch.from_numpy(reqs[w.n3:].view(np.uint8).copy()).to(b.dev)
    d_best = torch.zeros(3 * w.m, dtype=torch.int32, device=b.dev)

    def step(st=None):
        b.fill(stream=st)
        b.best_endpoint_requests(d_ul, w.m, False, d_best, d_tmpl.data_ptr(),
                                 h["d_reqs"].data_ptr() + w.n3 * TBREQ_DTYPE.itemsize, stream=st)
        b.traceback_device(h, stream=st)

    step()
    torch.cuda.synchronize()
    best = d_best.cpu().numpy().reshape(-1, 3)
    got_reqs = h["d_reqs"].cpu().numpy().view(TBREQ_DTYPE)
    counts = h["d_counts"].cpu().numpy().reshape(-1, 6)
    pairs = h["d_pairs"].cpu().numpy().view(PAIR_DTYPE)
    want_best = b.best_endpoint(w.ul)
    assert (best == want_best).all()
    for k in range(w.m):
        br, bc = int(best[k, 1]), int(best[k, 2])
        r = got_reqs[w.n3 + k]
        assert (r["r"], r["c"]) == (br, bc)
        assert r["fill"] == w.ul[k, 0 if bc >= br else 1]
    # every request's pairs against a one-request-at-a-time host traceback
    host_reqs = got_reqs.copy()
    hp, hc = b.traceback(host_reqs, genome=w.genome())
    for k in range(len(host_reqs)):
        off, n = int(host_reqs["pair_off"][k]), int(counts[k, 0])
        assert (counts[k] == hc[k]).all(), k
        assert (pairs[off:off + n] == hp[k]).all(), k
    # a captured graph of the whole step replays to the same bits
    want = (h["d_pairs"].clone(), h["d_counts"].clone(), d_best.clone())
    g = b.capture(step)
    try:
        for rep in range(3):
            for t in (h["d_pairs"], h["d_counts"], d_best):
                t.zero_()
            torch.cuda.synchronize()
            b.graph_launch(g)
            torch.cuda.synchronize()
            for x, y in zip((h["d_pairs"], h["d_counts"], d_best), want):
                assert torch.equal(x, y), rep
    finally:
        b.graph_destroy(g)


@pytest.mark.parametrize("keep", [True, False])
def test_wide_bands(oracle, keep):
    """Bands spanning 1000-2700 columns (Dynprog_compute_bands with
    widebandp over a short read and a long window, or the reverse) take the
    wide single sweep; bands past the matrix's edges are narrowed by the
    planner.  Cells, final scores and walks against the oracle."""
    specs = H.wide_specs(91 if keep else 92, 36)
    b = _run(oracle, specs, keep_scores=keep)
    assert b.kernels & 0x40, "no problem took the wide single sweep"
    _check_fills(oracle, specs, b)
    _check_tracebacks(oracle, specs, b, 93)
