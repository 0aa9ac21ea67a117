"""GPU parity of the device genome and the serving loop.

- gdp_genome_segments against the oracle's Genome_get_segment_right/left
  over the reference's own test chromosome (tests/golden/chr17test.fa.gz),
  in both genome formats, with out-of-bounds and reverse-complement cases;
- genome skips of the traceback read through a genomecomp image equal the
  oracle's get_genomic_nt;
- every served request equals the oracle's Dynprog_single_gap DP (segment,
  bands, kernel width, fill, final score, traceback), and the same request
  gives the same result alone and inside a mix of others.
"""
import gzip
import os

import numpy as np
import pytest

import genomecomp as GC
import helpers as H

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _chr17():
    fa = gzip.open(os.path.join(HERE, "golden", "chr17test.fa.gz")).read().decode()
    return "".join(line.strip() for line in fa.splitlines() if not line.startswith(">")).encode()


def _segments(rng, n, nbases):
    from gmapdp.lib import SEGMENT_DTYPE
    segs = np.zeros(n, SEGMENT_DTYPE)
    off = 0
    for k in range(n):
        length = int(rng.integers(0, 700))
        left = int(rng.integers(0, 2))
        coord = int(rng.integers(0, nbases + 50))
        bound = int(rng.integers(max(0, coord - 900), coord + 900)) if rng.random() < 0.3 else \
            (nbases if not left else 0)
        segs[k] = (coord, bound, length, left, int(rng.integers(0, 2)), 0, off)
        off += length + int(rng.integers(0, 5))
    return segs


def test_segments_match_oracle(oracle):
    from gmapdp import serve
    seq = _chr17()
    n = len(seq)
    gc = GC.pack(seq)
    genome = GC.decode(gc, n)
    rng = np.random.default_rng(21)
    segs = _segments(rng, 400, n)
    for fmt, data in ((1, gc), (0, np.frombuffer(seq, np.uint8))):
        dg = serve.DeviceGenome(data, n, fmt)
        got = dg.segments(segs)
        for k, s in enumerate(segs):
            L = int(s["length"])
            if s["left"]:
                want = oracle.segment_left(genome, int(s["coord"]), L, int(s["bound"]), bool(s["revcomp"]))
            else:
                want = oracle.segment_right(genome, int(s["coord"]), L, int(s["bound"]), bool(s["revcomp"]))
            assert got[k] == want, (fmt, k, s)


def test_traceback_genome_skips_read_the_packed_genome(oracle):
    """Full-band walks that end in a genome skip read its bases from the
    genome (get_genomic_nt, pairpool.c:1016-1063), on both strands and
    across the chromosome bounds."""
    from gmapdp.batch import DeviceBatch
    seq = _chr17()
    n = len(seq)
    gc = GC.pack(seq)
    genome = GC.decode(gc, n)
    rng = np.random.default_rng(5)
    specs, tb = [], []
    for i in range(120):
        rl = int(rng.integers(1, 60))
        gl = rl + int(rng.integers(1, 12))
        g = H.rand_seq(rng, gl)
        q = H.rand_seq(rng, rl)
        s = H.Spec(0, 16 if i % 2 else 8, q, g, g, 0, -8, -3, 2, gl - rl + 2, bool(i % 3), bool(i % 4 == 1))
        specs.append(s)
        goff = int(rng.integers(0, n))
        chroffset = int(rng.integers(0, n // 2))
        chrhigh = int(rng.integers(chroffset, n + 10))
        tb.append((goff, chroffset, chrhigh, bool(rng.integers(0, 2))))
    probs, qbuf, gbuf = H.pack_batch(specs)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    reqs = np.array([H.tb_request(i, s, s.rl, s.gl, 3, t[0], t[1], t[2], t[3]) for i, (s, t) in
                     enumerate(zip(specs, tb))])
    got_c, cnt_c = b.traceback(reqs, genome=gc, genome_format=1, genome_len=n)
    got_a, cnt_a = b.traceback(reqs, genome=np.frombuffer(seq, np.uint8), genome_format=0)
    skips = 0
    for i, (s, t) in enumerate(zip(specs, tb)):
        Ho, Do = H.oracle_fill(oracle, s)
        Do = np.where(H.computed_mask(s), Do, 0).astype(np.uint8)
        po, co = oracle.traceback(0, Do, s.rl, s.gl, s.q, s.q.upper(), s.g, s.galt, 3, t[0], s.revp, t[1], t[2],
                                  t[3], 0, genome=genome)
        for got, cnt in ((got_c, cnt_c), (got_a, cnt_a)):
            assert cnt[i, 0] == len(po), i
            assert (H.pairs_to_records(got[i]) == po).all(), i
            assert (cnt[i, 1:] == co).all(), i
        skips += int(((po[:, 2] == ord(" ")) & (po[:, 4] != ord("*"))).sum())
    assert skips > 50  # genome bases were read, not only '*'


def _serve_oracle(O, genome, nbases, r, read):
    """Dynprog_single_gap's DP (dynprog_single.c:544-629) on the CPU."""
    rl, gl = int(r["rl"]), int(r["gl"])
    chroffset, chrhigh, goffset = int(r["chroffset"]), int(r["chrhigh"]), int(r["goffset"])
    if r["watsonp"]:
        seg = O.segment_right(genome, chroffset + goffset, gl, chrhigh, False)
    else:
        seg = O.segment_left(genome, chrhigh - goffset + 1, gl, chroffset, True)
    eb = int(r["extraband"])
    if not r["widebandp"]:
        lb = ub = eb
    elif gl >= rl:
        lb, ub = eb, gl - rl + eb
    else:
        lb, ub = rl - gl + eb, eb
    u8 = O.use8p(int(r["mismatchtype"]))
    width = 8 if (rl < u8 and gl < u8) else 16
    s = H.Spec(0, width, read, seg, seg, int(r["mismatchtype"]), int(r["open"]), int(r["extend"]), lb, ub,
               bool(r["jump_late"]), False)
    Ho, Do = H.oracle_fill(O, s)
    Do = np.where(H.computed_mask(s), Do, 0).astype(np.uint8)
    po, co = O.traceback(0, Do, rl, gl, read, read.upper(), seg, seg, int(r["roffset"]), goffset, False,
                         chroffset, chrhigh, bool(r["watsonp"]), int(r["genestrand"]), genome=genome)
    return int(Ho[gl, rl]), width, po, co


def _read(reads, r):
    from gmapdp import serve
    k = int(r["read"]) - reads.ctypes.data
    return reads[k:k + int(r["rl"])].tobytes()


@pytest.fixture(scope="module")
def small_genome():
    from gmapdp import serve
    nbases = 3_000_017
    gc = serve.gen_genomecomp(1005, nbases)
    return gc, nbases, GC.decode(gc, nbases)


def _check_result(O, genome, nbases, req, read, res, pairs):
    fs, width, po, co = _serve_oracle(O, genome, nbases, req, read)
    tag = int(req["tag"])
    assert res["status"] == 0, tag
    assert int(res["width"]) == width, tag
    assert int(res["finalscore"]) == fs, (tag, int(res["finalscore"]), fs)
    assert int(res["counts"][0]) == len(po), tag
    assert (H.pairs_to_records(pairs) == po).all(), tag
    assert (res["counts"][1:] == co).all(), tag


@pytest.mark.parametrize("long_rows,plan", [(0, (112, 128)), (300, (112, 128)), (300, (240, 256, 8, 1))])
def test_server_matches_oracle(oracle, small_genome, long_rows, plan):
    """long_rows > 0 batches long and short reads apart (two queues); a
    plan with mix_span > 0 sweeps requests of neighbouring bands in shared
    waves."""
    from gmapdp import serve
    gc, nbases, genome = small_genome
    reqs, reads = serve.gen_requests(1005, 0, 300, gc, nbases)
    reqs["genestrand"] = np.arange(300) % 3
    reqs["roffset"] = np.arange(300) % 7
    dg = serve.DeviceGenome(gc, nbases)
    with serve.Server(dg, max_batch=64, slots=3, long_rows=long_rows, plan=plan) as srv:
        srv.submit(reqs)
        out = srv.collect(len(reqs))
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 5
    for r in reqs:
        res, pairs = out[int(r["tag"])]
        _check_result(oracle, genome, nbases, r, _read(reads, r), res, pairs)


@pytest.mark.parametrize("slots,plan", [(3, None), (12, (240, 256, 8, 1))])
def test_server_result_alone_equals_result_in_mix(oracle, small_genome, slots, plan):
    """Each request's result is the same served alone and served inside a
    stream of other requests arriving while batches run (slots form their
    batches on threads of their own)."""
    from gmapdp import serve
    gc, nbases, genome = small_genome
    reqs, reads = serve.gen_requests(77, 0, 2000, gc, nbases)
    dg = serve.DeviceGenome(gc, nbases)
    probe = [3, 50, 400, 1234, 1999]
    with serve.Server(dg, max_batch=256, slots=slots, plan=plan) as srv:
        alone = {}
        for k in probe:
            srv.submit(reqs[k:k + 1])
            alone.update(srv.collect(1))
        # the mix: all 2000 in chunks submitted while earlier chunks run
        for a in range(0, 2000, 100):
            srv.submit(reqs[a:a + 100])
        mixed = srv.collect(2000)
        st = srv.stats()
    assert st["inflight_peak"] >= 2, "batches never overlapped"
    for k in probe:
        ra, pa = alone[k]
        rm, pm = mixed[k]
        for f in ("status", "finalscore", "width"):
            assert ra[f] == rm[f], (k, f)
        assert (ra["counts"] == rm["counts"]).all(), k
        assert (pa == pm).all(), k
        _check_result(oracle, genome, nbases, reqs[k], _read(reads, reqs[k]), rm, pm)
    # and a sample of the mix against the oracle
    for k in range(0, 2000, 97):
        rm, pm = mixed[k]
        _check_result(oracle, genome, nbases, reqs[k], _read(reads, reqs[k]), rm, pm)


def test_server_scoring_by_defect_rate(oracle, small_genome):
    """scoring = 1 picks Dynprog_single_gap's mismatch type and penalties
    from defect_rate (dynprog_single.c:458-481); the results equal those of
    the same requests with the choice written out, and the oracle's."""
    from gmapdp import serve
    gc, nbases, genome = small_genome
    reqs, reads = serve.gen_requests(1007, 0, 240, gc, nbases)
    rates = np.array([0.0, 0.001, 0.0029999, 0.003, 0.01, 0.0139999, 0.014, 0.05])
    by_rate = reqs.copy()
    by_rate["scoring"] = 1
    by_rate["defect_rate"] = rates[np.arange(len(reqs)) % len(rates)]
    by_rate["mismatchtype"], by_rate["open"], by_rate["extend"] = 0, 0, 0  # ignored with scoring 1
    explicit = reqs.copy()
    hq, mq = by_rate["defect_rate"] < 0.003, by_rate["defect_rate"] < 0.014
    explicit["mismatchtype"] = np.where(hq, 0, np.where(mq, 1, 2))
    explicit["open"] = np.where(hq, -8, np.where(mq, -7, -6))
    explicit["extend"] = np.where(hq, -3, np.where(mq, -2, -1))
    dg = serve.DeviceGenome(gc, nbases)
    with serve.Server(dg, max_batch=64, slots=3) as srv:
        srv.submit(by_rate)
        got = srv.collect(len(reqs))
        srv.submit(explicit)
        want = srv.collect(len(reqs))
    for k in range(len(reqs)):
        tag = int(reqs["tag"][k])
        rg, pg = got[tag]
        rw, pw = want[tag]
        assert rg["status"] == 0 and rw["status"] == 0, k
        for f in ("finalscore", "width"):
            assert rg[f] == rw[f], (k, f)
        assert (rg["counts"] == rw["counts"]).all(), k
        assert (pg == pw).all(), k
    for k in range(0, len(reqs), 23):
        rg, pg = got[int(reqs["tag"][k])]
        _check_result(oracle, genome, nbases, explicit[k], _read(reads, explicit[k]), rg, pg)


def test_server_admission_by_direction_budget(oracle, small_genome):
    """A direction budget far below one batch's worth forces the server to
    split and serialise batches; results stay exact and the bytes in flight
    never exceed the budget."""
    from gmapdp import serve
    gc, nbases, genome = small_genome
    reqs, reads = serve.gen_requests(99, 0, 400, gc, nbases)
    budget = 1 << 20
    dg = serve.DeviceGenome(gc, nbases)
    with serve.Server(dg, dir_budget_bytes=budget, max_batch=4096, slots=4) as srv:
        srv.submit(reqs)
        out = srv.collect(len(reqs))
        st = srv.stats()
    assert 0 < st["dir_bytes_peak"] <= budget
    assert st["batches"] > 4
    for k in range(0, 400, 13):
        res, pairs = out[k]
        _check_result(oracle, genome, nbases, reqs[k], _read(reads, reqs[k]), res, pairs)


def test_server_refuses_what_the_reference_refuses(small_genome):
    """rl or gl <= 0 or above the server's limits: status GDP_ETOOBIG and
    traceback_score NEG_INFINITY_32 (dynprog_single.c:508-520)."""
    from gmapdp import serve
    gc, nbases, _ = small_genome
    reqs, reads = serve.gen_requests(5, 0, 4, gc, nbases)
    reqs["rl"][0] = 0
    reqs["gl"][1] = 0
    reqs["rl"][2] = 1001
    dg = serve.DeviceGenome(gc, nbases)
    with serve.Server(dg, max_batch=16, slots=1, max_rl=1000) as srv:
        srv.submit(reqs)
        out = srv.collect(4)
    for k in range(3):
        r, p = out[k]
        assert r["status"] == -2 and r["counts"][1] == -32768 and r["counts"][0] == 0 and len(p) == 0
    assert out[3][0]["status"] == 0


def _window(gc, nbases, r):
    """The oracle's genome model of the bases a request can read: its
    segment and the genome skips of its walk, with a margin."""
    gl, chroffset, chrhigh, goffset = int(r["gl"]), int(r["chroffset"]), int(r["chrhigh"]), int(r["goffset"])
    start = chroffset + goffset if r["watsonp"] else chrhigh - goffset - gl
    return GC.decode_window(gc, nbases, start - 64, gl + 128)


def test_server_on_the_whole_genome(oracle):
    """Config 5's genome: 3.1 Gbp of genomecomp on the device, cut into
    chromosomes of 2^27 bases, so that requests carry chromosome offsets past
    2^31.  300 requests served together, each checked against the oracle's
    Dynprog_single_gap DP over the bases it reads (decoded window by
    window)."""
    from gmapdp import serve
    nbases = 3_100_000_000
    gc = serve.gen_genomecomp(1005, nbases)
    reqs, reads = serve.gen_requests(1005, 0, 300, gc, nbases)
    reqs["roffset"] = np.arange(300) % 11
    high = reqs["chroffset"].astype(np.int64) >= 1 << 31
    assert high.sum() >= 50, "too few requests past 2^31"
    dg = serve.DeviceGenome(gc, nbases)
    with serve.Server(dg, max_batch=128, slots=2) as srv:
        srv.submit(reqs)
        out = srv.collect(len(reqs))
    del dg
    skipped = 0
    for r in reqs:
        res, pairs = out[int(r["tag"])]
        _check_result(oracle, _window(gc, nbases, r), nbases, r, _read(reads, r), res, pairs)
        skipped += int((pairs["cdna"] == ord(" ")).sum())
    assert skipped > 0, "no walk crossed a genome gap"
