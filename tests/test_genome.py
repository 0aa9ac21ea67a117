"""Device-genome codec on the host: GMAP's genomecomp image as gmap_build
writes it, pinned by the reference's own setup test (tests/ss.chr17test ->
tests/setup.genomecomp.ok, kept as tests/golden/chr17test.fa.gz and the
image's SHA-256 and sampled words), plus the serving workload generator.
No GPU calls."""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

import genomecomp as GC
from gmapdp import serve

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def chr17():
    fa = gzip.open(os.path.join(HERE, "golden", "chr17test.fa.gz")).read().decode()
    seq = "".join(line.strip() for line in fa.splitlines() if not line.startswith(">")).encode()
    gold = json.load(open(os.path.join(HERE, "golden", "genomecomp_golden.json")))
    return seq, gold


def _check_image(words, gold):
    raw = words.astype("<u4").tobytes()
    assert len(raw) == gold["bytes"]
    for i, w in gold["sample"].items():
        assert int(words[int(i)]) == w, ("word", i)
    assert hashlib.sha256(raw).hexdigest() == gold["sha256"]


def test_oracle_pack_is_gmap_build_image(chr17):
    seq, gold = chr17
    _check_image(GC.pack(seq), gold)


def test_library_pack_is_gmap_build_image(chr17):
    seq, gold = chr17
    _check_image(serve.genomecomp_pack(seq), gold)


def test_decode_round_trip(chr17):
    seq, _ = chr17
    gc = GC.pack(seq)
    codes, nflags = GC.decode(gc, len(seq))
    assert GC.to_ascii(codes, nflags).tobytes() == seq
    assert serve.decode(gc, len(seq)).tobytes() == seq


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pack_rules_edge_characters(seed):
    # lower case, N, X, IUPAC codes, punctuation and every tail length
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGTacgtNnXxRYKMSWBDHVU-.*", np.uint8)
    for n in list(range(0, 70)) + [1000, 4095, 4097]:
        s = alphabet[rng.integers(0, len(alphabet), n)].tobytes()
        a, b = GC.pack(s), serve.genomecomp_pack(s)
        assert (a == b).all(), n
        codes, nflags = GC.decode(a, n)
        want = np.array([c if c in b"ACGT" else ord("N") for c in s.upper()], np.uint8)
        assert (GC.to_ascii(codes, nflags) == want).all(), n
        # X keeps code 3 (T) under its flag; other flagged characters code 0
        for i, c in enumerate(s.upper()):
            if c == ord("X"):
                assert codes[i] == 3
            elif c not in b"ACGT":
                assert codes[i] == 0
        nb = (n + 31) // 32
        assert len(a) == 3 * nb + 2 and (a[-2:] == 0xFFFFFFFF).all()
        if n % 32:
            tail = a[3 * (nb - 1) + 2] >> (n % 32)
            assert tail == (0xFFFFFFFF >> (n % 32)), n  # padding X are flagged


def test_generated_genome_is_a_valid_image():
    n = 20_000_037
    gc = serve.gen_genomecomp(5, n)
    assert len(gc) == 3 * ((n + 31) // 32) + 2
    codes, nflags = GC.decode(gc, n)
    # blocks of N carry code 0, as the packer writes N
    assert (codes[nflags == 1] == 0).all()
    frac_n = nflags.mean()
    assert 0.0 <= frac_n < 0.05
    again = GC.pack(GC.to_ascii(codes, nflags).tobytes())
    assert (again == gc).all()
    # N comes in runs of at least 1 kb
    edges = np.flatnonzero(np.diff(np.concatenate([[0], nflags.astype(np.int8), [0]])))
    runs = edges[1::2] - edges[::2]
    assert len(runs) > 0 and runs.min() >= 1024 - 32
    counts = np.bincount(codes[nflags == 0], minlength=4) / (nflags == 0).sum()
    assert np.abs(counts - 0.25).max() < 0.01


def test_serving_requests_config3_mix():
    n = 1 << 24
    gc = serve.gen_genomecomp(5, n)
    reqs, reads = serve.gen_requests(1005, 0, 4000, gc, n)
    rl, gl = reqs["rl"], reqs["gl"]
    assert rl.min() >= 20 and rl.max() <= 1000
    assert (gl >= rl).all() and (gl <= np.round(rl * 1.5) + 1).all()
    assert reqs["extraband"].min() >= 3 and reqs["extraband"].max() <= 64
    assert 0.4 < reqs["watsonp"].mean() < 0.6
    assert (reqs["goffset"] >= 8).all() and (reqs["goffset"] + gl < n).all()
    # no request segment touches N
    ascii_all = serve.decode(gc, n)
    for k in range(0, 4000, 7):
        lo = reqs["goffset"][k] if reqs["watsonp"][k] else int(reqs["chrhigh"][k]) - reqs["goffset"][k] + 1 - gl[k]
        assert b"N" not in ascii_all[lo:lo + gl[k]].tobytes(), k
    assert (reqs["tag"] == np.arange(4000)).all()
    base = reads.ctypes.data
    assert (reqs["read"] == base + np.arange(4000) * serve.READ_STRIDE).all()
    # half of the problems are int8 under HIGHQ's use8p_size of 41 only when both lengths < 41
    assert ((rl < 41) & (gl < 41)).mean() > 0.05
    # a read is its segment (reverse complement on the minus strand) with one
    # gap and a few edits: most reads share their first 8 bases with it
    ascii_genome = ascii_all
    comp = bytes.maketrans(b"ACGTN", b"TGCAN")
    for strand in (1, 0):
        hits = 0
        ks = np.nonzero(reqs["watsonp"] == strand)[0][:200]
        for k in ks:
            if strand:
                seg = ascii_genome[reqs["goffset"][k]:reqs["goffset"][k] + gl[k]].tobytes()
            else:
                right = int(reqs["chrhigh"][k]) - int(reqs["goffset"][k]) + 1
                seg = ascii_genome[right - gl[k]:right].tobytes()[::-1].translate(comp)
            r = reads[k * serve.READ_STRIDE:k * serve.READ_STRIDE + rl[k]].tobytes()
            hits += r[:8] in seg
        assert hits > 0.6 * len(ks), (strand, hits)
