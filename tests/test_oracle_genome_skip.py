"""Genome skips read from a real genome: the oracle's restatement of
get_genomic_nt (pairpool.c:1016-1063) over Genome_get_char (genome.c:11003)
inside its tracebacks, pinned to the reference's own walks.

- against tests/golden/genome_skip_golden.npz, made by the reference's
  Dynprog_traceback_{8,16}(_upper) with a Genome_T over the genomecomp image
  of the reference's test chromosome (N runs put in), on both strands, with
  revp, and with chromosome bounds that some skipped bases fall outside;
- against the reference built in place (oracle/_ref, genome.c included) on
  fresh cases where it exists.
"""
import numpy as np
import pytest

import genomecomp as GC
import golden_io
import helpers as H


def _oracle_walk(oracle, genome, s, t):
    goff, chroffset, chrhigh, watsonp = t
    Ho, Do = H.oracle_fill(oracle, s)
    Do = np.where(H.computed_mask(s), Do, 0).astype(np.uint8)
    return oracle.traceback(s.kind, Do, s.rl, s.gl, s.q, s.q.upper(), s.g, s.galt, 3, goff, s.revp, chroffset,
                            chrhigh, watsonp, 0, genome=genome)


def test_oracle_genome_skips_match_golden(oracle):
    seq, gc, cases = golden_io.load_genome_skips()
    genome = GC.decode(gc, len(seq))
    read = star = gaps = short = 0
    for k, (s, t, want, wcnt) in enumerate(cases):
        po, co = _oracle_walk(oracle, genome, s, t)
        assert po.shape == want.shape and (po == want).all(), ("walk", k)
        assert (co == wcnt).all(), ("counters", k)
        a, b, c = golden_io.skip_kinds(want)
        read, star, gaps = read + a, star + b, gaps + c
        short += a + b > 0
    # the vectors exercise every branch: bases read, bases outside the bounds,
    # gap holders, over more than 500 walks
    assert len(cases) >= 500 and short > 250 and read > 500 and star > 100 and gaps > 150


@pytest.mark.ref
def test_oracle_genome_skips_match_reference(oracle, reference):
    seq, gc = golden_io.skip_genome_image()
    n = len(seq)
    genome = GC.decode(gc, n)
    for k, (s, t) in enumerate(H.genome_skip_cases(np.random.default_rng(77), 300, n)):
        goff, chroffset, chrhigh, watsonp = t
        if s.kind == 0:
            reference.fill_full(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp)
        else:
            reference.fill_upper(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.uband, s.late, s.revp)
        pr, cr = reference.traceback(s.kind, None, s.rl, s.gl, s.q, s.q.upper(), s.g, s.galt, 3, goff, s.revp,
                                     chroffset, chrhigh, watsonp, 0, genomecomp=(gc, n))
        po, co = _oracle_walk(oracle, genome, s, t)
        assert po.shape == pr.shape and (po == pr).all() and (co == cr).all(), ("walk", k)


def test_window_decode_equals_whole_decode():
    seq, gc = golden_io.skip_genome_image()
    n = len(seq)
    codes, nflags = GC.decode(gc, n)
    for start, count in ((0, 100), (12345, 777), (n - 50, 200), (199_990, 5)):
        c, f, base, length = GC.decode_window(gc, n, start, count)
        assert length == n and base == start
        assert (c == codes[start:start + count]).all() and (f == nflags[start:start + count]).all()
