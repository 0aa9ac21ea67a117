#!/usr/bin/env python3
"""Writes tests/golden/genome_skip_golden.npz: walks that end in a genome
skip, traced back by the reference itself with a real genome.

Run in a container where the reference tree is mounted, after
`make -C oracle ref`, which compiles the reference's dynprog_simd.c,
pairpool.c and genome.c in place into oracle/_ref/libgmapref.so.  The
genome is the reference's test chromosome (tests/golden/chr17test.fa.gz)
with N runs put in (helpers.skip_genome), packed as genomecomp; each walk's
skipped bases come from Pairpool_add_genomeskip -> get_genomic_nt ->
Genome_get_char over that image (pairpool.c:1016-1180, genome.c:11003),
with the walk's chromosome bounds, strand and genome offset.  The cases are
helpers.genome_skip_cases: skips of 1-8 bases (pushed base by base) and of
9-30 (a gap holder), full-band and upper fills, both widths, revp, both
strands, and bounds that some skipped bases fall outside of.

Stored per case: the fill's inputs, the walk's offsets and bounds, and the
reference's pair records (8 ints each) and counters.  The genome's SHA-256
is stored so that the reader can check it rebuilt the same image.
"""
import gzip
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "gmap-gsnap_amd")]

import genomecomp as GC  # noqa: E402
import helpers as H  # noqa: E402
import pyoracle as po  # noqa: E402

OUT = os.path.join(HERE, "genome_skip_golden.npz")
SEED, N = 2027, 640


def skip_image():
    fa = gzip.open(os.path.join(HERE, "chr17test.fa.gz")).read().decode()
    seq = H.skip_genome("".join(line.strip() for line in fa.splitlines() if not line.startswith(">")).encode())
    return seq, GC.pack(seq)


def main():
    seq, gc = skip_image()
    n = len(seq)
    R = po.Reference()
    cases = H.genome_skip_cases(np.random.default_rng(SEED), N, n)
    cols = {k: [] for k in ("kind", "width", "mt", "open", "ext", "lband", "uband", "late", "revp", "goff",
                            "chroffset", "chrhigh", "watsonp", "qoff", "rl", "gl", "goff_buf", "pair_off",
                            "npairs")}
    qbuf, gbuf, pairs, counters = [], [], [], []
    qo = go = po_ = 0
    for s, (goff, chroffset, chrhigh, watsonp) in cases:
        if s.kind == 0:
            R.fill_full(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp)
        else:
            R.fill_upper(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.uband, s.late, s.revp)
        pr, cr = R.traceback(s.kind, None, s.rl, s.gl, s.q, s.q.upper(), s.g, s.galt, 3, goff, s.revp, chroffset,
                             chrhigh, watsonp, 0, genomecomp=(gc, n))
        for k, v in (("kind", s.kind), ("width", s.width), ("mt", s.mt), ("open", s.open), ("ext", s.ext),
                     ("lband", s.lband), ("uband", s.uband), ("late", int(s.late)), ("revp", int(s.revp)),
                     ("goff", goff), ("chroffset", chroffset), ("chrhigh", chrhigh), ("watsonp", int(watsonp)),
                     ("qoff", qo), ("rl", s.rl), ("gl", s.gl), ("goff_buf", go), ("pair_off", po_),
                     ("npairs", len(pr))):
            cols[k].append(v)
        qbuf.append(s.q)
        gbuf.append(s.g)
        pairs.append(pr)
        counters.append(cr)
        qo += s.rl
        go += s.gl
        po_ += len(pr)
    np.savez_compressed(OUT, genome_sha256=np.frombuffer(hashlib.sha256(gc.tobytes()).digest(), np.uint8),
                        genome_bases=np.int64(n), qbuf=np.frombuffer(b"".join(qbuf), np.uint8),
                        gbuf=np.frombuffer(b"".join(gbuf), np.uint8), pairs=np.concatenate(pairs).astype(np.int32),
                        counters=np.array(counters, np.int32),
                        **{k: np.array(v, np.int64) for k, v in cols.items()})
    print("wrote %s: %d walks" % (OUT, len(cases)))


if __name__ == "__main__":
    main()
