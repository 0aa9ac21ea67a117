#!/usr/bin/env python3
"""Writes tests/golden/intron_bridge_golden.npz from the reference's intron
bridge.

Run in a container where the reference tree is mounted, after
`make -C oracle ref`, which compiles dynprog_genome.c in place (through
oracle/ref_genome_shim.c) and maxent_hr.c into oracle/_ref/libgmapref.so.
The cases are helpers.intron_gap_cases(seed 77) over helpers.intron_genome():
the first 600, and seven from the first 30000 of seeds 77, 78 and 79
(EXTRA) whose result turns on the dinucleotide rule's bound of 2 *
PROB_CEILING on the best join's probability sum (found by scanning them with
that bound lowered in a copy of the oracle), and three (625, 1480, 1642 of
seed 77) whose best join sits at the last column the upper triangle's loop
reaches, min(r + uband, glength-1) - 1.  Each case stores the
read, both segments and their alternate alleles, the call's arguments, the splice-site probabilities
the reference's Maxent_hr models give each column (what the site-level
bridge reads), and the reference's result: bridge_intron_gap_{8,16}_ud over
the oracle's four one-sided fills (the fills are pinned by the fill vectors,
so the reader recomputes them instead of storing score tables).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "gmap-gsnap_amd")]

import helpers as H  # noqa: E402
import pyoracle as po  # noqa: E402

OUT = os.path.join(HERE, "intron_bridge_golden.npz")
SEED, SCAN, FIRST = 77, 30000, 600
EXTRA = ((77, 9975), (78, 15066), (78, 28101), (78, 28278), (79, 3144), (79, 3582), (79, 9484),
         (77, 625), (77, 1480), (77, 1642))
FIELDS = ("width", "mt", "open", "ext", "rlength", "glengthL", "glengthR", "lbandL", "ubandL", "lbandR", "ubandR",
          "leftoffset", "rightoffset", "chroffset", "chrhigh", "watsonp", "cdna_direction", "finalp", "halfp", "late")


def main():
    seq, gc = H.intron_genome()
    O, R = po.Oracle(), po.Reference()
    R.genome_setup(gc, len(seq))
    by_seed = {sd: H.intron_gap_cases(np.random.default_rng(sd), SCAN, seq) for sd in sorted({SEED, *dict(EXTRA)})}
    pick = [(SEED, i) for i in range(FIRST)] + list(EXTRA)
    cols = {k: [] for k in FIELDS}
    qb, gb, pb, outs, probs = [], [], [], [], []
    for sd, i in pick:
        c = by_seed[sd][i]
        Hs = [H.oracle_fill(O, s)[0] for s in H.intron_gap_specs(c)]
        pL, pR = R.site_probs(c["glengthL"], c["glengthR"], c["leftoffset"], c["rightoffset"], c["chroffset"],
                              c["chrhigh"], c["cdna_direction"], c["watsonp"])
        res, pr = R.bridge_intron(c["width"], *Hs, c["gL"], c["gLalt"], c["gRseg"], c["gRsegalt"], c["rlength"],
                                  c["lbandL"], c["ubandL"], c["lbandR"], c["ubandR"], c["leftoffset"],
                                  c["rightoffset"], c["chroffset"], c["chrhigh"], c["cdna_direction"], c["watsonp"],
                                  c["finalp"], c["halfp"], c["late"])
        for k in FIELDS:
            cols[k].append(int(c[k]))
        qb.append(c["q"])
        gb.append(c["gL"] + c["gRseg"] + c["gLalt"] + c["gRsegalt"])
        pb.append(np.concatenate([pL, pR]))
        outs.append(res)
        probs.append(pr)
    np.savez_compressed(OUT, seed_index=np.array(pick, np.int32), qbuf=np.frombuffer(b"".join(qb), np.uint8),
                        gbuf=np.frombuffer(b"".join(gb), np.uint8), site_probs=np.concatenate(pb),
                        out=np.array(outs, np.int32), out_probs=np.array(probs, np.float64),
                        **{k: np.array(v, np.int32) for k, v in cols.items()})
    print("wrote %s: %d cases" % (OUT, len(pick)))


if __name__ == "__main__":
    main()
