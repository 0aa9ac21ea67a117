#!/usr/bin/env python3
"""Writes tests/golden/bridge_golden.npz from the reference's cDNA gap bridge.

Run in a container where the reference tree is mounted, after
`make -C oracle ref`, which compiles dynprog_cdna.c in place (through
oracle/ref_cdna_shim.c) into oracle/_ref/libgmapref.so.  Each case is the
input of bridge_cdna_gap_{8,16}_ud -- four score tables, lengths, bands,
penalties, offsets and the tie rule -- and the reference's output
(bestscore, bestcL, bestcR, bestrL, bestrR).  Half the cases take their
tables from the oracle's one-sided fills of random sequences (the bridge's
real input); the others are random tables, many of them flat, to exercise
ties.  Every case keeps rl >= gl-1 on both ends, where the reference reads
only rows its fills wrote.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "gmap-gsnap_amd")]

import helpers as H  # noqa: E402
import pyoracle as po  # noqa: E402

FIELDS = ("width", "gl", "rlL", "rlR", "lbandL", "ubandL", "lbandR", "ubandR", "open", "ext", "leftoffset",
          "rightoffset", "late")


def fill_tables(O, rng, width, gl, rlL, rlR, bands, open_, ext, late):
    """Score tables of the four one-sided fills of one gap, from the oracle."""
    g = H.rand_seq(rng, gl)
    qL = H.mutate(rng, g, 0.05, 0.02)[:rlL].ljust(rlL, b"A")
    qR = H.mutate(rng, g[::-1], 0.05, 0.02)[:rlR].ljust(rlR, b"C")
    lbL, ubL, lbR, ubR = bands
    LU = O.fill_upper(width, qL, g, g, 3, open_, ext, ubL, late, False)[0]
    LL = O.fill_lower(width, qL, g, g, 3, open_, ext, lbL, late, False)[0]
    RU = O.fill_upper(width, qR, g[::-1], g[::-1], 3, open_, ext, ubR, not late, False)[0]
    RL = O.fill_lower(width, qR, g[::-1], g[::-1], 3, open_, ext, lbR, not late, False)[0]
    return LU, LL, RU, RL


def main():
    O, R = po.Oracle(), po.Reference()
    rng = np.random.default_rng(4242)
    rows, tabs, offs, outs = [], [], [0], []
    for t in range(160):
        gl = int(rng.integers(1, 22))
        rlL = int(rng.integers(max(1, gl - 1), 26))
        rlR = int(rng.integers(max(1, gl - 1), 26))
        width = int(rng.choice([8, 16]))
        late = int(rng.integers(0, 2))
        bands = [int(x) for x in rng.integers(0, 10, 4)]
        open_, ext = int(rng.integers(-12, -5)), int(rng.integers(-3, 0))
        left = int(rng.integers(0, 1000))
        right = left + int(rng.integers(0, 50))
        if t % 2 == 0:
            A = fill_tables(O, rng, width, gl, rlL, rlR, bands, open_, ext, bool(late))
            A = [np.asarray(a, np.int16).reshape(-1) for a in A]
        else:
            lo, hi = (-128, 128) if width == 8 else (-400, 80)
            sh = [(gl + 1) * (rlL + 1), (rlL + 1) * (gl + 1), (gl + 1) * (rlR + 1), (rlR + 1) * (gl + 1)]
            div = 1 if t % 4 == 1 else 60
            A = [(rng.integers(lo, hi, z) // div).astype(np.int16) for z in sh]
        args = (width, *A, gl, rlL, rlR, *bands, open_, ext, left, right, late)
        outs.append(R.bridge_cdna(*args))
        rows.append((width, gl, rlL, rlR, *bands, open_, ext, left, right, late))
        for a in A:
            tabs.append(a)
            offs.append(offs[-1] + len(a))
    P = np.array(rows, np.int32)
    np.savez_compressed(os.path.join(HERE, "bridge_golden.npz"), **{f: P[:, i] for i, f in enumerate(FIELDS)},
                        tables=np.concatenate(tabs).astype(np.int16), table_off=np.array(offs, np.int64),
                        out=np.array(outs, np.int32))


if __name__ == "__main__":
    main()
