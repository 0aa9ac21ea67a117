#!/usr/bin/env python3
"""Writes tests/golden/std_golden.npz from the reference's scalar DP build.

Run in a container where the reference tree is mounted, after
`make -C oracle ref` (which compiles dynprog.c and its helpers in place,
without SIMD, into oracle/_ref_std/libgmapref_std.so).  The fixture holds
inputs and the reference's outputs only: for each problem, SHA-256 digests of
Dynprog_standard's dense direction nibbles (every cell) and of its scores
for r, c >= 1 (int32), the corner score, and the pair list and counters of
Dynprog_traceback_std from one start cell.  The scalar build clears its
matrices on every fill, so all of it is deterministic.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "gmap-gsnap_amd")]

import helpers as H  # noqa: E402
import pyoracle as po  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def edge_specs():
    """Band and length corners: single cells, zero-width bands, bands that
    miss the corner, the column-1 and row-1 special cases."""
    rng = np.random.default_rng(77)
    out = []
    for rl, gl, lb, ub in [(1, 1, 0, 0), (1, 6, 0, 5), (6, 1, 5, 0), (3, 40, 0, 37), (40, 3, 37, 0), (30, 30, 0, 0),
                           (50, 50, 1, 1), (20, 60, 2, 3), (60, 20, 3, 2), (17, 17, 16, 16)]:
        g = H.rand_seq(rng, gl)
        q = H.rand_seq(rng, rl)
        for late in (False, True):
            for sat, up, lo in ((-32768, True, True), (-10, True, False), (-32768, False, True)):
                out.append(H.Spec(3, 32, q, g, g, 1, -8, -2, lb, ub, late, late, saturation=sat, upperp=up,
                                  lowerp=lo))
    return out


def main():
    R = po.ReferenceStd()
    specs = H.random_std_specs(5151, 120, maxrl=120) + edge_specs()
    rng = np.random.default_rng(6)
    fields = {k: [] for k in ("rl", "gl", "qoff", "goff", "galtoff", "mt", "open", "ext", "lband", "uband", "late",
                              "revp", "saturation", "upperp", "lowerp", "final", "tb_r", "tb_c", "tb_n", "tb_off")}
    qbuf, gbuf, dig, tbp, tbc = bytearray(), bytearray(), [], [], []
    for s in specs:
        Hr, Dr = R.fill_std(s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp, s.saturation,
                            s.upperp, s.lowerp)
        B = po.band_mask_full(s.rl, s.gl, s.lband, s.uband)
        if B[s.gl, s.rl]:
            r, c = s.rl, s.gl
        else:
            cells = np.argwhere(B)
            c, r = (int(x) for x in cells[rng.integers(0, len(cells))]) if len(cells) else (-1, -1)
        n = 0
        if r >= 0:
            pairs, cnt = R.traceback(r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
            n = len(pairs)
            fields["tb_off"].append(sum(len(p) for p in tbp))
            tbp.append(pairs)
            tbc.append(cnt)
        else:
            fields["tb_off"].append(0)
            tbc.append(np.zeros(5, np.int32))
        vals = dict(rl=s.rl, gl=s.gl, qoff=len(qbuf), goff=len(gbuf), galtoff=len(gbuf) + s.gl, mt=s.mt, open=s.open,
                    ext=s.ext, lband=s.lband, uband=s.uband, late=int(s.late), revp=int(s.revp),
                    saturation=s.saturation, upperp=int(s.upperp), lowerp=int(s.lowerp),
                    final=int(Hr[s.gl, s.rl]) if B[s.gl, s.rl] else -(1 << 31), tb_r=r, tb_c=c, tb_n=n)
        for k, v in vals.items():
            fields[k].append(v)
        qbuf += s.q
        gbuf += s.g + s.galt
        dig.append(digest(Dr.astype(np.uint8)) + digest(Hr[1:, 1:].astype(np.int32)))
    out = {k: np.array(v, np.int32) for k, v in fields.items()}
    out["qbuf"] = np.frombuffer(bytes(qbuf), np.uint8)
    out["gbuf"] = np.frombuffer(bytes(gbuf), np.uint8)
    out["digests"] = np.array(dig)
    out["tb_pairs"] = np.concatenate(tbp).astype(np.int32)
    out["tb_counts"] = np.array(tbc, np.int32)
    np.savez_compressed(os.path.join(HERE, "std_golden.npz"), **out)
    print("wrote %d problems, %d traceback pairs" % (len(specs), len(out["tb_pairs"])))


if __name__ == "__main__":
    main()
