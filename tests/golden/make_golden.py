#!/usr/bin/env python3
"""Writes tests/golden/golden_dp.npz from the reference's own DP objects.

Run in a container where the reference tree is mounted, after
`make -C oracle ref` (which compiles the reference sources in place into
oracle/_ref/libgmapref.so).  The fixture holds inputs and the reference's
outputs only: in-band direction bytes and scores (dense for small problems,
SHA-256 digests for large ones), corner scores, and traceback pair lists
with their counters.  Tracebacks are kept only where the reference is
deterministic: each fill runs twice with different contents left in its
workspace (0x00 and 0xff), and a traceback whose output changes between the
two runs read memory the fill never wrote.

Sets: random adversarial problems of every kind/width/tie rule/strand, the
int8 length thresholds 23/24/40/41/62/63/126/127, and samples of the four
benchmark configurations.
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


def ref_fill(R, s, garbage):
    if s.kind == 0:
        return R.fill_full(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp,
                           garbage=garbage)
    if s.kind == 1:
        return R.fill_upper(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.uband, s.late, s.revp, garbage=garbage)
    return R.fill_lower(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.late, s.revp, garbage=garbage)


def threshold_specs():
    rng = np.random.default_rng(31)
    out = []
    for L in (23, 24, 40, 41, 62, 63, 126, 127):
        for mt, op, ex in ((0, -8, -3), (2, -6, -1), (3, -10, -2)):
            g = H.rand_seq(rng, L)
            q = H.mutate(rng, g, 0.3, 0.02)[:L]
            q = q + H.rand_seq(rng, L - len(q))
            for kind in (0, 1, 2):
                out.append(H.Spec(kind, 8, q, g, g, mt, op, ex, 5, 5, bool(rng.integers(0, 2)), False))
    return out


def config_specs():
    from gmapdp.workload import Workload
    out = []
    for cfg, n, variant in ((1, 12, 0), (1, 12, 1), (2, 12, 0), (3, 24, 0), (4, 12, 0)):
        w = Workload(cfg, n, genome_len=1 << 20, variant=variant)
        for p, m in zip(w.problems, w.meta):
            q = bytes(w.qbuf[p["qoff"]: p["qoff"] + p["rl"]])
            g = bytes(w.gbuf[p["goff"]: p["goff"] + p["gl"]])
            ga = bytes(w.gbuf[p["galtoff"]: p["galtoff"] + p["gl"]])
            revp = bool(m["revp"])
            if revp:  # the generator stores sweep order; the reference gets the reversed pointers
                q, g, ga = q[::-1], g[::-1], ga[::-1]
            out.append(H.Spec(int(p["kind"]), int(p["width"]), q, g, ga, int(p["mismatchtype"]), int(p["open"]),
                              int(p["extend"]), int(p["lband"]), int(p["uband"]), bool(p["jump_late"]), revp))
    return out


def main():
    R = po.Reference()
    O = po.Oracle()
    sets = [("adversarial", H.random_specs(4242, 240, maxrl8=40, maxrl16=80)),
            ("thresholds", threshold_specs()), ("configs", config_specs())]
    rng = np.random.default_rng(99)
    rec = {k: [] for k in ("set", "kind", "width", "mt", "open", "ext", "lband", "uband", "late", "revp", "rl", "gl",
                           "qoff", "goff", "galtoff", "tb_r", "tb_c", "tb_n", "tb_off", "dense_off", "final")}
    qs, gs, dense_d, dense_h, tb_pairs, tb_counts, dig = [], [], [], [], [], [], []
    qo = go = doff = toff = 0
    nondet = 0
    for si, (name, specs) in enumerate(sets):
        for s in specs:
            Ha, Da = ref_fill(R, s, 0)
            B = H.band_mask(s)
            # Traceback start: the corner for full-band problems, a random in-band cell otherwise.
            start = None
            if s.kind == 0 and s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl:
                start = (s.rl, s.gl)
            elif s.kind != 0:
                cells = np.argwhere(B)
                if len(cells):
                    a, b = cells[rng.integers(0, len(cells))]
                    start = (int(b), int(a)) if s.kind == 1 else (int(a), int(b))
            ta = R.traceback(s.kind, None, start[0], start[1], s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp) \
                if start else None
            Hb, Db = ref_fill(R, s, 255)
            tbb = R.traceback(s.kind, None, start[0], start[1], s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp) \
                if start else None
            assert (Ha[B] == Hb[B]).all() and (Da[B] == Db[B]).all()
            det = start is not None and ta[0].shape == tbb[0].shape and (ta[0] == tbb[0]).all() and \
                (ta[1] == tbb[1]).all()
            if start is not None and not det:
                nondet += 1
            rec["set"].append(si)
            for k, v in (("kind", s.kind), ("width", s.width), ("mt", s.mt), ("open", s.open), ("ext", s.ext),
                         ("lband", s.lband), ("uband", s.uband), ("late", int(s.late)), ("revp", int(s.revp)),
                         ("rl", s.rl), ("gl", s.gl)):
                rec[k].append(v)
            rec["qoff"].append(qo)
            rec["goff"].append(go)
            rec["galtoff"].append(go + s.gl)
            qs.append(s.q)
            gs.append(s.g + s.galt)
            qo += s.rl
            go += 2 * s.gl
            if s.kind == 0 and s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl:
                rec["final"].append(int(Ha[s.gl, s.rl]))
            else:
                rec["final"].append(-(1 << 31))
            if det:
                rec["tb_r"].append(start[0])
                rec["tb_c"].append(start[1])
                rec["tb_n"].append(len(ta[0]))
                rec["tb_off"].append(toff)
                tb_pairs.append(ta[0])
                tb_counts.append(ta[1])
                toff += len(ta[0])
            else:
                rec["tb_r"].append(-1)
                rec["tb_c"].append(-1)
                rec["tb_n"].append(0)
                rec["tb_off"].append(toff)
                tb_counts.append(np.zeros(5, np.int32))
            if B.sum() <= 6000:
                rec["dense_off"].append(doff)
                dense_d.append(Da[B])
                dense_h.append(Ha[B])
                doff += int(B.sum())
                dig.append("")
            else:
                rec["dense_off"].append(-1)
                dig.append(digest(Da[B]) + digest(Ha[B]))
            # The oracle must already agree (pinned separately in the tests).
            Ho, Do = H.oracle_fill(O, s)
            assert (Ho[B] == Ha[B]).all() and (Do[B] == Da[B]).all(), ("oracle disagrees", name)
    out = {k: np.array(v, np.int32) for k, v in rec.items()}
    out["qbuf"] = np.frombuffer(b"".join(qs), np.uint8)
    out["gbuf"] = np.frombuffer(b"".join(gs), np.uint8)
    out["dense_dirs"] = np.concatenate(dense_d).astype(np.uint8)
    out["dense_scores"] = np.concatenate(dense_h).astype(np.int16)
    out["tb_pairs"] = np.concatenate(tb_pairs).astype(np.int32) if tb_pairs else np.zeros((0, 8), np.int32)
    out["tb_counts"] = np.stack(tb_counts).astype(np.int32)
    out["digests"] = np.array(dig)
    path = os.path.join(HERE, "golden_dp.npz")
    np.savez_compressed(path, **out)
    print("wrote %s: %d problems, %d dense cells, %d traceback pairs, %d reference-nondeterministic tracebacks"
          % (path, len(rec["kind"]), doff, toff, nondet))


if __name__ == "__main__":
    main()
