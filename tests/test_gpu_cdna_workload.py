"""The cDNA gap workload as bench.py runs it (fills keeping every score, the
bridge writing both ends' traceback requests on the device, the tracebacks)
against the oracle: the joins, and each end's walk with the right end's
coordinates taken backwards (revp) from the cDNA's last base."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _expected(oracle, w, layout):
    """Per gap the oracle's join and, per end with a cell, (request index,
    fill, r, c, pair records, counters)."""
    joins, walks = [], []
    for k in range(w.n):
        p = w.problems[4 * k: 4 * k + 4]
        specs = []
        for j in range(4):
            q = bytes(w.qbuf[p["qoff"][j]: p["qoff"][j] + p["rl"][j]])
            g = bytes(w.gbuf[p["goff"][j]: p["goff"][j] + p["gl"][j]])
            specs.append(H.Spec(int(p["kind"][j]), int(p["width"][j]), q, g, g, 0, int(p["open"][j]),
                                int(p["extend"][j]), int(p["lband"][j]), int(p["uband"][j]),
                                bool(p["jump_late"][j]), False))
        fills = [H.oracle_fill(oracle, s) for s in specs]
        su = specs[0]
        want = oracle.bridge_cdna(su.width, *[f[0] for f in fills], su.gl, su.rl, specs[2].rl, specs[1].lband,
                                  su.uband, specs[3].lband, specs[2].uband, su.open, su.ext, 0,
                                  int(w.quads[k, 4]), int(su.late))
        joins.append(want)
        _, cL, cR, rL, rR = want
        tmpl = w.templates()
        for j, (r, c, base) in enumerate(((rR, cR, 2), (rL, cL, 0))):
            if r <= 0 and c <= 0:
                continue
            kind = 1 if c >= r else 2
            i = base + kind - 1
            s = specs[i]
            revp = base == 2
            Dir = np.where(H.computed_mask(s, layout[4 * k + i]), fills[i][1], 0).astype(np.uint8)
            # the oracle walks the reference's own orientation: the right end's
            # strings read backwards from their last base
            qs, gs = (s.q[::-1], s.g[::-1]) if revp else (s.q, s.g)
            q = tmpl[2 * k + j]
            po, co = oracle.traceback(kind, Dir, r, c, qs, qs.upper(), gs, gs, int(q["queryoffset"]),
                                      int(q["genomeoffset"]), revp)
            walks.append((2 * k + j, 4 * k + i, r, c, po, co))
    return joins, walks


def test_cdna_workload_pipeline(oracle):
    import torch
    from gmapdp.batch import DeviceBatch
    from gmapdp.lib import PAIR_DTYPE
    from gmapdp.workload import CdnaGaps
    w = CdnaGaps(150, seed=77, gl_range=(8, 90))
    b = DeviceBatch(w.problems, w.qbuf.tobytes(), w.gbuf.tobytes())
    b.fill()
    reqs = w.templates()
    h = b.prepare_traceback(reqs, genome=w.genome())
    d_quads = torch.from_numpy(w.quads.reshape(-1).copy()).to(b.dev)
    d_tmpl = torch.from_numpy(reqs.view(np.uint8).copy()).to(b.dev)
    d_out = torch.zeros(5 * w.n, dtype=torch.int32, device=b.dev)
    b.bridge_requests(d_quads, w.n, d_out, d_tmpl.data_ptr(), h["d_reqs"].data_ptr())
    b.traceback_device(h)
    out = d_out.cpu().numpy().reshape(-1, 5)
    got = h["d_reqs"].cpu().numpy().view(reqs.dtype)
    counts = h["d_counts"].cpu().numpy().reshape(-1, 6)
    pairs = h["d_pairs"].cpu().numpy().view(PAIR_DTYPE)
    joins, walks = _expected(oracle, w, b.layout)
    for k in range(w.n):
        assert tuple(out[k]) == joins[k], ("bridge", k)
    assert len(walks) > 150
    for ri, fill, r, c, po, co in walks:
        q = got[ri]
        assert (q["fill"], q["r"], q["c"]) == (fill, r, c), ("request", ri)
        off = int(q["pair_off"])
        gp = pairs[off: off + counts[ri, 0]]
        assert (H.pairs_to_records(gp) == po).all() and (counts[ri, 1:] == co).all(), ("walk", ri)
