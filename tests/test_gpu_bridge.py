"""cDNA gap bridge on the GPU (gdp_bridge_cdna, the batched form of
bridge_cdna_gap_{8,16}_ud, dynprog_cdna.c:123-644): four one-sided fills per
gap kept with every score, the bridge, and the two ends' tracebacks from
the requests the bridge writes, against the oracle."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _bands(rl, gl, extraband):
    """Dynprog_compute_bands with widebandp (dynprog.c:1245-1261): (lband, uband)."""
    return (extraband, gl - rl + extraband) if gl >= rl else (rl - gl + extraband, extraband)


def _gaps(rng, n):
    out = []
    for t in range(n):
        gl = int(rng.integers(2, 90))
        short = t % 7 == 3  # ends shorter than the gap: upper rows beyond rl are not candidates
        rlL = int(rng.integers(1, gl)) if short else int(rng.integers(gl - 1, gl + 30))
        rlR = int(rng.integers(1, gl)) if short else int(rng.integers(gl - 1, gl + 30))
        if t % 13 == 7:
            # tables between 4096 and 8192 LDS words: the full-size launch's table form
            gl = int(rng.integers(110, 150))
            rlL, rlR = gl + int(rng.integers(5, 15)), gl + int(rng.integers(5, 15))
        width = 8 if t % 3 == 0 else 16
        late = bool(t % 2)
        ext = 20 if t % 13 == 7 else 4 + t % 5
        lbL, ubL = _bands(rlL, gl, ext)
        lbR, ubR = _bands(rlR, gl, ext)
        if t % 11 == 5:
            lbL = ubL = 520  # a left band past the kernel's LDS column: the direct form
        g = H.rand_seq(rng, gl)
        qL = H.mutate(rng, g, 0.04, 0.02)[:rlL].ljust(rlL, b"G")
        qR = H.mutate(rng, g[::-1], 0.04, 0.02)[:rlR].ljust(rlR, b"T")
        gr = g[::-1]
        open_, extend = (-10, -2) if t % 4 else (-8, -3)
        specs = [H.Spec(1, width, qL, g, g, 0, open_, extend, lbL, ubL, late, False),
                 H.Spec(2, width, qL, g, g, 0, open_, extend, lbL, ubL, late, False),
                 # the right end runs backwards from the gap's far side, with the other tie rule
                 H.Spec(1, width, qR, gr, gr, 0, open_, extend, lbR, ubR, not late, False),
                 H.Spec(2, width, qR, gr, gr, 0, open_, extend, lbR, ubR, not late, False)]
        D = int(rng.integers(rlL + rlR - 5, rlL + rlR + 40)) if t % 5 else int(rng.integers(0, 12))
        out.append((specs, D))
    return out


BRIDGE_LDS, BRIDGE_LDS_SMALL, BRIDGE_LDS_TINY, BRIDGE_ROWS = 8192, 4096, 2048, 1024  # gdp_kernels.hip


def _form(sp):
    """The kernel's form for a gap, from the same sizes it tests: 'tiny'
    (table form in the 8 KB launch), 'small' (the 16 KB launch), 'table'
    (the 32 KB launch) or 'direct'."""
    gl, rlR = sp[0].gl, sp[2].rl
    lbL, ubL, lbR, ubR = sp[1].lband, sp[0].uband, sp[3].lband, sp[2].uband
    fast = gl + 1 <= BRIDGE_LDS and ubL + lbL + 1 <= BRIDGE_ROWS
    nRsum = sum(max(0, max(min(c - 1, rlR), min(rlR - 1, c + lbR)) - max(1, c - ubR) + 1) for c in range(gl))
    need = 2 * (gl + 1) + nRsum + 2 * (ubL + lbL + 1)
    if fast and need <= BRIDGE_LDS_TINY:
        return "tiny"
    if fast and need <= BRIDGE_LDS_SMALL:
        return "small"
    if fast and need <= BRIDGE_LDS:
        return "table"
    return "direct"


def test_bridge_and_end_tracebacks(oracle):
    from gmapdp.batch import DeviceBatch
    rng = np.random.default_rng(2024)
    gaps = _gaps(rng, 90)
    # every form of the three launches is taken by some gap of the batch
    forms = [_form(sp) for sp, _ in gaps]
    assert all(forms.count(f) > 0 for f in ("tiny", "small", "table", "direct")), forms
    specs = [s for sp, _ in gaps for s in sp]
    probs, qbuf, gbuf = H.pack_batch(specs, keep_scores=True)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    quads = np.array([[4 * k, 4 * k + 1, 4 * k + 2, 4 * k + 3, D] for k, (_, D) in enumerate(gaps)], np.int32)
    tmpl = []
    for k, (sp, _) in enumerate(gaps):
        tmpl.append(H.tb_request(4 * k + 2, sp[2], 0, 0, qoff=11, goff=13))
        tmpl.append(H.tb_request(4 * k, sp[0], 0, 0, qoff=3, goff=5))
        for q in tmpl[-2:]:
            q["max_pairs"] = sp[0].rl + sp[2].rl + sp[0].gl + 4
    got, reqs = b.bridge_cdna(quads, np.array(tmpl))
    assert (b.bridge_cdna(quads) == got).all()
    found = 0
    walks = []
    for k, (sp, D) in enumerate(gaps):
        Hs = [H.oracle_fill(oracle, s)[0] for s in sp]
        su = sp[0]
        want = oracle.bridge_cdna(su.width, *Hs, su.gl, su.rl, sp[2].rl, sp[1].lband, su.uband, sp[3].lband,
                                  sp[2].uband, su.open, su.ext, 0, D, int(su.late))
        assert tuple(got[k]) == want, ("bridge", k, tuple(got[k]), want)
        _, cL, cR, rL, rR = want
        if cL < 0:
            assert all(reqs[2 * k + j]["r"] == 0 and reqs[2 * k + j]["c"] == 0 for j in (0, 1))
            continue
        found += 1
        for j, (r, c, ui) in enumerate(((rR, cR, 2), (rL, cL, 0))):
            q = reqs[2 * k + j]
            kind = 1 if c >= r else 2
            assert (q["fill"], q["r"], q["c"]) == (4 * k + ui + (kind - 1), r, c)
            if r == 0 and c == 0:
                continue
            s = sp[ui + kind - 1]
            Dir = np.where(H.computed_mask(s, b.layout[4 * k + ui + kind - 1]), H.oracle_fill(oracle, s)[1],
                           0).astype(np.uint8)
            qo, go = int(q["queryoffset"]), int(q["genomeoffset"])
            walks.append((q, oracle.traceback(kind, Dir, r, c, s.q, s.q.upper(), s.g, s.galt, qo, go, s.revp)))
    assert found > 40
    pairs, counts = b.traceback(np.array([w[0] for w in walks]))
    for j, (_, (po, co)) in enumerate(walks):
        assert (H.pairs_to_records(pairs[j]) == po).all() and (counts[j, 1:] == co).all(), ("end walk", j)
