"""The oracle's end-cell scans and nogaps walk against the reference's own
static functions of dynprog_end.c (compiled in place through
oracle/ref_end_shim.c): find_best_endpoint_{8,16} (143-290),
find_best_endpoint_to_queryend_indels_{8,16} (358-508) and
traceback_nogaps (648-722)."""
import numpy as np

import helpers as H


def test_best_endpoint_matches_reference(oracle, reference):
    rng = np.random.default_rng(31)
    n = 0
    for t in range(600):
        # gl >= rl: the reference's lower-triangle loop is bounded by the
        # diagonal only (dynprog_end.c:178), so gl < r would leave the matrix
        rl = int(rng.integers(1, 40))
        gl = rl + int(rng.integers(0, 12))
        lb, ub = int(rng.integers(0, 10)), int(rng.integers(0, 10))
        width = 8 if t % 2 else 16
        span = 4 if t % 3 == 0 else (120 if width == 8 else 3000)
        Hu = rng.integers(-span, span, (gl + 1) * (rl + 1)).astype(np.int16)
        Hl = rng.integers(-span, span, (rl + 1) * (gl + 1)).astype(np.int16)
        for qi in (False, True):
            for late in (False, True):
                want = reference.best_endpoint(qi, Hu, Hl, width, rl, gl, lb, ub, late)
                assert oracle.best_endpoint(qi, Hu, Hl, width, rl, gl, lb, ub, late) == want, (t, qi, late)
                n += 1
    assert n == 2400


def test_traceback_nogaps_matches_reference(oracle, reference):
    rng = np.random.default_rng(32)
    for t in range(400):
        rl, gl = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        q = H.rand_seq(rng, rl, 0.05)
        if t % 4 == 0:
            q = q.lower()
        g = H.rand_seq(rng, gl, 0.05)
        ga = bytes(b if rng.random() > 0.1 else ord("R") for b in g)
        r, c = int(rng.integers(0, rl + 1)), int(rng.integers(0, gl + 1))
        # genestrand 0: the standard mode builds only that consistency table
        # (dynprog.c:1027-1037)
        revp, gs = bool(t % 2), 0
        qo, go = int(rng.integers(-3, 40)), int(rng.integers(-3, 40))
        po, co = oracle.traceback_nogaps(r, c, q, q.upper(), g, ga, qo, go, revp, gs)
        pr, cr = reference.traceback_nogaps(r, c, q, q.upper(), g, ga, qo, go, revp, gs)
        assert (po == pr).all() and (co[:3] == cr[:3]).all() and (co[3:] == 0).all(), ("nogaps", t)
