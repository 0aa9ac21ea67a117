"""The oracle's end-cell scans and nogaps walk against the reference's own
static functions of dynprog_end.c (compiled in place through
oracle/ref_end_shim.c): find_best_endpoint_{8,16} (143-290),
find_best_endpoint_to_queryend_indels_{8,16} (358-508) and
traceback_nogaps (648-722)."""
import numpy as np
import pytest

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


def _local_cases(rng, n):
    """One-sided fills of end-like shapes and local walks from random cells to
    random stop columns, both orientations and widths."""
    for t in range(n):
        rl = int(rng.integers(3, 45))
        gl = rl + int(rng.integers(0, 15))
        g = H.rand_seq(rng, gl, 0.02)
        q = H.mutate(rng, g, 0.08, 0.06)[:rl].ljust(rl, b"A")
        if t % 5 == 0:
            q = q.lower()
        kind = 1 + t % 2
        width = 8 if t % 3 == 0 else 16
        band = int(rng.integers(0, 12)) + (gl - rl if kind == 1 else 0)
        revp, late = bool(t % 4 >= 2), bool(t % 7 % 2)
        r, c = int(rng.integers(0, rl + 1)), int(rng.integers(0, gl + 1))
        endc = int(rng.integers(0, max(1, min(c, gl) + 3)))
        yield t, kind, width, q, g, band, revp, late, r, c, endc


def test_traceback_local_matches_reference(oracle, reference):
    rng = np.random.default_rng(33)
    walks = 0
    for t, kind, width, q, g, band, revp, late, r, c, endc in _local_cases(rng, 1500):
        lb, ub = (0, band) if kind == 1 else (band, 0)
        s = H.Spec(kind, width, q, g, g, 3, -10, -2, lb, ub, late, revp)
        Ho, Do = H.oracle_fill(oracle, s)
        if kind == 1:
            reference.fill_upper(width, q, g, g, 3, -10, -2, ub, late, revp)
        else:
            reference.fill_lower(width, q, g, g, 3, -10, -2, lb, late, revp)
        # walks stay inside the band's computed cells from a cell in the band
        if kind == 1 and not (r <= c <= r + ub):
            continue
        if kind == 2 and not (c <= r <= c + lb):
            continue
        po, co, eo = oracle.traceback_local(kind, Do, r, c, endc, q, q.upper(), g, g, 3, 11, revp)
        pr, cr, er = reference.traceback_local(kind, r, c, endc, q, q.upper(), g, g, 3, 11, revp)
        assert (po == pr).all() and (co == cr).all() and eo == er, ("local walk", t, kind, r, c, endc)
        walks += 1
    assert walks > 300


@pytest.mark.parametrize("five", [True, False])
@pytest.mark.parametrize("long_jumps", [False, True])
def test_splicejunction_matches_reference(oracle, reference, five, long_jumps):
    """The oracle's composition of Dynprog_end5_splicejunction (five) and
    Dynprog_end3_splicejunction against the reference's own callers; with
    long_jumps the known gap holder spans 32768-500000 bases, past int16."""
    rng = np.random.default_rng((35 if five else 37) + (100 if long_jumps else 0))
    ends = H.random_splice_ends(rng, 300, long_jumps)
    got = 0
    for k, e in enumerate(ends):
        want = reference.end5_splicejunction(e["rev_rseq"], e["rev_gseq"], e["rev_galt"], e["rev_roffset"],
                                             e["anchor"], e["far"], e["genestrand"], e["jump_late"],
                                             e["extraband"], e["defect_rate"], e["contlength"], five=five)
        mine = H.oracle_end5_splicejunction(oracle, e, five)
        assert (want is None) == (mine is None), ("null", k)
        if want is None:
            continue
        assert (mine[0] == want[0]).all() and (mine[1] == want[1]).all(), ("splice end", k)
        got += 1
    assert 150 < got < 300
