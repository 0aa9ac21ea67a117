"""Pins the CPU oracle: reference golden vectors, closed forms, brute force."""
import numpy as np

import golden_io
import helpers as H


def test_oracle_matches_reference_golden_vectors(oracle):
    g, specs = golden_io.load()
    ntb = 0
    for i, s in enumerate(specs):
        Hs, D = H.oracle_fill(oracle, s)
        golden_io.check_fill(g, i, s, D, Hs)
        exp = golden_io.traceback_expect(g, i)
        if exp is not None:
            r, c, pairs, counts = exp
            got, cnt = oracle.traceback(s.kind, D, r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
            assert got.shape == pairs.shape and (got == pairs).all(), ("golden pairs", i)
            assert (cnt == counts).all(), ("golden counters", i)
            ntb += 1
    assert ntb > 300


def test_oracle_matches_std_golden_vectors(oracle):
    """Dynprog_standard / Dynprog_traceback_std vectors of the reference's
    scalar build (tests/golden/make_std_golden.py)."""
    g, specs = golden_io.load_std()
    ntb = 0
    for i, s in enumerate(specs):
        Hs, D = H.oracle_fill(oracle, s)
        golden_io.check_std(g, i, D, Hs)
        exp = golden_io.std_traceback_expect(g, i)
        if exp is not None:
            r, c, pairs, counts = exp
            got, cnt = oracle.traceback(0, D, r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
            assert got.shape == pairs.shape and (got == pairs).all(), ("std golden pairs", i)
            assert (cnt == counts).all(), ("std golden counters", i)
            ntb += 1
    assert ntb > 150


def test_std_without_gaps_is_the_diagonal_sum(oracle):
    """With upperp and lowerp off no gap can win, and each in-band cell is a
    running sum of pair scores along its diagonal, started from the initial
    row or column value where the diagonal meets row 1 or column 1.  Stored
    cells are floored at the saturation; the sum carries the floored value
    forward except along the band's top edge (c - r = uband), whose chain
    runs unclamped (dynprog.c:1578)."""
    rng = np.random.default_rng(8)
    pd = [oracle.pairdistance(mt).astype(np.int64) for mt in range(4)]
    for _ in range(40):
        rl, gl = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        q, g = H.rand_seq(rng, rl, 0.05), H.rand_seq(rng, gl, 0.05)
        lb, ub = int(rng.integers(0, 70)), int(rng.integers(0, 70))
        mt, sat = int(rng.integers(0, 4)), int(rng.choice([-32768, -5]))
        Hs, D = oracle.fill_std(q, g, g, mt, -7, -2, lb, ub, bool(rng.integers(0, 2)), False, sat, False, False)
        assert not (D[1:, 1:] & 3).any()
        for c in range(1, gl + 1):
            for r in range(max(1, c - ub), min(rl, c + lb) + 1):
                k = min(r, c) - 1
                r0, c0 = r - k, c - k
                if r0 == 1 and c0 == 1:
                    acc = 0
                elif r0 == 1:
                    acc = -7 - 2 * (c0 - 1)  # H(0, c0 - 1)
                else:
                    acc = -7 - 2 * (r0 - 1)  # H(r0 - 1, 0)
                for j in range(k + 1):
                    if j > 0 and c - r != ub:
                        acc = max(acc, sat)
                    acc += int(pd[mt][q[r0 + j - 1], g[c0 + j - 1]])
                assert Hs[c, r] == max(acc, sat), (r, c, lb, ub)


def test_traceback_score_is_sum_of_pair_scores(oracle):
    # traceback_score = MATCH*nmatches + MISMATCH*nmismatches + opens*(-3) + indels*(-1) (scores.h:4-9)
    for s in H.random_specs(5, 150, kinds=(0,)):
        if not (s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl):
            continue
        Hs, D = H.oracle_fill(oracle, s)
        qoff, goff = (s.rl, s.gl) if s.revp else (0, 0)  # keep every pushed coordinate >= 0
        pairs, cnt = oracle.traceback(0, D, s.rl, s.gl, s.q, s.q.upper(), s.g, s.galt, qoff, goff, s.revp)
        score, nm, nmm, no, ni = (int(x) for x in cnt)
        assert score == nm - 3 * nmm - 3 * no - ni
        comps = pairs[:, 3]
        gap = pairs[:, 6] == 1
        assert nm == int(((comps == ord("*")) | (comps == ord(":"))).sum())
        assert nmm == int(((comps == ord(" ")) & ~gap).sum())
        assert ni == int((comps == ord("-")).sum()) + int(pairs[gap, 7].sum()) - \
            sum(int(j) for j in pairs[gap, 7])  # gap holders carry no indel count
        # the pair list reconstructs both sequences in order (end to start)
        qpos = pairs[~gap, 0]
        assert (np.diff(qpos[qpos >= 0]) * (1 if s.revp else -1) >= 0).all()


def test_fill_matches_textbook_gotoh_in_the_interior(oracle):
    # Wide band, int16, no saturation: inside the band the block semantics
    # reduce to Gotoh's recurrence with the reference's boundary (H(0,0)=0,
    # column 0 by vertical gaps, row 0 by horizontal gaps from (0,0)).
    rng = np.random.default_rng(3)
    pd = oracle.pairdistance(0)
    for _ in range(30):
        rl, gl = int(rng.integers(5, 40)), int(rng.integers(5, 40))
        q, g = H.rand_seq(rng, rl), H.rand_seq(rng, gl)
        o, e = -8, -3
        Hs, D = oracle.fill_full(16, q, g, g, 0, o, e, rl + 5, gl + 5, True, False)
        NEG = -10 ** 9
        Hm = np.full((rl + 1, gl + 1), NEG, np.int64)
        Em = np.full((rl + 1, gl + 1), NEG, np.int64)
        Fm = np.full((rl + 1, gl + 1), NEG, np.int64)
        Hm[0, 0] = 0
        for r in range(rl + 1):
            for c in range(gl + 1):
                if r == 0 and c == 0:
                    continue
                if c > 0:
                    Em[r, c] = max(Em[r, c - 1], Hm[r, c - 1] + o) + e
                if r > 0:
                    Fm[r, c] = max(Fm[r - 1, c], Hm[r - 1, c] + o) + e
                d = Hm[r - 1, c - 1] + pd[q[r - 1], g[c - 1]] if r > 0 and c > 0 else NEG
                Hm[r, c] = max(d, Em[r, c], Fm[r, c])
        assert (Hs[1:, 1:].T == Hm[1:, 1:]).all()


def test_best_endpoint_is_row_major_argmax(oracle):
    rng = np.random.default_rng(8)
    for late in (False, True):
        for _ in range(40):
            # gl >= rl: the reference's lower-triangle loop is bounded by the
            # diagonal only (dynprog_end.c:178), so gl < r would leave the matrix.
            rl = int(rng.integers(2, 30))
            gl = rl + int(rng.integers(0, 10))
            lb, ub = int(rng.integers(0, 8)), int(rng.integers(0, 8))
            Hu = rng.integers(-6, 6, (gl + 1, rl + 1)).astype(np.int16)
            Hl = rng.integers(-6, 6, (rl + 1, gl + 1)).astype(np.int16)
            for qi in (False, True):
                fs, br, bc = oracle.best_endpoint(qi, Hu, Hl, 16, rl, gl, lb, ub, late)
                best, where = (-32768, (rl, 0)) if qi else (0, (0, 0))
                for r in range(rl if qi else 1, rl + 1):
                    for c in range(max(1, r - lb), min(gl, r + ub) + 1):
                        v = int(Hl[r, c]) if c < r else int(Hu[c, r])
                        if (v >= best) if late else (v > best):
                            best, where = v, (r, c)
                assert (fs, br, bc) == (best, where[0], where[1])


def test_genome_segments_match_slicing(oracle):
    rng = np.random.default_rng(4)
    n = 500
    codes = rng.integers(0, 4, n).astype(np.uint8)
    nf = (rng.random(n) < 0.05).astype(np.uint8)
    seq = bytes(ord("N") if nf[i] else b"ACGT"[codes[i]] for i in range(n))
    comp = bytes.maketrans(b"ACGTN", b"TGCAN")
    for _ in range(100):
        left, length = int(rng.integers(0, n - 1)), int(rng.integers(1, 60))
        chrhigh = int(rng.integers(left + 1, n + 1))
        want = (seq[left:min(left + length, chrhigh)] + b"*" * max(0, left + length - chrhigh))[:length]
        assert oracle.segment_right((codes, nf), left, length, chrhigh, False) == want
        assert oracle.segment_right((codes, nf), left, length, chrhigh, True) == want.translate(comp)[::-1]
        right = int(rng.integers(1, n))
        chroffset = int(rng.integers(0, right))
        lo = right - length
        want = (b"*" * max(0, chroffset - lo) + seq[max(lo, chroffset):right])[-length:] if length <= right else None
        if want is not None and len(want) == length:
            assert oracle.segment_left((codes, nf), right, length, chroffset, False) == want


def test_cell_counts_brute_force(oracle):
    rng = np.random.default_rng(6)
    for _ in range(50):
        rl, gl, lb, ub = (int(x) for x in rng.integers(1, 60, 4))
        r = np.arange(rl + 1)[None, :]
        c = np.arange(gl + 1)[:, None]
        full = ((c >= 1) & (r >= 1) & (r >= c - ub) & (r <= c + lb)).sum()
        assert oracle.cells_full(rl, gl, lb, ub) == full
        assert oracle.cells_upper(rl, gl, ub) == ((c >= 1) & (r >= 1) & (r >= c - ub) & (r <= c)).sum()
        assert oracle.cells_lower(rl, gl, lb) == ((c >= 1) & (r >= 1) & (c >= r - lb) & (c <= r)).sum()


def test_traceback_nogaps_is_the_diagonal(oracle):
    """traceback_nogaps (dynprog_end.c:648-722): pairs are exactly the cells
    (r-k, c-k) while both stay positive, comp '*' on equal upper-case
    characters, ':' on consistent ones, ' ' otherwise; '*' genome ends it
    without a pair.  Restated here from the definition, per pair."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        rl, gl = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        q = H.rand_seq(rng, rl, 0.1)
        q = bytes(ch + 32 if rng.random() < 0.2 else ch for ch in q)
        g = H.rand_seq(rng, gl, 0.05)
        ga = bytes(ch if rng.random() < 0.9 else H.ALPH[rng.integers(0, 4)] for ch in g)
        revp = bool(rng.integers(0, 2))
        gs = int(rng.integers(0, 3))
        r, c = min(rl, gl), min(rl, gl)
        got, cnt = oracle.traceback_nogaps(r, c, q, q.upper(), g, ga, 3, 11, revp, gs)
        want, score, nm, nmm = [], 0, 0, 0
        for k in range(min(r, c)):
            rr, cc = r - k, c - k
            # revp: the caller passes pointers to the last character and the walk reads [1 - r]
            c1, c2, c2a = (q[rl - rr], g[gl - cc], ga[gl - cc]) if revp else (q[rr - 1], g[cc - 1], ga[cc - 1])
            c1uc = ord(chr(c1).upper())
            comp = ord('*') if c1uc in (c2, c2a) else (
                ord(':') if oracle.consistent(gs, c1uc, c2) or oracle.consistent(gs, c1uc, c2a) else ord(' '))
            qc, gc = (-(rr - 1), -(cc - 1)) if revp else (rr - 1, cc - 1)
            want.append([3 + qc, 11 + gc, c1, comp, c2, c2a, 0, 0])
            if comp == ord(' '):
                score, nmm = score - 3, nmm + 1
            else:
                score, nm = score + 1, nm + 1
        want = np.array(want, np.int32).reshape(-1, 8)
        keep = (want[:, 0] >= 0) & (want[:, 1] >= 0)  # Pairpool_push drops negative coordinates
        assert (got == want[keep]).all()
        assert list(cnt[:3]) == [score, nm, nmm]
