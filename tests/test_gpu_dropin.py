"""The drop-in library (libgmapdp_dynprog.so, include/gmapdp_dynprog.h) called
with the reference's own signatures, workspaces and pair pools.

The reference build (oracle/_ref/libgmapref.so) supplies what GMAP's
dynprog.o and pairpool.o supply in a real link: Dynprog_new workspaces, the
score tables and Pairpool_*.  Each problem goes through the drop-in's
Dynprog_simd_* fill and Dynprog_traceback_*; the returned tables (read
through the returned row pointers, as the callers read them) and the
returned List_T of pairs must equal the oracle's, which is pinned to the
reference by the oracle tests.
"""
import ctypes
import os

import numpy as np
import pytest

import helpers as H
import pyoracle as po

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

c_p, c_i, c_b, c_u = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_uint


@pytest.fixture(scope="module")
def libs(reference):
    from gmapdp import lib
    lib.load()  # torch's HIP runtime first, then the kernels library
    ref = ctypes.CDLL(po.REF_SO, mode=os.RTLD_LAZY | os.RTLD_GLOBAL)
    path = os.path.join(os.path.dirname(lib.library_path()), "libgmapdp_dynprog.so")
    mine = ctypes.CDLL(path, mode=os.RTLD_LAZY)
    for w in ("8", "16"):
        for sfx in ("", "_upper", "_lower"):
            getattr(mine, "Dynprog_simd_%s%s" % (w, sfx)).restype = c_p
            getattr(mine, "Dynprog_traceback_%s%s" % (w, sfx)).restype = c_p
    ref.refh_new.restype = c_p
    ref.refh_new.argtypes = [c_i, c_i, c_i]
    ref.refh_pool_new.restype = c_p
    ref.refh_pool_reset.argtypes = [c_p]
    ref.refh_flatten.argtypes = [c_p, c_p, c_i]
    return ref, mine, ref.refh_new(400, 800, 0), ref.refh_new(400, 800, 1), ref.refh_pool_new()


def _read(tab, outer, inner, width):
    T = ctypes.c_int8 if width == 8 else ctypes.c_int16
    t = ctypes.cast(tab, ctypes.POINTER(ctypes.POINTER(T)))
    out = np.zeros((outer + 1, inner + 1), np.int32)
    for o in range(outer + 1):
        out[o] = np.ctypeslib.as_array(t[o], shape=(inner + 1,))
    return out


def _enc(nogap, egap, fgap):
    # dynprog.h:95-97: DIAG 0, HORIZ -1, anything else is VERT
    return ((nogap == -1) * 1 + ((nogap != 0) & (nogap != -1)) * 2 + (egap != 0) * 4 + (fgap != 0) * 8).astype(np.uint8)


def _fill(mine, s, ws, qq, gg, gga):
    no, eg, fg = c_p(), c_p(), c_p()
    W = s.width
    common = [c_p(ws), c_p(qq.ptr), c_p(gg.ptr), c_p(gga.ptr), c_i(s.rl), c_i(s.gl), c_i(s.mt), c_i(s.open),
              c_i(s.ext)]
    if s.kind == 0:
        m = getattr(mine, "Dynprog_simd_%d" % W)(ctypes.byref(no), ctypes.byref(eg), ctypes.byref(fg), *common,
                                                  c_i(s.lband), c_i(s.uband), c_b(s.late), c_b(s.revp))
    elif s.kind == 1:
        m = getattr(mine, "Dynprog_simd_%d_upper" % W)(ctypes.byref(no), ctypes.byref(eg), *common, c_i(s.uband),
                                                        c_b(s.late), c_b(s.revp))
    else:
        m = getattr(mine, "Dynprog_simd_%d_lower" % W)(ctypes.byref(no), ctypes.byref(eg), *common, c_i(s.lband),
                                                        c_b(s.late), c_b(s.revp))
    return m, no.value, eg.value, fg.value


def test_dropin_matches_oracle(oracle, libs):
    ref, mine, ws_full, ws_two, pool = libs
    specs = [s for s in H.random_specs(77, 200) if s.rl <= 150 and s.gl <= 200]
    rng = np.random.default_rng(77)
    ntb = 0
    for i, s in enumerate(specs):
        qq, gg, gga = po._Seq(s.q, s.revp), po._Seq(s.g, s.revp), po._Seq(s.galt, s.revp)
        quc = po._Seq(s.q.upper(), s.revp)
        W = s.width
        m, no, eg, fg = _fill(mine, s, ws_full if s.kind == 0 else ws_two, qq, gg, gga)
        outer, inner = (s.rl, s.gl) if s.kind == 2 else (s.gl, s.rl)
        Hm = _read(m, outer, inner, W)
        F = _read(fg, outer, inner, W) if s.kind == 0 else np.zeros((outer + 1, inner + 1), np.int32)
        Dm = _enc(_read(no, outer, inner, W), _read(eg, outer, inner, W), F)
        Ho, Do = H.oracle_fill(oracle, s)
        B = H.band_mask(s)
        assert (Hm[B] == Ho[B]).all(), ("drop-in scores", i, s.kind, W)
        assert (Dm[B] == Do[B]).all(), ("drop-in directions", i, s.kind, W)

        if s.kind == 0:
            if not (s.uband >= s.gl - s.rl and s.lband >= s.rl - s.gl):
                continue
            r, c = s.rl, s.gl
        else:
            cells = np.argwhere(B)
            if not len(cells):
                continue
            a, b = cells[rng.integers(0, len(cells))]
            r, c = (int(b), int(a)) if s.kind == 1 else (int(a), int(b))
        ref.refh_pool_reset(c_p(pool))
        cnt = [c_i(0) for _ in range(5)]
        args = [c_p(None)] + [ctypes.byref(x) for x in cnt] + [c_p(no), c_p(eg)]
        if s.kind == 0:
            args.append(c_p(fg))
        args += [c_i(r), c_i(c), c_p(qq.ptr), c_p(quc.ptr), c_p(gg.ptr), c_p(gga.ptr), c_i(5), c_i(7)]
        if s.kind == 2:
            args += [c_p(pool), c_i(0), c_b(s.revp), c_i(0)]
            fn = "Dynprog_traceback_%d_lower" % W
        else:
            args += [c_p(None), c_p(None), c_p(pool), c_b(s.revp), c_u(0), c_u(0), c_b(True), c_i(0), c_i(0)]
            fn = "Dynprog_traceback_%d%s" % (W, "" if s.kind == 0 else "_upper")
        lst = getattr(mine, fn)(*args)
        out = np.zeros((s.rl + s.gl + 8, 8), np.int32)
        n = ref.refh_flatten(c_p(lst), out.ctypes.data, len(out))
        Dmask = np.where(H.computed_mask(s), Do, 0).astype(np.uint8)
        want, wcnt = oracle.traceback(s.kind, Dmask, r, c, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
        assert n == len(want) and (out[:n] == want).all(), ("drop-in pairs", i, s.kind, W)
        assert [x.value for x in cnt] == list(wcnt), ("drop-in counters", i, s.kind, W)
        ntb += 1
    assert ntb > 50


def test_dropin_wide_bands(oracle, libs):
    """GMAP's end callers chop the genome window to the workspace's
    max_glength (at least 2000, dynprog.c:608-623) and widen the band to the
    far corner, so a short read end meets a band of about 2000 columns: the
    drop-in fills it on the GPU (wide single sweep) like any other."""
    ref, mine, _, _, pool = libs
    ws = ref.refh_new(100, 2100, 0)
    rng = np.random.default_rng(5)
    for t in range(6):
        rl, gl = int(rng.integers(30, 60)), int(rng.integers(1900, 2001))
        g = H.rand_seq(rng, gl)
        q = H.mutate(rng, g[-rl - 5:], 0.03, 0.01)[:rl].ljust(rl, b"C")
        eb = 6
        s = H.Spec(0, 16, q, g, g, 3, -10, -2, eb, gl - rl + eb, bool(t % 2), bool(t % 3 == 0))
        qq, gg, gga = po._Seq(s.q, s.revp), po._Seq(s.g, s.revp), po._Seq(s.galt, s.revp)
        quc = po._Seq(s.q.upper(), s.revp)
        m, no, eg, fg = _fill(mine, s, ws, qq, gg, gga)
        Hm = _read(m, s.gl, s.rl, 16)
        Dm = _enc(_read(no, s.gl, s.rl, 16), _read(eg, s.gl, s.rl, 16), _read(fg, s.gl, s.rl, 16))
        Ho, Do = H.oracle_fill(oracle, s)
        B = H.band_mask(s)
        assert (Hm[B] == Ho[B]).all() and (Dm[B] == Do[B]).all(), ("wide drop-in fill", t)
        ref.refh_pool_reset(c_p(pool))
        cnt = [c_i(0) for _ in range(5)]
        args = [c_p(None)] + [ctypes.byref(x) for x in cnt] + [c_p(no), c_p(eg), c_p(fg)]
        args += [c_i(s.rl), c_i(s.gl), c_p(qq.ptr), c_p(quc.ptr), c_p(gg.ptr), c_p(gga.ptr), c_i(5), c_i(7)]
        args += [c_p(None), c_p(None), c_p(pool), c_b(s.revp), c_u(0), c_u(0), c_b(True), c_i(0), c_i(0)]
        lst = mine.Dynprog_traceback_16(*args)
        out = np.zeros((s.rl + s.gl + 8, 8), np.int32)
        n = ref.refh_flatten(c_p(lst), out.ctypes.data, len(out))
        Dmask = np.where(H.computed_mask(s), Do, 0).astype(np.uint8)
        want, wcnt = oracle.traceback(0, Dmask, s.rl, s.gl, s.q, s.q.upper(), s.g, s.galt, 5, 7, s.revp)
        assert n == len(want) and (out[:n] == want).all(), ("wide drop-in pairs", t)
        assert [x.value for x in cnt] == list(wcnt), ("wide drop-in counters", t)
