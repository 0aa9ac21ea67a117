"""Dynprog_simd_16_noalloc (include/gmapdp_dynprog.h): the allocation-free
variant of the drop-in's Dynprog_simd_16, whose device buffers are parts of
a fixed scratch array in the library.  Called through the C ABI with the
reference's own workspaces and pair pools (oracle/_ref/libgmapref.so plays
the host program), its tables and its traceback's pair list must equal
those of Dynprog_simd_16 on the same problem, bit for bit, on two calls in a
row (the second smaller, so it reuses the array the first call left
behind), and a problem too large for the array is refused with a message.
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
    L = lib.load()  # torch's HIP runtime first, then the kernels library
    ref = ctypes.CDLL(po.REF_SO, mode=os.RTLD_LAZY | os.RTLD_GLOBAL)
    path = os.path.join(os.path.dirname(lib.library_path()), "libgmapdp_dynprog.so")
    mine = ctypes.CDLL(path, mode=os.RTLD_LAZY)
    for name in ("Dynprog_simd_16", "Dynprog_simd_16_noalloc", "Dynprog_traceback_16"):
        getattr(mine, name).restype = c_p
    ref.refh_new.restype = c_p
    ref.refh_new.argtypes = [c_i, c_i, c_i]
    ref.refh_pool_new.restype = c_p
    ref.refh_pool_reset.argtypes = [c_p]
    ref.refh_flatten.argtypes = [c_p, c_p, c_i]
    return L, ref, mine


def _tables(mine, fn, s, ws, seqs):
    """One full-band int16 fill through `fn`; returns (score table pointer,
    dense scores, dense direction codes, the three direction tables)."""
    qq, gg, gga = seqs
    no, eg, fg = c_p(), c_p(), c_p()
    m = getattr(mine, fn)(ctypes.byref(no), ctypes.byref(eg), ctypes.byref(fg), c_p(ws), c_p(qq.ptr), c_p(gg.ptr),
                          c_p(gga.ptr), c_i(s.rl), c_i(s.gl), c_i(s.mt), c_i(s.open), c_i(s.ext), c_i(s.lband),
                          c_i(s.uband), c_b(s.late), c_b(s.revp))
    if not m:
        return None
    T = ctypes.POINTER(ctypes.POINTER(ctypes.c_int16))
    out = []
    for tab in (m, no.value, eg.value, fg.value):
        t = ctypes.cast(tab, T)
        out.append(np.stack([np.ctypeslib.as_array(t[c], shape=(s.rl + 1,)).copy() for c in range(s.gl + 1)]))
    return m, out, (no.value, eg.value, fg.value)


def _walk(ref, mine, pool, s, dirs, seqs):
    qq, gg, gga = seqs
    quc = po._Seq(s.q.upper(), s.revp)
    ref.refh_pool_reset(c_p(pool))
    cnt = [c_i(0) for _ in range(5)]
    args = [c_p(None)] + [ctypes.byref(x) for x in cnt] + [c_p(d) for d in dirs]
    args += [c_i(s.rl), c_i(s.gl), c_p(qq.ptr), c_p(quc.ptr), c_p(gg.ptr), c_p(gga.ptr), c_i(5), c_i(7)]
    args += [c_p(None), c_p(None), c_p(pool), c_b(s.revp), c_u(0), c_u(0), c_b(True), c_i(0), c_i(0)]
    lst = mine.Dynprog_traceback_16(*args)
    out = np.zeros((s.rl + s.gl + 8, 8), np.int32)
    n = ref.refh_flatten(c_p(lst), out.ctypes.data, len(out))
    return out[:n].copy(), [x.value for x in cnt]


def _problem(rng, rl, gl, eb):
    g = H.rand_seq(rng, gl, 0.01)
    q = H.mutate(rng, g, 0.05, 0.02)[:rl].ljust(rl, b"T")
    return H.Spec(0, 16, q, g, g, int(rng.integers(0, 3)), -8, -3, eb, gl - rl + eb, bool(rng.integers(0, 2)),
                  bool(rng.integers(0, 2)))


def test_noalloc_matches_allocating_fill(libs):
    L, ref, mine = libs
    ws_a, ws_b = ref.refh_new(1100, 1600, 0), ref.refh_new(1100, 1600, 0)
    pool = ref.refh_pool_new()
    rng = np.random.default_rng(4242)
    # two calls in a row on the same array, the second smaller; then a wide band
    for rl, gl, eb in ((900, 1100, 40), (120, 150, 8), (40, 1500, 6)):
        s = _problem(rng, rl, gl, eb)
        seqs = (po._Seq(s.q, s.revp), po._Seq(s.g, s.revp), po._Seq(s.galt, s.revp))
        _, want, wdirs = _tables(mine, "Dynprog_simd_16", s, ws_a, seqs)
        got = _tables(mine, "Dynprog_simd_16_noalloc", s, ws_b, seqs)
        assert got is not None, L.gdp_last_error().decode()
        _, have, hdirs = got
        B = H.band_mask(s)
        for k, name in enumerate(("scores", "nogap", "Egap", "Fgap")):
            assert (have[k][B] == want[k][B]).all(), (name, rl, gl)
        # both walks, on their own fills' tables, give the same pair list
        pw, cw = _walk(ref, mine, pool, s, wdirs, seqs)
        ph, ch = _walk(ref, mine, pool, s, hdirs, seqs)
        assert len(pw) > 0 and (ph.shape == pw.shape) and (ph == pw).all() and ch == cw, ("pairs", rl, gl)


def test_noalloc_refuses_what_does_not_fit(libs):
    L, ref, mine = libs
    ws = ref.refh_new(100, 100, 0)  # refused before the workspace is touched
    rng = np.random.default_rng(7)
    s = _problem(rng, 4000, 4400, 100)  # over 100 MB of reference-layout tables alone
    seqs = (po._Seq(s.q, s.revp), po._Seq(s.g, s.revp), po._Seq(s.galt, s.revp))
    assert _tables(mine, "Dynprog_simd_16_noalloc", s, ws, seqs) is None
    assert b"scratch" in L.gdp_last_error()
