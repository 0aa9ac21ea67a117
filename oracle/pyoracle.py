"""ctypes front end for the CPU checkers (test infrastructure only).

Two libraries share one plain-array interface:
  * ``Oracle`` -- oracle/liboracle.so, this project's restatement
    (gdp_oracle.c), which also travels to the GPU host;
  * ``Reference`` -- oracle/_ref/libgmapref.so, the reference's own DP
    objects built in place by oracle/Makefile (only where the reference tree
    exists);
  * ``ReferenceStd`` -- oracle/_ref_std/libgmapref_std.so, the same objects
    built without SIMD (Dynprog_standard and Dynprog_traceback_std).

Sequences are passed the way the reference receives them: for ``revp`` the
pointer addresses the *last* character and the kernels read ``s[1-pos]``
(dynprog_end.c:1415-1451).  Helpers here take ordinary Python ``bytes`` and
compute those pointers.  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgmapref.so")
REF_STD_SO = os.path.join(HERE, "_ref_std", "libgmapref_std.so")

KIND_FULL, KIND_UPPER, KIND_LOWER = 0, 1, 2
PAIR_FIELDS = 8

_i = ctypes.c_int
_u = ctypes.c_uint
_p = ctypes.c_void_p


class _Seq:
    """A NUL-padded ctypes copy of a sequence plus the pointer the kernels get."""

    def __init__(self, data: bytes, revp: bool):
        self.buf = ctypes.create_string_buffer(b"\0" + data + b"\0")
        base = ctypes.addressof(self.buf) + 1
        self.ptr = base + (len(data) - 1 if revp else 0)


def _np_ptr(a):
    return a.ctypes.data_as(_p)


class _Common:
    lib = None

    def _seqs(self, rseq, gseq, galt, revp):
        q = _Seq(rseq, revp)
        g = _Seq(gseq, revp)
        ga = g if galt is None or galt is gseq else _Seq(galt, revp)
        return q, g, ga


def _bridge_cdna_call(fn, width, LU, LL, RU, RL, gl, rlL, rlR, lbandL, ubandL, lbandR, ubandR, open_, ext,
                      leftoffset, rightoffset, late):
    arrs = [np.ascontiguousarray(a, np.int16) for a in (LU, LL, RU, RL)]
    assert arrs[0].size == (gl + 1) * (rlL + 1) and arrs[1].size == (rlL + 1) * (gl + 1)
    assert arrs[2].size == (gl + 1) * (rlR + 1) and arrs[3].size == (rlR + 1) * (gl + 1)
    out = np.zeros(5, np.int32)
    fn(int(width), *[_np_ptr(a) for a in arrs], gl, rlL, rlR, lbandL, ubandL, lbandR, ubandR, open_, ext,
       leftoffset, rightoffset, int(late), _np_ptr(out))
    return tuple(int(x) for x in out)


def _bridge_intron_args(width, LU, LL, RU, RL, gL, gLalt, gR, gRalt, rlength, lbandL, ubandL, lbandR, ubandR):
    gl_L, gl_R = len(gL), len(gR)
    arrs = [np.ascontiguousarray(a, np.int16) for a in (LU, LL, RU, RL)]
    assert arrs[0].size == (gl_L + 1) * (rlength + 1) and arrs[1].size == (rlength + 1) * (gl_L + 1)
    assert arrs[2].size == (gl_R + 1) * (rlength + 1) and arrs[3].size == (rlength + 1) * (gl_R + 1)
    assert len(gLalt) == gl_L and len(gRalt) == gl_R
    bufs = [ctypes.create_string_buffer(bytes(x), len(x) + 1) for x in (gL, gLalt, gR, gRalt)]
    return arrs, bufs, gl_L, gl_R


class Oracle(_Common):
    """CPU restatement (gdp_oracle.c)."""

    def __init__(self, mode: int = 0, path: str = ORACLE_SO):
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_init.argtypes = [_i]
        L.orc_pairdistance.restype = ctypes.POINTER(ctypes.c_short)
        L.orc_pairdistance.argtypes = [_i]
        L.orc_nt_to_int.restype = ctypes.POINTER(ctypes.c_int)
        L.orc_use8p_size.argtypes = [_i]
        L.orc_consistent.argtypes = [_i, _i, _i]
        L.orc_fill_full.argtypes = [_i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p]
        L.orc_fill_upper.argtypes = [_i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p]
        L.orc_fill_lower.argtypes = [_i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p]
        L.orc_traceback.argtypes = [_i, _p, _i, _i, _i, _i, _p, _p, _p, _p, _i, _i, _p, _i,
                                    _u, _u, _i, _i, _i, _p, _p, _i]
        L.orc_traceback.restype = _i
        L.orc_traceback_nogaps.argtypes = [_i, _i, _p, _p, _p, _p, _i, _i, _i, _i, _p, _p, _i]
        L.orc_traceback_nogaps.restype = _i
        L.orc_best_endpoint.argtypes = [_i, _p, _p, _i, _i, _i, _i, _i, _i, _p, _p, _p]
        L.orc_fill_std.argtypes = [_p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p]
        L.orc_bridge_cdna.argtypes = [_i, _p, _p, _p, _p] + [_i] * 12 + [_p]
        L.orc_bridge_intron.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p] + [_i] * 12 + [_p, _p]
        L.orc_traceback_local.argtypes = [_i, _p, _i, _i, _i, _i, _i, _p, _p, _p, _p, _i, _i, _i, _i, _p, _p, _i, _p]
        L.orc_traceback_local.restype = _i
        L.orc_segment_right.argtypes = [_p, _p, _u, _u, _u, _i]
        L.orc_segment_left.argtypes = [_p, _p, _u, _u, _u, _i]
        for f in ("orc_cells_full",):
            getattr(L, f).restype = ctypes.c_longlong
            getattr(L, f).argtypes = [_i, _i, _i, _i]
        for f in ("orc_cells_upper", "orc_cells_lower"):
            getattr(L, f).restype = ctypes.c_longlong
            getattr(L, f).argtypes = [_i, _i, _i]
        L.orc_init(mode)
        self._genome_keep = None

    # tables -----------------------------------------------------------
    def pairdistance(self, mt):
        p = self.lib.orc_pairdistance(mt)
        return np.ctypeslib.as_array(p, shape=(128 * 128,)).reshape(128, 128).copy()

    def nt_to_int(self):
        return np.ctypeslib.as_array(self.lib.orc_nt_to_int(), shape=(128,)).copy()

    def use8p(self, mt):
        return self.lib.orc_use8p_size(mt)

    def consistent(self, gs, c, g):
        return self.lib.orc_consistent(gs, c, g)

    # fills ------------------------------------------------------------
    def fill_full(self, width, rseq, gseq, galt, mt, open_, ext, lband, uband, late, revp):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((gl + 1, rl + 1), np.int16)
        D = np.zeros((gl + 1, rl + 1), np.uint8)
        self.lib.orc_fill_full(width, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext, lband, uband,
                               int(late), int(revp), _np_ptr(H), _np_ptr(D))
        return H, D

    def fill_upper(self, width, rseq, gseq, galt, mt, open_, ext, uband, late, revp):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((gl + 1, rl + 1), np.int16)
        D = np.zeros((gl + 1, rl + 1), np.uint8)
        self.lib.orc_fill_upper(width, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext, uband,
                                int(late), int(revp), _np_ptr(H), _np_ptr(D))
        return H, D

    def fill_lower(self, width, rseq, gseq, galt, mt, open_, ext, lband, late, revp):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((rl + 1, gl + 1), np.int16)
        D = np.zeros((rl + 1, gl + 1), np.uint8)
        self.lib.orc_fill_lower(width, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext, lband,
                                int(late), int(revp), _np_ptr(H), _np_ptr(D))
        return H, D

    def fill_std(self, rseq, gseq, galt, mt, open_, ext, lband, uband, late, revp, saturation=-32768,
                 upperp=True, lowerp=True):
        """Dynprog_standard: int32 H and nibble dirs, both dense [c][r]."""
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((gl + 1, rl + 1), np.int32)
        D = np.zeros((gl + 1, rl + 1), np.uint8)
        self.lib.orc_fill_std(q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext, lband, uband, int(late), int(revp),
                              saturation, int(upperp), int(lowerp), _np_ptr(H), _np_ptr(D))
        return H, D

    # traceback ----------------------------------------------------------
    def traceback(self, kind, dirs, r, c, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0,
                  revp=False, chroffset=0, chrhigh=0, watsonp=True, genestrand=0, genome=None,
                  maxpairs=8192):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        dirs = np.ascontiguousarray(dirs, np.uint8)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        gptr = None
        if genome is not None:
            gptr = self._genome_struct(genome)
        n = self.lib.orc_traceback(kind, _np_ptr(dirs), rl, gl, r, c, q.ptr, quc.ptr, g.ptr, ga.ptr,
                                   queryoffset, genomeoffset, gptr, int(revp), chroffset, chrhigh,
                                   int(watsonp), genestrand, 0, _np_ptr(cnt), _np_ptr(out), maxpairs)
        return out[:n].copy(), cnt

    def traceback_nogaps(self, r, c, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0, revp=False,
                         genestrand=0, maxpairs=8192):
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        n = self.lib.orc_traceback_nogaps(r, c, q.ptr, quc.ptr, g.ptr, ga.ptr, queryoffset, genomeoffset, int(revp),
                                          genestrand, _np_ptr(cnt), _np_ptr(out), maxpairs)
        return out[:n].copy(), cnt

    def _genome_struct(self, genome):
        """genome: (codes, nflags) of a whole genome, or (codes, nflags,
        base, length) of positions base .. base + len(codes) - 1 of a genome
        of `length` bases (genomecomp.decode_window)."""
        if len(genome) == 2:
            (codes, nflags), base, length = genome, 0, len(genome[0])
        else:
            codes, nflags, base, length = genome

        class G(ctypes.Structure):
            _fields_ = [("codes", _p), ("nflags", _p), ("length", _u), ("base", _u), ("count", _u)]

        codes = np.ascontiguousarray(codes, np.uint8)
        nflags = np.ascontiguousarray(nflags, np.uint8)
        st = G(codes.ctypes.data, nflags.ctypes.data, int(length), int(base), len(codes))
        self._genome_keep = (codes, nflags, st)
        return ctypes.addressof(st)

    def best_endpoint(self, queryend_indels, Hupper, Hlower, width, rl, gl, lband, uband, late):
        fs, br, bc = _i(), _i(), _i()
        Hu = np.ascontiguousarray(Hupper, np.int16)
        Hl = np.ascontiguousarray(Hlower, np.int16)
        self.lib.orc_best_endpoint(int(queryend_indels), _np_ptr(Hu), _np_ptr(Hl), width, rl, gl,
                                   lband, uband, int(late), ctypes.byref(fs), ctypes.byref(br),
                                   ctypes.byref(bc))
        return fs.value, br.value, bc.value

    def traceback_local(self, kind, dirs, r, c, endc, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0,
                        revp=False, genestrand=0, maxpairs=8192):
        """orc_traceback_local: (pair records, counters[5], (r, c) where the walk stopped)."""
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        dirs = np.ascontiguousarray(dirs, np.uint8)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        end = np.zeros(2, np.int32)
        n = self.lib.orc_traceback_local(kind, _np_ptr(dirs), rl, gl, r, c, endc, q.ptr, quc.ptr, g.ptr, ga.ptr,
                                         queryoffset, genomeoffset, int(revp), genestrand, _np_ptr(cnt),
                                         _np_ptr(out), maxpairs, _np_ptr(end))
        return out[:n].copy(), cnt, (int(end[0]), int(end[1]))

    def bridge_cdna(self, *args):
        """orc_bridge_cdna: (bestscore, bestcL, bestcR, bestrL, bestrR)."""
        return _bridge_cdna_call(self.lib.orc_bridge_cdna, *args)

    def bridge_intron(self, width, LU, LL, RU, RL, gL, gLalt, gR, gRalt, probL, probR, rlength, lbandL, ubandL,
                      lbandR, ubandR, leftoffset, rightoffset, cdna_direction, finalp, halfp):
        """orc_bridge_intron.  gL / gR are each end's genome in fill order
        (the right one reversed); probL / probR per column.  Returns
        ((finalscore, bestrL, bestrR, bestcL, bestcR), (left_prob, right_prob))."""
        arrs, bufs, gl_L, gl_R = _bridge_intron_args(width, LU, LL, RU, RL, gL, gLalt, gR, gRalt, rlength,
                                                     lbandL, ubandL, lbandR, ubandR)
        pL = np.ascontiguousarray(probL, np.float64)
        pR = np.ascontiguousarray(probR, np.float64)
        assert pL.size >= gl_L and pR.size >= gl_R
        out = np.zeros(5, np.int32)
        pr = np.zeros(2, np.float64)
        self.lib.orc_bridge_intron(int(width), *[_np_ptr(a) for a in arrs], *bufs, _np_ptr(pL), _np_ptr(pR),
                                   rlength, gl_L, gl_R, lbandL, ubandL, lbandR, ubandR, leftoffset, rightoffset,
                                   int(cdna_direction), int(finalp), int(halfp), _np_ptr(out), _np_ptr(pr))
        return tuple(int(x) for x in out), (float(pr[0]), float(pr[1]))

    def segment_right(self, genome, left, length, chrhigh, revcomp):
        out = ctypes.create_string_buffer(length + 1)
        self.lib.orc_segment_right(out, self._genome_struct(genome), left, length, chrhigh, int(revcomp))
        return out.raw[:length]

    def segment_left(self, genome, right, length, chroffset, revcomp):
        out = ctypes.create_string_buffer(length + 1)
        self.lib.orc_segment_left(out, self._genome_struct(genome), right, length, chroffset, int(revcomp))
        return out.raw[:length]

    def cells_full(self, rl, gl, lband, uband):
        return self.lib.orc_cells_full(rl, gl, lband, uband)

    def cells_upper(self, rl, gl, uband):
        return self.lib.orc_cells_upper(rl, gl, uband)

    def cells_lower(self, rl, gl, lband):
        return self.lib.orc_cells_lower(rl, gl, lband)


class Reference(_Common):
    """The reference's AVX2 objects (oracle/_ref/libgmapref.so)."""

    def __init__(self, mode: int = 0, path: str = REF_SO, max_rl: int = 1100, max_gl: int = 2200):
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        L = self.lib
        L.refh_init.argtypes = [_i]
        L.refh_new.restype = _p
        L.refh_new.argtypes = [_i, _i, _i]
        L.refh_use8p.argtypes = [_i]
        L.refh_pairdistance.argtypes = [_i, _p]
        L.refh_consistent.argtypes = [_i, _i, _i]
        L.refh_fill_full.argtypes = [_p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p, _i]
        L.refh_fill_upper.argtypes = [_p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p, _i]
        L.refh_fill_lower.argtypes = [_p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p, _i]
        L.refh_traceback.argtypes = [_i, _i, _i, _i, _p, _p, _p, _p, _i, _i, _i, _u, _u, _i, _i, _i,
                                     _p, _p, _i]
        L.refh_traceback.restype = _i
        L.refh_traceback_genome.argtypes = [_p, _u] + list(L.refh_traceback.argtypes)
        L.refh_traceback_genome.restype = _i
        L.refc_bridge_cdna.argtypes = [_i, _p, _p, _p, _p] + [_i] * 12 + [_p]
        L.refe_best_endpoint.argtypes = [_i, _i, _p, _p, _i, _i, _i, _i, _i, _p]
        L.refe_traceback_nogaps.argtypes = [_i, _i, _p, _p, _p, _p, _i, _i, _i, _i, _p, _p, _i]
        L.refe_traceback_nogaps.restype = _i
        L.refe_traceback_local.argtypes = [_i, _i, _i, _i, _i, _p, _p, _p, _p, _i, _i, _i, _i, _p, _p, _i, _p]
        L.refe_traceback_local.restype = _i
        L.refe_end5_splicejunction.argtypes = [_p, _p, _p, _p, _p] + [_i] * 8 + [ctypes.c_double, _i, _p, _p, _i]
        L.refe_end5_splicejunction.restype = _i
        L.refe_end3_splicejunction.argtypes = L.refe_end5_splicejunction.argtypes
        L.refe_end3_splicejunction.restype = _i
        L.refg_setup.argtypes = [_p, _u]
        L.refg_site_probs.argtypes = [_p, _p, _i, _i, _u, _u, _u, _u, _i, _i]
        L.refg_bridge_intron.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p, _p] + [_i] * 9 + [_u] * 4 + [_i] * 3 + \
            [_p, _p]
        L.refh_init(mode)
        self.ws_single = L.refh_new(max_rl, max_gl, 0)
        self.ws_double = L.refh_new(max_rl, max_gl, 1)
        self._last = {}

    def pairdistance(self, mt):
        out = np.zeros(128 * 128, np.int16)
        self.lib.refh_pairdistance(mt, _np_ptr(out))
        return out.reshape(128, 128)

    def use8p(self, mt):
        return self.lib.refh_use8p(mt)

    def consistent(self, gs, c, g):
        return self.lib.refh_consistent(gs, c, g)

    def fill_full(self, width, rseq, gseq, galt, mt, open_, ext, lband, uband, late, revp, garbage=0):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((gl + 1, rl + 1), np.int16)
        D = np.zeros((gl + 1, rl + 1), np.uint8)
        self.lib.refh_fill_full(self.ws_single, width, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext,
                                lband, uband, int(late), int(revp), _np_ptr(H), _np_ptr(D), garbage)
        self._last[KIND_FULL] = width
        return H, D

    def fill_upper(self, width, rseq, gseq, galt, mt, open_, ext, uband, late, revp, garbage=0):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((gl + 1, rl + 1), np.int16)
        D = np.zeros((gl + 1, rl + 1), np.uint8)
        self.lib.refh_fill_upper(self.ws_double, width, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext,
                                 uband, int(late), int(revp), _np_ptr(H), _np_ptr(D), garbage)
        self._last[KIND_UPPER] = width
        return H, D

    def fill_lower(self, width, rseq, gseq, galt, mt, open_, ext, lband, late, revp, garbage=0):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((rl + 1, gl + 1), np.int16)
        D = np.zeros((rl + 1, gl + 1), np.uint8)
        self.lib.refh_fill_lower(self.ws_double, width, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext,
                                 lband, int(late), int(revp), _np_ptr(H), _np_ptr(D), garbage)
        self._last[KIND_LOWER] = width
        return H, D

    def best_endpoint(self, queryend_indels, Hupper, Hlower, width, rl, gl, lband, uband, late):
        """The reference's find_best_endpoint{,_to_queryend_indels}_{8,16}, as Oracle.best_endpoint."""
        Hu = np.ascontiguousarray(Hupper, np.int16)
        Hl = np.ascontiguousarray(Hlower, np.int16)
        assert Hu.size == (gl + 1) * (rl + 1) and Hl.size == (rl + 1) * (gl + 1)
        out = np.zeros(3, np.int32)
        self.lib.refe_best_endpoint(int(queryend_indels), width, _np_ptr(Hu), _np_ptr(Hl), rl, gl, lband, uband,
                                    int(late), _np_ptr(out))
        return tuple(int(x) for x in out)

    def traceback_nogaps(self, r, c, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0, revp=False,
                         genestrand=0, maxpairs=8192):
        """The reference's traceback_nogaps, as Oracle.traceback_nogaps."""
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        n = self.lib.refe_traceback_nogaps(r, c, q.ptr, quc.ptr, g.ptr, ga.ptr, queryoffset, genomeoffset,
                                           int(revp), genestrand, _np_ptr(cnt), _np_ptr(out), maxpairs)
        return out[:n].copy(), cnt

    def traceback_local(self, kind, r, c, endc, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0,
                        revp=False, genestrand=0, maxpairs=8192):
        """The reference's traceback_local_* over the *last* fill of this kind (1 upper, 2 lower)."""
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        end = np.zeros(2, np.int32)
        n = self.lib.refe_traceback_local(kind, self._last[kind], r, c, endc, q.ptr, quc.ptr, g.ptr, ga.ptr,
                                          queryoffset, genomeoffset, int(revp), genestrand, _np_ptr(cnt),
                                          _np_ptr(out), maxpairs, _np_ptr(end))
        return out[:n].copy(), cnt, (int(end[0]), int(end[1]))

    def end5_splicejunction(self, rev_rseq, rev_gseq, rev_galt, rev_roffset, anchor, far, genestrand, late,
                            extraband_end, defect_rate, contlength, maxpairs=8192, five=True):
        """The reference's Dynprog_end5_splicejunction (five) or
        Dynprog_end3_splicejunction; strings in the caller's orientation (the
        5' form reads them backwards from their last base).  Returns None for
        a NULL list, else (the returned list's pair records from its last
        element to its head, {traceback_score, missscore, nmatches,
        nmismatches, nopens, nindels})."""
        q, g, ga = self._seqs(rev_rseq, rev_gseq, rev_galt, five)
        quc = _Seq(rev_rseq.upper(), five)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        res = np.zeros(6, np.int32)
        fn = self.lib.refe_end5_splicejunction if five else self.lib.refe_end3_splicejunction
        n = fn(self.ws_double, q.ptr, quc.ptr, g.ptr, ga.ptr, len(rev_rseq),
                                              len(rev_gseq), rev_roffset, anchor, far, genestrand, int(late),
                                              extraband_end, float(defect_rate), contlength, _np_ptr(res),
                                              _np_ptr(out), maxpairs)
        if n < 0:
            return None
        return out[:n].copy(), res

    def bridge_cdna(self, *args):
        """The reference's bridge_cdna_gap_{8,16}_ud, same arguments as Oracle.bridge_cdna."""
        return _bridge_cdna_call(self.lib.refc_bridge_cdna, *args)

    def genome_setup(self, image, nbases):
        """Dynprog_genome_setup without a splicing IIT, and the genomecomp
        image the splice-site models read (kept alive here)."""
        self._genome_image = np.ascontiguousarray(image, np.uint32)
        self.lib.refg_setup(_np_ptr(self._genome_image), ctypes.c_uint(int(nbases)))

    def site_probs(self, glengthL, glengthR, leftoffset, rightoffset, chroffset, chrhigh, cdna_direction, watsonp):
        """The per-column splice-site probabilities the site-level bridge
        computes (Maxent_hr over the genome of genome_setup)."""
        left = np.zeros(glengthL, np.float64)
        right = np.zeros(glengthR, np.float64)
        self.lib.refg_site_probs(_np_ptr(left), _np_ptr(right), glengthL, glengthR, leftoffset, rightoffset,
                                 chroffset, chrhigh, int(cdna_direction), int(watsonp))
        return left, right

    def bridge_intron(self, width, LU, LL, RU, RL, gL, gLalt, gRseg, gRsegalt, rlength, lbandL, ubandL, lbandR,
                      ubandR, leftoffset, rightoffset, chroffset, chrhigh, cdna_direction, watsonp, finalp, halfp,
                      late):
        """The reference's bridge_intron_gap_{8,16}_ud.  gRseg is the right
        segment in genome order (the bridge reads it backwards).  Returns as
        Oracle.bridge_intron."""
        arrs, bufs, gl_L, gl_R = _bridge_intron_args(width, LU, LL, RU, RL, gL, gLalt, gRseg, gRsegalt, rlength,
                                                     lbandL, ubandL, lbandR, ubandR)
        out = np.zeros(5, np.int32)
        pr = np.zeros(2, np.float64)
        self.lib.refg_bridge_intron(int(width), *[_np_ptr(a) for a in arrs], *bufs, rlength, gl_L, gl_R,
                                    int(cdna_direction), int(watsonp), lbandL, ubandL, lbandR, ubandR, leftoffset,
                                    rightoffset, chroffset, chrhigh, int(halfp), int(finalp), int(late),
                                    _np_ptr(out), _np_ptr(pr))
        return tuple(int(x) for x in out), (float(pr[0]), float(pr[1]))

    def traceback(self, kind, dirs, r, c, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0,
                  revp=False, chroffset=0, chrhigh=0, watsonp=True, genestrand=0, genome=None,
                  maxpairs=8192, genomecomp=None):
        """Traceback over the *last* fill of this kind (dirs is ignored).
        genomecomp = (image, nbases): genome skips read that genomecomp
        image through the reference's Genome_get_char; without it every
        genome-skip character is '*' (chrhigh = 0)."""
        assert genome is None, "the reference reads a genome as genomecomp=(image, nbases)"
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        if genomecomp is not None:
            image = np.ascontiguousarray(genomecomp[0], np.uint32)
            n = self.lib.refh_traceback_genome(_np_ptr(image), ctypes.c_uint(int(genomecomp[1])), kind,
                                               self._last[kind], r, c, q.ptr, quc.ptr, g.ptr, ga.ptr, queryoffset,
                                               genomeoffset, int(revp), ctypes.c_uint(chroffset),
                                               ctypes.c_uint(chrhigh), int(watsonp), genestrand, 0, _np_ptr(cnt),
                                               _np_ptr(out), maxpairs)
            return out[:n].copy(), cnt
        n = self.lib.refh_traceback(kind, self._last[kind], r, c, q.ptr, quc.ptr, g.ptr, ga.ptr,
                                    queryoffset, genomeoffset, int(revp), chroffset, chrhigh,
                                    int(watsonp), genestrand, 0, _np_ptr(cnt), _np_ptr(out), maxpairs)
        return out[:n].copy(), cnt


class ReferenceStd(_Common):
    """The reference's scalar objects (oracle/_ref_std/libgmapref_std.so)."""

    def __init__(self, mode: int = 0, path: str = REF_STD_SO, max_rl: int = 1100, max_gl: int = 2200):
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        L = self.lib
        L.refs_init.argtypes = [_i]
        L.refs_new.restype = _p
        L.refs_new.argtypes = [_i, _i]
        L.refs_fill.argtypes = [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p]
        L.refs_traceback.argtypes = [_i, _i, _p, _p, _p, _p, _i, _i, _i, _i, _p, _p, _i]
        L.refs_traceback.restype = _i
        L.refs_init(mode)
        self.ws = L.refs_new(max_rl, max_gl)

    def fill_std(self, rseq, gseq, galt, mt, open_, ext, lband, uband, late, revp, saturation=-32768,
                 upperp=True, lowerp=True):
        rl, gl = len(rseq), len(gseq)
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        H = np.zeros((gl + 1, rl + 1), np.int32)
        D = np.zeros((gl + 1, rl + 1), np.uint8)
        self.lib.refs_fill(self.ws, q.ptr, g.ptr, ga.ptr, rl, gl, mt, open_, ext, lband, uband, int(late),
                           int(revp), saturation, int(upperp), int(lowerp), _np_ptr(H), _np_ptr(D))
        return H, D

    def traceback(self, r, c, rseq, rsequc, gseq, galt, queryoffset=0, genomeoffset=0, revp=False,
                  genestrand=0, maxpairs=8192):
        """Dynprog_traceback_std over the last fill (chrhigh 0: genome skips give '*')."""
        q, g, ga = self._seqs(rseq, gseq, galt, revp)
        quc = _Seq(rsequc, revp)
        out = np.zeros((maxpairs, PAIR_FIELDS), np.int32)
        cnt = np.zeros(5, np.int32)
        n = self.lib.refs_traceback(r, c, q.ptr, quc.ptr, g.ptr, ga.ptr, queryoffset, genomeoffset, int(revp),
                                    genestrand, _np_ptr(cnt), _np_ptr(out), maxpairs)
        return out[:n].copy(), cnt


def band_mask_full(rl, gl, lband, uband):
    """In-band cells B (r>=1, c>=1) as a [c][r] boolean mask."""
    c = np.arange(gl + 1)[:, None]
    r = np.arange(rl + 1)[None, :]
    return (c >= 1) & (r >= 1) & (r >= c - uband) & (r <= c + lband)


def band_mask_upper(rl, gl, uband):
    c = np.arange(gl + 1)[:, None]
    r = np.arange(rl + 1)[None, :]
    return (c >= 1) & (r >= 1) & (r >= c - uband) & (r <= c)


def band_mask_lower(rl, gl, lband):
    """[r][c] mask of the lower triangle r >= c >= max(1, r-lband)."""
    r = np.arange(rl + 1)[:, None]
    c = np.arange(gl + 1)[None, :]
    return (c >= 1) & (r >= 1) & (c >= r - lband) & (c <= r)
