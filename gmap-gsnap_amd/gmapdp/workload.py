"""Synthetic workloads of the benchmark configurations (generated by the
library's C++ generator, gdp_gen.cpp)."""
import ctypes

import numpy as np

from .lib import INTRON_GAP_DTYPE, PROBLEM_DTYPE, TBREQ_DTYPE, load

META_DTYPE = np.dtype([("queryoffset", "<i4"), ("genomeoffset", "<i4"), ("chroffset", "<u4"), ("chrhigh", "<u4"),
                       ("watsonp", "i1"), ("revp", "i1"), ("pad", "i1", (2,)), ("pair", "<i4")])
assert META_DTYPE.itemsize == 24

SEEDS = {1: 1001, 2: 1002, 3: 1003, 4: 1004, 5: 1005, "cdna": 1006, "intron": 1007}


def _bind(L):
    if getattr(L, "_gen_bound", False):
        return L
    L.gdp_gen_genome.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int64]
    L.gdp_gen_genome.restype = None
    L.gdp_gen_max_bytes.argtypes = [ctypes.c_int]
    L.gdp_gen_max_bytes.restype = ctypes.c_int32
    L.gdp_gen_workload.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    L.gdp_gen_workload.restype = ctypes.c_int
    L._gen_bound = True
    return L


class Workload:
    """problems (PROBLEM_DTYPE), meta (META_DTYPE), qbuf and gbuf (uint8) of
    one configuration; gbuf starts with the genome."""

    def __init__(self, config, count, genome_len=1 << 24, variant=0, seed=None, first=0, genome_seed=None):
        L = _bind(load())
        self.config = config
        seed = SEEDS[config] if seed is None else seed
        genome_seed = seed if genome_seed is None else genome_seed
        stride = L.gdp_gen_max_bytes(config)
        nprob = 2 * count if config == 4 else count
        self.genome_len = genome_len
        extra = count * stride if config == 4 else 0
        self.gbuf = np.zeros(genome_len + extra + 8, np.uint8)
        L.gdp_gen_genome(genome_seed, self.gbuf.ctypes.data, genome_len)
        self.qbuf = np.zeros(count * stride + 8, np.uint8)
        self.problems = np.zeros(nprob, PROBLEM_DTYPE)
        self.meta = np.zeros(nprob, META_DTYPE)
        rc = L.gdp_gen_workload(config, variant, seed, first, count, self.gbuf.ctypes.data, genome_len,
                                self.problems.ctypes.data, self.meta.ctypes.data, self.qbuf.ctypes.data,
                                self.gbuf.ctypes.data, genome_len)
        if rc != 0:
            raise RuntimeError("gdp_gen_workload failed")

    def genome(self):
        return self.gbuf[: self.genome_len]

    def tb_requests(self, start_cells=None):
        """Traceback requests: full-band problems from (rl, gl); for end
        problems pass start_cells[k] = (fill index, r, c)."""
        p, m = self.problems, self.meta
        if start_cells is None:
            idx = np.arange(len(p))
            r, c = p["rl"], p["gl"]
        else:
            idx, r, c = start_cells[:, 0], start_cells[:, 1], start_cells[:, 2]
        req = np.zeros(len(idx), TBREQ_DTYPE)
        req["fill"], req["r"], req["c"] = idx, r, c
        req["queryoffset"] = m["queryoffset"][idx]
        req["genomeoffset"] = m["genomeoffset"][idx]
        req["chroffset"] = m["chroffset"][idx]
        req["chrhigh"] = m["chrhigh"][idx]
        req["watsonp"] = m["watsonp"][idx]
        req["revp"] = m["revp"][idx]
        req["max_pairs"] = p["rl"][idx] + p["gl"][idx] + 4
        off = np.zeros(len(idx), np.int64)
        off[1:] = np.cumsum(req["max_pairs"].astype(np.int64))[:-1]
        req["pair_off"] = off
        return req


class EndMix:
    """Config 4: end-extension problems (an upper and a lower fill each,
    best end cell over both, traceback in the triangle holding it) mixed 1:1
    with config-3 single-gap problems, over one genome.  problems[:n3] are
    the single-gap fills, problems[n3:] the (upper, lower) pairs.

    keep: keep_scores of the end fills -- 2 (GDP_KEEP_BEST, the default: each
    fill hands its triangle's best end cell to gdp_best_endpoint, for the
    BEST_LOCAL / QUERYEND_GAP scan), 3 (the same for queryend_indels), or 1
    (every score kept and scanned afterwards)."""

    def __init__(self, n_end, n_single, genome_len=1 << 24, keep=2):
        e = Workload(4, n_end, genome_len=genome_len)
        e.problems["keep_scores"] = keep
        s = Workload(3, n_single, genome_len=genome_len, genome_seed=SEEDS[4])
        self.n3, self.m = n_single, n_end
        self.genome_len = genome_len
        qs, gs = len(s.qbuf), len(s.gbuf)
        pe = e.problems.copy()
        pe["qoff"] += qs
        pe["goff"] += gs
        pe["galtoff"] += gs
        self.problems = np.concatenate([s.problems, pe])
        self.meta = np.concatenate([s.meta, e.meta])
        self.qbuf = np.concatenate([s.qbuf, e.qbuf])
        self.gbuf = np.concatenate([s.gbuf, e.gbuf])
        self._genome = e.genome().copy()
        self.ul = np.stack([n_single + np.arange(0, 2 * n_end, 2), n_single + np.arange(1, 2 * n_end, 2)],
                           1).astype(np.int32)

    def genome(self):
        return self._genome

    def single_requests(self):
        p, m = self.problems[: self.n3], self.meta[: self.n3]
        req = np.zeros(self.n3, TBREQ_DTYPE)
        req["fill"], req["r"], req["c"] = np.arange(self.n3), p["rl"], p["gl"]
        for f in ("queryoffset", "genomeoffset", "chroffset", "chrhigh", "watsonp", "revp"):
            req[f] = m[f]
        req["max_pairs"] = p["rl"] + p["gl"] + 4
        req["pair_off"][1:] = np.cumsum(req["max_pairs"].astype(np.int64))[:-1]
        return req

    def end_templates(self):
        """Traceback request templates of the end pairs (fill, r, c left to
        gdp_best_endpoint_ex)."""
        up = self.ul[:, 0]
        p, m = self.problems[up], self.meta[up]
        req = np.zeros(self.m, TBREQ_DTYPE)
        for f in ("queryoffset", "genomeoffset", "chroffset", "chrhigh", "watsonp", "revp"):
            req[f] = m[f]
        req["max_pairs"] = p["rl"] + p["gl"] + 4
        return req


class CdnaGaps:
    """cDNA gaps as stage 3 hands them to Dynprog_cdna_gap
    (stage3.c:9307-9322): a genomic gap of gl bases, both ends' query length
    rl = gl + 8 (extramaterial_paired, gmap.c:283), extraband_paired 14
    (gmap.c:292), the cDNA penalties -10/-7 (dynprog_cdna.c:31-37) under
    HIGHQ scoring, and the width by the rule of dynprog_cdna.c:906-913 (int8
    when gl or both ends are under use8p_size).  The cDNA between the two
    anchors is the genomic gap with 3 % substitutions and 10-60 bases of
    extra material at a random point.

    problems[4k .. 4k+3]: the gap's left upper, left lower, right upper and
    right lower fills, kept with every score (the right end runs over the
    reversed cDNA and genome, with the other tie rule); quads: the rows of
    gdp_bridge_cdna; genome(): the gaps' genomic bases end to end, gap k at
    goffset[k]."""

    OPEN, EXTEND, EXTRABAND, EXTRAMATERIAL = -10, -7, 14, 8

    def __init__(self, n, seed=None, gl_range=(16, 160), jump_late=False):
        L = load()
        L.gdp_host_tables(0)  # GMAP's score tables on the host, for use8p_size
        rng = np.random.default_rng(SEEDS["cdna"] if seed is None else seed)
        use8 = int(L.gdp_use8p_size(0))
        acgt = np.frombuffer(b"ACGT", np.uint8)
        gls = rng.integers(gl_range[0], gl_range[1] + 1, n).astype(np.int64)
        rls = gls + self.EXTRAMATERIAL
        self.goffset = np.zeros(n, np.int64)
        self.goffset[1:] = np.cumsum(gls)[:-1]
        self._genome = acgt[rng.integers(0, 4, int(gls.sum()))]
        cuts = rng.integers(0, gls + 1)
        ins = rng.integers(10, 61, n)  # extra cDNA bases: the gap's query span is gl + ins
        extra = acgt[rng.integers(0, 4, int(ins.sum()))]
        eoff = np.zeros(n + 1, np.int64)
        eoff[1:] = np.cumsum(ins)
        # per gap: left query, right query (reversed cDNA); genome, reversed genome
        qbuf = np.empty(int(2 * rls.sum()) + 8, np.uint8)
        gbuf = np.empty(int(2 * gls.sum()) + 8, np.uint8)
        qbuf[-8:] = 0
        gbuf[-8:] = 0
        qo = np.zeros(n, np.int64)
        qo[1:] = np.cumsum(2 * rls)[:-1]
        go = 2 * self.goffset
        for k in range(n):
            gl, rl, c = int(gls[k]), int(rls[k]), int(cuts[k])
            g = self._genome[self.goffset[k]: self.goffset[k] + gl]
            cdna = np.concatenate([g[:c], extra[eoff[k]: eoff[k + 1]], g[c:]])
            sub = rng.random(len(cdna)) < 0.03
            cdna[sub] = acgt[rng.integers(0, 4, int(sub.sum()))]
            q0 = int(qo[k])
            qbuf[q0: q0 + rl] = cdna[:rl]
            qbuf[q0 + rl: q0 + 2 * rl] = cdna[::-1][:rl]
            g0 = int(go[k])
            gbuf[g0: g0 + gl] = g
            gbuf[g0 + gl: g0 + 2 * gl] = g[::-1]
        probs = np.zeros(4 * n, PROBLEM_DTYPE)
        late = int(jump_late)
        for j in range(4):
            p = probs[j::4]
            right = j >= 2
            p["qoff"] = qo + (rls if right else 0)
            p["goff"] = go + (gls if right else 0)
            p["galtoff"] = p["goff"]
            p["rl"], p["gl"] = rls, gls
            # Dynprog_compute_bands with widebandp, rl > gl (dynprog.c:1245-1261)
            p["lband"], p["uband"] = rls - gls + self.EXTRABAND, self.EXTRABAND
            p["kind"] = 1 + (j & 1)
            p["width"] = np.where((gls < use8) | (rls < use8), 8, 16)
            p["jump_late"] = 1 - late if right else late
            p["open"], p["extend"], p["keep_scores"] = self.OPEN, self.EXTEND, 1
        self.problems = probs
        self.qbuf, self.gbuf = qbuf, gbuf
        # the left end starts at query offset 0 (roffsetL) of the gap's gl + ins
        # cDNA bases, the right end at its last base (rev_roffsetR); the bridge
        # keeps the ends' rows below rev_roffsetR - roffsetL
        self.rev_roffset = gls + ins - 1
        self.quads = np.stack([4 * np.arange(n) + j for j in range(4)] + [self.rev_roffset], 1).astype(np.int32)
        self.n = n

    def genome(self):
        return self._genome

    def templates(self):
        """Traceback request templates, right end then left end per gap
        (fill, r and c are set by gdp_bridge_cdna); pair_off laid out."""
        p = self.problems
        req = np.zeros(2 * self.n, TBREQ_DTYPE)
        rl, gl = p["rl"][0::4], p["gl"][0::4]
        req["queryoffset"][0::2] = self.rev_roffset  # walked with revp
        req["queryoffset"][1::2] = 0  # roffsetL
        req["genomeoffset"][0::2] = self.goffset + gl - 1  # rev_goffset
        req["genomeoffset"][1::2] = self.goffset
        req["revp"][0::2] = 1
        req["chroffset"], req["chrhigh"], req["watsonp"] = 0, len(self._genome), 1
        req["max_pairs"] = np.repeat(rl + gl + 4, 2)
        req["pair_off"][1:] = np.cumsum(req["max_pairs"].astype(np.int64))[:-1]
        return req


class IntronGaps:
    """Intron gaps as stage 3 hands them to Dynprog_genome_gap
    (stage3.c:9538-9576): a read gap of rl = U{40..250} bases, both genome
    segments rl + 8 long (extramaterial_paired, gmap.c:283), extraband_paired
    14 (gmap.c:292), the paired HIGHQ penalties -8/-3 (dynprog.h:69-75) and
    the width rule of dynprog_genome.c:3504-3508.  The read is the end of a
    left exon (before a GT) and the start of a right exon (after an AG) with
    2 % substitutions; the segments are random bases around them.  The
    splice-site probabilities are synthetic per-column values in steps of
    0.05 (the caller's site model stands behind them in GMAP).

    problems[4k .. 4k+3]: left upper, left lower (the read forward against the
    left segment), right upper, right lower (the read backwards against the
    right segment read backwards, the other tie rule), kept with every
    score; gaps: the rows of gdp_bridge_intron; probs: the site
    probabilities; genome(): the segments end to end, gap k's left one at
    goffset[k] and its right one just after."""

    OPEN, EXTEND, EXTRABAND, EXTRAMATERIAL, INTRON = -8, -3, 14, 8, 5000

    def __init__(self, n, seed=None, rl_range=(40, 250)):
        L = load()
        L.gdp_host_tables(0)
        rng = np.random.default_rng(SEEDS["intron"] if seed is None else seed)
        use8 = int(L.gdp_use8p_size(0))
        acgt = np.frombuffer(b"ACGT", np.uint8)
        rls = rng.integers(rl_range[0], rl_range[1] + 1, n).astype(np.int64)
        gls = rls + self.EXTRAMATERIAL
        a = rng.integers(1, rls)  # read bases from the left exon
        seg = acgt[rng.integers(0, 4, int(2 * gls.sum()))]
        self.goffset = np.zeros(n, np.int64)
        self.goffset[1:] = np.cumsum(2 * gls)[:-1]
        qbuf = np.empty(int(2 * rls.sum()) + 8, np.uint8)
        gbuf = np.empty(int(2 * gls.sum()) + 8, np.uint8)
        qbuf[-8:] = 0
        gbuf[-8:] = 0
        qo = np.zeros(n, np.int64)
        qo[1:] = np.cumsum(2 * rls)[:-1]
        for k in range(n):
            rl, gl, ak, g0 = int(rls[k]), int(gls[k]), int(a[k]), int(self.goffset[k])
            left, right = seg[g0: g0 + gl], seg[g0 + gl: g0 + 2 * gl]
            left[ak: ak + 2] = (ord("G"), ord("T"))[: max(0, min(2, gl - ak))]
            b = rl - ak
            if gl - b - 2 >= 0:
                right[gl - b - 2: gl - b] = (ord("A"), ord("G"))
            read = np.concatenate([left[:ak], right[gl - b:]])
            sub = rng.random(rl) < 0.02
            read[sub] = acgt[rng.integers(0, 4, int(sub.sum()))]
            q0 = int(qo[k])
            qbuf[q0: q0 + rl] = read
            qbuf[q0 + rl: q0 + 2 * rl] = read[::-1]
            gbuf[g0: g0 + gl] = left
            gbuf[g0 + gl: g0 + 2 * gl] = right[::-1]
        self._genome = seg
        probs = np.zeros(4 * n, PROBLEM_DTYPE)
        for j in range(4):
            p = probs[j::4]
            right = j >= 2
            p["qoff"] = qo + (rls if right else 0)
            p["goff"] = self.goffset + (gls if right else 0)
            p["galtoff"] = p["goff"]
            p["rl"], p["gl"] = rls, gls
            # Dynprog_compute_bands with widebandp, gl >= rl (dynprog.c:1245-1261)
            p["lband"], p["uband"] = self.EXTRABAND, gls - rls + self.EXTRABAND
            p["kind"] = 1 + (j & 1)
            p["width"] = np.where((rls < use8) | (gls < use8), 8, 16)
            p["jump_late"] = 1 if right else 0
            p["open"], p["extend"], p["keep_scores"] = self.OPEN, self.EXTEND, 1
        self.problems = probs
        self.qbuf, self.gbuf = qbuf, gbuf
        self.probs = (rng.integers(0, 21, int(2 * gls.sum())) / 20.0).astype(np.float64)
        gaps = np.zeros(n, INTRON_GAP_DTYPE)
        gaps["fills"] = np.stack([4 * np.arange(n) + j for j in range(4)], 1)
        gaps["span"] = rls + self.INTRON  # rev_goffsetR - goffsetL
        gaps["cdna_direction"] = 1
        gaps["prob_off"] = 2 * np.concatenate([[0], np.cumsum(gls)[:-1]])
        self.gaps = gaps
        self.n = n

    def genome(self):
        return self._genome

    def templates(self):
        """Traceback request templates, right end then left end per gap
        (fill, r and c are set by gdp_bridge_intron); pair_off laid out."""
        p = self.problems
        req = np.zeros(2 * self.n, TBREQ_DTYPE)
        rl, gl = p["rl"][0::4], p["gl"][0::4]
        req["queryoffset"][0::2] = rl - 1  # rev_roffset, walked with revp
        req["queryoffset"][1::2] = 0  # roffset
        req["genomeoffset"][0::2] = self.goffset + 2 * gl - 1  # rev_goffsetR
        req["genomeoffset"][1::2] = self.goffset
        req["revp"][0::2] = 1
        req["chroffset"], req["chrhigh"], req["watsonp"] = 0, len(self._genome), 1
        req["max_pairs"] = np.repeat(rl + gl + 4, 2)
        req["pair_off"][1:] = np.cumsum(req["max_pairs"].astype(np.int64))[:-1]
        return req
