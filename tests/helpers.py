"""Shared generators and comparison helpers for the tests."""
import os

import numpy as np

import pyoracle as po
from gmapdp.lib import PROBLEM_DTYPE, TBREQ_DTYPE

ALPH = np.frombuffer(b"ACGT", np.uint8)


class Spec:
    """One DP problem as the reference's callers would pass it."""

    def __init__(self, kind, width, q, g, galt, mt, open_, ext, lband, uband, late, revp, start=None,
                 saturation=-32768, upperp=True, lowerp=True):
        self.kind, self.width = kind, width
        self.q, self.g, self.galt = q, g, galt
        self.mt, self.open, self.ext = mt, open_, ext
        self.lband, self.uband, self.late, self.revp = lband, uband, late, revp
        self.start = start  # (r, c) traceback start
        # kind 3 (Dynprog_standard) only
        self.saturation, self.upperp, self.lowerp = saturation, upperp, lowerp

    @property
    def rl(self):
        return len(self.q)

    @property
    def gl(self):
        return len(self.g)


def rand_seq(rng, n, nfrac=0.0):
    s = ALPH[rng.integers(0, 4, n)].copy()
    if nfrac:
        s[rng.random(n) < nfrac] = ord("N")
    return s.tobytes()


def mutate(rng, src, sub, ind, maxind=3):
    out = bytearray()
    i = 0
    while i < len(src):
        u = rng.random()
        if u < ind / 2:
            out += rand_seq(rng, int(rng.integers(1, maxind + 1)))
            continue
        if u < ind:
            i += int(rng.integers(1, maxind + 1))
            continue
        ch = src[i]
        i += 1
        if rng.random() < sub:
            ch = rand_seq(rng, 1)[0]
        out.append(ch)
    return bytes(out)


def random_specs(seed, n, kinds=(0, 1, 2), widths=(8, 16), maxrl8=60, maxrl16=200, lowercase=True):
    """Adversarial random problems: all kinds, widths, tie rules, strands,
    IUPAC/N characters, lower case, alternate alleles and narrow bands."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        width = int(rng.choice(widths))
        rl = int(rng.integers(1, maxrl8 if width == 8 else maxrl16))
        gl = int(rng.integers(max(1, rl - 10), rl + 40))
        g = rand_seq(rng, gl, 0.02)
        q = mutate(rng, g, float(rng.choice([0.02, 0.1, 0.3, 0.75])), 0.03)[:rl]
        if len(q) < rl:
            q = q + rand_seq(rng, rl - len(q))
        if lowercase and rng.random() < 0.2:
            q = bytes(ch + 32 if (65 <= ch <= 90 and rng.random() < 0.3) else ch for ch in q)
        if rng.random() < 0.1:
            q = bytes(b"RYKMSWBDHVNX"[rng.integers(0, 12)] if rng.random() < 0.1 else ch for ch in q)
        galt = g if rng.random() < 0.8 else bytes(ch if rng.random() < 0.9 else ALPH[rng.integers(0, 4)] for ch in g)
        mt = int(rng.integers(0, 4))
        op = int(rng.choice([-8, -7, -6, -10, -20]))
        ex = int(rng.choice([-3, -2, -1, -5]))
        late = bool(rng.integers(0, 2))
        revp = bool(rng.integers(0, 2))
        eb = int(rng.integers(0, 25))
        lb, ub = (eb, gl - rl + eb) if gl >= rl else (rl - gl + eb, eb)
        kind = int(rng.choice(kinds))
        specs.append(Spec(kind, width, q, g, galt, mt, op, ex, lb, ub, late, revp))
    return specs


def wide_specs(seed, n):
    """Problems whose band spans more than 1000 columns (the wide single
    sweep): short reads against long genome windows and the reverse, with
    Dynprog_compute_bands' widened bands, both one-sided kinds, and bands
    past the matrix's edges (narrowed by the planner, no cell changes)."""
    rng = np.random.default_rng(seed)
    specs = []
    for t in range(n):
        shape = t % 6
        eb = int(rng.integers(3, 30))
        if shape in (0, 2):  # a short read against a long window
            rl, gl = int(rng.integers(10, 80)), int(rng.integers(1100, 2600))
        elif shape in (1, 3):  # a long read against a short window
            rl, gl = int(rng.integers(1100, 2000)), int(rng.integers(10, 60))
        else:  # bands past both edges of a small matrix
            rl, gl = int(rng.integers(20, 200)), int(rng.integers(20, 200))
        g = rand_seq(rng, gl, 0.01)
        src = g if gl >= rl else g + rand_seq(rng, rl)
        q = mutate(rng, src, 0.05, 0.02)[:rl].ljust(rl, b"A")
        lb, ub = (eb, gl - rl + eb) if gl >= rl else (rl - gl + eb, eb)
        if shape in (0, 1):
            kind, width = 0, 16
        elif shape == 2:
            kind, width = 1, int(rng.choice([8, 16]))
        elif shape == 3:
            kind, width = 2, int(rng.choice([8, 16]))
        else:
            kind, width = 0, 16
            lb, ub = rl + int(rng.integers(1, 40)), gl + int(rng.integers(1, 900))
        late, revp = bool(t % 2), bool((t // 2) % 2)
        specs.append(Spec(kind, width, q, g, g, int(rng.integers(0, 4)), -8, -3, lb, ub, late, revp))
    return specs


def random_std_specs(seed, n, maxrl=260):
    """Problems for the scalar standard fill (kind 3): the shapes of
    random_specs plus its switches -- upperp / lowerp off, a saturation floor
    above -inf, narrow and skewed bands (some missing the corner cell)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in random_specs(seed, n, kinds=(0,), widths=(16,), maxrl16=maxrl):
        lb, ub = s.lband, s.uband
        u = rng.random()
        if u < 0.15:
            lb, ub = int(rng.integers(0, 6)), int(rng.integers(0, 6))
        elif u < 0.25:
            lb = ub = min(s.rl + s.gl, 263)
        out.append(Spec(3, 32, s.q, s.g, s.galt, s.mt, s.open, s.ext, lb, ub, s.late, s.revp,
                        saturation=int(rng.choice([-32768, -32768, -200, -40, 0])),
                        upperp=bool(rng.random() > 0.15), lowerp=bool(rng.random() > 0.15)))
    return out


def oracle_fill(O, s):
    if s.kind == 3:
        return O.fill_std(s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp, s.saturation,
                          s.upperp, s.lowerp)
    if s.kind == 0:
        return O.fill_full(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.uband, s.late, s.revp)
    if s.kind == 1:
        return O.fill_upper(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.uband, s.late, s.revp)
    return O.fill_lower(s.width, s.q, s.g, s.galt, s.mt, s.open, s.ext, s.lband, s.late, s.revp)


def band_mask(s):
    if s.kind in (0, 3):
        return po.band_mask_full(s.rl, s.gl, s.lband, s.uband)
    if s.kind == 1:
        return po.band_mask_upper(s.rl, s.gl, s.uband)
    return po.band_mask_lower(s.rl, s.gl, s.lband)


def computed_mask(s, layout=0):
    """Cells the fill computes deterministically, as dense [col][row] of the
    sweep (rows = query for full/upper, genome for lower): the band plus the
    out-of-band prefix of each emulated 32/16-row block; a narrow one-sided
    sweep (layout 2) computes the band's cells from the diagonal on only."""
    S = 1 if (s.kind == 3 or layout == 2) else (32 if s.width == 8 else 16)
    R, C = (s.gl, s.rl) if s.kind == 2 else (s.rl, s.gl)
    r = np.arange(R + 1)[None, :]
    c = np.arange(C + 1)[:, None]
    blk = (r // S) * S
    if s.kind in (0, 3):
        cs = np.maximum(0, blk - s.lband)
        ce = np.minimum(C, r + s.uband)
    else:
        ub = s.uband if s.kind == 1 else s.lband
        cs = blk
        ce = np.minimum(C, r + ub)
    return (c >= cs) & (c <= ce)


def pack_batch(specs, keep_scores=False):
    """Problems plus query/genome buffers in sweep order."""
    probs = np.zeros(len(specs), PROBLEM_DTYPE)
    qparts, gparts = [], []
    qo = go = 0
    for i, s in enumerate(specs):
        q, g, ga = (s.q[::-1], s.g[::-1], s.galt[::-1]) if s.revp else (s.q, s.g, s.galt)
        p = probs[i]
        p["qoff"], p["goff"], p["galtoff"] = qo, go, go + len(g)
        p["rl"], p["gl"], p["lband"], p["uband"] = s.rl, s.gl, s.lband, s.uband
        p["kind"], p["width"], p["mismatchtype"], p["jump_late"] = s.kind, s.width, s.mt, int(s.late)
        p["open"], p["extend"], p["keep_scores"] = s.open, s.ext, int(keep_scores)
        if s.kind == 3:
            p["std_flags"] = int(s.upperp) | (int(s.lowerp) << 1)
            p["saturation"] = s.saturation
        qparts.append(q)
        gparts.append(g + ga)
        qo += len(q)
        go += len(g) + len(ga)
    return probs, b"".join(qparts), b"".join(gparts)


def pairs_to_records(pairs):
    """GPU pair records -> oracle 8-int records."""
    out = np.zeros((len(pairs), 8), np.int32)
    out[:, 0] = pairs["querypos"]
    out[:, 1] = pairs["genomepos"]
    out[:, 2] = pairs["cdna"]
    out[:, 3] = pairs["comp"]
    out[:, 4] = pairs["genome"]
    out[:, 5] = pairs["genomealt"]
    out[:, 6] = pairs["gapp"]
    out[:, 7] = pairs["genomejump"]
    return out


def tb_request(i, s, r, c, qoff=5, goff=7, chroffset=0, chrhigh=0, watsonp=True, genestrand=0):
    req = np.zeros(1, TBREQ_DTYPE)
    req["fill"], req["r"], req["c"] = i, r, c
    req["queryoffset"], req["genomeoffset"] = qoff, goff
    req["chroffset"], req["chrhigh"], req["watsonp"] = chroffset, chrhigh, int(watsonp)
    req["revp"], req["genestrand"] = int(s.revp), genestrand
    req["max_pairs"] = s.rl + s.gl + 4
    return req[0]


def oracle_end5_splicejunction(O, e, five=True):
    """Dynprog_end5_splicejunction (dynprog_end.c:1652-1922) composed from the
    oracle's fills, last-row best cell and local walks; e as
    gmapdp.ends.end5_splicejunction takes it.  None or (pair records,
    {traceback_score, missscore, nmatches, nmismatches, nopens, nindels})."""
    from gmapdp.ends import end_penalties, wide_bands
    q, g, ga = e["rev_rseq"], e["rev_gseq"], e["rev_galt"]
    rl, gl = len(q), len(g)
    open_, ext = end_penalties(e["defect_rate"])
    lb, ub = wide_bands(rl, gl, e["extraband"])
    u8 = O.use8p(3)
    width = 8 if (rl < u8 or gl < u8) else 16
    late = bool(e["jump_late"]) != five  # the 5' form fills with the opposite tie rule
    su = Spec(1, width, q, g, ga, 3, open_, ext, lb, ub, late, five)
    sl = Spec(2, width, q, g, ga, 3, open_, ext, lb, ub, late, five)
    Hu, Du = oracle_fill(O, su)
    Hl, Dl = oracle_fill(O, sl)
    fs, br, bc = O.best_endpoint(True, Hu, Hl, width, rl, gl, lb, ub, late)
    if fs < 0:
        return None
    walk = []
    r, c = br, bc
    for endc, goff in ((e["contlength"], e["far"]), (0, e["anchor"])):
        kind = 1 if c >= r else 2
        p, cnt, (r, c) = O.traceback_local(kind, Du if kind == 1 else Dl, r, c, endc, q, q.upper(), g, ga,
                                           e["rev_roffset"], goff, five, e["genestrand"])
        walk.append((p, cnt))
    # the known-site gap holder (pairpool.c:374-452, knownp): gapp 2
    gap = np.array([[-1, -1, 32, 32, 32, 32, 2, (e["anchor"] - e["far"]) * (1 if five else -1)]], np.int32)
    pairs = np.concatenate([walk[0][0], gap, walk[1][0]])
    cnt = walk[0][1] + walk[1][1]
    k = 0
    while k < len(pairs) and pairs[k, 3] == ord("-"):
        k += 1
    pairs = pairs[k:] if five else pairs[k:][::-1].copy()  # the 3' form returns its trimmed list as it is
    return pairs, np.array([cnt[0], cnt[0] - rl * 3, cnt[1], cnt[2], cnt[3], cnt[4]], np.int32)


def random_splice_ends(rng, n, long_jumps=False):
    """Splice-junction end cases: a read end against a genome window that
    holds its continuation past the far column.  long_jumps draws the
    anchor-to-far distance (the intron the known gap holder spans) from
    32768-500000 instead of 20-400."""
    out = []
    for t in range(n):
        rl = int(rng.integers(8, 70))
        gl = rl + int(rng.integers(0, 25))
        g = rand_seq(rng, gl, 0.01)
        q = mutate(rng, g, 0.05, 0.03)[:rl].ljust(rl, b"C")
        if t % 6 == 0:
            q = rand_seq(rng, rl)  # no match: a negative best score, no walk
        ga = bytes(x if rng.random() > 0.05 else ord("R") for x in g)
        anchor = int(rng.integers(600000, 2000000) if long_jumps else rng.integers(200, 5000))
        jump = int(rng.integers(32768, 500001) if long_jumps else rng.integers(20, 400))
        out.append({"rev_rseq": q, "rev_gseq": g, "rev_galt": ga, "rev_roffset": int(rng.integers(rl, rl + 300)),
                    "anchor": anchor, "far": anchor + jump * (-1 if t % 2 else 1),
                    "genestrand": 0,
                    "jump_late": bool(t % 2), "extraband": int(rng.integers(3, 10)),
                    "defect_rate": float(rng.choice([0.001, 0.01, 0.05])),
                    "contlength": int(rng.integers(0, min(rl, gl) + 2))})
    return out


def skip_genome(seq):
    """The reference's test chromosome with N runs put in (40 bases in every
    seventh kilobase), so that genome skips also read flagged bases."""
    a = np.frombuffer(bytes(seq), np.uint8).copy()
    pos = np.arange(len(a))
    a[((pos // 1000) % 7 == 3) & (pos % 1000 < 40)] = ord("N")
    return a.tobytes()


def genome_skip_cases(rng, n, nbases):
    """Walks that end in a genome skip read from the genome
    (Pairpool_add_genomeskip, pairpool.c:1067-1180): the read is the tail of
    its genome window, so the walk from (rl, gl) reaches row 0 at column L,
    1-8 (bases pushed one by one through get_genomic_nt) or 9-30 (a gap
    holder).  Full-band and upper fills, both widths, both strands, revp,
    and chromosome bounds placed so that some skipped bases fall outside
    them.  Returns [(Spec, (genomeoffset, chroffset, chrhigh, watsonp))]."""
    out = []
    for t in range(n):
        L = int(rng.integers(1, 9)) if t % 5 < 3 else int(rng.integers(9, 31))
        rl = int(rng.integers(3, 60))
        gl = rl + L
        g = rand_seq(rng, gl)
        q = g[L:]
        if t % 4 == 1:  # a mismatch or two inside the read
            q = bytes(ch if rng.random() > 0.04 else ALPH[rng.integers(0, 4)] for ch in q)
        kind = 0 if t % 3 else 1
        width = 8 if t % 2 else 16
        eb = int(rng.integers(1, 6))
        lb, ub = eb, L + eb
        revp = t % 7 == 2
        late = bool((t // 2) % 2)
        s = Spec(kind, width, q, g, g, int(rng.integers(0, 3)), -8, -3, lb, ub, late, revp)
        watsonp = bool(rng.integers(0, 2))
        u = rng.random()
        if u < 0.5:  # inside the chromosome
            chroffset = int(rng.integers(0, nbases // 2))
            chrhigh = int(rng.integers(chroffset + 200, nbases))
            goff = int(rng.integers(40, chrhigh - chroffset - 40)) if watsonp else \
                int(rng.integers(40, chrhigh - chroffset - 40))
        elif u < 0.75:  # the skip crosses chrhigh (plus) or falls past it (minus)
            chroffset = int(rng.integers(0, nbases // 2))
            chrhigh = int(rng.integers(chroffset + 100, nbases))
            goff = (chrhigh - chroffset) - int(rng.integers(0, L + 1)) if watsonp else int(rng.integers(-L, 2))
        else:  # the skip crosses chroffset
            chroffset = int(rng.integers(100, nbases // 2))
            chrhigh = int(rng.integers(chroffset + 100, nbases))
            goff = -int(rng.integers(0, L + 1)) if watsonp else (chrhigh - chroffset) - int(rng.integers(-L, 2))
        if revp:
            goff += gl  # the reversed walk's coordinates run down from genomeoffset
        out.append((s, (goff, chroffset, chrhigh, watsonp)))
    return out


_COMPL = bytes.maketrans(b"ACGTNacgtn", b"TGCANtgcan")


def revcomp(b):
    return bytes(b).translate(_COMPL)[::-1]


def caller_cases(rng, genome_seq, n_single, n_end, chromosomes=2):
    """Calls of the reference's own callers over a genome: Dynprog_single_gap
    with the single-gap generator (100-nt reads from 150-nt windows, 2 %
    substitutions, 0.2 % indels of 1-3, extraband 10, widened bands, HIGHQ
    and LOWQ defect rates, so int16 and int8 fills), a few windows of about
    2000 bases against short reads (bands of about 2000 columns), and
    windows as long as the read (the ungapped shortcut); Dynprog_end5_gap and
    Dynprog_end3_gap with the end-extension generator (reads of 10-150
    against windows 0-20 longer, 3 % substitutions, 1 % indels of 1-5, 1 % N
    in the read, extraband 6) under all four end modes, a few windows past
    the workspace's 2000 columns (chopped by the caller).  The genome is cut
    into `chromosomes` equal chromosomes.  Each case is a dict of the call's
    arguments; the read is taken from the segment the caller extracts."""
    nb = len(genome_seq)
    clen = nb // chromosomes
    cases = []

    def chrom():
        k = int(rng.integers(0, chromosomes))
        return k * clen, k * clen + clen - 1  # chroffset, chrhigh

    def seg_right(left, length, chrhigh):
        s = bytearray(genome_seq[left:left + length])
        for j in range(len(s)):
            if left + j >= chrhigh:
                s[j] = ord("*")
        return bytes(s)

    for t in range(n_single + n_end):
        single = t < n_single
        watsonp = bool(rng.integers(0, 2))
        chroffset, chrhigh = chrom()
        if single:
            u = rng.random()
            if u < 0.06:
                rl, gl = int(rng.integers(30, 60)), int(rng.integers(1900, 2001))
            elif u < 0.12:
                rl = int(rng.integers(20, 120))
                gl = rl
            else:
                rl, gl = 100, int(rng.integers(100, 151))
        else:
            rl = int(rng.integers(10, 151))
            gl = rl + int(rng.integers(0, 21)) if rng.random() > 0.05 else int(rng.integers(2001, 2100))
        five = (not single) and t % 2 == 0
        goffset = int(rng.integers(10 + (gl if five else 0), clen - gl - 10))
        # the segment the caller extracts (dynprog_single.c:544-550,
        # dynprog_end.c:1393-1399 for the 5' end, 1990-1996 for the 3' end)
        if five:
            left = chroffset + goffset + 1 - gl if watsonp else chrhigh - goffset
        else:
            left = chroffset + goffset if watsonp else chrhigh - goffset + 1 - gl
        seg = genome_seq[left:left + gl]
        if not watsonp:
            seg = revcomp(seg)
        if single:
            if gl == rl:
                q = mutate(rng, seg, 0.02, 0.0)[:rl]
            else:
                a = int(rng.integers(0, rl + 1))
                src = seg[:a] + seg[a + gl - rl:]
                q = mutate(rng, src, 0.02, 0.002, 3)[:rl]
        else:
            # the 5' end reads its segment backwards from the last base: the
            # read is anchored at the segment's end; the 3' read at its start
            src = seg[gl - rl:] if five else seg[:rl]
            q = mutate(rng, src, 0.03, 0.01, 5)[:rl]
            q = bytes(ord("N") if rng.random() < 0.01 else ch for ch in q)
        q = q.ljust(rl, b"A")
        case = {"single": single, "five": five, "rseq": q, "rl": rl, "gl": gl,
                "roffset": int(rng.integers(0, 50)), "goffset": goffset, "chroffset": chroffset,
                "chrhigh": chrhigh, "watsonp": watsonp, "genestrand": 0, "jump_late": bool(rng.integers(0, 2)),
                "extraband": 10 if single else 6, "widebandp": True,
                "defect_rate": float(rng.choice([0.001, 0.02] if single else [0.001, 0.005, 0.02])),
                "endalign": int(rng.choice([3, 0, 1, 2], p=[0.4, 0.4, 0.1, 0.1])) if not single else 0}
        cases.append(case)
    return cases


def intron_genome():
    """The intron-bridge genome: the reference's test chromosome followed by
    6000 bases of a 12-base repeat holding GT, AG, CT and AC, where equal
    splice-site contexts give equal probabilities (ties the bridge breaks by
    scan order).  ASCII and genomecomp."""
    import gzip
    import genomecomp as GC
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chr17test.fa.gz")
    fa = gzip.open(path).read().decode()
    seq = "".join(line.strip() for line in fa.splitlines() if not line.startswith(">")).encode()
    seq += b"GTAAGCTACAGT" * 500
    return seq, GC.pack(seq)


_SITES = {1: (b"GT", b"AG"), -1: (b"CT", b"AC"), 0: (b"GT", b"AG")}


def intron_gap_cases(rng, n, genome_seq, repeat_start=200000, long_reads=False):
    """Intron gaps set up as Dynprog_genome_gap does (dynprog_genome.c:
    3287-3547, called from stage3.c:9567 with glengthL = glengthR = queryjump +
    extramaterial_paired): a read made of the end of a left exon and the start
    of a right exon (2 % substitutions, 0.5 % indels), the left segment from
    goffsetL forward and the right segment ending at rev_goffsetR, both
    longer than the read, on either strand (the minus strand's coordinate x
    reads the complement of base chrhigh - x).  Exon ends sit before the
    canonical dinucleotides of cdna_direction (+1 GT..AG, -1 CT..AC, 0 either)
    in most cases; introns of 1-20 bases, and segments that overlap, make
    the bridge's rightoffset-leftoffset bound bite; some gaps lie in the repeat after
    repeat_start; one read in ten has 60 % substitutions (joins that score
    below 0); one gap in four has SNP alleles on 5 % of its segments' bases.
    long_reads draws reads of 1000-1499 bases instead.  Widths follow the caller's use8p rule (3504-3508);
    penalties PAIRED_*_{HIGHQ,MEDQ,LOWQ} (dynprog.h:69-75).  Returns dicts
    with the bridge's arguments and the read."""
    nb = len(genome_seq)
    G = np.frombuffer(genome_seq, np.uint8)
    comp = np.frombuffer(bytes.maketrans(b"ACGTN", b"TGCAN"), np.uint8)
    use8p = {0: 41, 1: 63, 2: 127}
    pen = {0: (-8, -3), 1: (-7, -2), 2: (-6, -1)}
    out = []
    for t in range(n):
        watsonp = bool(t % 2 == 0) if t < 8 else bool(rng.integers(0, 2))
        cdna_direction = [1, -1, 0][t % 3]
        u = rng.random()
        rlength = int(rng.integers(2, 40)) if u < 0.3 else (int(rng.integers(40, 200)) if u < 0.93 else
                                                            int(rng.integers(200, 520)))
        if long_reads:
            rlength = int(rng.integers(1000, 1500))
        extra = int(rng.integers(1, 26))
        glength = rlength + extra
        intron = int(rng.integers(1, 21)) if t % 6 == 5 else int(rng.integers(30, 3000))
        if t % 12 == 11:
            intron = -int(rng.integers(0, 12))  # segments that overlap (rightoffset-leftoffset < rlength)
        in_repeat = t % 9 == 4
        chroffset, chrhigh = (int(rng.integers(0, 500)), nb - int(rng.integers(64, 400)))
        span = chrhigh - chroffset

        def S(x0, length):
            if watsonp:
                return G[chroffset + x0: chroffset + x0 + length].tobytes()
            idx = chrhigh - np.arange(x0, x0 + length)
            return comp[G[idx]].tobytes()

        a = int(rng.integers(1, rlength)) if rlength > 1 else 1
        b = rlength - a
        if in_repeat:
            intron = min(intron, 300)
        # room for the segments, the dinucleotide search (200 bases each side)
        # and the splice models' context
        lo = 200 + glength
        hi = span - intron - glength - 400
        if in_repeat and watsonp:
            lo = repeat_start - chroffset + 200 + glength
        elif in_repeat:  # the minus strand meets the repeat at its start
            hi = chrhigh - repeat_start - intron - glength - 400
        e1 = int(rng.integers(lo, hi))  # last base of the left exon
        if rng.random() < 0.8:
            want = _SITES[cdna_direction][0]
            for d in range(200):
                if S(e1 + 1 + d, 2) == want:
                    e1 += d
                    break
        i2 = e1 + 1 + intron  # first base of the right exon
        if rng.random() < 0.8 and intron >= 4:
            want = _SITES[cdna_direction][1]
            for d in range(200):
                if S(i2 - 2 + d, 2) == want:
                    i2 += d
                    break
        goffsetL = e1 + 1 - a
        rev_goffsetR = i2 + b - 1
        gL = S(goffsetL, glength)
        gRseg = S(rev_goffsetR + 1 - glength, glength)
        src = S(goffsetL, a) + S(rev_goffsetR + 1 - b, b)
        sub = 0.6 if t % 10 == 7 else 0.02  # some reads too poor for a positive join
        q = mutate(rng, src, sub, 0.005, 3)[:rlength].ljust(rlength, b"C")
        gLalt, gRsegalt = gL, gRseg
        if t % 4 == 3:  # SNP alleles: the fills score them, the sites read either allele
            gLalt = bytes(ch if rng.random() > 0.05 else int(ALPH[rng.integers(0, 4)]) for ch in gL)
            gRsegalt = bytes(ch if rng.random() > 0.05 else int(ALPH[rng.integers(0, 4)]) for ch in gRseg)
        mt = int(rng.integers(0, 3))
        width = 8 if rlength < use8p[mt] or glength < use8p[mt] else 16
        eb = int(rng.integers(2, 13))
        lb, ub = eb, glength - rlength + eb  # Dynprog_compute_bands, widebandp (dynprog.c:1245-1261)
        out.append({"width": width, "mt": mt, "open": pen[mt][0], "ext": pen[mt][1], "q": q, "gL": gL,
                    "gRseg": gRseg, "gLalt": gLalt, "gRsegalt": gRsegalt, "rlength": rlength, "glengthL": glength, "glengthR": glength, "lbandL": lb,
                    "ubandL": ub, "lbandR": lb, "ubandR": ub, "leftoffset": goffsetL, "rightoffset": rev_goffsetR,
                    "chroffset": chroffset, "chrhigh": chrhigh, "watsonp": watsonp,
                    "cdna_direction": cdna_direction, "finalp": bool(rng.integers(0, 2)),
                    "halfp": bool(t % 4 == 1), "late": bool(rng.integers(0, 2))})
    return out


def intron_gap_specs(c):
    """The four one-sided fills of an intron gap (dynprog_genome.c:3510-3547):
    left upper and lower over the read forward against the left segment,
    right upper and lower over the read backwards against the right segment
    read backwards, with the other tie rule."""
    q, gL, gLa, gRf, gRfa = c["q"], c["gL"], c["gLalt"], c["gRseg"][::-1], c["gRsegalt"][::-1]
    w, mt, op, ex = c["width"], c["mt"], c["open"], c["ext"]
    return [Spec(1, w, q, gL, gLa, mt, op, ex, c["lbandL"], c["ubandL"], c["late"], False),
            Spec(2, w, q, gL, gLa, mt, op, ex, c["lbandL"], c["ubandL"], c["late"], False),
            Spec(1, w, q[::-1], gRf, gRfa, mt, op, ex, c["lbandR"], c["ubandR"], not c["late"], False),
            Spec(2, w, q[::-1], gRf, gRfa, mt, op, ex, c["lbandR"], c["ubandR"], not c["late"], False)]


def genome_gap_calls(rng, genome_seq, n, chromosomes=2):
    """Calls of Dynprog_genome_gap as stage 3 makes them (stage3.c:9538-9576:
    glengthL = glengthR = queryjump + extramaterial_paired, goffsetL =
    genomedp5, rev_goffsetR = genomedp3): a read of 10-300 bases made of the
    end of a left exon and the start of a right exon across an intron of
    30-4000 bases (1 % substitutions, 0.2 % indels of 1-2 bases, each exon
    part keeping its anchor; one read in twelve unrelated to the genome), exon ends before the canonical dinucleotides of
    cdna_direction in most calls, on either strand of one of `chromosomes`
    equal chromosomes, extramaterial 6-30, extraband 3-10, all three defect
    classes (so both widths and, for preliminary calls under DEFECT_MEDQ,
    genome_gap_simple), finalp and halfp both ways; a few reads past the
    workspace's 500 rows (chopped by the caller).  Dicts of the call's
    arguments."""
    nb = len(genome_seq)
    clen = nb // chromosomes
    G = np.frombuffer(genome_seq, np.uint8)
    comp = np.frombuffer(bytes.maketrans(b"ACGTN", b"TGCAN"), np.uint8)
    out = []
    for t in range(n):
        k = int(rng.integers(0, chromosomes))
        chroffset, chrhigh = k * clen, k * clen + clen - 1
        watsonp = bool(rng.integers(0, 2))
        cdna_direction = [1, -1, 0][t % 3]
        rl = int(rng.integers(10, 300)) if t % 40 else int(rng.integers(501, 520))
        extra = int(rng.integers(6, 31))
        gl = rl + extra
        intron = int(rng.integers(30, 4000))

        def S(x0, length):
            if watsonp:
                return G[chroffset + x0: chroffset + x0 + length].tobytes()
            return comp[G[chrhigh - np.arange(x0, x0 + length)]].tobytes()

        a = int(rng.integers(1, rl))
        b = rl - a
        e1 = int(rng.integers(gl + 300, clen - intron - gl - 700))
        if rng.random() < 0.8:
            for d in range(200):
                if S(e1 + 1 + d, 2) == _SITES[cdna_direction][0]:
                    e1 += d
                    break
        i2 = e1 + 1 + intron
        if rng.random() < 0.8:
            for d in range(200):
                if S(i2 - 2 + d, 2) == _SITES[cdna_direction][1]:
                    i2 += d
                    break
        goffsetL, rev_goffsetR = e1 + 1 - a, i2 + b - 1
        # each exon part keeps its anchor: the left from the read's start,
        # the right at its end (the gap's two ends are aligned bases)
        left = mutate(rng, S(goffsetL, a), 0.01, 0.002, 2)[:a].ljust(a, b"A")
        right = mutate(rng, S(rev_goffsetR + 1 - b, b), 0.01, 0.002, 2)[-b:].rjust(b, b"A") if b else b""
        q = left + right if t % 12 != 5 else rand_seq(rng, rl)
        out.append({"rseq": q, "rl": rl, "glL": gl, "glR": gl, "roffset": int(rng.integers(0, 50)),
                    "goffsetL": goffsetL, "rev_goffsetR": rev_goffsetR, "chroffset": chroffset, "chrhigh": chrhigh,
                    "cdna_direction": cdna_direction, "watsonp": watsonp, "genestrand": 0,
                    "jump_late": bool(rng.integers(0, 2)), "extraband": int(rng.integers(3, 11)),
                    "defect_rate": float(rng.choice([0.001, 0.005, 0.02])), "maxpeelback": 20,
                    "halfp": bool(t % 5 == 2), "finalp": bool(rng.integers(0, 2))})
    return out


def cdna_gap_calls(rng, genome_seq, n, chromosomes=2, flank=260, block=0):
    """Calls of Dynprog_cdna_gap as stage 3 makes them (stage3.c:9274-9321: a
    query gap longer than its genome gap by more than MININTRONLEN, both ends
    read rlengthL = rlengthR = genomejump + extramaterial_paired bases): a
    genome gap of 2-150 columns on either strand, and a query holding the
    genome gap's bases with an insertion of 10-60 random bases at a random
    point (1 % substitutions), between `flank` bases of the genome on each
    side (the ends read past the query gap into them).  Extramaterial 4-30,
    extraband 3-10, all three defect classes; a few genome gaps past the
    workspace's 2000 columns.  Dicts of the call's arguments; query holds the
    whole read, qdp5 / qdp3 the query gap's first and last bases."""
    nb = len(genome_seq)
    clen = nb // chromosomes
    G = np.frombuffer(genome_seq, np.uint8)
    comp = np.frombuffer(bytes.maketrans(b"ACGTN", b"TGCAN"), np.uint8)
    out = []
    for t in range(n):
        k = int(rng.integers(0, chromosomes))
        chroffset, chrhigh = k * clen, k * clen + clen - 1
        watsonp = bool(rng.integers(0, 2))
        glength = int(rng.integers(2, 151)) if t % 50 else int(rng.integers(2001, 2050))
        ins = int(rng.integers(10, 61))
        extra = int(rng.integers(4, 31))
        goffset = int(rng.integers(flank + 10, clen - glength - flank - 10))

        def S(x0, length):
            if watsonp:
                return G[chroffset + x0: chroffset + x0 + length].tobytes()
            return comp[G[chrhigh - np.arange(x0, x0 + length)]].tobytes()

        a = int(rng.integers(0, glength + 1))
        seg = S(goffset, glength)
        if block:
            # no insertion: `block` bases of the gap replaced by others, which
            # a join can leave to both sides as cDNA and genome insertions
            a = min(a, max(glength - block, 0))
            swap = bytes(b"ACGT"[(b"ACGT".find(bytes([x])) + 1 + int(rng.integers(0, 3))) % 4]
                         for x in seg[a:a + block])
            mid = seg[:a] + swap + seg[a + block:]
        else:
            mid = mutate(rng, seg[:a], 0.01, 0.0) + rand_seq(rng, ins) + mutate(rng, seg[a:], 0.01, 0.0)
        query = S(goffset - flank, flank) + mid + S(goffset + glength, flank)
        qdp5 = flank
        qdp3 = flank + len(mid) - 1
        rlen = min(glength + extra, flank)
        out.append({"query": query, "qdp5": qdp5, "qdp3": qdp3, "rlengthL": rlen, "rlengthR": rlen,
                    "glength": glength, "goffset": goffset, "chroffset": chroffset, "chrhigh": chrhigh,
                    "watsonp": watsonp, "genestrand": 0, "jump_late": bool(rng.integers(0, 2)),
                    "extraband": int(rng.integers(3, 11)), "defect_rate": float(rng.choice([0.001, 0.005, 0.02]))})
    return out
