"""Reader of tests/golden/golden_dp.npz (reference outputs, see make_golden.py)."""
import hashlib
import os

import numpy as np

import helpers as H

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_dp.npz")


def load():
    z = np.load(PATH, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    specs = []
    for i in range(len(g["kind"])):
        qo, go, gao = int(g["qoff"][i]), int(g["goff"][i]), int(g["galtoff"][i])
        rl, gl = int(g["rl"][i]), int(g["gl"][i])
        q = bytes(g["qbuf"][qo:qo + rl])
        gg = bytes(g["gbuf"][go:go + gl])
        ga = bytes(g["gbuf"][gao:gao + gl])
        specs.append(H.Spec(int(g["kind"][i]), int(g["width"][i]), q, gg, ga, int(g["mt"][i]), int(g["open"][i]),
                            int(g["ext"][i]), int(g["lband"][i]), int(g["uband"][i]), bool(g["late"][i]),
                            bool(g["revp"][i])))
    return g, specs


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_fill(g, i, s, D, Hs, scores=True):
    """D, Hs: dense [col][row] directions and scores of problem i (oracle
    layout).  scores=False: the fill kept no scores; directions only."""
    B = H.band_mask(s)
    off = int(g["dense_off"][i])
    if off >= 0:
        n = int(B.sum())
        assert (D[B] == g["dense_dirs"][off:off + n]).all(), ("golden directions", i)
        if scores:
            assert (Hs[B] == g["dense_scores"][off:off + n]).all(), ("golden scores", i)
    elif scores:
        assert digest(D[B].astype(np.uint8)) + digest(Hs[B].astype(np.int16)) == str(g["digests"][i]), ("golden digest", i)
    else:
        assert str(g["digests"][i]).startswith(digest(D[B].astype(np.uint8))), ("golden direction digest", i)
    if scores and g["final"][i] != -(1 << 31):
        assert int(Hs[s.gl, s.rl]) == int(g["final"][i]), ("golden corner score", i)


def traceback_expect(g, i):
    n = int(g["tb_n"][i])
    if g["tb_r"][i] < 0:
        return None
    off = int(g["tb_off"][i])
    return int(g["tb_r"][i]), int(g["tb_c"][i]), g["tb_pairs"][off:off + n], g["tb_counts"][i]


STD_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "std_golden.npz")


def load_std():
    """Scalar standard-fill vectors (make_std_golden.py): specs of kind 3."""
    z = np.load(STD_PATH, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    specs = []
    for i in range(len(g["rl"])):
        qo, go, gao = int(g["qoff"][i]), int(g["goff"][i]), int(g["galtoff"][i])
        rl, gl = int(g["rl"][i]), int(g["gl"][i])
        specs.append(H.Spec(3, 32, bytes(g["qbuf"][qo:qo + rl]), bytes(g["gbuf"][go:go + gl]),
                            bytes(g["gbuf"][gao:gao + gl]), int(g["mt"][i]), int(g["open"][i]), int(g["ext"][i]),
                            int(g["lband"][i]), int(g["uband"][i]), bool(g["late"][i]), bool(g["revp"][i]),
                            saturation=int(g["saturation"][i]), upperp=bool(g["upperp"][i]),
                            lowerp=bool(g["lowerp"][i])))
    return g, specs


def check_std(g, i, D, Hs):
    """D: dense [c][r] direction nibbles (every cell); Hs: scores, of which
    the cells r, c >= 1 are compared (int32)."""
    assert digest(D.astype(np.uint8)) + digest(Hs[1:, 1:].astype(np.int32)) == str(g["digests"][i]), \
        ("std golden digest", i)
    if g["final"][i] != -(1 << 31):
        s_gl, s_rl = int(g["gl"][i]), int(g["rl"][i])
        assert int(Hs[s_gl, s_rl]) == int(g["final"][i]), ("std golden corner", i)


def std_traceback_expect(g, i):
    if g["tb_r"][i] < 0:
        return None
    off, n = int(g["tb_off"][i]), int(g["tb_n"][i])
    return int(g["tb_r"][i]), int(g["tb_c"][i]), g["tb_pairs"][off:off + n], g["tb_counts"][i]


BRIDGE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bridge_golden.npz")
BRIDGE_FIELDS = ("width", "gl", "rlL", "rlR", "lbandL", "ubandL", "lbandR", "ubandR", "open", "ext", "leftoffset",
                 "rightoffset", "late")


def load_bridge():
    """cDNA gap bridge vectors (make_bridge_golden.py): a list of
    (args for Oracle.bridge_cdna, the reference's 5-tuple)."""
    z = np.load(BRIDGE_PATH, allow_pickle=False)
    tabs, offs, out = z["tables"], z["table_off"], z["out"]
    cases = []
    for i in range(len(out)):
        f = {k: int(z[k][i]) for k in BRIDGE_FIELDS}
        A = [tabs[offs[4 * i + j]: offs[4 * i + j + 1]] for j in range(4)]
        args = (f["width"], *A, f["gl"], f["rlL"], f["rlR"], f["lbandL"], f["ubandL"], f["lbandR"], f["ubandR"],
                f["open"], f["ext"], f["leftoffset"], f["rightoffset"], f["late"])
        cases.append((args, tuple(int(x) for x in out[i])))
    return cases


SKIP_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "genome_skip_golden.npz")


def skip_genome_image():
    """The genome of the genome-skip vectors: the reference's test
    chromosome with N runs (helpers.skip_genome), as ASCII and genomecomp."""
    import gzip
    import genomecomp as GC
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chr17test.fa.gz")
    fa = gzip.open(path).read().decode()
    seq = H.skip_genome("".join(line.strip() for line in fa.splitlines() if not line.startswith(">")).encode())
    return seq, GC.pack(seq)


def load_genome_skips():
    """tests/golden/genome_skip_golden.npz (make_genome_skip_golden.py):
    (seq, genomecomp image, [(Spec, (genomeoffset, chroffset, chrhigh,
    watsonp), pair records, counters)])."""
    z = np.load(SKIP_PATH, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    seq, gc = skip_genome_image()
    assert len(seq) == int(g["genome_bases"])
    assert hashlib.sha256(gc.tobytes()).digest() == g["genome_sha256"].tobytes(), "genome image differs"
    cases = []
    for i in range(len(g["kind"])):
        qo, go, rl, gl = int(g["qoff"][i]), int(g["goff_buf"][i]), int(g["rl"][i]), int(g["gl"][i])
        q, gg = bytes(g["qbuf"][qo:qo + rl]), bytes(g["gbuf"][go:go + gl])
        s = H.Spec(int(g["kind"][i]), int(g["width"][i]), q, gg, gg, int(g["mt"][i]), int(g["open"][i]),
                   int(g["ext"][i]), int(g["lband"][i]), int(g["uband"][i]), bool(g["late"][i]), bool(g["revp"][i]))
        po_, n = int(g["pair_off"][i]), int(g["npairs"][i])
        cases.append((s, (int(g["goff"][i]), int(g["chroffset"][i]), int(g["chrhigh"][i]), bool(g["watsonp"][i])),
                      g["pairs"][po_:po_ + n], g["counters"][i]))
    return seq, gc, cases


def skip_kinds(pairs):
    """Counts in one walk's records: skipped bases read from the genome
    (an A/C/G/T/N character), read as outside the chromosome ('*'), and gap
    holders."""
    sk = pairs[(pairs[:, 2] == ord(" ")) & (pairs[:, 3] == ord("-"))]
    star = int((sk[:, 4] == ord("*")).sum())
    return len(sk) - star, star, int((pairs[:, 6] == 1).sum())


INTRON_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "intron_bridge_golden.npz")
INTRON_FIELDS = ("width", "mt", "open", "ext", "rlength", "glengthL", "glengthR", "lbandL", "ubandL", "lbandR",
                 "ubandR", "leftoffset", "rightoffset", "chroffset", "chrhigh", "watsonp", "cdna_direction",
                 "finalp", "halfp", "late")


def load_intron_bridge():
    """Intron bridge vectors (make_intron_bridge_golden.py): a list of
    (case dict as helpers.intron_gap_cases makes it, plus "probL"/"probR",
    the reference's (finalscore, bestrL, bestrR, bestcL, bestcR) and its
    (left_prob, right_prob))."""
    z = np.load(INTRON_PATH, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    out = []
    qo = go = po_ = 0
    for i in range(len(g["out"])):
        c = {k: int(g[k][i]) for k in INTRON_FIELDS}
        for k in ("watsonp", "finalp", "halfp", "late"):
            c[k] = bool(c[k])
        rl, glL, glR = c["rlength"], c["glengthL"], c["glengthR"]
        c["q"] = bytes(g["qbuf"][qo:qo + rl])
        c["gL"] = bytes(g["gbuf"][go:go + glL])
        c["gRseg"] = bytes(g["gbuf"][go + glL:go + glL + glR])
        c["gLalt"] = bytes(g["gbuf"][go + glL + glR:go + 2 * glL + glR])
        c["gRsegalt"] = bytes(g["gbuf"][go + 2 * glL + glR:go + 2 * glL + 2 * glR])
        c["probL"] = g["site_probs"][po_:po_ + glL]
        c["probR"] = g["site_probs"][po_ + glL:po_ + glL + glR]
        qo += rl
        go += 2 * (glL + glR)
        po_ += glL + glR
        out.append((c, tuple(int(x) for x in g["out"][i]), tuple(float(x) for x in g["out_probs"][i])))
    return out
