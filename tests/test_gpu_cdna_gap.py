"""Dynprog_cdna_gap batched on the device (gdp_cd_run, gmapdp.callers.cdna_gap)
against the reference's own caller: the 600 calls of
tests/golden/callers_cdna_golden.npz (make_callers_golden.py ran
Dynprog_cdna_gap, dynprog_cdna.c:786-1180, over the reference's SIMD fills
on the genome-skip test genome, with GMAP's stage-3 workspaces).  Genome gaps
of 2-150 columns on either strand under a query gap holding a 10-60-base
insertion, all three defect classes (both fill widths), and a few gaps past
the workspace's 2000 columns.  The returned lists (the right walk, the
insertion pairs or the gap holder, the left walk), traceback_score,
incompletep and whether dynprogindex advanced must match.  The 300 calls of
tests/golden/callers_cdna_insert_golden.npz hold query gaps as long as their
genome gaps with 9 bases replaced, so that most joins take the reference's
insertion-pairs branch (dynprog_cdna.c:1007-1030: 9 cDNA-insertion and 9
genome-insertion pairs in place of the gap holder)."""
import os

import numpy as np
import pytest

import golden_io
from test_gpu_single_gap import _records

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDS = {"gaps": "callers_cdna_golden.npz", "insert_pairs": "callers_cdna_insert_golden.npz"}


@pytest.mark.parametrize("which", sorted(GOLDS))
def test_cdna_gap_callers_match_the_reference(which):
    from gmapdp import callers, serve
    from gmapdp.lib import CD_CALL_DTYPE
    z = np.load(os.path.join(HERE, "golden", GOLDS[which]), allow_pickle=False)
    g = {k: z[k] for k in z.files}
    n = len(g["qlen"])
    c = np.zeros(n, CD_CALL_DTYPE)
    c["qoffL"] = g["qoff"] + g["qdp5"]
    c["qoffR"] = g["qoff"] + g["qdp3"]
    c["roffsetL"], c["rev_roffsetR"] = g["qdp5"], g["qdp3"]
    for f in ("rlengthL", "rlengthR", "glength", "goffset", "chroffset", "chrhigh", "extraband", "watsonp",
              "genestrand", "jump_late", "defect_rate"):
        c[f] = g[f]
    seq, gc = golden_io.skip_genome_image()
    dg = serve.DeviceGenome(gc, len(seq))
    out, res = callers.cdna_gap(c, g["qbuf"], dg, max_rlength=500, max_glength=2000)
    starts = np.zeros(n + 1, np.int64)
    starts[1:] = np.cumsum(np.maximum(g["out_npairs"], 0))
    kinds = {"null": 0, "joined": 0, "gapholder": 0, "inserted": 0}
    for k in range(n):
        want_n = int(g["out_npairs"][k])
        idx, score, incomplete = (int(x) for x in g["out_counters"][k])
        assert int(res["advanced"][k]) == idx - 1, ("dynprogindex", k)
        assert int(res["traceback_score"][k]) == score, ("traceback_score", k, int(res["traceback_score"][k]), score)
        assert int(res["incomplete"][k]) == incomplete, ("incompletep", k)
        if want_n < 0:
            assert out[k] is None, ("a list where the reference returns NULL", k)
            kinds["null"] += 1
            continue
        assert out[k] is not None, ("NULL where the reference returns a list", k)
        want = g["out_pairs"][starts[k]:starts[k] + want_n]
        got = _records(out[k][0])
        assert got.shape == want.shape, ("length", k, got.shape, want.shape)
        bad = np.nonzero((got != want).any(1))[0]
        assert len(bad) == 0, ("pairs", k, bad[:4], got[bad[:2]].tolist(), want[bad[:2]].tolist())
        kinds["joined"] += 1
        kinds["gapholder" if incomplete else "inserted"] += 1
    # the golden set holds NULLs, joins around a gap holder and lists
    if which == "gaps":
        assert kinds["null"] > 0 and kinds["gapholder"] > 100, kinds
    else:
        assert kinds["inserted"] > 100 and kinds["gapholder"] > 50, kinds
