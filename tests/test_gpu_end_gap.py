"""Dynprog_end5_gap / Dynprog_end3_gap batched on the device (gdp_eg_run,
gmapdp.callers.end_gap) against the reference's own callers: the 1000 end
calls of tests/golden/callers_golden.npz (make_callers_golden.py ran
Dynprog_end5_gap, dynprog_end.c:1293-1650, and Dynprog_end3_gap,
1923-2243, over the reference's SIMD fills on the genome-skip test
genome).  Both ends, all four end modes (QUERYEND_GAP, QUERYEND_INDELS,
QUERYEND_NOGAPS, BEST_LOCAL), windows past max_glength chopped by the
caller, and both strands.  The returned lists and counters must match;
the early NULLs leave the caller's dynprogindex."""
import os

import numpy as np
import pytest

import golden_io
from test_gpu_single_gap import _records

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "callers_golden.npz")


def _golden():
    z = np.load(GOLD, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    starts = np.zeros(len(g["single"]) + 1, np.int64)
    starts[1:] = np.cumsum(np.maximum(g["out_npairs"], 0))
    return g, starts


def _calls(g, idx):
    from gmapdp.lib import EG_CALL_DTYPE
    c = np.zeros(len(idx), EG_CALL_DTYPE)
    c["qoff"] = g["roff"][idx]
    for f in ("rl", "gl", "roffset", "goffset", "chroffset", "chrhigh", "extraband", "watsonp", "genestrand",
              "jump_late", "endalign", "defect_rate"):
        c[f] = g[f][idx]
    return c


@pytest.mark.parametrize("five", [1, 0])
def test_end_gap_callers_match_the_reference(five):
    from gmapdp import callers, serve
    g, starts = _golden()
    idx = np.nonzero((g["single"] == 0) & (g["five"] == five))[0]
    seq, gc = golden_io.skip_genome_image()
    dg = serve.DeviceGenome(gc, len(seq))
    out, res = callers.end_gap(five, _calls(g, idx), g["rbuf"].tobytes(), dg, max_rlength=500, max_glength=2000)
    modes = set()
    for j, k in enumerate(idx):
        want_n = int(g["out_npairs"][k])
        cnt = g["out_counters"][k]
        assert list(res["counters"][j]) == list(cnt[1:6]), ("counters", k, list(res["counters"][j]), list(cnt[1:6]))
        assert int(res["advanced"][j]) == int(cnt[0]) - 1, ("dynprogindex", k)
        if want_n < 0:
            assert out[j] is None, ("a list where the reference returns NULL", k)
            continue
        assert out[j] is not None, ("NULL where the reference returns a list", k)
        want = g["out_pairs"][starts[k]:starts[k] + want_n]
        got = _records(out[j][0])
        assert got.shape == want.shape and (got == want).all(), ("pairs", k, int(g["endalign"][k]))
        modes.add(int(g["endalign"][k]))
    assert modes == {0, 1, 2, 3}


def test_end_gap_early_nulls_and_require_pos_score():
    from gmapdp import callers, serve
    g, starts = _golden()
    idx = np.nonzero((g["single"] == 0) & (g["five"] == 1))[0][:24]
    seq, gc = golden_io.skip_genome_image()
    dg = serve.DeviceGenome(gc, len(seq))
    calls = _calls(g, idx)
    calls["rl"][0] = 0
    calls["gl"][1] = 0
    calls["goffset"][2] = -1   # no room to extend 5'
    out, res = callers.end_gap(1, calls, g["rbuf"].tobytes(), dg)
    for j in range(3):
        assert out[j] is None and res["advanced"][j] == 0 and (res["counters"][j] == 0).all(), j
    assert all(res["advanced"][3:] == 1)
    # require_pos_score_p: NULL with every counter 0 outside QUERYEND_NOGAPS
    out, res = callers.end_gap(1, calls[3:], g["rbuf"].tobytes(), dg, require_pos_score=True)
    for j in range(len(out)):
        if calls["endalign"][3 + j] != 2:
            assert out[j] is None and (res["counters"][j] == 0).all(), j
