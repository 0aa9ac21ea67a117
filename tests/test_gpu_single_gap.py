"""Dynprog_single_gap batched on the device (gdp_sg_run, gmapdp.callers)
against the reference's own caller: the 2000 single-gap calls of
tests/golden/callers_golden.npz (make_callers_golden.py ran
Dynprog_single_gap, dynprog_single.c:428-675, over the reference's SIMD
fills, on the genome-skip test genome cut into two chromosomes).  HIGHQ
int16 and LOWQ int8 fills, about 2000-column bands, both strands, and 168
windows as long as their read, which single_gap_simple takes when it finds
at most one mismatch.  The whole returned lists (pair records from the
head, gap holders included) and counters must match; lengths past the
workspace give NULL with NEG_INFINITY_32."""
import os

import numpy as np
import pytest

import golden_io

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "callers_golden.npz")


def _golden_single():
    z = np.load(GOLD, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    idx = np.nonzero(g["single"] == 1)[0]
    starts = np.zeros(len(g["single"]) + 1, np.int64)
    starts[1:] = np.cumsum(np.maximum(g["out_npairs"], 0))
    return g, idx, starts


def _calls(g, idx):
    from gmapdp.lib import SG_CALL_DTYPE
    c = np.zeros(len(idx), SG_CALL_DTYPE)
    c["qoff"] = g["roff"][idx]
    for f in ("rl", "gl", "roffset", "goffset", "chroffset", "chrhigh", "extraband", "watsonp", "genestrand",
              "widebandp", "defect_rate"):
        c[f] = g[f][idx]
    c["jump_late"] = g["jump_late"][idx]
    return c


def _records(pairs):
    """PAIR_DTYPE records as the golden's 8 ints {querypos, genomepos, cdna,
    comp, genome, genomealt, gapp, genomejump}."""
    out = np.zeros((len(pairs), 8), np.int32)
    for j, f in enumerate(("querypos", "genomepos", "cdna", "comp", "genome", "genomealt", "gapp", "genomejump")):
        out[:, j] = pairs[f]
    return out


def test_single_gap_callers_match_the_reference():
    from gmapdp import callers, serve
    g, idx, starts = _golden_single()
    seq, gc = golden_io.skip_genome_image()
    dg = serve.DeviceGenome(gc, len(seq))
    calls = _calls(g, idx)
    out, res = callers.single_gap(calls, g["rbuf"].tobytes(), dg, max_rlength=500, max_glength=2000)
    nsimple = 0
    for j, k in enumerate(idx):
        want_n = int(g["out_npairs"][k])
        cnt = g["out_counters"][k]
        if want_n < 0:
            assert out[j] is None, k
            continue
        assert out[j] is not None, ("NULL where the reference returns a list", k)
        pairs, counters, simple = out[j]
        nsimple += simple
        assert list(counters) == list(cnt[1:6]), ("counters", k, list(counters), list(cnt[1:6]))
        want = g["out_pairs"][starts[k]:starts[k] + want_n]
        got = _records(pairs)
        assert got.shape == want.shape and (got == want).all(), ("pairs", k)
    assert nsimple > 50, "single_gap_simple never taken"
    assert (calls["gl"] == calls["rl"]).sum() >= nsimple


def test_single_gap_lengths_past_the_workspace():
    from gmapdp import callers, serve
    g, idx, starts = _golden_single()
    seq, gc = golden_io.skip_genome_image()
    dg = serve.DeviceGenome(gc, len(seq))
    calls = _calls(g, idx[:40])
    calls["rl"][::4] = 0          # empty read
    calls["gl"][1::4] = 2001      # window past max_glength
    out, res = callers.single_gap(calls, g["rbuf"].tobytes(), dg, max_rlength=500, max_glength=2000)
    for j in range(len(calls)):
        if j % 4 in (0, 1):
            assert out[j] is None and res["counters"][j][0] == -32768, j
        else:
            assert out[j] is not None, j
