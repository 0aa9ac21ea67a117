"""The HIP path against the reference's golden vectors (tests/golden)."""
import numpy as np
import pytest

import golden_io
import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("keep", [True, False])
def test_gpu_matches_reference_golden_vectors(keep):
    # keep=False: full-band int16 problems take the paired sweep
    from gmapdp.batch import DeviceBatch
    g, specs = golden_io.load()
    probs, qbuf, gbuf = H.pack_batch(specs, keep_scores=keep)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    fin = b.final_scores()
    reqs, idx = [], []
    for i, s in enumerate(specs):
        D, S = b.export(i)
        golden_io.check_fill(g, i, s, D, S, scores=keep)
        if g["final"][i] != -(1 << 31):
            assert fin[i] == g["final"][i], ("corner score", i)
        exp = golden_io.traceback_expect(g, i)
        if exp is not None:
            reqs.append(H.tb_request(i, s, exp[0], exp[1]))
            idx.append(i)
    pairs, counts = b.traceback(np.array(reqs))
    for k, i in enumerate(idx):
        r, c, want_pairs, want_counts = golden_io.traceback_expect(g, i)
        got = H.pairs_to_records(pairs[k])
        assert got.shape == want_pairs.shape and (got == want_pairs).all(), ("pairs", i)
        assert (counts[k, 1:] == want_counts).all(), ("counters", i)
    assert len(idx) > 300
