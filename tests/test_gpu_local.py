"""Local end walks on the GPU (GDP_TB_LOCAL through gdp_traceback_ends:
traceback_local_{8,16}_{upper,lower}, dynprog_end.c:727-1135) against the
oracle's restatement, which tests/test_oracle_end.py pins against the
reference's own functions: pairs, counters and the cell where each walk
stopped."""
import numpy as np
import pytest

import helpers as H
from test_oracle_end import _local_cases

pytestmark = pytest.mark.gpu

GDP_TB_LOCAL = 4


def test_local_walks(oracle):
    from gmapdp.batch import DeviceBatch
    rng = np.random.default_rng(34)
    specs, walks = [], []
    for t, kind, width, q, g, band, revp, late, r, c, endc in _local_cases(rng, 1200):
        lb, ub = (0, band) if kind == 1 else (band, 0)
        if kind == 1 and not (r <= c <= r + ub):
            continue
        if kind == 2 and not (c <= r <= c + lb):
            continue
        walks.append((len(specs), r, c, endc))
        specs.append(H.Spec(kind, width, q, g, g, 3, -10, -2, lb, ub, late, revp))
    assert len(walks) > 300
    probs, qbuf, gbuf = H.pack_batch(specs)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    reqs = []
    for i, r, c, endc in walks:
        q = H.tb_request(i, specs[i], r, c, qoff=3, goff=11)
        q["flags"], q["endc"] = GDP_TB_LOCAL, endc
        reqs.append(q)
    pairs, counts, ends = b.traceback(np.array(reqs), ends=True)
    for j, (i, r, c, endc) in enumerate(walks):
        s = specs[i]
        _, D = H.oracle_fill(oracle, s)
        D = np.where(H.computed_mask(s, b.layout[i]), D, 0).astype(np.uint8)
        po, co, eo = oracle.traceback_local(s.kind, D, r, c, endc, s.q, s.q.upper(), s.g, s.galt, 3, 11, s.revp)
        assert (H.pairs_to_records(pairs[j]) == po).all(), ("local walk pairs", j, s.kind, r, c, endc)
        assert (counts[j, 1:] == co).all() and tuple(ends[j]) == eo, ("local walk counters", j)
    # a negative stop column is refused (column -1 is no cell)
    i, r, c, _ = walks[0]
    q = H.tb_request(i, specs[i], r, c)
    q["flags"], q["endc"] = GDP_TB_LOCAL, -1
    _, cneg = b.traceback(np.array([q]))
    assert cneg[0, 0] == -1
    # a full-band fill refuses the flag
    full = [H.Spec(0, 16, s.q, s.g, s.g, 3, -10, -2, 3, 3, False, False) for s in specs[:1]]
    probs, qbuf, gbuf = H.pack_batch(full)
    bf = DeviceBatch(probs, qbuf, gbuf)
    bf.fill()
    q = H.tb_request(0, full[0], full[0].rl, full[0].gl)
    q["flags"] = GDP_TB_LOCAL
    _, cf = bf.traceback(np.array([q]))
    assert cf[0, 0] == -1


@pytest.mark.parametrize("five", [True, False])
@pytest.mark.parametrize("long_jumps", [False, True])
def test_splicejunction(oracle, five, long_jumps):
    """Dynprog_end5/3_splicejunction composed from the batched kernels
    (gmapdp.ends) against the oracle's composition, which
    tests/test_oracle_end.py pins against the reference's own callers; with
    long_jumps the known gap holders span 32768-500000 bases."""
    from gmapdp.ends import end3_splicejunction, end5_splicejunction
    rng = np.random.default_rng((36 if five else 38) + (100 if long_jumps else 0))
    ends = H.random_splice_ends(rng, 400, long_jumps)
    # an empty read end and one past the workspace's 1000 rows: NULL lists
    # (dynprog_end.c:1702-1717), as is every end whose best cell scores below 0
    edge = [dict(ends[0], rev_rseq=b""), dict(ends[1], rev_rseq=H.rand_seq(rng, 1001), rev_gseq=H.rand_seq(rng, 1010),
                                              rev_galt=H.rand_seq(rng, 1010))]
    got = (end5_splicejunction if five else end3_splicejunction)(ends + edge)
    assert got[-2] is None and got[-1] is None
    walked = 0
    for k, e in enumerate(ends):
        want = H.oracle_end5_splicejunction(oracle, e, five)
        assert (got[k] is None) == (want is None), ("null", k)
        if want is None:
            continue
        assert (H.pairs_to_records(got[k][0]) == want[0]).all() and (got[k][1] == want[1]).all(), ("end", k)
        walked += 1
    assert walked > 200
