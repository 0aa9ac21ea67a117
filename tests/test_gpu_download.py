"""gdp_download_results: each request's counts and only the pair records its
walk wrote reach pinned host memory at the request's offsets, equal to the
device buffers; the records past a walk's count are left as they were."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def test_download_results_matches_the_device_buffers():
    import torch
    from gmapdp.batch import DeviceBatch
    from gmapdp.lib import PAIR_DTYPE
    specs = H.random_specs(77, 120, kinds=(0,))
    probs, qbuf, gbuf = H.pack_batch(specs)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    reqs = np.array([H.tb_request(k, s, s.rl, s.gl) for k, s in enumerate(specs)])
    off = np.concatenate([[0], np.cumsum(reqs["max_pairs"].astype(np.int64))[:-1]])
    reqs["pair_off"] = off
    h = b.prepare_traceback(reqs)
    b.traceback_device(h)
    torch.cuda.synchronize()
    want_cnt = h["d_counts"].cpu().numpy().reshape(-1, 6)
    want_pairs = h["d_pairs"].cpu().numpy().view(PAIR_DTYPE)
    h_cnt = torch.full((len(reqs) * 6,), -7, dtype=torch.int32, pin_memory=True)
    h_pairs = torch.full((h["d_pairs"].numel(),), 0x5a, dtype=torch.uint8, pin_memory=True)
    rc = b.L.gdp_download_results(h["d_reqs"].data_ptr(), len(reqs), h["d_counts"].data_ptr(),
                                  h["d_pairs"].data_ptr(), h_cnt.data_ptr(), h_pairs.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
    assert rc == 0, b.L.gdp_last_error()
    torch.cuda.synchronize()
    got_cnt = h_cnt.numpy().reshape(-1, 6)
    got_pairs = h_pairs.numpy().view(PAIR_DTYPE)
    assert (got_cnt == want_cnt).all()
    untouched = np.ones(len(got_pairs), bool)
    walked = 0
    for k in range(len(reqs)):
        n = int(want_cnt[k, 0])
        if n <= 0:
            continue
        walked += 1
        o = int(off[k])
        assert (got_pairs[o:o + n] == want_pairs[o:o + n]).all(), k
        untouched[o:o + n] = False
    assert walked > 100
    assert (h_pairs.numpy().reshape(-1, 16)[untouched] == 0x5a).all()
    # unpinned host memory is refused
    plain = np.zeros(16, np.int32)
    rc = b.L.gdp_download_results(h["d_reqs"].data_ptr(), 1, h["d_counts"].data_ptr(), h["d_pairs"].data_ptr(),
                                  plain.ctypes.data, h_pairs.data_ptr(), None)
    assert rc != 0
    # and the refusal leaves no HIP error behind for the next launch to report
    rc = b.L.gdp_download_results(h["d_reqs"].data_ptr(), len(reqs), h["d_counts"].data_ptr(),
                                  h["d_pairs"].data_ptr(), h_cnt.data_ptr(), h_pairs.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
    assert rc == 0, b.L.gdp_last_error()
    b.fill()
    torch.cuda.synchronize()
