"""The HIP traceback's genome skips against the reference's own walks
(tests/golden/genome_skip_golden.npz: Dynprog_traceback_{8,16}(_upper) with a
Genome_T over a genomecomp image, pairpool.c:1016-1180), with the device
genome in both formats: the genomecomp image itself and ASCII bases."""
import numpy as np
import pytest

import golden_io
import helpers as H

pytestmark = pytest.mark.gpu


def test_gpu_genome_skips_match_golden():
    from gmapdp.batch import DeviceBatch
    seq, gc, cases = golden_io.load_genome_skips()
    specs = [c[0] for c in cases]
    probs, qbuf, gbuf = H.pack_batch(specs)
    b = DeviceBatch(probs, qbuf, gbuf)
    b.fill()
    reqs = np.array([H.tb_request(i, s, s.rl, s.gl, 3, t[0], t[1], t[2], t[3]) for i, (s, t, _, _) in
                     enumerate(cases)])
    for fmt, data, glen in ((1, gc, len(seq)), (0, np.frombuffer(seq, np.uint8), None)):
        pairs, counts = b.traceback(reqs, genome=data, genome_format=fmt, genome_len=glen)
        for i, (s, t, want, wcnt) in enumerate(cases):
            assert counts[i, 0] == len(want), (fmt, i)
            assert (H.pairs_to_records(pairs[i]) == want).all(), (fmt, i, s.kind, s.width, t)
            assert (counts[i, 1:] == wcnt).all(), (fmt, i)
