#!/usr/bin/env python3
"""Writes tests/golden/chr17test.fa.gz and tests/golden/genomecomp_golden.json
from the reference's own genome-packing test data: tests/ss.chr17test (the
FASTA its setup test feeds to gmap_build) and tests/setup.genomecomp.ok (the
.genomecomp image gmap_build must write for it).  The image is kept as its
SHA-256, size and a sample of words; the FASTA is kept whole (gzip).

Run where the reference tree is mounted: python3 tests/golden/make_genomecomp_golden.py
"""
import gzip
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/tests"

fa = open(os.path.join(REF, "ss.chr17test"), "rb").read()
ok = open(os.path.join(REF, "setup.genomecomp.ok"), "rb").read()
with gzip.GzipFile(os.path.join(HERE, "chr17test.fa.gz"), "wb", mtime=0) as f:
    f.write(fa)
words = np.frombuffer(ok, "<u4")
rng = np.random.default_rng(17)
idx = sorted(set(range(0, 48)) | set(range(len(words) - 48, len(words))) |
             set(int(i) for i in rng.integers(0, len(words), 256)))
json.dump({"sha256": hashlib.sha256(ok).hexdigest(), "bytes": len(ok),
           "sample": {str(i): int(words[i]) for i in idx}},
          open(os.path.join(HERE, "genomecomp_golden.json"), "w"), indent=0)
