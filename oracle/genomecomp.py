"""CPU restatement of GMAP's genomecomp genome image (test infrastructure:
only the tests and the benchmark's CPU leg use this module; the product
packs and reads genomes in gmap-gsnap_amd/csrc).

pack   -- Compress_compress (compress-write.c:535-650): 32 bases per block
          of three words {high, low, flags}; each base is shifted in at the
          top of `high`, so after 32 bases base k sits at bits 2k of `low`
          (k < 16) or 2(k-16) of `high` and at flag bit k.  A/C/G/T (either
          case) give codes 0-3, N a flag (code 0), X code 3 with a flag,
          anything else a flag (code 0).  The last block is padded with X
          and two 0xFFFFFFFF words close the image.
decode -- uncompress_one_char (genome.c:9723-9760) with flag character
          'N', returned as (codes, nflags) per base, the genome model of
          pyoracle's traceback and segment functions.
"""
import numpy as np


def pack(seq):
    a = np.frombuffer(bytes(seq), np.uint8)
    n = len(a)
    nb = (n + 31) // 32
    padded = np.full(nb * 32, ord("X"), np.uint8)
    padded[:n] = a
    low5 = padded | 0x20
    code = np.zeros(nb * 32, np.uint32)
    code[low5 == ord("c")] = 1
    code[low5 == ord("g")] = 2
    code[low5 == ord("t")] = 3
    code[low5 == ord("x")] = 3
    acgt = np.isin(low5, np.frombuffer(b"acgt", np.uint8))
    flag = (~acgt).astype(np.uint32)
    code = code.reshape(nb, 32)
    flag = flag.reshape(nb, 32)
    sh = (2 * np.arange(16, dtype=np.uint32))[None, :]
    low = np.bitwise_or.reduce(code[:, :16] << sh, axis=1) if nb else np.zeros(0, np.uint32)
    high = np.bitwise_or.reduce(code[:, 16:] << sh, axis=1) if nb else np.zeros(0, np.uint32)
    flags = np.bitwise_or.reduce(flag << np.arange(32, dtype=np.uint32)[None, :], axis=1) if nb else \
        np.zeros(0, np.uint32)
    out = np.empty(3 * nb + 2, np.uint32)
    out[0:3 * nb:3] = high
    out[1:3 * nb:3] = low
    out[2:3 * nb:3] = flags
    out[3 * nb:] = 0xFFFFFFFF
    return out


def decode(gc, nbases):
    """(codes, nflags) of bases 0..nbases-1 of a genomecomp image."""
    pos = np.arange(nbases, dtype=np.int64)
    blk = (pos // 32) * 3
    k = (pos % 32).astype(np.uint32)
    word = np.where(k < 16, gc[blk + 1], gc[blk])
    codes = ((word >> (2 * (k & 15))) & 3).astype(np.uint8)
    nflags = ((gc[blk + 2] >> k) & 1).astype(np.uint8)
    return codes, nflags


def decode_window(gc, nbases, start, count):
    """The oracle's genome model of positions start .. start + count - 1 of
    a genome of nbases bases (clipped to the genome): (codes, nflags, base,
    nbases) for pyoracle, without decoding the rest."""
    start = max(0, min(int(start), nbases))
    end = max(start, min(int(start) + int(count), nbases))
    pos = np.arange(start, end, dtype=np.int64)
    blk = (pos // 32) * 3
    k = (pos % 32).astype(np.uint32)
    word = np.where(k < 16, gc[blk + 1], gc[blk])
    codes = ((word >> (2 * (k & 15))) & 3).astype(np.uint8)
    nflags = ((gc[blk + 2] >> k) & 1).astype(np.uint8)
    return codes, nflags, start, nbases


def to_ascii(codes, nflags):
    ch = np.frombuffer(b"ACGT", np.uint8)[codes]
    return np.where(nflags != 0, ord("N"), ch).astype(np.uint8)
