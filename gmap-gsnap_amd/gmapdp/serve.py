"""Device genome (GMAP's genomecomp image) and the serving loop of
libgmapdp.so (gdp_genomecomp_*, gdp_genome_segments, gdp_server_*).

torch is used only as the device allocator; extraction, fills and
tracebacks run in the library's HIP kernels and the request loop in its C++
server thread.
"""
import ctypes

import numpy as np

from .batch import _check, _torch, init
from .lib import (ALIGN_REQ_DTYPE, ALIGN_RES_DTYPE, GENOME_COMP, PAIR_DTYPE, SEGMENT_DTYPE, SERVER_CFG_DTYPE,
                  SERVER_STATS_DTYPE, load)

READ_STRIDE = 1008  # bytes per read slot written by gdp_gen_serve_requests


def genomecomp_pack(seq):
    """Packs ASCII bases into a genomecomp image (gdp_genomecomp_pack)."""
    L = load()
    seq = np.frombuffer(bytes(seq), np.uint8) if not isinstance(seq, np.ndarray) else np.ascontiguousarray(seq)
    out = np.zeros(int(L.gdp_genomecomp_words(len(seq))), np.uint32)
    _check(L, L.gdp_genomecomp_pack(seq.ctypes.data, len(seq), out.ctypes.data), "gdp_genomecomp_pack")
    return out


def _bind_gen(L):
    if getattr(L, "_serve_gen_bound", False):
        return L
    L.gdp_gen_genomecomp.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
    L.gdp_gen_genomecomp.restype = None
    L.gdp_gen_serve_requests.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    L.gdp_gen_serve_requests.restype = ctypes.c_int
    L.gdp_serve_drive.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                  ctypes.c_uint64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    L.gdp_serve_drive.restype = ctypes.c_int64
    L._serve_gen_bound = True
    return L


def gen_genomecomp(seed, nbases):
    """Random genome of nbases bases as a genomecomp image (gdp_gen_genomecomp)."""
    L = _bind_gen(load())
    out = np.zeros(int(L.gdp_genomecomp_words(nbases)), np.uint32)
    L.gdp_gen_genomecomp(seed, nbases, out.ctypes.data)
    return out


def gen_requests(seed, first, count, gc, nbases):
    """Serving requests with the config-3 mix over a genomecomp genome.
    Returns (requests, reads); requests['read'] points into reads."""
    L = _bind_gen(load())
    reqs = np.zeros(count, ALIGN_REQ_DTYPE)
    reads = np.zeros(count * READ_STRIDE + 16, np.uint8)
    rc = L.gdp_gen_serve_requests(seed, first, count, gc.ctypes.data, nbases, reqs.ctypes.data, reads.ctypes.data)
    if rc != 0:
        raise RuntimeError("gdp_gen_serve_requests failed")
    return reqs, reads


def decode(gc, nbases, start=0, length=None):
    """Host decode of genomecomp bases (uncompress_one_char with flag 'N')."""
    length = nbases - start if length is None else length
    pos = np.arange(start, start + length, dtype=np.int64)
    blk = gc[(pos // 32) * 3 + 0], gc[(pos // 32) * 3 + 1], gc[(pos // 32) * 3 + 2]
    k = (pos % 32).astype(np.uint32)
    word = np.where(k < 16, blk[1], blk[0])
    code = (word >> (2 * (k & 15))) & 3
    ch = np.frombuffer(b"ACGT", np.uint8)[code]
    flag = (blk[2] >> k) & 1
    ch = np.where((flag == 1) | (pos >= nbases), ord("N"), ch).astype(np.uint8)
    return ch


class DeviceGenome:
    """A genome image resident in device memory (format GENOME_COMP or GENOME_ASCII)."""

    def __init__(self, data, nbases, fmt=GENOME_COMP, device="cuda"):
        torch = _torch()
        arr = np.ascontiguousarray(data)
        if not arr.flags.writeable:
            arr = arr.copy()
        self.t = torch.from_numpy(arr.view(np.uint8)).to(device)
        self.nbases, self.fmt = nbases, fmt

    @property
    def ptr(self):
        return self.t.data_ptr()

    def segments(self, segs):
        """Extracts segments (SEGMENT_DTYPE) on the device; returns bytes per segment."""
        torch = _torch()
        L = init()
        segs = np.ascontiguousarray(segs, SEGMENT_DTYPE)
        total = int((segs["out_off"] + np.maximum(segs["length"], 0)).max()) if len(segs) else 0
        d_segs = torch.from_numpy(segs.view(np.uint8)).to(self.t.device)
        d_out = torch.zeros(total + 16, dtype=torch.uint8, device=self.t.device)
        _check(L, L.gdp_genome_segments(self.ptr, self.fmt, self.nbases, d_segs.data_ptr(), len(segs),
                                        d_out.data_ptr(), None), "gdp_genome_segments")
        torch.cuda.synchronize()
        out = d_out.cpu().numpy()
        return [out[o:o + n].tobytes() for o, n in zip(segs["out_off"], segs["length"])]


class Server:
    """gdp_server over a DeviceGenome."""

    def __init__(self, genome, dir_budget_bytes=1 << 30, max_batch=8192, slots=3, max_rl=1000, max_gl=1500,
                 pair_capacity=None, plan=None, result_buffers=0, long_rows=0):
        L = init()
        self.L, self.genome = L, genome
        cfg = np.zeros(1, SERVER_CFG_DTYPE)
        cfg["dir_budget_bytes"] = dir_budget_bytes
        cfg["max_batch"], cfg["slots"], cfg["max_rl"], cfg["max_gl"] = max_batch, slots, max_rl, max_gl
        cfg["pair_capacity"] = pair_capacity or max_batch * (max_rl + max_gl + 2) // 2
        cfg["result_buffers"] = result_buffers
        cfg["long_rows"] = long_rows
        if plan is not None:  # (g16_max_span, g64_min_rows[, mix_span[, mix_order]])
            cfg["plan"]["g16_max_span"], cfg["plan"]["g64_min_rows"] = plan[0], plan[1]
            if len(plan) > 2:
                cfg["plan"]["mix_span"] = plan[2]
            if len(plan) > 3:
                cfg["plan"]["mix_order"] = plan[3]
        self.h = L.gdp_server_create(cfg.ctypes.data, genome.ptr, genome.fmt, genome.nbases)
        if not self.h:
            raise RuntimeError("gdp_server_create failed: %s" % L.gdp_last_error().decode())

    def submit(self, reqs):
        reqs = np.ascontiguousarray(reqs, ALIGN_REQ_DTYPE)
        _check(self.L, self.L.gdp_server_submit(self.h, reqs.ctypes.data, len(reqs)), "gdp_server_submit")

    def poll(self, max_results=4096, wait_us=1000):
        """Returns [(result, pairs)] of finished requests; the pairs are
        copied out of the server's pinned buffers before the next poll."""
        res = np.zeros(max_results, ALIGN_RES_DTYPE)
        n = self.L.gdp_server_poll(self.h, res.ctypes.data, max_results, wait_us)
        if n < 0:
            raise RuntimeError("gdp_server_poll failed: %s" % self.L.gdp_last_error().decode())
        out = []
        for r in res[:n]:
            m = int(r["counts"][0])
            if m > 0:
                buf = (ctypes.c_char * (m * PAIR_DTYPE.itemsize)).from_address(int(r["pairs"]))
                pairs = np.frombuffer(bytes(buf), PAIR_DTYPE)
            else:
                pairs = np.zeros(0, PAIR_DTYPE)
            out.append((r.copy(), pairs))
        return out

    def collect(self, count, timeout_s=60.0):
        """Polls until `count` results arrived; returns {tag: (result, pairs)}."""
        import time
        out = {}
        t_end = time.time() + timeout_s
        while len(out) < count:
            if time.time() > t_end:
                raise TimeoutError("gdp_server: %d of %d results after %.0f s" % (len(out), count, timeout_s))
            for r, pairs in self.poll():
                out[int(r["tag"])] = (r, pairs)
        return out

    def stats(self):
        st = np.zeros(1, SERVER_STATS_DTYPE)
        self.L.gdp_server_get_stats(self.h, st.ctypes.data)
        return st[0]

    def drive(self, reqs, rate, seed=5):
        """Open-loop Poisson load (gdp_serve_drive); returns (latency_s per request, elapsed_s, ok)."""
        L = _bind_gen(self.L)
        reqs = np.ascontiguousarray(reqs, ALIGN_REQ_DTYPE)
        lat = np.full(len(reqs), np.nan)
        el = ctypes.c_double()
        ok = L.gdp_serve_drive(self.h, reqs.ctypes.data, len(reqs), float(rate), seed, lat.ctypes.data,
                               ctypes.byref(el))
        return lat, el.value, int(ok)

    def close(self):
        if self.h:
            self.L.gdp_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
