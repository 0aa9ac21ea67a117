"""Device-resident batches of DP problems driven through libgmapdp.so.

torch is used only as the device allocator and stream provider; every
computation runs in the library's HIP kernels.
"""
import ctypes

import numpy as np

from .lib import INTRON_GAP_DTYPE, PAIR_DTYPE, PLAN_PARAMS_DTYPE, PROBLEM_DTYPE, TBREQ_DTYPE, load

_initialized_mode = None


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("gmapdp needs a HIP device (torch.cuda.is_available() is False)")
    return torch


def init(mode=0):
    """Builds and uploads the score tables once per process (gdp_init)."""
    global _initialized_mode
    L = load()
    if _initialized_mode != mode:
        _torch()
        rc = L.gdp_init(mode)
        if rc != 0:
            raise RuntimeError("gdp_init failed: %s" % L.gdp_last_error().decode())
        _initialized_mode = mode
    return L


def _check(L, rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, L.gdp_last_error().decode()))


PLAN_SINGLE = 1  # GDP_PLAN_SINGLE: every problem swept on its own


def plan_params(g16_max_span=240, g64_min_rows=8001, mix_span=0, mix_order=0):
    """gdp_plan_params: lane-group choice of the paired sweep, the band
    spread allowed in mixed waves (0: one band per wave) and how mixed waves
    are formed (1: binned by band, in order of rows within a bin)."""
    p = np.zeros(1, PLAN_PARAMS_DTYPE)
    p["g16_max_span"], p["g64_min_rows"], p["mix_span"], p["mix_order"] = g16_max_span, g64_min_rows, mix_span, mix_order
    return p


def plan(problems, flags=0, params=None):
    """Host planning (no device needed): descriptors and buffer sizes.
    params: (g16_max_span, g64_min_rows[, mix_span]) or None for the defaults."""
    L = load()
    problems = np.ascontiguousarray(problems, PROBLEM_DTYPE)
    n = len(problems)
    descs = np.zeros(int(L.gdp_plan_bytes(n)), np.uint8)
    dw, ss = ctypes.c_int64(), ctypes.c_int64()
    pp = None if params is None else plan_params(*params)
    _check(L, L.gdp_plan_tuned(problems.ctypes.data, n, flags, None if pp is None else pp.ctypes.data,
                               descs.ctypes.data, ctypes.byref(dw), ctypes.byref(ss)), "gdp_plan")
    return descs, dw.value, ss.value


def cells(problems):
    L = load()
    problems = np.ascontiguousarray(problems, PROBLEM_DTYPE)
    return np.array([L.gdp_cells(problems[i:i + 1].ctypes.data) for i in range(len(problems))], np.int64)


class DeviceBatch:
    """Problems, sequences and outputs resident in device memory."""

    def __init__(self, problems, qbuf, gbuf, mode=0, device="cuda", single=False, params=None):
        """single=True plans every problem on its own sweep (required for
        fill_one, the one-launch-per-problem form); params = (g16_max_span,
        g64_min_rows) of the paired sweep (None: defaults)."""
        torch = _torch()
        self.torch = torch
        self.L = init(mode)
        self.problems = np.ascontiguousarray(problems, PROBLEM_DTYPE)
        self.n = len(self.problems)
        self.single = single
        descs, dw, ss = plan(self.problems, PLAN_SINGLE if single else 0, params)
        self.kernels = int(self.L.gdp_plan_kernels(descs.ctypes.data, self.n))
        # the sweep the plan chose per problem (gdp_dev_problem byte 58): 0
        # single, 1 paired, 2 narrow one-sided
        self.layout = descs[: self.n * 96].reshape(self.n, 96)[:, 58].copy() if self.n else np.zeros(0, np.uint8)
        self.dir_words, self.score_slots = dw, ss
        dev = torch.device(device)
        self.d_descs = torch.from_numpy(descs).to(dev)
        # 16 bytes of padding: the traceback reads sequences in aligned 16-byte windows
        self.d_q = torch.from_numpy(np.frombuffer(bytes(qbuf) + b"\0" * 16, np.uint8).copy()).to(dev)
        self.d_g = torch.from_numpy(np.frombuffer(bytes(gbuf) + b"\0" * 16, np.uint8).copy()).to(dev)
        self.d_dirs = torch.zeros(max(dw, 1), dtype=torch.int32, device=dev)
        self.d_scores = torch.zeros(max(ss, 1), dtype=torch.int16, device=dev)
        self.d_final = torch.zeros(max(self.n, 1), dtype=torch.int32, device=dev)
        self.dev = dev

    @property
    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def fill(self, stream=None):
        t = self
        _check(t.L, t.L.gdp_fill_ex(t.d_descs.data_ptr(), t.n, t.kernels, t.d_q.data_ptr(), t.d_g.data_ptr(),
                                    t.d_dirs.data_ptr(), t.d_scores.data_ptr(), t.d_final.data_ptr(),
                                    t.stream if stream is None else stream), "gdp_fill")

    def final_scores(self):
        return self.d_final[: self.n].cpu().numpy()

    def export(self, idx):
        """Dense directions and scores of one problem ([col][row] of its sweep)."""
        p = self.problems[idx]
        R, C = (int(p["gl"]), int(p["rl"])) if p["kind"] == 2 else (int(p["rl"]), int(p["gl"]))
        dd = self.torch.zeros((C + 1) * (R + 1), dtype=self.torch.uint8, device=self.dev)
        ds = self.torch.zeros((C + 1) * (R + 1), dtype=self.torch.int16, device=self.dev)
        _check(self.L, self.L.gdp_export_dense(self.d_descs.data_ptr(), idx, self.d_dirs.data_ptr(),
                                               self.d_scores.data_ptr(), dd.data_ptr(), ds.data_ptr(), self.stream),
               "gdp_export_dense")
        return (dd.cpu().numpy().reshape(C + 1, R + 1), ds.cpu().numpy().reshape(C + 1, R + 1))

    def traceback(self, reqs, genome=None, genome_format=0, genome_len=None, ends=False):
        """Runs the requests; returns (pairs per request, counts[n,6]), and
        with ends=True also the cells [n,2] where the walks stopped
        (gdp_traceback_ends).  genome: ASCII bases (genome_format 0) or a
        genomecomp image of genome_len bases (genome_format 1)."""
        torch = self.torch
        reqs = np.ascontiguousarray(reqs, TBREQ_DTYPE).copy()
        off = 0
        for k in range(len(reqs)):
            reqs["pair_off"][k] = off
            off += int(reqs["max_pairs"][k])
        d_reqs = torch.from_numpy(reqs.view(np.uint8)).to(self.dev)
        d_pairs = torch.zeros(max(off, 1) * PAIR_DTYPE.itemsize, dtype=torch.uint8, device=self.dev)
        d_counts = torch.zeros(max(len(reqs), 1) * 6, dtype=torch.int32, device=self.dev)
        gptr, glen = None, 0
        if genome is not None:
            if genome_format == 0:
                d_gen = torch.from_numpy(np.array(genome, np.uint8, copy=True)).to(self.dev)
                glen = len(genome)
            else:
                d_gen = torch.from_numpy(np.ascontiguousarray(genome, np.uint32).view(np.uint8)).to(self.dev)
                glen = genome_len
            gptr = d_gen.data_ptr()
        d_ends = torch.full((max(len(reqs), 1) * 2,), -1, dtype=torch.int32, device=self.dev) if ends else None
        _check(self.L, self.L.gdp_traceback_ends(self.d_descs.data_ptr(), d_reqs.data_ptr(), len(reqs),
                                                 self.d_q.data_ptr(), self.d_g.data_ptr(), self.d_dirs.data_ptr(),
                                                 gptr, genome_format, glen, d_pairs.data_ptr(), d_counts.data_ptr(),
                                                 d_ends.data_ptr() if ends else None, self.stream), "gdp_traceback")
        counts = d_counts.cpu().numpy()[: 6 * len(reqs)].reshape(-1, 6)
        pairs = d_pairs.cpu().numpy().view(PAIR_DTYPE)
        out = []
        for k in range(len(reqs)):
            n = min(int(counts[k, 0]), int(reqs["max_pairs"][k]))
            out.append(pairs[int(reqs["pair_off"][k]): int(reqs["pair_off"][k]) + n].copy())
        if ends:
            return out, counts, d_ends.cpu().numpy()[: 2 * len(reqs)].reshape(-1, 2)
        return out, counts

    def best_endpoint(self, ul, queryend_indels=False):
        torch = self.torch
        ul = np.ascontiguousarray(ul, np.int32).reshape(-1)
        m = len(ul) // 2
        d_ul = torch.from_numpy(ul).to(self.dev)
        d_out = torch.zeros(max(m, 1) * 3, dtype=torch.int32, device=self.dev)
        _check(self.L, self.L.gdp_best_endpoint(self.d_descs.data_ptr(), d_ul.data_ptr(), m, int(queryend_indels),
                                                self.d_scores.data_ptr(), d_out.data_ptr(), self.stream),
               "gdp_best_endpoint")
        return d_out.cpu().numpy()[: 3 * m].reshape(-1, 3)

    def bridge_cdna(self, quads, tb_template=None):
        """gdp_bridge_cdna over rows {left upper, left lower, right upper,
        right lower, rightoffset - leftoffset}: (m, 5) rows {bestscore,
        bestcL, bestcR, bestrL, bestrR}; with tb_template (2m requests) also
        the completed traceback requests (right end, then left end)."""
        torch = self.torch
        quads = np.ascontiguousarray(quads, np.int32).reshape(-1, 5)
        m = len(quads)
        d_q = torch.from_numpy(quads.reshape(-1).copy()).to(self.dev)
        d_out = torch.zeros(max(m, 1) * 5, dtype=torch.int32, device=self.dev)
        d_tmpl = d_tb = None
        if tb_template is not None:
            tmpl = np.ascontiguousarray(tb_template, TBREQ_DTYPE)
            assert len(tmpl) == 2 * m
            d_tmpl = torch.from_numpy(tmpl.view(np.uint8).copy()).to(self.dev)
            d_tb = torch.zeros_like(d_tmpl)
        _check(self.L, self.L.gdp_bridge_cdna(self.d_descs.data_ptr(), d_q.data_ptr(), m, self.d_scores.data_ptr(),
                                              d_out.data_ptr(), d_tmpl.data_ptr() if d_tmpl is not None else None,
                                              d_tb.data_ptr() if d_tb is not None else None, self.stream),
               "gdp_bridge_cdna")
        out = d_out.cpu().numpy()[: 5 * m].reshape(-1, 5)
        if tb_template is None:
            return out
        return out, d_tb.cpu().numpy().view(TBREQ_DTYPE)

    def bridge_intron(self, gaps, site_probs, tb_template=None):
        """gdp_bridge_intron over INTRON_GAP_DTYPE rows, with the per-column
        splice-site probabilities site_probs (float64, indexed by each gap's
        prob_off): ((m, 5) rows {finalscore, bestrL, bestrR, bestcL, bestcR},
        (m, 2) rows {left_prob, right_prob}); with tb_template (2m requests)
        also the completed traceback requests (right end, then left end)."""
        torch = self.torch
        gaps = np.ascontiguousarray(gaps, INTRON_GAP_DTYPE)
        m = len(gaps)
        d_gaps = torch.from_numpy(gaps.view(np.uint8).copy()).to(self.dev)
        probs = np.ascontiguousarray(site_probs, np.float64)
        d_p = torch.from_numpy(probs.copy() if probs.size else np.zeros(1)).to(self.dev)
        d_out = torch.zeros(max(m, 1) * 5, dtype=torch.int32, device=self.dev)
        d_outp = torch.zeros(max(m, 1) * 2, dtype=torch.float64, device=self.dev)
        d_tmpl = d_tb = None
        if tb_template is not None:
            tmpl = np.ascontiguousarray(tb_template, TBREQ_DTYPE)
            assert len(tmpl) == 2 * m
            d_tmpl = torch.from_numpy(tmpl.view(np.uint8).copy()).to(self.dev)
            d_tb = torch.zeros_like(d_tmpl)
        _check(self.L, self.L.gdp_bridge_intron(self.d_descs.data_ptr(), d_gaps.data_ptr(), m,
                                                self.d_scores.data_ptr(), self.d_g.data_ptr(), d_p.data_ptr(),
                                                d_out.data_ptr(), d_outp.data_ptr(),
                                                d_tmpl.data_ptr() if d_tmpl is not None else None,
                                                d_tb.data_ptr() if d_tb is not None else None, self.stream),
               "gdp_bridge_intron")
        out = d_out.cpu().numpy()[: 5 * m].reshape(-1, 5)
        outp = d_outp.cpu().numpy()[: 2 * m].reshape(-1, 2)
        if tb_template is None:
            return out, outp
        return out, outp, d_tb.cpu().numpy().view(TBREQ_DTYPE)

    def bridge_intron_requests(self, d_gaps, m, d_probs, d_out, d_outp, d_tmpl, d_tbout, stream=None):
        """gdp_bridge_intron on device buffers: joins into d_out / d_outp and
        both ends' traceback requests (templates at address d_tmpl completed)
        to address d_tbout."""
        _check(self.L, self.L.gdp_bridge_intron(self.d_descs.data_ptr(), d_gaps.data_ptr(), m, self.d_scores.data_ptr(),
                                                self.d_g.data_ptr(), d_probs.data_ptr(), d_out.data_ptr(),
                                                d_outp.data_ptr(), d_tmpl, d_tbout,
                                                self.stream if stream is None else stream), "gdp_bridge_intron")

    def bridge_requests(self, d_quads, m, d_out, d_tmpl, d_tbout, stream=None):
        """gdp_bridge_cdna on device buffers: joins into d_out and both ends'
        traceback requests (templates at address d_tmpl completed) to address
        d_tbout."""
        _check(self.L, self.L.gdp_bridge_cdna(self.d_descs.data_ptr(), d_quads.data_ptr(), m, self.d_scores.data_ptr(),
                                              d_out.data_ptr(), d_tmpl, d_tbout,
                                              self.stream if stream is None else stream), "gdp_bridge_cdna")

    # -- device-resident traceback (timed paths) ---------------------------
    def prepare_traceback(self, reqs, genome=None):
        """Uploads requests and allocates outputs once; returns a handle."""
        torch = self.torch
        reqs = np.ascontiguousarray(reqs, TBREQ_DTYPE)
        npairs = int((reqs["pair_off"].astype(np.int64) + reqs["max_pairs"]).max()) if len(reqs) else 1
        h = {"n": len(reqs), "reqs": reqs,
             "d_reqs": torch.from_numpy(reqs.view(np.uint8).copy()).to(self.dev),
             "d_pairs": torch.zeros(max(npairs, 1) * PAIR_DTYPE.itemsize, dtype=torch.uint8, device=self.dev),
             "d_counts": torch.zeros(max(len(reqs), 1) * 6, dtype=torch.int32, device=self.dev),
             "d_genome": None, "glen": 0}
        if genome is not None:
            h["d_genome"] = torch.from_numpy(np.ascontiguousarray(genome, np.uint8)).to(self.dev)
            h["glen"] = len(genome)
        return h

    def traceback_device(self, h, first=0, count=None, stream=None):
        count = h["n"] - first if count is None else count
        gptr = h["d_genome"].data_ptr() if h["d_genome"] is not None else None
        _check(self.L, self.L.gdp_traceback(self.d_descs.data_ptr(),
                                            h["d_reqs"].data_ptr() + first * TBREQ_DTYPE.itemsize, count,
                                            self.d_q.data_ptr(), self.d_g.data_ptr(), self.d_dirs.data_ptr(),
                                            gptr, h["glen"], h["d_pairs"].data_ptr(),
                                            h["d_counts"].data_ptr() + first * 24,
                                            self.stream if stream is None else stream), "gdp_traceback")

    def best_endpoint_requests(self, d_ul, m, queryend_indels, d_out, d_tmpl, d_tbout, stream=None):
        """gdp_best_endpoint_ex on device buffers: best cells into d_out and
        traceback requests (templates at address d_tmpl completed) to
        address d_tbout."""
        _check(self.L, self.L.gdp_best_endpoint_ex(self.d_descs.data_ptr(), d_ul.data_ptr(), m, int(queryend_indels),
                                                   self.d_scores.data_ptr(), d_out.data_ptr(), d_tmpl, d_tbout,
                                                   self.stream if stream is None else stream), "gdp_best_endpoint")

    def capture(self, fn):
        """Records fn(stream) into one HIP graph (gdp_graph_begin/end)."""
        g, st = ctypes.c_void_p(), ctypes.c_void_p()
        _check(self.L, self.L.gdp_graph_begin(ctypes.byref(g), ctypes.byref(st)), "gdp_graph_begin")
        try:
            fn(st)
        except Exception:
            self.L.gdp_graph_end(g)
            self.L.gdp_graph_destroy(g)
            raise
        rc = self.L.gdp_graph_end(g)
        if rc != 0:
            self.L.gdp_graph_destroy(g)
            _check(self.L, rc, "gdp_graph_end")
        return g

    def graph(self, h):
        """Captures fill + traceback (requests of prepare_traceback's handle h)
        into one HIP graph; returns a handle for graph_launch."""
        gptr = h["d_genome"].data_ptr() if h["d_genome"] is not None else None
        g = self.L.gdp_graph_create(self.d_descs.data_ptr(), self.n, self.kernels, self.d_q.data_ptr(),
                                    self.d_g.data_ptr(), self.d_dirs.data_ptr(), self.d_scores.data_ptr(),
                                    self.d_final.data_ptr(), h["d_reqs"].data_ptr(), h["n"], gptr, 0, h["glen"],
                                    h["d_pairs"].data_ptr(), h["d_counts"].data_ptr())
        if not g:
            raise RuntimeError("gdp_graph_create failed: %s" % self.L.gdp_last_error().decode())
        return g

    def graph_launch(self, g):
        _check(self.L, self.L.gdp_graph_launch(g, self.stream), "gdp_graph_launch")

    def graph_destroy(self, g):
        self.L.gdp_graph_destroy(g)

    def traceback_results(self, h):
        counts = h["d_counts"].cpu().numpy()[: 6 * h["n"]].reshape(-1, 6)
        return counts

    def fill_one(self, i):
        """One problem per launch (the per-call form)."""
        if not self.single:
            raise RuntimeError("fill_one needs a batch planned with single=True")
        desc = self.d_descs.data_ptr() + i * self.L.gdp_descriptor_bytes()
        _check(self.L, self.L.gdp_fill(desc, 1, self.d_q.data_ptr(), self.d_g.data_ptr(), self.d_dirs.data_ptr(),
                                       self.d_scores.data_ptr(), self.d_final.data_ptr() + 4 * i, self.stream),
               "gdp_fill")

    def traceback_one(self, h, i):
        self.traceback_device(h, i, 1)

    def align_each(self, h, stream=None):
        """gdp_align_each: every problem filled and walked on its own, with a
        stream synchronize after each (request k of handle h walks problem
        k).  Needs a batch planned with single=True."""
        if not self.single:
            raise RuntimeError("align_each needs a batch planned with single=True")
        if h["n"] != self.n:
            raise ValueError("align_each needs one request per problem")
        descs = self.d_descs.cpu().numpy()
        gptr = h["d_genome"].data_ptr() if h["d_genome"] is not None else None
        _check(self.L, self.L.gdp_align_each(descs.ctypes.data, self.d_descs.data_ptr(), self.n, self.d_q.data_ptr(),
                                             self.d_g.data_ptr(), self.d_dirs.data_ptr(), self.d_scores.data_ptr(),
                                             self.d_final.data_ptr(), h["d_reqs"].data_ptr(), gptr, 0, h["glen"],
                                             h["d_pairs"].data_ptr(), h["d_counts"].data_ptr(),
                                             self.stream if stream is None else stream), "gdp_align_each")
