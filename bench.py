#!/usr/bin/env python3
"""Throughput benchmark of the banded affine-gap DP core on one MI355X.

A step is one pass of the hot path over one batch: every problem of the
workload is filled (one batched launch) and traced back (one batched
launch).  Inputs are resident in HBM before the timed region.  The default
workload is the uneven 2x10^6-problem single-gap batch (rl log-uniform
20-1000, gl = rl x U(1, 1.5), extraband 3-64, int8 and int16), on which the
targets are quoted; --workload c2 selects the 10^6-problem batch of rl 150.

After the batch, the serving loop runs (unless --no-serve): a 3.1-Gbp random
genome resident on the device as a genomecomp image, alignment requests of
the uneven mix, first all at once (saturation throughput), then as Poisson
arrivals at fractions of that rate, with p50/p99 latency per rate.

Prints one JSON line (the fields the driver reads) with extra objects:
  roofline      -- the fill kernel against the packed-int16 VALU peak, its
                   average duration measured with HIP events on its stream;
  cpu_baseline  -- the reference's own AVX2 DP objects (oracle/_ref, built
                   from the reference sources) or, where absent, the CPU
                   restatement, timed on a bounded sample on the host cores
                   (fill + traceback, and fill alone);
  percall_gcups -- the one-launch-per-problem GPU form (gdp_align_each: a
                   C loop of one fill launch, one traceback launch and a
                   stream synchronize per problem) on the first problems;
  e2e_gcups     -- the step with its PCIe transfers: inputs uploaded from
                   pinned host memory, fill and traceback, counts and pair
                   records downloaded.
"""
import argparse
import concurrent.futures as cf
import ctypes
import json
import subprocess
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gmap-gsnap_amd"))

import numpy as np  # noqa: E402

# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2 (packed int16) x 2.4 GHz.
PEAK_INT16_TOPS = 78.6
OPS_PER_CELL = 13  # algorithmic int16 ops per in-band cell (E 4, F 4, H 1+4)
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c1": (1, 2_000, "single-gap 100-nt reads from 2%-substituted, 0.2%-indel genome spans, extraband 10, "
                     "int16 HIGHQ, 2000 problems (the CPU-runnable case)"),
    "c1lowq": (1, 2_000, "as c1 with LOWQ scoring: int8 fills, 2000 problems"),
    "c2": (2, 1_000_000, "single-gap rl=150 gl=U{160..200} extraband=U{5..20} int16, 1e6 problems"),
    "c3": (3, 2_000_000, "single-gap rl=logU[20,1000] gl=rl*U(1,1.5) extraband=U{3..64} mixed int8/int16, 2e6 problems"),
    "c4": (4, 2_000_000, "end extensions rl=U{10..150} gl=rl+U{0..20} ENDQ (upper+lower fill, best end cell, "
                         "traceback) mixed 1:1 with single-gap logU[20,1000] problems, 2e6 + 2e6"),
    "cdna": ("cdna", 200_000, "cDNA gaps gl=U{16..160}, both ends rl=gl+8, extraband 14, open/extend -10/-7 "
                              "(four one-sided fills keeping every score, the bridge, two end tracebacks), 2e5 gaps"),
    "intron": ("intron", 200_000, "intron gaps rl=U{40..250}, both segments gl=rl+8, extraband 14, open/extend -8/-3 "
                                  "(four one-sided fills keeping every score, the site-level intron bridge over "
                                  "per-column splice-site probabilities, two end tracebacks), 2e5 gaps"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, sample_n, threads):
    """Times the CPU DP on a bounded sample: fill + traceback, then the same
    sample filled only."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    probs = wl.problems[:sample_n]
    qbuf, gbuf = wl.qbuf, wl.gbuf
    cells = float(batch_cells(probs))
    if os.path.exists(pyoracle.REF_SO):
        L = ctypes.CDLL(pyoracle.REF_SO, mode=os.RTLD_LAZY)
        L.refh_init.argtypes = [ctypes.c_int]
        L.refh_init(0)
        L.refh_run_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
        L.refh_run_batch.restype = ctypes.c_longlong
        maxrl, maxgl = int(probs["rl"].max()) + 8, int(probs["gl"].max()) + 8
        run = lambda p, tb=1: L.refh_run_batch(p.ctypes.data, len(p), qbuf.ctypes.data, gbuf.ctypes.data, tb, maxrl,
                                               maxgl)
        kind = "reference"
    else:
        L = ctypes.CDLL(pyoracle.ORACLE_SO)
        L.orc_init.argtypes = [ctypes.c_int]
        L.orc_init(0)
        L.orc_run_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_run_batch.restype = ctypes.c_longlong
        run = lambda p, tb=1: L.orc_run_batch(p.ctypes.data, len(p), qbuf.ctypes.data, gbuf.ctypes.data, tb)
        kind = "port"
    parts = np.array_split(np.ascontiguousarray(probs), threads)

    def timed(tb):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda p: run(p, tb), parts))
        return time.perf_counter() - t0

    dt = timed(1)
    dtf = timed(0)
    return {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": kind,
            "host_cores": os.cpu_count(), "fill_only_value": round(cells / dtf / 1e9, 4),
            "sample": "%d problems of the same workload (the first %d), fill + traceback in %.1f s, fill alone "
                      "in %.1f s, %d threads (the host share of one GPU), one workspace and pair pool each"
                      % (sample_n, sample_n, dt, dtf, threads)}


def _splice_ends(n, seed=1007):
    """Splice-junction ends of the end-extension mix: read ends of 10-150
    bases against their genome window plus 0-20 bases (3 % substitutions),
    extraband 6, continuation columns up to half the read length."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    ends = []
    for k in range(n):
        rl = int(rng.integers(10, 151))
        gl = rl + int(rng.integers(0, 21))
        g = acgt[rng.integers(0, 4, gl)]
        q = g[:rl].copy()
        sub = rng.random(rl) < 0.03
        q[sub] = acgt[rng.integers(0, 4, int(sub.sum()))]
        anchor = int(rng.integers(1000, 1 << 30))
        ends.append({"rev_rseq": q.tobytes(), "rev_gseq": g.tobytes(), "rev_galt": g.tobytes(),
                     "rev_roffset": rl + 100, "anchor": anchor, "far": anchor - int(rng.integers(50, 5000)),
                     "genestrand": 0, "jump_late": bool(k % 2), "extraband": 6, "defect_rate": 0.001,
                     "contlength": int(rng.integers(0, rl // 2 + 1))})
    return ends


def callers_bench(args):
    """GMAP's DP callers batched in the library, from host buffers in to host
    results out (PCIe included): Dynprog_end5_splicejunction (gdp_sj_run),
    Dynprog_single_gap (gdp_sg_run), Dynprog_end5_gap (gdp_eg_run) and
    Dynprog_cdna_gap (gdp_cd_run) over a device genome.  Each is run once
    untimed, then timed over 5 runs."""
    from gmapdp import callers, ends, serve
    from gmapdp.lib import EG_CALL_DTYPE, SG_CALL_DTYPE
    out = {}

    def timed(f, reps=5):
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = f()
        return (time.perf_counter() - t0) / reps, r

    n = args.callers_n
    arr, qbuf, gbuf = ends.pack_ends(_splice_ends(n), True)
    sb = ends.SpliceBatch()
    dt, (res, pairs) = timed(lambda: sb.run(True, arr, qbuf, gbuf, copy=False))
    out["end5_splicejunction"] = {"calls": n, "seconds": round(dt, 4), "calls_per_s": round(n / dt, 1),
                                  "pairs": int(len(pairs)), "walked": int((res["npairs"] >= 0).sum())}
    log("callers: end5_splicejunction %d ends in %.4f s = %.0f ends/s" % (n, dt, n / dt))

    nb = 100_000_000
    gc = serve.gen_genomecomp(1005, nb)
    dg = serve.DeviceGenome(gc, nb)
    reqs, reads = serve.gen_requests(1005, 0, n, gc, nb)
    # the reads packed end to end, as a caller hands them over
    src = reqs["read"].astype(np.int64) - reads.ctypes.data
    rl = reqs["rl"].astype(np.int64)
    qoff = np.concatenate([[0], np.cumsum(rl)[:-1]])
    qbuf = np.zeros(int(rl.sum()) + 16, np.uint8)
    for k in range(n):
        qbuf[qoff[k]:qoff[k] + rl[k]] = reads[src[k]:src[k] + rl[k]]
    del reads
    sg = np.zeros(n, SG_CALL_DTYPE)
    sg["qoff"] = qoff
    for f in ("rl", "gl", "roffset", "goffset", "chroffset", "chrhigh", "extraband", "watsonp", "genestrand",
              "jump_late", "widebandp"):
        sg[f] = reqs[f]
    sg["defect_rate"] = np.where(np.arange(n) % 2 == 0, 0.001, 0.02)
    b = callers.SingleGapBatch()
    dt, (res, pairs) = timed(lambda: b.run(sg, qbuf, dg, max_rlength=1000, max_glength=1500, copy=False))
    out["single_gap"] = {"calls": n, "seconds": round(dt, 4), "calls_per_s": round(n / dt, 1),
                         "pairs": int(len(pairs)), "simple": int(res["simple"].sum()),
                         "sample": "the serving mix (rl log-uniform 20-1000) on a 1e8-base device genome, "
                                   "HIGHQ and LOWQ alternating, max_rlength 1000, max_glength 1500"}
    log("callers: single_gap %d calls in %.4f s = %.0f calls/s" % (n, dt, n / dt))
    eg = np.zeros(n, EG_CALL_DTYPE)
    eg["qoff"] = qoff
    for f in ("rl", "roffset", "chroffset", "chrhigh", "extraband", "watsonp", "genestrand", "jump_late"):
        eg[f] = reqs[f]
    eg["gl"] = reqs["rl"] + 20
    eg["goffset"] = reqs["goffset"] + reqs["gl"]
    eg["endalign"] = np.arange(n) % 4
    eg["defect_rate"] = 0.001
    b2 = callers.EndGapBatch()
    dt, (res, pairs) = timed(lambda: b2.run(True, eg, qbuf, dg, max_rlength=1000, max_glength=1500, copy=False))
    out["end5_gap"] = {"calls": n, "seconds": round(dt, 4), "calls_per_s": round(n / dt, 1), "pairs": int(len(pairs)),
                       "sample": "read ends of the serving mix against windows 20 bases longer, the four end "
                                 "modes in turn"}
    log("callers: end5_gap %d calls in %.4f s = %.0f calls/s" % (n, dt, n / dt))

    # Dynprog_cdna_gap: genome gaps of 16-160 columns on the plus strand under
    # query gaps holding 10-60 extra bases, both ends reading gl + 8 bases
    from gmapdp.lib import CD_CALL_DTYPE
    rng = np.random.default_rng(1006)
    region, flank = 20_000_000, 200
    G = serve.decode(gc, nb, 0, region)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    gls = rng.integers(16, 161, n)
    ins = rng.integers(10, 61, n)
    cuts = rng.integers(0, gls + 1)
    goff = rng.integers(flank + 10, region - 600, n)
    parts, qlen = [], np.zeros(n, np.int64)
    for k in range(n):
        g0, gl, a = int(goff[k]), int(gls[k]), int(cuts[k])
        q = np.concatenate([G[g0 - flank:g0 + a], acgt[rng.integers(0, 4, int(ins[k]))], G[g0 + a:g0 + gl + flank]])
        parts.append(q)
        qlen[k] = len(q)
    cq = np.concatenate(parts + [np.zeros(16, np.uint8)])
    qstart = np.concatenate([[0], np.cumsum(qlen)[:-1]])
    mid = gls + ins
    cd = np.zeros(n, CD_CALL_DTYPE)
    cd["qoffL"], cd["qoffR"] = qstart + flank, qstart + flank + mid - 1
    cd["roffsetL"], cd["rev_roffsetR"] = flank, flank + mid - 1
    cd["rlengthL"] = cd["rlengthR"] = np.minimum(gls + 8, flank)
    cd["glength"], cd["goffset"] = gls, goff
    cd["chroffset"], cd["chrhigh"], cd["watsonp"] = 0, region, 1
    cd["extraband"], cd["defect_rate"] = 14, 0.001
    b3 = callers.CdnaGapBatch()
    dt, (res, pairs) = timed(lambda: b3.run(cd, cq, dg, max_rlength=500, max_glength=2000, copy=False))
    out["cdna_gap"] = {"calls": n, "seconds": round(dt, 4), "calls_per_s": round(n / dt, 1), "pairs": int(len(pairs)),
                       "joined": int((res["npairs"] >= 0).sum()),
                       "sample": "genome gaps of 16-160 columns under query gaps with 10-60 extra bases, ends of "
                                 "gl + 8 bases, on a 1e8-base device genome"}
    log("callers: cdna_gap %d calls in %.4f s = %.0f calls/s" % (n, dt, n / dt))
    return out


def serving_bench(args):
    """Open-loop Poisson load against the request loop over a device genome."""
    from gmapdp import serve
    from gmapdp.lib import PROBLEM_DTYPE
    t0 = time.time()
    nb = args.serve_genome
    gc = serve.gen_genomecomp(1005, nb)
    dg = serve.DeviceGenome(gc, nb)
    nreq = args.serve_requests
    reqs, reads = serve.gen_requests(1005, 0, nreq, gc, nb)
    del gc
    probs = np.zeros(nreq, PROBLEM_DTYPE)
    probs["rl"], probs["gl"] = reqs["rl"], reqs["gl"]
    probs["lband"], probs["uband"] = reqs["extraband"], reqs["gl"] - reqs["rl"] + reqs["extraband"]
    from gmapdp.batch import cells as cells_of
    cells = cells_of(probs)
    log("serving: genome %d bases on the device, %d requests, %.1fs" % (nb, nreq, time.time() - t0))
    # lane groups for latency: the autotuner times fills of the longest
    # problems of a sample with the requests' length mix
    from gmapdp.tune import autotune
    from gmapdp.workload import Workload
    tw = Workload(3, 50_000, genome_len=1 << 22)
    lane, table = autotune(tw.problems, tw.qbuf.tobytes(), tw.gbuf.tobytes(), objective="latency")
    del tw
    log("serving lane groups (latency autotuner): g16_max_span=%d g64_min_rows=%d" % lane)
    knob_table = None
    if args.serve_batch and args.serve_long >= 0:
        knob = (args.serve_slots, args.serve_batch, args.serve_long, max(args.serve_mix, 0), args.serve_mix_order)
    else:
        # slots, batch size, length split and wave mixing: the highest
        # saturation whose p99 at half of it stays within 90 % of 5 ms
        # (measured; see gmapdp.tune.tune_server)
        from gmapdp.tune import tune_server
        knob, knob_table = tune_server(dg, reqs, lane, dir_budget_bytes=args.serve_budget << 20, max_rl=1000,
                                       max_gl=1500)
        log("serving knobs (slots, max_batch, long_rows, mix_span, mix_order): %s" % (knob,))
    slots, max_batch, long_rows, mix_span, mix_order = knob
    plan = tuple(lane[:2]) + (mix_span, mix_order)
    srv = serve.Server(dg, dir_budget_bytes=args.serve_budget << 20, max_batch=max_batch, slots=slots,
                       max_rl=1000, max_gl=1500, pair_capacity=max_batch * 700, plan=plan, long_rows=long_rows)
    try:
        srv.drive(reqs[:min(nreq, 20000)], 0.0)  # warm-up
        lat, el, ok = srv.drive(reqs, 0.0)
        if ok != nreq:
            raise RuntimeError("serving: %d of %d requests failed" % (nreq - ok, nreq))
        sat = nreq / el
        log("serving saturation: %d requests in %.3f s = %.0f req/s, %.1f GCUPS" % (nreq, el, sat, cells.sum() / el / 1e9))
        points = []
        for frac in (None, 0.1, 0.25, 0.5, 0.75, 0.9):
            rate = 1e4 if frac is None else frac * sat
            m = int(min(nreq, max(2000, rate * args.serve_seconds)))
            # the half-load point is run five times (different arrival
            # seeds) and reports the run with the median p99, all five listed:
            # a single run moves with the shared host's load
            runs = []
            if frac == 0.5:
                # one unmeasured run at this rate first: the first run after a
                # rate change has measured up to 50 % slower at p99
                srv.drive(reqs[:m // 2], rate, seed=99)
            for rep in range(5 if frac == 0.5 else 1):
                lat, el, ok = srv.drive(reqs[:m], rate, seed=7 + rep)
                runs.append((float(np.percentile(lat, 99)) * 1e3, float(np.percentile(lat, 50)) * 1e3, el, ok))
            runs_sorted = sorted(runs)
            p99, p50, el, ok = runs_sorted[len(runs) // 2]
            pt = {"rate_req_s": round(rate, 1), "fraction_of_saturation": None if frac is None else frac,
                  "requests": m, "achieved_req_s": round(m / el, 1),
                  "p50_ms": round(p50, 3), "p99_ms": round(p99, 3), "ok": ok}
            if len(runs) > 1:
                pt["p99_ms_runs"] = [round(r[0], 3) for r in runs]
            log("serving", pt)
            points.append(pt)
        st = srv.stats()
    finally:
        srv.close()
    half = [p for p in points if p["fraction_of_saturation"] == 0.5][0]
    return {"genome_bases": nb, "genome_format": "genomecomp (device-resident)",
            "requests": nreq, "mean_cells_per_request": round(float(cells.mean()), 1),
            "saturation_req_s": round(sat, 1), "saturation_gcups": round(cells.sum() / (nreq / sat) / 1e9, 2),
            "p50_ms_at_50pct": half["p50_ms"], "p99_ms_at_50pct": half["p99_ms"], "points": points,
            "batches": int(st["batches"]), "max_batch": max_batch, "slots": slots, "long_rows": long_rows,
            "mix_span": mix_span, "mix_order": mix_order,
            "per_batch_ms": {k: round(float(st[k + "_seconds"]) / max(int(st["batches"]), 1) * 1e3, 3)
                             for k in ("prep", "retire", "batch", "device")},
            "knob_candidates": None if knob_table is None else
            [{"slots": k[0], "max_batch": k[1], "long_rows": k[2], "mix_span": k[3], "mix_order": k[4],
              "saturation_req_s": round(sat_, 1),
              "p99_ms_at_50pct": round(p99_, 3)} for k, sat_, p99_ in knob_table],
            "lane_groups": {"g16_max_span": lane[0], "g64_min_rows": lane[1],
                            "candidates_ms": [[list(p), round(ms, 3)] for p, ms in table]},
            "dir_budget_mib": args.serve_budget, "dir_bytes_peak": int(st["dir_bytes_peak"])}


def batch_cells(probs):
    from gmapdp.batch import cells
    return int(cells(probs).sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the problem count")
    ap.add_argument("--cpu-sample", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--percall", type=int, default=2000, help="problems timed one launch each (0 = skip)")
    ap.add_argument("--no-e2e", action="store_true", help="skip the PCIe-inclusive step")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="before the timed region, copy MIB MiB once on the device (a known-size kernel that "
                         "calibrates HBM byte counters in a profiled run)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-graph", action="store_true", help="launch fill and traceback per step instead of "
                    "replaying one captured HIP graph")
    ap.add_argument("--no-tune", action="store_true", help="default lane groups instead of the autotuner's")
    ap.add_argument("--lane-groups", default="", metavar="SPAN,ROWS",
                    help="fixed lane groups (g16_max_span,g64_min_rows) instead of the autotuner's")
    ap.add_argument("--mix-span", type=int, default=-1,
                    help="band spread allowed in mixed paired waves of the batch (0: one band per wave; default: "
                         "the autotuner's choice, or 8 with --lane-groups / --no-tune)")
    ap.add_argument("--mix-order", type=int, default=1, choices=(0, 1),
                    help="how mixed waves are formed with --mix-span (gdp_plan_params.mix_order): 0 leftovers in "
                         "order of band, 1 every problem binned by band and taken in order of rows within a bin")
    ap.add_argument("--serve-mix", type=int, default=8,
                    help="mix_span of the server's batches with --serve-batch and --serve-long (0: one band per "
                         "wave)")
    ap.add_argument("--serve-mix-order", type=int, default=1, choices=(0, 1),
                    help="mix_order of the server's batches with --serve-batch and --serve-long")
    ap.add_argument("--callers", action="store_true",
                    help="also time GMAP's batched DP callers (splice-junction ends, single gaps, end gaps, "
                         "cDNA gaps)")
    ap.add_argument("--callers-n", type=int, default=200_000, help="calls of each caller with --callers")
    ap.add_argument("--serve-genome", type=int, default=3_100_000_000, help="bases of the device genome")
    ap.add_argument("--serve-requests", type=int, default=600_000)
    ap.add_argument("--serve-batch", type=int, default=0,
                    help="largest batch the server forms (0: measured choice with --serve-long)")
    ap.add_argument("--serve-slots", type=int, default=15, help="batches in flight (streams) with --serve-batch")
    ap.add_argument("--serve-long", type=int, default=-1,
                    help="reads of at least this many bases are batched apart (0: one class; -1: measured)")
    ap.add_argument("--serve-budget", type=int, default=8192, help="direction budget in MiB")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="target duration of each load point")
    ap.add_argument("--serve-only", action="store_true",
                    help="run the serving measurement alone and print its JSON (what the default run starts in a "
                         "process of its own)")
    args = ap.parse_args()
    if args.serve_only:
        print(json.dumps(serving_bench(args)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (problems are independent, so a batch "
              "would split by length bin across GPUs, each with its own copy of the genome).  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from gmapdp.batch import DeviceBatch
    from gmapdp.workload import Workload

    config, n_default, desc = WORKLOADS[args.workload]
    n = args.n or n_default
    t0 = time.time()
    if config == 4:
        from gmapdp.workload import EndMix
        wl = EndMix(n, n)
    elif config == "cdna":
        from gmapdp.workload import CdnaGaps
        wl = CdnaGaps(n)
        args.percall, args.no_cpu = 0, True  # neither form covers the bridge
    elif config == "intron":
        from gmapdp.workload import IntronGaps
        wl = IntronGaps(n)
        args.percall, args.no_cpu = 0, True
    else:
        wl = Workload(config, n, variant=1 if args.workload == "c1lowq" else 0)
    probs = wl.problems
    total_cells = batch_cells(probs)
    log("workload %s: %d problems, %.3e cells, generated in %.1fs" % (args.workload, len(probs), total_cells,
                                                                      time.time() - t0))

    tuning = None
    params = None
    if args.lane_groups:
        params = tuple(int(x) for x in args.lane_groups.split(","))
    elif not args.no_tune:
        from gmapdp.tune import autotune
        t1 = time.time()
        params, table = autotune(probs, wl.qbuf.tobytes(), wl.gbuf.tobytes(), sample=min(len(probs), 1_000_000))
        tuning = {"objective": "throughput", "sample": min(len(probs), 1_000_000), "chosen": list(params),
                  "candidates": [[list(p), round(ms, 3)] for p, ms in table], "seconds": round(time.time() - t1, 1)}
        log("autotuner: g16_max_span=%d g64_min_rows=%d mix_span=%d mix_order=%d (%.1fs)" % (params + (time.time() - t1,)))
    if args.mix_span >= 0:
        params = tuple(params or (240, 8001))[:2] + (args.mix_span, args.mix_order)
    elif params is None or len(params) == 2:
        params = tuple(params or (240, 8001))[:2] + (8, 1)
    b = DeviceBatch(probs, wl.qbuf.tobytes(), wl.gbuf.tobytes(), params=params)
    stream = torch.cuda.current_stream()
    graph = None
    if config == 4:
        from gmapdp.lib import TBREQ_DTYPE
        reqs = np.concatenate([wl.single_requests(), wl.end_templates()])
        off = np.zeros(len(reqs), np.int64)
        off[1:] = np.cumsum(reqs["max_pairs"].astype(np.int64))[:-1]
        reqs["pair_off"] = off
        d_ul = torch.from_numpy(wl.ul.reshape(-1).copy()).to(b.dev)
        d_tmpl = torch.from_numpy(reqs[wl.n3:].view(np.uint8).copy()).to(b.dev)
        d_best = torch.zeros(3 * wl.m, dtype=torch.int32, device=b.dev)
    elif config == "cdna":
        reqs = wl.templates()
        d_quads = torch.from_numpy(wl.quads.reshape(-1).copy()).to(b.dev)
        d_tmpl = torch.from_numpy(reqs.view(np.uint8).copy()).to(b.dev)
        d_best = torch.zeros(5 * wl.n, dtype=torch.int32, device=b.dev)
    elif config == "intron":
        reqs = wl.templates()
        d_gaps = torch.from_numpy(wl.gaps.view(np.uint8).copy()).to(b.dev)
        d_sprobs = torch.from_numpy(wl.probs).to(b.dev)
        d_tmpl = torch.from_numpy(reqs.view(np.uint8).copy()).to(b.dev)
        d_best = torch.zeros(5 * wl.n, dtype=torch.int32, device=b.dev)
        d_bestp = torch.zeros(2 * wl.n, dtype=torch.float64, device=b.dev)
    else:
        reqs = wl.tb_requests()
    reqs_dev = b.prepare_traceback(reqs, genome=wl.genome())

    def enqueue_rest(st=None):
        # everything after the fill: end cells (config 4) or the cDNA bridge,
        # then the tracebacks
        if config == 4:
            b.best_endpoint_requests(d_ul, wl.m, False, d_best, d_tmpl.data_ptr(),
                                     reqs_dev["d_reqs"].data_ptr() + wl.n3 * TBREQ_DTYPE.itemsize, stream=st)
        elif config == "cdna":
            b.bridge_requests(d_quads, wl.n, d_best, d_tmpl.data_ptr(), reqs_dev["d_reqs"].data_ptr(), stream=st)
        elif config == "intron":
            b.bridge_intron_requests(d_gaps, wl.n, d_sprobs, d_best, d_bestp, d_tmpl.data_ptr(),
                                     reqs_dev["d_reqs"].data_ptr(), stream=st)
        b.traceback_device(reqs_dev, stream=st)

    def step(ev=None, tev=None):
        if graph is not None:
            # the whole step replayed from one captured graph; the fill and
            # traceback split is measured on the eager launches below
            b.graph_launch(graph)
            return
        if ev is not None:
            ev[0].record(stream)
        b.fill()
        if ev is not None:
            ev[1].record(stream)
        enqueue_rest()
        if tev is not None:
            tev.record(stream)

    if args.calib_copy:
        src = torch.ones(args.calib_copy << 18, dtype=torch.int32, device=b.dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read, %d MiB written" % (args.calib_copy, args.calib_copy))
        del src, dst
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # fill and traceback durations from HIP events around eager launches
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    tevs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    for k in range(args.steps):
        step(evs[k], tevs[k])
    torch.cuda.synchronize()
    fill_ms = float(np.mean([a.elapsed_time(z) for a, z in evs]))
    tb_ms = float(np.mean([z.elapsed_time(t) for (a, z), t in zip(evs, tevs)]))
    if not args.no_graph:
        graph = b.capture(lambda st: (b.fill(stream=st), enqueue_rest(st)))
        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k], tevs[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    gcups = total_cells * args.steps / dt / 1e9
    achieved = total_cells * OPS_PER_CELL / (fill_ms * 1e-3) / 1e12
    log("fill %.2f ms, traceback %.2f ms, step %.2f ms, %.1f GCUPS end-to-end, fill-only %.1f GCUPS"
        % (fill_ms, tb_ms, ms_per_step, gcups, total_cells / (fill_ms * 1e-3) / 1e9))

    bridge_ms = None
    if config == "cdna":
        # the bridge alone, HIP events on the stream it runs on (torch's current one)
        bev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        bev[0].record(stream)
        for _ in range(args.steps):
            b.bridge_requests(d_quads, wl.n, d_best, d_tmpl.data_ptr(), reqs_dev["d_reqs"].data_ptr())
        bev[1].record(stream)
        torch.cuda.synchronize()
        bridge_ms = bev[0].elapsed_time(bev[1]) / args.steps
        log("cDNA bridge: %.3f ms for %d gaps" % (bridge_ms, wl.n))
    elif config == "intron":
        bev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        bev[0].record(stream)
        for _ in range(args.steps):
            b.bridge_intron_requests(d_gaps, wl.n, d_sprobs, d_best, d_bestp, d_tmpl.data_ptr(),
                                     reqs_dev["d_reqs"].data_ptr())
        bev[1].record(stream)
        torch.cuda.synchronize()
        bridge_ms = bev[0].elapsed_time(bev[1]) / args.steps
        log("intron bridge: %.3f ms for %d gaps" % (bridge_ms, wl.n))

    # PCIe-inclusive step: inputs from pinned host memory, the step, and each
    # request's counts and the pair records its walk wrote back to pinned
    # memory (gdp_download_results).
    e2e = None
    if not args.no_e2e and config not in (4, "cdna", "intron"):
        h_in = [torch.empty(t.numel(), dtype=t.dtype, pin_memory=True) for t in (b.d_descs, b.d_q, b.d_g,
                                                                                  reqs_dev["d_reqs"])]
        for h, t in zip(h_in, (b.d_descs, b.d_q, b.d_g, reqs_dev["d_reqs"])):
            h.copy_(t)
        h_cnt = torch.empty(reqs_dev["d_counts"].numel(), dtype=torch.int32, pin_memory=True)
        h_pairs = torch.empty(reqs_dev["d_pairs"].numel(), dtype=torch.uint8, pin_memory=True)
        torch.cuda.synchronize()
        reps = 2
        t1 = time.perf_counter()
        for _ in range(reps):
            for h, t in zip(h_in, (b.d_descs, b.d_q, b.d_g, reqs_dev["d_reqs"])):
                t.copy_(h, non_blocking=True)
            step()
            rc = b.L.gdp_download_results(reqs_dev["d_reqs"].data_ptr(), len(reqs), reqs_dev["d_counts"].data_ptr(),
                                          reqs_dev["d_pairs"].data_ptr(), h_cnt.data_ptr(), h_pairs.data_ptr(),
                                          stream.cuda_stream)
            if rc != 0:
                raise RuntimeError("gdp_download_results: %s" % b.L.gdp_last_error().decode())
        torch.cuda.synchronize()
        dte = (time.perf_counter() - t1) / reps
        up = sum(h.numel() * h.element_size() for h in h_in)
        npairs = h_cnt.view(-1, 6)[:, 0].clamp(min=0).sum().item()
        down = h_cnt.numel() * 4 + npairs * 16
        e2e = {"gcups": round(total_cells / dte / 1e9, 3), "ms_per_step": round(dte * 1e3, 3),
               "upload_bytes": int(up), "download_bytes": int(down),
               "method": "descriptors, sequences and traceback requests uploaded from pinned memory, the step, "
                         "then each request's counts and the pair records its walk wrote written to pinned memory "
                         "by a kernel (gdp_download_results), one synchronize"}
        log("PCIe-inclusive step: %.2f ms, %.1f GCUPS (%.2f GB up, %.2f GB down)" % (dte * 1e3, e2e["gcups"],
                                                                                    up / 1e9, down / 1e9))
        del h_in, h_cnt, h_pairs

    # Plain one-launch-per-problem form on the same inputs: a C loop
    # (gdp_align_each) of one fill launch of the problem's own kernel, one
    # traceback launch and a stream synchronize per problem.
    percall = None
    if args.percall:
        m = min(args.percall, n)
        sub = DeviceBatch(probs[:m], wl.qbuf.tobytes(), wl.gbuf.tobytes(), single=True)  # one sweep per problem
        # (config 4: its single-gap half, problems[:n3])
        sreq = sub.prepare_traceback((wl.single_requests() if config == 4 else wl.tb_requests())[:m],
                                     genome=wl.genome())
        sub.align_each(sreq)  # untimed: first launches of each kernel
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sub.align_each(sreq)
        torch.cuda.synchronize()
        dtp = time.perf_counter() - t1
        percall = batch_cells(probs[:m]) / dtp / 1e9
        log("per-call form: %d problems, %.3f GCUPS (%.1f us per problem)" % (m, percall, dtp / m * 1e6))
        del sub, sreq

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_fill")

    serving = None
    if not args.no_serve or args.callers:
        del b, reqs_dev
        torch.cuda.empty_cache()
    if not args.no_serve:
        # The server runs in a process of its own, as a deployment would run
        # it, with a hardware queue per batch slot (GPU_MAX_HW_QUEUES; HIP's
        # default of 4 is shared by every stream of a process, and slots
        # sharing a queue run one after another): measured with three slots,
        # p99 at half load 5.2-5.7 ms in this process against 3.9-4.2 ms in a
        # fresh one; with 16 queues, 15 slots reach 5.8e6 requests/s.
        cmd = [sys.executable, os.path.abspath(__file__), "--serve-only"]
        for flag in ("serve_mix", "serve_mix_order", "serve_genome", "serve_requests", "serve_batch", "serve_slots",
                     "serve_long", "serve_budget", "serve_seconds"):
            cmd.append("--%s=%s" % (flag.replace("_", "-"), getattr(args, flag)))
        env = dict(os.environ, GPU_MAX_HW_QUEUES="16")
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError("serving measurement failed (exit %d)" % r.returncode)
        serving = json.loads(r.stdout.strip().splitlines()[-1])
    callers_out = callers_bench(args) if args.callers else None

    cpu = None
    if not args.no_cpu:
        # the host share of one GPU on these machines is 16 cores
        threads = min(16, os.cpu_count() or 1)
        sample = args.cpu_sample or (700000 if args.workload == "c2" else 40000)
        cpu = cpu_baseline(wl, min(sample, n), threads)
        log("cpu baseline:", cpu)

    out = {
        "metric": "bit-exact GCUPS: batched persistent vs per-call GPU DP; serving p50/p99 ms",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"c1lowq": "int8", "c3": "int16+int8", "c4": "int16+int8", "cdna": "int16+int8",
                  "intron": "int16+int8"}.get(
            args.workload, "int16"),
        "data": "synthetic (seeded generator of the workload)",
        "config": {"workload": desc, "problems": len(probs), "cells": total_cells,
                   "step": {4: "fill + best end cell + traceback", "cdna": "fill + bridge + traceback",
                            "intron": "fill + intron bridge + traceback"}.get(
                       config, "fill + traceback"),
                   "launch": "eager" if graph is None else "one HIP graph per step",
                   "lane_groups": {"g16_max_span": params[0], "g64_min_rows": params[1]},
                   "mix_span": params[2], "mix_order": params[3]},
        "fill_ms": round(fill_ms, 3),
        "traceback_ms": round(tb_ms, 3),
        "fill_gcups": round(total_cells / (fill_ms * 1e-3) / 1e9, 3),
        "percall_gcups": None if percall is None else round(percall, 4),
        "percall_method": None if percall is None else
        "gdp_align_each over the first %d problems: per problem one fill launch of its own kernel, one traceback "
        "launch, one stream synchronize (a C loop)" % min(args.percall, n),
        "e2e": e2e,
        "batched_vs_percall": None if percall is None else round(gcups / percall, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT16_TOPS, "unit": "TOP/s",
                     "frac": round(achieved / PEAK_INT16_TOPS, 4), "traffic": traffic,
                     "kernel": ("gdp_fill_kernel (one fill of one-sided problems keeping every score)"
                                if config in ("cdna", "intron") else
                                "gdp_fill2_kernel (one fill: every launch of gdp_fill, HIP events on its stream)"),
                     "ops_per_cell": OPS_PER_CELL},
        "cpu_baseline": cpu,
        "bridge_ms": None if bridge_ms is None else round(bridge_ms, 3),
        "serving": serving,
        "callers": callers_out,
        "tuning": tuning,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
