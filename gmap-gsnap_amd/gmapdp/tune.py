"""Autotuner of the paired sweep's plan over a measured batch.

The plan's knobs (gdp_plan_params) trade padding against sweep length:
16-lane groups pack eight problems into a wave with little padding, 64-lane
groups sweep a long problem in about a quarter of the steps, and mixed waves
(problems of neighbouring bands and rows in one wave) fill waves whose
problems rarely share a band.  The tuner times real fills of a seeded sample
of the caller's problems on the device for each candidate and keeps the
best: the whole sample's fill time for throughput (the sample large enough
that the slowest wave does not decide it), or the fill time of a batch of
the longest reads for latency (the time a request waits on the slowest
problem of its batch).  A second tuner picks the request server's batch
slots, batch size, length split and wave mixing by measured saturation and
p99.
"""
import numpy as np

from .batch import DeviceBatch, _torch

SPANS = (240, 112)
ROWS = {"throughput": (8001, 512, 256), "latency": (8001, 512, 256, 128)}
# Throughput candidates (g16_max_span, g64_min_rows, mix_span, mix_order):
# mixed waves binned by band and rows (mix_order 1) against the leftover
# mixing of one band per wave first (mix_order 0).  Measured on the c3 mix
# (2e6 problems): (240, 8001, 4, 1) 71.9 ms, (240, 8001, 8, 1) 73.6 ms,
# (240, 8001, 16, 1) 75.2 ms, (240, 512, 16, 1) 80.1 ms, (240, 512, 16, 0)
# 76.8 ms, (112, 8001, 8, 1) 85.5 ms; 16-lane groups cost less per step
# than 64-lane ones, whose band edges stretch over four times as many steps.
THROUGHPUT = ((240, 8001, 4, 1), (240, 8001, 2, 1), (240, 8001, 8, 1), (240, 8001, 16, 1), (240, 512, 4, 1),
              (240, 512, 16, 0))


def _time_fill(b, reps):
    torch = _torch()
    b.fill()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, z in ev:
        a.record()
        b.fill()
        z.record()
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(z) for a, z in ev]))


def autotune(problems, qbuf, gbuf, sample=1_000_000, objective="throughput", reps=2, seed=0, longest=4096):
    """Returns (params, table) where table lists (params, milliseconds) of
    every candidate: params = (g16_max_span, g64_min_rows, mix_span,
    mix_order) for throughput, (g16_max_span, g64_min_rows) for latency."""
    rng = np.random.default_rng(seed)
    n = len(problems)
    if objective == "latency":
        # a batch of long reads (a server batches reads of 256 bases and more
        # apart), whatever their bands: the slowest sweeps set a serving
        # batch's time, and the lane groups decide how long they are
        long_idx = np.nonzero(problems["rl"] >= 256)[0]
        idx = np.sort(rng.choice(long_idx, size=min(longest, len(long_idx)), replace=False))
    else:
        idx = np.sort(rng.choice(n, size=min(sample, n), replace=False))
    sub = np.ascontiguousarray(problems[idx])
    table = []
    cands = THROUGHPUT if objective == "throughput" else [(s, r) for s in SPANS for r in ROWS[objective]]
    for prm in cands:
        b = DeviceBatch(sub, qbuf, gbuf, params=prm)
        table.append((prm, _time_fill(b, reps)))
        del b
    best = min(table, key=lambda t: t[1])[0]
    return best, table


# (slots, max_batch, long_rows, mix_span, mix_order) candidates of the
# request server: batches in flight, requests per batch, the query length
# from which reads are batched apart, and how the requests of a batch share
# paired waves (gdp_plan_params).  At half load a request waits for a batch
# slot and then for its batch, whose time the longest reads set.  Measured
# on the configuration-5 mix (saturation, then 3 runs at half of it) with a
# hardware queue per slot (GPU_MAX_HW_QUEUES 16) and a thread per slot
# forming its batches (with one thread for all slots, planning held the
# server at 3.3e6 requests/s): waves of one band only, (6, 8192, 300) 4.39e6
# requests/s at p99 4.2-4.3 ms; mixed waves binned by 4 columns, (6, 8192,
# 300) 4.86e6 at 4.8 ms, (12, 8192, 300) 5.29e6-5.50e6 at 4.0-4.3 ms,
# (15, 8192, 300) 5.53e6 at 3.7-3.9 ms, (12, 16384, 300) 6.01e6 at 4.4-4.5
# ms, (15, 4096, 300) 5.03e6 at 3.6-3.7 ms; binned by 8 columns,
# (15, 8192, 300) 5.81e6 at 3.9-4.0 ms.  With the latency tuner's lane
# groups 112,128: (12, 16384, 300, 4) 5.80e6 at 3.6 ms, (12, 24576, 300, 4)
# 5.86e6 at 3.7 ms, (15, 16384, 300, 8) 5.99e6 at 3.5-3.6 ms, (15, 12288,
# 300, 4) 5.69e6 at 3.4-3.5 ms: saturation levels off near 2.8e2 GCUPS.
SERVER_KNOBS = ((15, 16384, 300, 8, 1), (15, 8192, 300, 8, 1), (12, 16384, 300, 4, 1), (15, 8192, 300, 4, 1))


def tune_server(genome, reqs, plan, target_p99_ms=5.0, knobs=SERVER_KNOBS, seconds=1.5, margin=0.9, **server_kw):
    """Picks (slots, max_batch, long_rows, mix_span, mix_order) for the
    request server by measurement, with the lane groups plan[:2]:
    for each candidate, the saturation throughput (all requests at once) and
    the median p99 latency of three runs of Poisson arrivals at half of it.
    Keeps the candidate with the highest saturation whose p99 is within
    margin * target (p99 of a short run moves by several percent from run to
    run), else within the target, else the one with the lowest p99.
    Returns (knob, table of (knob, saturation req/s, p99 ms))."""
    from .serve import Server
    table = []
    for slots, mb, lr, mix, order in knobs:
        with Server(genome, max_batch=mb, slots=slots, pair_capacity=mb * 700, plan=tuple(plan[:2]) + (mix, order),
                    long_rows=lr, **server_kw) as srv:
            srv.drive(reqs[:min(len(reqs), 20000)], 0.0)
            lat, el, ok = srv.drive(reqs, 0.0)
            sat = len(reqs) / el
            m = int(min(len(reqs), max(2000, 0.5 * sat * seconds)))
            p99 = []
            for rep in range(3):  # the median of three runs
                lat, el, ok = srv.drive(reqs[:m], 0.5 * sat, seed=11 + rep)
                p99.append(float(np.percentile(lat, 99)) * 1e3)
            table.append(((slots, mb, lr, mix, order), sat, float(np.median(p99))))
    good = [t for t in table if t[2] <= margin * target_p99_ms] or [t for t in table if t[2] <= target_p99_ms]
    best = max(good, key=lambda t: t[1]) if good else min(table, key=lambda t: t[2])
    return best[0], table
