#!/usr/bin/env python3
"""Benchmark of the batched transform path on one MI355X.

Workload (the batched-uneven configuration): 20000 independent ComplexF32
1-D transforms, lengths drawn log-uniform in [16, 65536] (seed 0), 20% snapped
to primes (Bluestein), 30% to powers of 3, the rest left as drawn; fft/bfft
50/50; entries N(0,1) (synthetic, generated on the device).  One step = one
execution of the whole batch (a captured hipGraph of grouped line-pass
launches) with inputs already resident in HBM.

Reported (one JSON line):
  value      GFLOP/s = sum 5 N log2 N / step time
  roofline   algorithmic HBM bytes (2 N sizeof(ComplexF32) per transform) per
             step / step device time (HIP events on the launch stream), vs
             8 TB/s
  error      worst relative L2 error / (2e-6 log2 N) over a sample, against the
             oracle in FP64 (FFTA's algorithm)
  cpu_baseline  the oracle (C++ restatement of FFTA's algorithm, FP32 like
             FFTA's ComplexF32 path) timed on a sample of the same workload,
             one transform per thread.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ffta.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def _primes_upto(n: int) -> np.ndarray:
    sieve = np.ones(n + 1, dtype=bool)
    sieve[:2] = False
    for p in range(2, int(n ** 0.5) + 1):
        if sieve[p]:
            sieve[p * p::p] = False
    return np.nonzero(sieve)[0]


def batched_uneven(count: int = 20000, seed: int = 0, lo: int = 16, hi: int = 65536):
    """Lengths and directions of the batched-uneven workload."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(lo), np.log(hi), count)
    n = np.clip(np.rint(np.exp(u)).astype(np.int64), lo, hi)
    cat = rng.uniform(0, 1, count)
    primes = _primes_upto(hi)
    primes = primes[primes >= lo]
    pow3 = np.array([3 ** k for k in range(1, 12) if lo <= 3 ** k <= hi], dtype=np.int64)
    out = n.copy()
    for i in range(count):
        if cat[i] < 0.2:
            j = np.searchsorted(primes, n[i])
            out[i] = primes[min(j, primes.size - 1)]
        elif cat[i] < 0.5:
            out[i] = pow3[np.argmin(np.abs(np.log(pow3) - np.log(n[i])))]
    dirs = np.where(rng.uniform(0, 1, count) < 0.5, -1, 1).astype(np.int32)
    return out, dirs


def flops_of(lengths) -> float:
    L = np.asarray(lengths, dtype=np.float64)
    return float(np.sum(np.where(L > 1, 5.0 * L * np.log2(np.maximum(L, 1)), 0.0)))


def cpu_baseline(lengths, dirs, seconds_target: float = 15.0, seed: int = 1):
    """Time the oracle on a bounded sample of the workload (one transform per thread)."""
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(lengths))
    # calibrate on a small sample, then size the sample to ~seconds_target
    take = 64
    while True:
        idx = order[:take]
        L = lengths[idx]
        offs = np.concatenate([[0], np.cumsum(L)[:-1]]).astype(np.int64)
        x = (rng.standard_normal(int(L.sum())) + 1j * rng.standard_normal(int(L.sum()))).astype(np.complex64)
        t0 = time.perf_counter()
        oracle.fft_batch(x, L, offs, dirs[idx], nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= seconds_target * 0.5 or take >= len(lengths):
            break
        grow = max(2.0, min(8.0, seconds_target / max(dt, 1e-3)))
        take = min(len(lengths), int(take * grow))
    return {
        "value": round(flops_of(L) / dt / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{take} of {len(lengths)} transforms (random subset, seed {seed}), ComplexF32, "
                  f"{dt:.1f} s wall on {threads} threads, one transform per thread",
    }


def parity_sample(torch, batch, x, y, lengths, dirs, offsets, k: int = 200, seed: int = 2):
    import oracle

    rng = np.random.default_rng(seed)
    idx = rng.choice(len(lengths), size=min(k, len(lengths)), replace=False)
    xs = x.cpu().numpy()
    ys = y.cpu().numpy()
    worst = 0.0
    for i in idx:
        n, o = int(lengths[i]), int(offsets[i])
        xi = xs[o:o + n].astype(np.complex128)
        ref = oracle.fft(xi, int(dirs[i]))
        err = oracle.relerr(ys[o:o + n], ref)
        bound = 2e-6 * math.log2(n) if n > 1 else 0.0
        worst = max(worst, err / bound if bound else (0.0 if err == 0 else float("inf")))
    return worst, len(idx)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a hipGraph")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(independent transforms shard across GPUs with no exchange)", file=sys.stderr)
        return 2

    import torch

    from fftahip.plan import Batch

    lengths, dirs = batched_uneven(args.count)
    t0 = time.perf_counter()
    batch = Batch(lengths, dirs, precision=32)
    plan_s = time.perf_counter() - t0
    total = int(lengths.sum())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    x = torch.randn(total, dtype=torch.complex64, device="cuda", generator=gen)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step():
        batch.execute_device(x.data_ptr(), y.data_ptr(), sp, graph=not args.no_graph)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps

    flops = flops_of(lengths)
    bytes_step = 2.0 * total * 8
    achieved = bytes_step / (ms * 1e-3) / 1e9
    worst, nsample = parity_sample(torch, batch, x, y, lengths, dirs, batch.offsets)
    desc = batch.describe()

    line = {
        "metric": "batched FFT GFLOP/s (5Nlog2N), %HBM roofline; serving p99 ms @ relerr<=2e-6",
        "value": round(flops / (ms * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c32 (fp32 butterflies, fp64-generated twiddles/chirps)",
        "data": "synthetic",
        "config": {
            "workload": f"batched uneven 1-D: {args.count} ComplexF32 transforms, lengths log-uniform [16,65536] "
                        f"(seed 0), 20% primes, 30% powers of 3, fft/bfft 50/50",
            "transforms": int(args.count),
            "elements": total,
            "distinct_lengths": int(np.unique(lengths).size),
            "launches_per_step": len(desc.get("phases", [])),
            "graph": not args.no_graph,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "error": {"worst_relerr_over_bound": round(worst, 4), "sample": nsample,
                  "bound": "2e-6*log2(N) vs FP64 oracle"},
        "plan_seconds": round(plan_s, 2),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lengths, dirs)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
