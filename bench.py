#!/usr/bin/env python3
"""Benchmark of the batched transform path on one MI355X.

Workload (the batched-uneven configuration): 20000 independent ComplexF32
1-D transforms, lengths drawn log-uniform in [16, 65536] (seed 0), 20% snapped
to primes (Bluestein), 30% to powers of 3, the rest left as drawn; fft/bfft
50/50; entries N(0,1) (synthetic, generated on the device).  One step = one
execution of the whole batch with inputs already resident in HBM: a replay of
the batch's hipGraph (its grouped line-pass launches as kernel nodes, their
dependencies as edges); with --eager, the same launch DAG issued over four
streams.  The other mode is timed beside it.

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
  serving    Poisson replay through the native request loop: p50/p99 latency,
             throughput.
  percall_fp64  the same library called once per transform in ComplexF64 on
             a sample of the workload from a native loop
             (ffta.jl_amd/bin/fftahip_percall; transforms/s), and the batched
             FP32 rate over it.
  per_call_configs  single transforms (4096 CF64 fft/bfft; 177147,
             31104000, 1000003; 4096x4096, 3000x2187; CF32 and CF64): GPU
             time per mul!, GFLOP/s, fraction of HBM peak, error against the
             FP64 oracle, and the oracle timed on one host core.
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


def serving_catalogue():
    """Request shapes of the serving configuration (fixed catalogues, so that a
    warm plan cache covers them): 1-D powers of two 64..4096; 12 composite
    lengths log-spaced in [1e3, 1e6]; 8 primes log-spaced up to 2^20; 2-D
    squares 256..2048."""
    pow2 = [(1 << k,) for k in range(6, 13)]
    comp = []
    for v in np.round(np.logspace(3, 6, 12)).astype(np.int64):
        v = int(v) + (int(v) % 2)  # even, hence composite
        comp.append((v,))
    primes = _primes_upto(1 << 20)
    pr = []
    for v in np.round(np.logspace(np.log10(1 << 10), np.log10(1 << 20), 8)).astype(np.int64):
        j = min(np.searchsorted(primes, v), primes.size - 1)
        pr.append((int(primes[j]),))
    squares = [(n, n) for n in (256, 512, 1024, 2048)]
    return pow2, comp, pr, squares


def serving_trace(rate: float, seconds: float, seed: int = 0):
    """Poisson arrivals at `rate` req/s: 70% pow2, 20% composite, 5% prime,
    5% 2-D; 80% ComplexF32 / 20% ComplexF64; fft/bfft 50/50."""
    rng = np.random.default_rng(seed)
    n = max(1, int(rate * seconds))
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    pow2, comp, pr, squares = serving_catalogue()
    cat = rng.uniform(0, 1, n)
    shapes = []
    for c in cat:
        if c < 0.70:
            shapes.append(pow2[rng.integers(len(pow2))])
        elif c < 0.90:
            shapes.append(comp[rng.integers(len(comp))])
        elif c < 0.95:
            shapes.append(pr[rng.integers(len(pr))])
        else:
            shapes.append(squares[rng.integers(len(squares))])
    prec = np.where(rng.uniform(0, 1, n) < 0.8, 32, 64)
    dirs = np.where(rng.uniform(0, 1, n) < 0.5, -1, 1)
    return t - t[0], shapes, prec, dirs


def run_serving(torch, srv, rate: float = 50000.0, seconds: float = 0.5, seed: int = 0, ring: int = 8):
    """Replay the serving trace against a Server with device-resident payloads."""
    t, shapes, prec, dirs = serving_trace(rate, seconds, seed)
    bufs = {}
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    for sh, p in set(zip(shapes, prec.tolist())):
        n = int(np.prod(sh))
        dt = torch.complex64 if p == 32 else torch.complex128
        x = torch.randn(n, dtype=dt, device="cuda", generator=gen)
        bufs[(sh, p)] = (x, [torch.empty_like(x) for _ in range(ring)])
    # warm the plan cache: one request per (shape, precision, direction)
    warm = []
    for key in set(zip(shapes, prec.tolist(), dirs.tolist())):
        sh, p, d = key
        x, ys = bufs[(sh, p)]
        warm.append(srv.submit(p, d, sh, x.data_ptr(), ys[0].data_ptr()))
    for w in warm:
        srv.wait(w)
    srv.drain()
    srv.reset_stats()
    ins, outs, count = [], [], {}
    for sh, p in zip(shapes, prec.tolist()):
        x, ys = bufs[(sh, p)]
        k = count.get((sh, p), 0)
        count[(sh, p)] = k + 1
        ins.append(x.data_ptr())
        outs.append(ys[k % ring].data_ptr())
    dims3 = [list(sh) + [0] * (3 - len(sh)) for sh in shapes]
    tickets = srv.replay(t, prec.tolist(), dirs.tolist(), dims3, ins, outs)
    srv.drain()
    st = srv.stats()
    st["requests"] = len(tickets)
    st["offered_rate"] = rate
    return st


def device_report(torch, stream, mib: int = 1024, reps: int = 10) -> dict:
    """The device as the runtime reports it (name, CUs, LDS per workgroup,
    memory) and the HBM rate of a plain device copy of `mib` MiB (read +
    write bytes over the time of one copy, events on `stream`): the practical
    ceiling that the roofline's datasheet peak stands beside."""
    from fftahip import device_info
    prop = torch.cuda.get_device_properties(0)
    info = device_info()
    lds, cus, warp = info["lds_per_block"], info["cu_count"], info["warp_size"]
    n = mib * (1 << 20) // 8
    a = torch.empty(n, dtype=torch.complex64, device="cuda")
    a.zero_()
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        b.copy_(a)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / reps
    del a, b
    copy_gbs = 2.0 * n * 8 / (ms * 1e-3) / 1e9
    return {"name": prop.name, "compute_units": int(cus), "lds_per_workgroup": int(lds), "wavefront": int(warp),
            "hbm_bytes": int(prop.total_memory), "copy_GBps": round(copy_gbs, 1),
            "copy_frac_of_peak": round(copy_gbs / HBM_PEAK_GBS, 3), "copy": f"torch device copy of {mib} MiB"}


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
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": round(flops_of(L) / dt / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu_model": model,
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


PER_CALL_CONFIGS = [
    # (name, Julia shape, precision, direction, input distribution)
    ("fft 4096", (4096,), 64, -1, "uniform"),
    ("bfft 4096", (4096,), 64, 1, "uniform"),
    ("fft 177147", (177147,), 64, -1, "normal"),
    ("fft 177147", (177147,), 32, -1, "normal"),
    ("fft 31104000", (31104000,), 64, -1, "normal"),
    ("fft 31104000", (31104000,), 32, -1, "normal"),
    ("fft 1000003", (1000003,), 64, -1, "normal"),
    ("fft 1000003", (1000003,), 32, -1, "normal"),
    ("fft 4096x4096", (4096, 4096), 64, -1, "normal"),
    ("fft 4096x4096", (4096, 4096), 32, -1, "normal"),
    ("fft 3000x2187", (3000, 2187), 64, -1, "normal"),
    ("fft 3000x2187", (3000, 2187), 32, -1, "normal"),
]


def per_call_configs(torch, cpu: bool = True, seed: int = 7, cpu_budget_s: float = 3.0, reps_fixed: int = 0,
                     only=None):
    """The single-transform configurations: GPU time per mul! with the arrays
    resident in HBM (HIP events on the launch stream), error against the FP64
    oracle, and the oracle in the plan's precision timed single-threaded on
    the host (FFTA runs one transform on one core)."""
    import oracle
    from fftahip.plan import Plan

    rows = []
    stream = torch.cuda.current_stream()
    for name, shape, prec, direction, dist in PER_CALL_CONFIGS:
        if only and not any(o == f"{'x'.join(map(str, shape))}:{prec}" for o in only):
            continue
        rng = np.random.default_rng(seed)
        n = int(np.prod(shape))
        if dist == "uniform":
            x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        else:
            x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        dt = np.complex64 if prec == 32 else np.complex128
        tdt = torch.complex64 if prec == 32 else torch.complex128
        x = x.astype(dt)
        tx = torch.from_numpy(x).to("cuda")
        ty = torch.empty_like(tx)
        t0 = time.perf_counter()
        plan = Plan(shape, prec, None, direction)
        plan_s = time.perf_counter() - t0

        def run():
            plan.execute_device(shape, tx.data_ptr(), ty.data_ptr(), stream.cuda_stream)

        run()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run()
        e1.record(stream)
        torch.cuda.synchronize()
        est = max(e0.elapsed_time(e1), 1e-3)
        reps = reps_fixed if reps_fixed > 0 else int(max(5, min(200, 200.0 / est)))
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        y = ty.cpu().numpy()
        xf = np.asfortranarray(x.reshape(shape, order="F"))
        ref = oracle.fftn(xf.astype(np.complex128), None, direction) if len(shape) > 1 else \
            oracle.fft(x.astype(np.complex128), direction)
        err = oracle.relerr(y.reshape(shape, order="F") if len(shape) > 1 else y, ref)
        bound = (1e-14 if prec == 64 else 2e-6) * math.log2(n)
        esz = 8 if prec == 32 else 16
        row = {
            "config": name, "shape": list(shape), "dtype": "c32" if prec == 32 else "c64",
            "gpu_us": round(ms * 1e3, 2), "gflops": round(5 * n * math.log2(n) / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(2 * n * esz / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "relerr_vs_fp64_oracle": float(f"{err:.3e}"), "bound": float(f"{bound:.3e}"),
            "plan_ms": round(plan_s * 1e3, 1),
        }
        if cpu:
            xs = x if len(shape) == 1 else xf
            samples = []
            while not samples or (sum(samples) < cpu_budget_s and len(samples) < 20):
                t0 = time.perf_counter()
                if len(shape) > 1:
                    oracle.fftn(xs, None, direction)
                else:
                    oracle.fft(xs, direction)
                samples.append(time.perf_counter() - t0)
            cpu_s = float(np.median(samples))
            row["cpu_us"] = round(cpu_s * 1e6, 1)
            row["cpu_samples"] = len(samples)
            row["gpu_vs_cpu"] = round(cpu_s / (ms * 1e-3), 1)
        rows.append(row)
        del tx, ty, plan
    return rows


PERCALL_BIN = os.path.join(ROOT, "ffta.jl_amd", "bin", "fftahip_percall")


def per_call_fp64(torch, lengths, dirs, sample: int = 2000, seed: int = 3):
    """The plain per-call FP64 path on a random sample of the batched
    workload: one mul! per transform (ComplexF64, plans built beforehand,
    arrays resident), timed over the whole sample by a native loop
    (ffta.jl_amd/bin/fftahip_percall: fftahip_execute per transform on one
    stream, device time between events and host wall time).  The same loop
    through the Python binding is timed beside it."""
    import subprocess
    import tempfile

    from fftahip.plan import Plan

    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(lengths), size=min(sample, len(lengths)), replace=False))
    L = lengths[idx].astype(np.int64)
    D = dirs[idx].astype(np.int32)
    with tempfile.NamedTemporaryFile(suffix=".calls", delete=False) as f:
        f.write(np.int64(L.size).tobytes() + L.tobytes() + D.tobytes())
        calls_path = f.name
    try:
        r = subprocess.run([PERCALL_BIN, calls_path, "64", "3"], capture_output=True, text=True, timeout=300)
    finally:
        os.unlink(calls_path)
    if r.returncode != 0:
        raise RuntimeError(f"fftahip_percall failed ({r.returncode}): {r.stderr[-2000:]}")
    native = json.loads(r.stdout.strip().splitlines()[-1])

    offs = np.concatenate([[0], np.cumsum(L)[:-1]]).astype(np.int64)
    total = int(L.sum())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    x = torch.randn(total, dtype=torch.complex128, device="cuda", generator=gen)
    y = torch.empty_like(x)
    plans = {}
    for n, d in zip(L.tolist(), D.tolist()):
        if (n, d) not in plans:
            plans[(n, d)] = Plan((n,), 64, None, d)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    xp, yp = x.data_ptr(), y.data_ptr()
    calls = [(plans[(n, d)], (n,), xp + 16 * o, yp + 16 * o) for n, d, o in zip(L.tolist(), D.tolist(), offs.tolist())]

    def run():
        for p, dims, a, b in calls:
            p.execute_device(dims, a, b, sp)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    py_wall = time.perf_counter() - t0
    return {"transforms": int(idx.size), "distinct_plans": native["distinct_plans"],
            "transforms_per_s": native["transforms_per_s"],
            "wall_transforms_per_s": native["wall_transforms_per_s"],
            "python_binding_transforms_per_s": round(idx.size / py_wall, 1),
            "sample": f"{idx.size} of {len(lengths)} transforms of the batched workload (seed {seed}), ComplexF64, "
                      f"one fftahip_execute per transform on one stream from a native loop (3 passes over the "
                      f"sample; device time between events, wall time with issue)"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "line_pass_batched_uneven_c32_traffic.json")
# radix schedules tuned on an MI355X for this workload and strided tile shapes
# for the single-transform configurations (python bench.py --autotune)
TUNING_DB = os.path.join(ROOT, "ffta.jl_amd", "tuning", "mi355x_batched_uneven_c32.txt")


def measured_traffic(desc):
    """HBM bytes per step from the committed rocprofv3 --pmc measurement
    (FETCH_SIZE + WRITE_SIZE over one step's launches, scaled by the same
    counters' reading for a 1 GiB copy), if it was taken on this schedule."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("launches_per_step") != len(desc.get("phases", [])):
        return None
    return t.get("bytes_per_step")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--eager", action="store_true",
                    help="time the eager launch DAG on four streams instead of replays of the batch's hipGraph (the "
                         "same DAG as kernel nodes and edges); the other mode is timed beside it")
    ap.add_argument("--serve-seconds", type=float, default=0.5, help="length of the serving replay (0: skip)")
    ap.add_argument("--no-configs", action="store_true", help="skip the single-transform configurations")
    ap.add_argument("--no-percall", action="store_true", help="skip the per-call FP64 comparison")
    ap.add_argument("--only-configs", action="store_true",
                    help="run only the single-transform configurations (profiling their kernels)")
    ap.add_argument("--config-reps", type=int, default=0, help="timed calls per configuration (0: about 200 ms)")
    ap.add_argument("--config", action="append", default=None,
                    help="restrict --only-configs to a shape:precision, e.g. 4096x4096:32 (repeatable)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the workload's in-LDS radix schedules on this device, write TUNING_DB and exit")
    ap.add_argument("--autotune-configs", action="store_true",
                    help="tune the strided tile shapes of the single-transform configurations on this device "
                         "(keeping TUNING_DB's radix schedules), write TUNING_DB and exit")
    ap.add_argument("--no-tuning", action="store_true", help="ignore TUNING_DB (default radix rules and tiles)")
    ap.add_argument("--cutoff", type=int, default=73,
                    help="BLUESTEIN_CUTOFF of the batch (FFTA's keyword; default 73 as FFTA)")
    ap.add_argument("--probe-copy", action="store_true",
                    help="after the steps, run one line pass of known bytes (256 MiB each way; calibrates the HBM "
                         "byte counters under rocprofv3 --pmc)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(independent transforms shard across GPUs with no exchange)", file=sys.stderr)
        return 2

    import torch

    from fftahip import autotune
    from fftahip.plan import Batch

    tuned_entries = 0
    if args.autotune or args.autotune_configs:
        t0 = time.perf_counter()
        res = {}
        if args.autotune:
            lengths, dirs = batched_uneven(args.count)
            res = autotune.tune_batch(lengths, dirs, 32)
        elif os.path.exists(TUNING_DB):
            autotune.load_table(TUNING_DB)
        # strided tile shapes of the single-transform configurations
        res["config_tiles"] = {}
        for shape, prec in sorted({(shape, prec) for _, shape, prec, _, _ in PER_CALL_CONFIGS}):
            res["config_tiles"][f"{'x'.join(map(str, shape))}:{prec}"] = [
                autotune.tune_shape(shape, prec), [autotune.tile(prec, n) for n in shape]]
        autotune.save(TUNING_DB)
        print(json.dumps({"autotune": res, "seconds": round(time.perf_counter() - t0, 1), "db": TUNING_DB}),
              flush=True)
        return 0
    if not args.no_tuning and os.path.exists(TUNING_DB):
        tuned_entries = autotune.load_table(TUNING_DB)
    if args.only_configs:
        rows = per_call_configs(torch, cpu=not args.no_cpu, reps_fixed=args.config_reps, only=args.config)
        print(json.dumps({"per_call_configs": rows}), flush=True)
        return 0
    lengths, dirs = batched_uneven(args.count)
    t0 = time.perf_counter()
    batch = Batch(lengths, dirs, precision=32, cutoff=args.cutoff)
    plan_s = time.perf_counter() - t0
    total = int(lengths.sum())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    x = torch.randn(total, dtype=torch.complex64, device="cuda", generator=gen)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def timed(graph: bool, steps: int, warmup: int) -> float:
        for _ in range(warmup):
            batch.execute_device(x.data_ptr(), y.data_ptr(), sp, graph=graph)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            batch.execute_device(x.data_ptr(), y.data_ptr(), sp, graph=graph)
        end.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(end) / steps

    # the step: the batch's launch DAG issued on four streams (eager), or with
    # --graph the captured hipGraph; the other mode is timed beside it
    graph = not args.eager
    ms = timed(graph, args.steps, args.warmup)
    other_ms = timed(not graph, args.steps, args.warmup)

    if args.probe_copy:
        # a launch of known HBM bytes: one in-LDS line pass over 32768
        # transforms of length 1024 (256 MiB read, 256 MiB written, plus a
        # 16 KiB twiddle table)
        kb = Batch(np.full(32768, 1024, dtype=np.int64), np.full(32768, -1, dtype=np.int32), precision=32)
        a = torch.randn(32768 * 1024, dtype=torch.complex64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        kb.execute_device(a.data_ptr(), b.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        del a, b, kb

    device = device_report(torch, stream)
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
            "launch": "hipGraph of the launch DAG" if graph else "launch DAG on 4 streams (eager)",
            ("eager_dag_ms_per_step" if graph else "graph_ms_per_step"): round(other_ms, 4),
            "autotuned_radix_schedules": tuned_entries,
            "bluestein_cutoff": args.cutoff,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(desc),
        },
        "error": {"worst_relerr_over_bound": round(worst, 4), "sample": nsample,
                  "bound": "2e-6*log2(N) vs FP64 oracle"},
        "plan_seconds": round(plan_s, 2),
        "device": device,
    }
    if args.serve_seconds > 0:
        from fftahip import Server

        # three replays of the trace (seeds 0, 1, 2) on one warm server:
        # p50/p99 are reported per replay and as their medians
        srv = Server(pool_bytes=64 << 30, nstreams=4)
        runs = [run_serving(torch, srv, rate=50000.0, seconds=args.serve_seconds, seed=sd) for sd in range(3)]
        srv.close()
        med = lambda k: round(float(np.median([r[k] for r in runs])), 4)  # noqa: E731
        line["serving"] = {
            "workload": "Poisson 50k req/s: 70% 1-D pow2 64..4096, 20% composite 1e3..1e6, 5% prime <= 2^20, "
                        "5% 2-D squares 256..2048; 80% ComplexF32; device-resident payloads; 64 GB pool; "
                        "3 replays of the trace (seeds 0-2)",
            "requests": sum(r["requests"] for r in runs), "completed": sum(r["completed"] for r in runs),
            "rejected": sum(r["rejected"] for r in runs),
            "p50_ms": med("p50_ms"), "p99_ms": med("p99_ms"),
            "p50_ms_runs": [round(r["p50_ms"], 4) for r in runs], "p99_ms_runs": [round(r["p99_ms"], 4) for r in runs],
            "max_ms": round(max(r["max_ms"] for r in runs), 4),
            "throughput_req_s": med("requests_per_s"), "payload_GBps": med("payload_gbps"),
            "coalesced_requests": sum(r["coalesced"] for r in runs), "coalesced_groups": sum(r["groups"] for r in runs),
            "pool_peak_bytes": max(r["pool_peak_bytes"] for r in runs),
        }
    if not args.no_percall:
        pc = per_call_fp64(torch, lengths, dirs)
        batched_tps = args.count / (ms * 1e-3)
        pc["batched_fp32_transforms_per_s"] = round(batched_tps, 1)
        pc["batched_fp32_over_percall_fp64"] = round(batched_tps / pc["transforms_per_s"], 2)
        line["percall_fp64"] = pc
    if not args.no_configs:
        line["per_call_configs"] = per_call_configs(torch, cpu=not args.no_cpu)
        line["per_call_cpu"] = {"kind": "port", "cores": 1, "threads": "Julia threads: n/a (FFTA is single-threaded)",
                                "what": "oracle (C++ restatement of FFTA's algorithm) in the plan's precision, "
                                        "median of up to 20 calls within ~3 s"}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lengths, dirs)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
