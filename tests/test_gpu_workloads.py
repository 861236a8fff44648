"""GPU tests at the benchmark configurations themselves.

* Batched uneven 1-D: the full 20000-transform ComplexF32 batch (lengths
  log-uniform in [16, 65536], seed 0, 20% primes, 30% powers of 3, fft/bfft
  50/50) with the shipped MI355X tuning table loaded, exactly as bench.py
  runs it.  One eager execution (the launch DAG on four streams) and three
  graph replays; every transform is checked against the FP64 oracle
  (relative L2 error <= 2e-6 log2 N), and every replay must equal the eager
  result bit for bit.
* Serving: the Poisson trace at 50k req/s for 0.5 s (70% 1-D pow2 64..4096,
  20% composite 1e3..1e6, 5% prime <= 2^20, 5% 2-D squares 256..2048; 80%
  ComplexF32) replayed through the request loop with a 64 GB pool.  Every
  request writes its own output buffer; each result must match the first
  result of its (shape, precision, direction) within the tolerance, and that
  one is checked against the FP64 oracle.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from conftest import tol

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _per_transform_relerr(y, ref, lengths, offsets):
    """Relative L2 error of every transform of a packed batch (FP64 sums)."""
    d = np.abs(y.astype(np.complex128) - ref) ** 2
    r = np.abs(ref) ** 2
    num = np.add.reduceat(d, offsets)
    den = np.add.reduceat(r, offsets)
    return np.sqrt(num / np.where(den > 0, den, 1.0))


def test_batched_uneven_config(torch_cuda):
    import bench
    from fftahip import autotune
    from fftahip.plan import Batch

    torch = torch_cuda
    entries = autotune.load_table(bench.TUNING_DB)
    assert entries > 0
    lengths, dirs = bench.batched_uneven(20000)
    print(f"case batched uneven: 20000 transforms, {int(lengths.sum())} elements, tuning entries {entries}",
          flush=True)
    b = Batch(lengths, dirs, precision=32)
    total = int(lengths.sum())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    x = torch.randn(total, dtype=torch.complex64, device="cuda", generator=gen)
    y = torch.empty_like(x)
    sp = torch.cuda.current_stream().cuda_stream
    b.execute_device(x.data_ptr(), y.data_ptr(), sp)
    torch.cuda.synchronize()
    eager = y.clone()
    replays_equal = []
    for _ in range(3):
        y.zero_()
        b.execute_device(x.data_ptr(), y.data_ptr(), sp, graph=True)
        torch.cuda.synchronize()
        replays_equal.append(bool(torch.equal(y, eager)))
    print(f"  graph replays equal to eager: {replays_equal}", flush=True)
    assert all(replays_equal)
    xh = x.cpu().numpy()
    yh = eager.cpu().numpy()
    offs = b.offsets
    nthreads = max(1, min(16, os.cpu_count() or 1))
    ref = oracle.fft_batch(xh.astype(np.complex128), lengths, offs, dirs, nthreads=nthreads)
    err = _per_transform_relerr(yh, ref, lengths, offs)
    bound = 2e-6 * np.log2(lengths.astype(np.float64))
    worst = int(np.argmax(err / bound))
    print(f"  worst relerr/bound {err[worst] / bound[worst]:.4f} at N={int(lengths[worst])}", flush=True)
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, [(int(lengths[i]), int(dirs[i]), float(err[i])) for i in bad[:10]]


def test_serving_config_replay(torch_cuda):
    import bench
    from fftahip import Server

    torch = torch_cuda
    rate, seconds, seed = 50000.0, 0.5, 0
    t, shapes, prec, dirs = bench.serving_trace(rate, seconds, seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    inputs = {}
    for sh, p in sorted(set(zip(shapes, prec.tolist()))):
        n = int(np.prod(sh))
        dt = torch.complex64 if p == 32 else torch.complex128
        inputs[(sh, p)] = torch.randn(n, dtype=dt, device="cuda", generator=gen)
    outs = [torch.empty_like(inputs[(sh, p)]) for sh, p in zip(shapes, prec.tolist())]
    nbytes = sum(o.numel() * o.element_size() for o in outs)
    print(f"case serving replay: {len(shapes)} requests at {rate:.0f}/s, {nbytes / 1e9:.1f} GB of outputs",
          flush=True)
    srv = Server(pool_bytes=64 << 30, nstreams=4)
    ins = [inputs[(sh, p)].data_ptr() for sh, p in zip(shapes, prec.tolist())]
    dims3 = [list(sh) + [0] * (3 - len(sh)) for sh in shapes]
    tickets = srv.replay(t, prec.tolist(), dirs.tolist(), dims3, ins, [o.data_ptr() for o in outs])
    srv.drain()
    st = srv.stats()
    srv.close()
    print(f"  completed {st['completed']}/{len(tickets)}  p50 {st['p50_ms']:.4f} ms  p99 {st['p99_ms']:.4f} ms  "
          f"{st['requests_per_s']:.0f} req/s", flush=True)
    assert st["completed"] == len(tickets) and st["failed"] == 0 and st["rejected"] == 0
    print(f"  coalesced {st['coalesced']} requests in {st['groups']} groups", flush=True)
    torch.cuda.synchronize()
    # every result against the first result of its (shape, precision,
    # direction), on the device: a request that ran inside a coalesced group
    # may round differently from one that ran alone, within the tolerance
    first = {}
    bad, exact = [], 0
    for i, (sh, p, d) in enumerate(zip(shapes, prec.tolist(), dirs.tolist())):
        key = (sh, p, d)
        if key not in first:
            first[key] = i
            continue
        a, b = outs[i], outs[first[key]]
        if torch.equal(a, b):
            exact += 1
            continue
        rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        if not rel <= tol(int(np.prod(sh)), p):
            bad.append((i, key, rel))
    assert not bad, bad[:10]
    print(f"  {len(first)} distinct (shape, precision, direction); {exact}/{len(shapes) - len(first)} repeats "
          f"bitwise equal to the first, all within tolerance", flush=True)

    def check(key):
        sh, p, d = key
        i = first[key]
        x = inputs[(sh, p)].cpu().numpy().astype(np.complex128)
        yv = outs[i].cpu().numpy()
        if len(sh) == 1:
            ref = oracle.fft(x, d)
        else:
            ref = oracle.fftn(x.reshape(sh, order="F"), None, d).ravel(order="F")
        n = int(np.prod(sh))
        return key, oracle.relerr(yv, ref), tol(n, p)

    with ThreadPoolExecutor(max_workers=16) as ex:
        results = list(ex.map(check, sorted(first)))
    bad = [(k, e, b) for k, e, b in results if not e <= b]
    assert not bad, bad[:10]
