"""GPU tests of the serving loop: results in a concurrent mix equal the same
transform run alone, host payloads, pool admission/rejection, and a short
Poisson replay."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import tol
from gpu_util import empty_device, to_device, to_host

pytestmark = pytest.mark.gpu

MIX = [((64,), 32, -1), ((4096,), 32, 1), ((1000,), 32, -1), ((3000,), 64, 1), ((1009,), 32, -1),
       ((65537,), 32, 1), ((100000,), 32, -1), ((2 * 4099,), 64, -1), ((256, 256), 32, -1), ((64, 80), 64, 1),
       ((128,), 64, -1), ((1024,), 32, 1), ((10007,), 32, -1), ((177147,), 32, 1)]


def _alone(torch, x, prec, direction):
    from fftahip.plan import Plan, mul_

    p = Plan(x.shape, prec, None, direction)
    tx, dx = to_device(torch, x)
    ty, dy = empty_device(torch, x.shape, x.dtype)
    mul_(dy, p, dx)
    return to_host(torch, ty, x.shape)


def test_mix_equals_alone(torch_cuda):
    from fftahip import Server

    srv = Server(pool_bytes=2 << 30, nstreams=4)
    rng = np.random.default_rng(11)
    reqs = []
    for rep in range(3):
        for shape, prec, direction in MIX:
            dt = np.complex64 if prec == 32 else np.complex128
            x = np.asfortranarray((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt))
            tx, _ = to_device(torch_cuda, x)
            ty, _ = empty_device(torch_cuda, shape, dt)
            reqs.append((x, tx, ty, shape, prec, direction))
    torch_cuda.cuda.synchronize()
    print(f"case serving mix of {len(reqs)} requests", flush=True)
    tickets = [srv.submit(prec, d, shape, tx.data_ptr(), ty.data_ptr()) for (x, tx, ty, shape, prec, d) in reqs]
    for t in tickets:
        srv.wait(t)
    exact = 0
    for (x, tx, ty, shape, prec, d) in reqs:
        y = to_host(torch_cuda, ty, shape)
        alone = _alone(torch_cuda, x, prec, d)
        n = int(np.prod(shape))
        ref = oracle.fftn(x.astype(np.complex128), None, d) if len(shape) > 1 else oracle.fft(x.astype(np.complex128), d)
        assert oracle.relerr(y, ref) <= tol(n, prec), (shape, prec, d)
        assert oracle.relerr(y, alone.astype(np.complex128)) <= tol(n, prec)
        exact += int(np.array_equal(y, alone))
    print(f"  {exact}/{len(reqs)} bitwise identical to the transform run alone", flush=True)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["rejected"] == 0
    srv.close()


def test_host_payloads(torch_cuda):
    from fftahip import Server

    srv = Server(pool_bytes=256 << 20, nstreams=2)
    rng = np.random.default_rng(3)
    jobs = []
    for n in (256, 4096, 3000, 1009):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        y = np.empty_like(x)
        jobs.append((x, y, srv.submit(32, -1, (n,), x.ctypes.data, y.ctypes.data, host=True)))
    print("case serving host payloads", flush=True)
    for x, y, t in jobs:
        srv.wait(t)
        assert oracle.relerr(y, oracle.fft(x.astype(np.complex128))) <= tol(x.size, 32)
    srv.close()


def test_pool_rejects_oversize(torch_cuda):
    from fftahip import FFTAHipError, Server

    srv = Server(pool_bytes=1 << 20, nstreams=1)
    n = 1 << 20
    tx = torch_cuda.zeros(n, dtype=torch_cuda.complex64, device="cuda")
    ty = torch_cuda.empty_like(tx)
    print("case serving oversize rejection", flush=True)
    # a host-payload request stages 2 x 8 MiB through the pool: larger than the pool
    x = np.zeros(n, dtype=np.complex64)
    y = np.empty_like(x)
    t = srv.submit(32, -1, (n,), x.ctypes.data, y.ctypes.data, host=True)
    with pytest.raises(FFTAHipError):
        srv.wait(t)
    # a small device-payload request still runs afterwards
    t2 = srv.submit(32, -1, (64,), tx.data_ptr(), ty.data_ptr())
    srv.wait(t2)
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_replay_short_trace(torch_cuda):
    import bench
    from fftahip import Server

    srv = Server(pool_bytes=4 << 30, nstreams=4)
    print("case serving replay", flush=True)
    res = bench.run_serving(torch_cuda, srv, rate=20000.0, seconds=0.1, seed=5)
    assert res["completed"] == res["requests"] and res["failed"] == 0
    assert 0 < res["p50_ms"] <= res["p99_ms"]
    srv.close()


def test_coalesced_groups_match_alone(torch_cuda):
    """Requests of one 1-D plan that pile up behind others run as coalesced
    groups (one set of work-list launches for up to 32 of them); each result
    matches the oracle and the same transform run alone."""
    from fftahip import Server

    torch = torch_cuda
    srv = Server(pool_bytes=4 << 30, nstreams=2)
    rng = np.random.default_rng(21)
    shapes = [((4096,), 32, -1), ((1000,), 64, 1), ((10007,), 32, 1), ((262144,), 32, -1), ((97 * 8,), 32, -1),
              ((12289,), 64, -1)]
    reqs = []
    # a large 2-D request first keeps the streams busy while the burst queues
    big = np.asfortranarray((rng.standard_normal((2048, 2048)) + 1j * rng.standard_normal((2048, 2048)))
                            .astype(np.complex64))
    tb, _ = to_device(torch, big)
    tby, _ = empty_device(torch, big.shape, np.complex64)
    for rep in range(40):
        for shape, prec, d in shapes:
            dt = np.complex64 if prec == 32 else np.complex128
            x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)
            tx, _ = to_device(torch, x)
            ty, _ = empty_device(torch, shape, dt)
            reqs.append((x, tx, ty, shape, prec, d))
    torch.cuda.synchronize()
    print(f"case serving coalesced burst of {len(reqs)} requests", flush=True)
    tickets = [srv.submit(32, -1, big.shape, tb.data_ptr(), tby.data_ptr())]
    tickets += [srv.submit(prec, d, shape, tx.data_ptr(), ty.data_ptr()) for (x, tx, ty, shape, prec, d) in reqs]
    for tk in tickets:
        srv.wait(tk)
    st = srv.stats()
    print(f"  coalesced {st['coalesced']} requests in {st['groups']} groups", flush=True)
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["groups"] > 0 and st["coalesced"] >= 2 * st["groups"]
    alone = {}
    exact = 0
    for (x, tx, ty, shape, prec, d) in reqs:
        y = to_host(torch, ty, shape)
        n = int(np.prod(shape))
        ref = oracle.fft(x.astype(np.complex128), d)
        assert oracle.relerr(y, ref) <= tol(n, prec), (shape, prec, d)
        key = (shape, prec, d)
        if key not in alone:
            alone[key] = True
        a = _alone(torch, x, prec, d)
        assert oracle.relerr(y, a.astype(np.complex128)) <= tol(n, prec)
        exact += int(np.array_equal(y, a))
    print(f"  {exact}/{len(reqs)} bitwise identical to the transform run alone", flush=True)
    yb = to_host(torch, tby, big.shape)
    assert oracle.relerr(yb, oracle.fftn(big.astype(np.complex128), None, -1)) <= tol(big.size, 32)
    srv.close()
