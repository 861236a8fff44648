"""GPU tests of the batched path: uneven-length batches against the oracle,
hipGraph replays against eager execution, a transform alone versus inside a
mixed batch, and in-place plan execution."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import tol
from gpu_util import empty_device, to_device, to_host

pytestmark = pytest.mark.gpu


def _lengths(seed, count):
    rng = np.random.default_rng(seed)
    base = np.rint(np.exp(rng.uniform(np.log(1), np.log(70000), count))).astype(np.int64)
    special = np.array([1, 2, 3, 16, 17, 97, 243, 4096, 4099, 8192, 8198, 10007, 16384, 17408, 20011, 59049, 65521,
                        65536, 34466, 2 * 4099 * 3])
    return np.concatenate([special, base])


def _run_batch(torch, lengths, dirs, precision, graph=False, reps=1):
    from fftahip.plan import Batch

    dt = np.complex64 if precision == 32 else np.complex128
    b = Batch(lengths, dirs, precision)
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(b.total) + 1j * rng.standard_normal(b.total)).astype(dt)
    tx, _ = to_device(torch, x)
    ty, _ = empty_device(torch, (b.total,), dt)
    outs = []
    for _ in range(reps):
        b.execute_device(tx.data_ptr(), ty.data_ptr(), torch.cuda.current_stream().cuda_stream, graph=graph)
        outs.append(to_host(torch, ty, (b.total,)).copy())
    return b, x, outs


@pytest.mark.parametrize("precision", [32, 64])
def test_batch_vs_oracle(torch_cuda, precision):
    lengths = _lengths(1, 120)
    rng = np.random.default_rng(2)
    dirs = np.where(rng.uniform(size=lengths.size) < 0.5, -1, 1)
    print(f"case batch prec={precision} count={lengths.size} total={lengths.sum()}", flush=True)
    b, x, (y,) = _run_batch(torch_cuda, lengths, dirs, precision)
    for n, d, o in zip(b.lengths, dirs, b.offsets):
        ref = oracle.fft(x[o:o + n].astype(np.complex128), int(d))
        err = oracle.relerr(y[o:o + n], ref)
        assert err <= tol(int(n), precision), (int(n), int(d), err)


def test_batch_graph_replays_match_eager(torch_cuda):
    lengths = _lengths(3, 60)
    dirs = np.where(np.arange(lengths.size) % 2 == 0, -1, 1)
    print("case batch graph replay", flush=True)
    _, _, eager = _run_batch(torch_cuda, lengths, dirs, 32, graph=False, reps=1)
    _, _, replays = _run_batch(torch_cuda, lengths, dirs, 32, graph=True, reps=3)
    for r in replays:
        assert np.array_equal(r, eager[0])


def test_alone_and_in_mix(torch_cuda):
    """A transform's result is the same alone (a plan) and inside a mixed batch."""
    from fftahip.plan import Plan, mul_

    lengths = _lengths(4, 40)
    dirs = np.where(np.arange(lengths.size) % 3 == 0, 1, -1)
    print("case alone vs mix", flush=True)
    b, x, (y,) = _run_batch(torch_cuda, lengths, dirs, 32)
    for i in range(0, lengths.size, 3):
        n, d, o = int(b.lengths[i]), int(dirs[i]), int(b.offsets[i])
        xi = np.ascontiguousarray(x[o:o + n])
        p = Plan((n,), 32, 1, d)
        tx, dx = to_device(torch_cuda, xi)
        ty, dy = empty_device(torch_cuda, (n,), np.complex64)
        mul_(dy, p, dx)
        alone = to_host(torch_cuda, ty, (n,))
        ref = oracle.fft(xi.astype(np.complex128), d)
        assert oracle.relerr(alone, ref) <= tol(n, 32)
        assert oracle.relerr(y[o:o + n], alone.astype(np.complex128)) <= tol(n, 32)


@pytest.mark.parametrize("n", [64, 4096, 65536, 10007, 20011, 8198, 3000])
def test_in_place(torch_cuda, n):
    from fftahip.plan import DeviceArray, Plan, mul_

    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    print(f"case in-place n={n}", flush=True)
    tx, dx = to_device(torch_cuda, x)
    p = Plan((n,), 32, 1, -1)
    mul_(DeviceArray(dx.ptr, (n,), 32), p, dx)
    y = to_host(torch_cuda, tx, (n,))
    assert oracle.relerr(y, oracle.fft(x.astype(np.complex128))) <= tol(n, 32)
