"""GPU tests of the drop-in paths a Julia caller takes.

* `p * x` and `mul!(y, p, x)` on host arrays go through fftahip_execute_host
  (what FFTAHip.jl's `*`/`mul!` call for Julia `Array`s), 1-D and 2-D, both
  precisions and directions, checked against the FP64 oracle.
* The reference's bitwise invariants (test/onedim/complex_forward.jl:4-12,
  27-33; test/twodim/complex_forward.jl:29-43): a 1-D plan applied along any
  dimension of an N-D array gives exactly (==) the 1-D transform of each
  slice, and an N-D plan over a sub-region gives exactly the N-D transform of
  each sub-array.  The device keeps this because a line's arithmetic depends
  only on its length's radix schedule and stage twiddles, not on where the
  line sits in memory: lines that fit one in-LDS pass take the same pass
  whatever their stride.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import tol
from gpu_util import empty_device, to_device, to_host

pytestmark = pytest.mark.gpu


def _rand(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return np.asfortranarray(x.astype(dtype))


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("shape", [(4096,), (1000,), (10007,), (64, 80), (256, 243), (33, 17, 5)])
def test_host_star_and_mul(torch_cuda, shape, precision):
    """p * x and mul!(y, p, x) on host (numpy, column-major) arrays."""
    from fftahip import plan as P

    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(shape, dt, seed=sum(shape))
    n = int(np.prod(shape))
    for direction, mk in ((-1, P.plan_fft), (1, P.plan_bfft)):
        print(f"case host p*x shape={shape} prec={precision} dir={direction}", flush=True)
        p = mk(x)
        y = p * x
        assert y.dtype == dt and y.shape == x.shape and y.flags.f_contiguous
        ref = oracle.fftn(x.astype(np.complex128), None, direction) if len(shape) > 1 else \
            oracle.fft(x.astype(np.complex128), direction)
        assert oracle.relerr(y, ref) <= tol(n, precision)
        y2 = np.empty_like(x, order="F")
        P.mul_(y2, p, x)
        assert np.array_equal(y2, y)


@pytest.mark.parametrize("precision", [32, 64])
def test_host_1d_plan_on_nd_array(torch_cuda, precision):
    """fft(x, r) on a host N-D array (the 1-D plan on an N-D array, src/plan.jl:170-207)."""
    from fftahip import plan as P

    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand((12, 13, 14), dt, seed=5)
    for r in (1, 2, 3):
        print(f"case host fft(x, {r}) prec={precision}", flush=True)
        y = P.fft(x, r)
        ref = oracle.fftn(x.astype(np.complex128), (r,), -1)
        assert oracle.relerr(y, ref) <= tol(x.shape[r - 1], precision)


@pytest.mark.parametrize("N", [8, 11, 15, 16, 27, 100, 275])
def test_reshape_equalities(torch_cuda, N):
    """test/onedim/complex_forward.jl:4-12: fft(ones(N)) == N*delta (approx) and
    y == fft(reshape(x,1,1,N),3)[:] == fft(reshape(x,N,1),1)[:] exactly."""
    from fftahip import plan as P

    x = np.ones(N, dtype=np.complex128)
    print(f"case reshape equalities N={N}", flush=True)
    y = P.fft(x)
    ref = np.zeros(N, dtype=np.complex128)
    ref[0] = N
    assert np.max(np.abs(y - ref)) <= 1e-12
    assert np.array_equal(y, P.fft(x.reshape((1, 1, N), order="F"), 3).ravel(order="F"))
    assert np.array_equal(y, P.fft(x.reshape((N, 1), order="F"), 1).ravel(order="F"))
    xr = _rand((N,), np.complex128, seed=N)
    yr = P.fft(xr)
    assert np.array_equal(yr, P.fft(xr.reshape((1, 1, N), order="F"), 3).ravel(order="F"))
    assert np.array_equal(yr, P.fft(xr.reshape((N, 1), order="F"), 1).ravel(order="F"))


def _device_fft(torch, x, region, direction=-1):
    from fftahip.plan import Plan, mul_

    prec = 32 if x.dtype == np.complex64 else 64
    p = Plan(x.shape, prec, region, direction)
    tx, dx = to_device(torch, x)
    ty, dy = empty_device(torch, x.shape, x.dtype)
    mul_(dy, p, dx)
    return to_host(torch, ty, x.shape)


def _slices_as_columns(x, r):
    """The 1-D slices of x along dimension r (1-based) as the columns of a
    column-major (n_r, K) matrix, in mapslices order, and its inverse."""
    a = np.moveaxis(x, r - 1, 0)
    cols = np.asfortranarray(a.reshape((x.shape[r - 1], -1), order="F"))

    def back(c):
        return np.asfortranarray(np.moveaxis(c.reshape(a.shape, order="F"), 0, r - 1))

    return cols, back


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 16, 17, 24, 31, 37, 47, 61, 64])
def test_mapslices_1d_bitwise(torch_cuda, n, precision):
    """test/onedim/complex_forward.jl:27-33: fft(x, r) == mapslices(fft, x; dims=r)
    for x of size (n, n+1, n+2) -- exact equality, as the reference tests it.
    mapslices runs the 1-D transform on each slice as a separate vector; here
    every slice goes through the 1-D plan as the columns of one contiguous
    matrix (the same single-vector line pass), and a few slices additionally
    one vector at a time."""
    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand((n, n + 1, n + 2), dt, seed=100 + n)
    rng = np.random.default_rng(n)
    for r in (1, 2, 3):
        print(f"case mapslices n={n} r={r} prec={precision}", flush=True)
        y = _device_fft(torch_cuda, x, r)
        cols, back = _slices_as_columns(x, r)
        ycols = _device_fft(torch_cuda, cols, 1)
        assert np.array_equal(y, back(ycols)), f"fft(x, {r}) != mapslices over columns"
        for k in rng.choice(cols.shape[1], size=min(4, cols.shape[1]), replace=False):
            v = np.ascontiguousarray(cols[:, k])
            assert np.array_equal(_device_fft(torch_cuda, v, 1), ycols[:, k]), f"slice {k}"
        ref = oracle.fftn(x.astype(np.complex128), (r,), -1)
        assert oracle.relerr(y, ref) <= tol(x.shape[r - 1], precision)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n", [1, 4, 9, 16, 23, 32])
@pytest.mark.parametrize("r", [(1, 2), (1, 3), (2, 3)])
def test_mapslices_2d_bitwise(torch_cuda, n, r, precision):
    """test/twodim/complex_forward.jl:29-43: fft(x, r) == mapslices(fft, x; dims=r)
    for a 2-D region r of x of size (n, n+1, n+2): every 2-D sub-array is
    transformed on its own as a contiguous matrix (a few picked at random)."""
    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand((n, n + 1, n + 2), dt, seed=200 + n)
    print(f"case mapslices-2d n={n} r={r} prec={precision}", flush=True)
    y = _device_fft(torch_cuda, x, r)
    other = ({1, 2, 3} - set(r)).pop()
    rng = np.random.default_rng(n + other)
    for k in rng.choice(x.shape[other - 1], size=min(3, x.shape[other - 1]), replace=False):
        idx = [slice(None)] * 3
        idx[other - 1] = k
        sub = np.asfortranarray(x[tuple(idx)])
        ysub = _device_fft(torch_cuda, sub, None)
        assert np.array_equal(y[tuple(idx)], ysub), f"sub-array {k} along dim {other}"
    ref = oracle.fftn(x.astype(np.complex128), r, -1)
    assert oracle.relerr(y, ref) <= tol(x.shape[r[0] - 1] * x.shape[r[1] - 1], precision)
