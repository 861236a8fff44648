"""GPU tests of the real transforms (plan_rfft / plan_brfft, src/plan.jl:102-150
and 386-583) against the oracle's complex transform of the same data, with
the reference's own test cases (test/onedim/real_*.jl, test/twodim/real_*.jl):
closed forms, every length 1..64, round trips through irfft, N-D arrays along
each dimension and pairs of dimensions, and the error messages.

Tolerances: the complex-transform ones (CF64 1e-14 log2 N, CF32 2e-6 log2 N,
relative L2 against the FP64 oracle)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import tol

pytestmark = pytest.mark.gpu


def _ref_rfft(x, region):
    """The oracle's complex transform of the real data, first region dim cut to n/2+1."""
    xc = np.asfortranarray(x.astype(np.complex128))
    reg = (region,) if isinstance(region, int) else tuple(region)
    y = oracle.fftn(xc, reg, -1) if (xc.ndim > 1 or len(reg) > 1 or reg != (1,)) else oracle.fft(xc, -1)
    d1 = reg[0] - 1
    h = x.shape[d1] // 2 + 1
    return np.take(y, np.arange(h), axis=d1)


def _ref_brfft(y, n, region):
    """brfft by the Hermitian extension along the first region dim, then the oracle's bfft, real part."""
    reg = (region,) if isinstance(region, int) else tuple(region)
    d1 = reg[0] - 1
    h = y.shape[d1]
    idx = np.array([t if t < h else n - t for t in range(n)])
    full = np.take(y.astype(np.complex128), idx, axis=d1)
    conj_mask = np.array([t >= h for t in range(n)])
    shape = [1] * y.ndim
    shape[d1] = n
    full = np.where(conj_mask.reshape(shape), np.conj(full), full)
    if len(reg) == 2:
        # the Hermitian extension along d1 also reverses d2 for the conjugated half
        d2 = reg[1] - 1
        rev = np.take(full, (-np.arange(full.shape[d2])) % full.shape[d2], axis=d2)
        full = np.where(conj_mask.reshape(shape), rev, full)
    out = oracle.fftn(np.asfortranarray(full), reg, 1)
    return out.real


def _rand_real(shape, dt, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape).astype(dt))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", list(range(1, 65)) + [100, 1000, 4096, 4099, 8192, 20011, 65536, 100000])
def test_rfft_1d(torch_cuda, n, dt):
    from fftahip import rfft

    x = _rand_real((n,), dt, seed=n)
    print(f"case rfft n={n} {np.dtype(dt)}", flush=True)
    y = rfft(x)
    assert y.shape == (n // 2 + 1,)
    ref = _ref_rfft(x, 1)
    prec = 32 if dt == np.float32 else 64
    assert oracle.relerr(y, ref) <= tol(n, prec)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", list(range(1, 65)) + [100, 1000, 4096, 4099, 8192, 20011, 65536, 100000])
def test_brfft_irfft_1d(torch_cuda, n, dt):
    from fftahip import brfft, irfft, rfft

    x = _rand_real((n,), dt, seed=n + 1)
    print(f"case brfft n={n} {np.dtype(dt)}", flush=True)
    y = rfft(x)
    z = brfft(y, n)
    assert z.shape == (n,) and z.dtype == dt
    prec = 32 if dt == np.float32 else 64
    ref = _ref_brfft(y, n, 1)
    assert oracle.relerr(z, ref) <= tol(n, prec)
    back = irfft(y, n)
    assert oracle.relerr(back, x.astype(np.float64)) <= 2 * tol(n, prec)


@pytest.mark.parametrize("N", [8, 11, 15, 16, 27, 100])
def test_closed_forms(torch_cuda, N):
    """test/onedim/real_forward.jl:3-11, real_backward.jl:3-13, twodim/real_*.jl:3-12."""
    from fftahip import brfft, rfft

    y = rfft(np.ones(N))
    ref = np.zeros(N // 2 + 1, dtype=complex)
    ref[0] = N
    assert np.allclose(y, ref, atol=1e-12)
    assert np.array_equal(y, rfft(np.ones((1, 1, N)), 3).ravel())
    z = brfft(np.ones(N, dtype=complex), 2 * (N - 1))
    zr = np.zeros(2 * (N - 1))
    zr[0] = 2 * (N - 1)
    assert np.allclose(z, zr, atol=1e-12) and z.dtype == np.float64
    Y = rfft(np.ones((N, N)))
    ref2 = np.zeros((N // 2 + 1, N), dtype=complex)
    ref2[0, 0] = N * N
    assert np.allclose(Y, ref2, atol=1e-10)
    Z = brfft(np.ones((N, N), dtype=complex), 2 * (N - 1))
    zr2 = np.zeros((2 * (N - 1), N))
    zr2[0, 0] = N * 2 * (N - 1)
    assert np.allclose(Z, zr2, atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 13, 16, 33, 64])
@pytest.mark.parametrize("r", [1, 2, 3])
def test_nd_along_each_dim(torch_cuda, n, r):
    """test/onedim/real_*.jl 1D plan on ND arrays: equals the per-slice transform; round trip."""
    from fftahip import brfft, irfft, rfft

    x = _rand_real((n, n + 1, n + 2), np.float64, seed=10 * n + r)
    print(f"case rfft nd n={n} r={r}", flush=True)
    y = rfft(x, r)
    assert y.shape[r - 1] == x.shape[r - 1] // 2 + 1
    assert oracle.relerr(y, _ref_rfft(x, r)) <= tol(x.shape[r - 1], 64)
    z = brfft(y, x.shape[r - 1], r)
    assert oracle.relerr(z, _ref_brfft(y, x.shape[r - 1], r)) <= tol(x.shape[r - 1], 64)
    back = irfft(y, x.shape[r - 1], r)
    assert oracle.relerr(back, x) <= 2 * tol(x.shape[r - 1], 64)


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (7, 7), (8, 8), (9, 8), (16, 15), (33, 64), (64, 65), (256, 256),
                                   (300, 218), (1024, 1000)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_2d(torch_cuda, shape, dt):
    """test/twodim/real_*.jl: 2D plan on 2D arrays vs the full complex transform; irfft round trip."""
    from fftahip import brfft, irfft, rfft

    x = _rand_real(shape, dt, seed=sum(shape))
    print(f"case rfft2 {shape} {np.dtype(dt)}", flush=True)
    y = rfft(x)
    assert y.shape == (shape[0] // 2 + 1, shape[1])
    prec = 32 if dt == np.float32 else 64
    n = shape[0] * shape[1]
    assert oracle.relerr(y, _ref_rfft(x, (1, 2))) <= tol(n, prec)
    z = brfft(y, shape[0])
    assert oracle.relerr(z, _ref_brfft(y, shape[0], (1, 2))) <= tol(n, prec)
    back = irfft(y, shape[0])
    assert oracle.relerr(back, x.astype(np.float64)) <= 2 * tol(n, prec)


@pytest.mark.parametrize("r", [(1, 2), (1, 3), (2, 3)])
@pytest.mark.parametrize("n", [3, 8, 17])
def test_2d_plan_nd_array(torch_cuda, n, r):
    from fftahip import irfft, rfft

    x = _rand_real((n, n + 1, n + 2), np.float64, seed=n)
    y = rfft(x, r)
    assert oracle.relerr(y, _ref_rfft(x, r)) <= tol(x.size, 64)
    back = irfft(y, x.shape[r[0] - 1], r)
    assert oracle.relerr(back, x) <= 2 * tol(x.size, 64)


def test_errors(torch_cuda):
    from fftahip import ArgumentError, DimensionMismatch, brfft, plan_rfft, rfft

    with pytest.raises(DimensionMismatch):
        rfft(np.zeros(0))
    with pytest.raises(DimensionMismatch):
        brfft(np.zeros(0, dtype=complex), 0)
    with pytest.raises(ArgumentError, match="only supports 1D and 2D FFTs"):
        plan_rfft(np.zeros((2, 3, 4)), (1, 2, 3))
    with pytest.raises(DimensionMismatch, match="should have size"):
        brfft(np.zeros((5, 3), dtype=complex), 20, 1)
