"""GPU parity at the full sizes of the single-transform configurations:
1-D N = 3^11, 2^10*3^5*5^3 = 31104000, 1000003 (prime) and 2-D 4096x4096,
3000x2187, in ComplexF32 and ComplexF64, fft and bfft, against the oracle on
the same inputs (N(0,1) entries), plus a size-independent round trip."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import tol
from gpu_util import empty_device, to_device, to_host

pytestmark = pytest.mark.gpu


def _relerr64(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _rand(shape, dt, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape, dtype=np.float64) + 1j * rng.standard_normal(shape, dtype=np.float64)
    return np.asfortranarray(x.astype(dt))


def _gpu(torch, x, direction):
    from fftahip.plan import Plan, mul_

    prec = 32 if x.dtype == np.complex64 else 64
    p = Plan(x.shape, prec, None, direction)
    tx, dx = to_device(torch, x)
    ty, dy = empty_device(torch, x.shape, x.dtype)
    mul_(dy, p, dx)
    return to_host(torch, ty, x.shape)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n", [177147, 1000003, 31104000])
def test_config_1d(torch_cuda, n, precision):
    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand((n,), dt, seed=n % 1000)
    for direction in (-1, 1):
        print(f"case config 1d n={n} prec={precision} dir={direction}", flush=True)
        y = _gpu(torch_cuda, x, direction)
        ref = oracle.fft(x.astype(np.complex128), direction)
        err = _relerr64(y, ref)
        print(f"  relerr {err:.3e} bound {tol(n, precision):.3e}", flush=True)
        assert err <= tol(n, precision)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("shape", [(4096, 4096), (3000, 2187)])
def test_config_2d(torch_cuda, shape, precision):
    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(shape, dt, seed=shape[0])
    n = shape[0] * shape[1]
    for direction in (-1, 1):
        print(f"case config 2d shape={shape} prec={precision} dir={direction}", flush=True)
        y = _gpu(torch_cuda, x, direction)
        ref = oracle.fftn(x.astype(np.complex128), None, direction)
        err = _relerr64(y, ref)
        print(f"  relerr {err:.3e} bound {tol(n, precision):.3e}", flush=True)
        assert err <= tol(n, precision)


@pytest.mark.parametrize("n", [31104000, 1000003])
def test_round_trip_full_size(torch_cuda, n):
    """bfft(fft(x)) / N == x at the full size (ComplexF32)."""
    x = _rand((n,), np.complex64, seed=5)
    print(f"case round trip n={n}", flush=True)
    y = _gpu(torch_cuda, x, -1)
    z = _gpu(torch_cuda, y, 1) / n
    assert _relerr64(z, x) <= 2 * tol(n, 32)


@pytest.mark.parametrize("table", ["explicit", "shipped"])
@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("shape", [(4096, 4096), (3000, 2187), (31104000,), (177147,), (1000003,)])
def test_config_tuned_tiles(torch_cuda, tmp_path, shape, precision, table):
    """The configuration shapes with non-default strided tile shapes against
    the oracle, both directions: from an explicit table (one strided pass of
    4 lines x 4096 in 140 KB; 8 lines for the others) and from the table
    shipped for MI355X (tuned by bench.py --autotune-configs)."""
    import bench
    from fftahip import autotune

    if table == "explicit":
        src = tmp_path / "tiles.txt"
        src.write_text("".join(f"S {p} {n} 143360 {4 if n == 4096 else 8}\n" for p in (32, 64)
                               for n in (4096, 2187, 3000, 31104000, 177147, 1000003)))
        nent = 12
    else:
        src = bench.TUNING_DB
        nent = None
    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(shape, dt, seed=11)
    n = int(np.prod(shape))
    autotune.clear()
    try:
        got = autotune.load_table(str(src))
        assert got == nent if nent else got > 0
        for direction in (-1, 1):
            print(f"case tuned tiles shape={shape} prec={precision} dir={direction}", flush=True)
            y = _gpu(torch_cuda, x, direction)
            ref = oracle.fftn(x.astype(np.complex128), None, direction) if len(shape) > 1 else \
                oracle.fft(x.astype(np.complex128), direction)
            err = _relerr64(y, ref)
            print(f"  relerr {err:.3e} bound {tol(n, precision):.3e}", flush=True)
            assert err <= tol(n, precision)
    finally:
        autotune.clear()


def test_autotune_shape_then_parity(torch_cuda):
    """fftahip_autotune_shape times the tile candidates of a shape on the
    device; plans built afterwards (whatever it chose) match the oracle."""
    from fftahip import autotune

    shape = (512, 1536)
    x = _rand(shape, np.complex64, seed=3)
    autotune.clear()
    try:
        changed = autotune.tune_shape(shape, 32)
        b, c = autotune.tile(32, 1536)
        print(f"case autotune shape {shape}: changed {changed}, tile of 1536 ({b}, {c})", flush=True)
        assert changed in (0, 2) and (b == 0) == (changed == 0)
        y = _gpu(torch_cuda, x, -1)
        ref = oracle.fftn(x.astype(np.complex128), None, -1)
        assert _relerr64(y, ref) <= tol(x.size, 32)
    finally:
        autotune.clear()
