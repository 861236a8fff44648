"""GPU parity of the HIP transform path against the oracle (FFTA's algorithm on the CPU).

Tolerances (relative L2 error against the FP64 oracle on the same inputs):
  ComplexF64: <= 1e-14 * log2 N;  ComplexF32: <= 2e-6 * log2 N;  N = 1 exact.
For a CF32 result the oracle runs in FP64 on the CF32 inputs (the reference's
Float64 path is the bar, as the reference's own accuracy test does,
test/onedim/accuracy.jl:48-51).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import tol
from gpu_util import empty_device, to_device, to_host

pytestmark = pytest.mark.gpu

SIZES_1D = [1, 2, 3, 4, 5, 7, 8, 9, 11, 12, 13, 15, 16, 17, 27, 32, 49, 64, 73, 100, 103, 121, 128, 169, 243, 256,
            275, 729, 1000, 1024, 2187, 3000, 4096, 5329, 8192, 10007, 16384, 65536, 177147, 262144]


def _rand(n_or_shape, dtype, seed):
    rng = np.random.default_rng(seed)
    shape = (n_or_shape,) if isinstance(n_or_shape, int) else tuple(n_or_shape)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return np.asfortranarray(x.astype(dtype))


def _gpu_transform(torch, x, direction, region=None):
    from fftahip.plan import Plan, mul_

    prec = 32 if x.dtype == np.complex64 else 64
    p = Plan(x.shape, prec, region, direction)
    tx, dx = to_device(torch, x)
    ty, dy = empty_device(torch, x.shape, x.dtype)
    mul_(dy, p, dx)
    return to_host(torch, ty, x.shape)


def _oracle(x, direction, region=None):
    x64 = np.asfortranarray(x.astype(np.complex128))
    if x.ndim == 1 and region is None:
        return oracle.fft(x64, direction)
    return oracle.fftn(x64, region, direction)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("direction", [-1, 1])
@pytest.mark.parametrize("n", SIZES_1D)
def test_1d_vs_oracle(torch_cuda, n, direction, precision):
    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(n, dt, seed=n * 7 + (direction > 0))
    print(f"case 1d n={n} dir={direction} prec={precision}", flush=True)
    y = _gpu_transform(torch_cuda, x, direction)
    ref = _oracle(x, direction)
    err = oracle.relerr(y, ref)
    assert err <= tol(n, precision), f"n={n} dir={direction} CF{precision}: relerr {err:.3e} > {tol(n, precision):.3e}"


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("shape", [(8, 11), (11, 8), (64, 64), (65, 64), (100, 103), (27, 128), (256, 243),
                                   (1, 5), (5, 1), (300, 218)])
def test_2d_vs_oracle(torch_cuda, shape, precision):
    dt = np.complex64 if precision == 32 else np.complex128
    for direction in (-1, 1):
        x = _rand(shape, dt, seed=sum(shape) + direction)
        print(f"case 2d shape={shape} dir={direction} prec={precision}", flush=True)
        y = _gpu_transform(torch_cuda, x, direction)
        ref = _oracle(x, direction)
        n = int(np.prod(shape))
        err = oracle.relerr(y, ref)
        assert err <= tol(n, precision), f"{shape} dir={direction}: relerr {err:.3e}"


@pytest.mark.parametrize("shape,region", [((5, 6, 7), 1), ((5, 6, 7), 2), ((5, 6, 7), 3), ((5, 6, 7), (1, 3)),
                                          ((5, 6, 7), (2, 3)), ((4, 14, 9), None), ((103, 5, 13), None),
                                          ((26, 33, 35, 4), (2, 4)), ((3, 3, 3, 3, 3), None)])
def test_nd_regions_vs_oracle(torch_cuda, shape, region):
    x = _rand(shape, np.complex128, seed=len(shape))
    for direction in (-1, 1):
        print(f"case nd shape={shape} region={region} dir={direction}", flush=True)
        y = _gpu_transform(torch_cuda, x, direction, region)
        ref = _oracle(x, direction, region)
        reg = (region,) if isinstance(region, int) else (region or tuple(range(1, len(shape) + 1)))
        n = int(np.prod([shape[r - 1] for r in reg]))
        assert oracle.relerr(y, ref) <= tol(n, 64)


def test_ones_delta_closed_form(torch_cuda):
    # test/onedim/complex_forward.jl:3-12 and complex_backward.jl:4-11
    for n in [8, 11, 15, 16, 27, 100, 103, 275]:
        for direction in (-1, 1):
            x = np.ones(n, dtype=np.complex128)
            print(f"case ones n={n} dir={direction}", flush=True)
            y = _gpu_transform(torch_cuda, x, direction)
            ref = np.zeros(n, dtype=np.complex128)
            ref[0] = n
            assert np.max(np.abs(y - ref)) <= 1e-12


def test_ndim_333_table(torch_cuda):
    # test/ndim/minimal_complex.jl:9-26 (delta at (2,2,2), 1-based)
    import os

    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "ndim_333.npz"))["y_ref"]
    y = np.zeros((3, 3, 3), dtype=np.complex128, order="F")
    y[1, 1, 1] = 1
    print("case ndim_333", flush=True)
    out = _gpu_transform(torch_cuda, y, -1)
    assert np.max(np.abs(out - ref)) <= 1e-14


# ---------------------------------------------------------------- scaled plans (ifft)

@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("shape,region", [((4096,), None), ((1009,), None), ((20011,), None), ((3199,), None),
                                          ((12297,), None), ((64, 80), None), ((5, 6, 7), (1, 3)),
                                          ((3000, 9), 2), ((1,), None)])
def test_plan_ifft_fused_scale(torch_cuda, shape, region, precision):
    """plan_ifft = plan_bfft with 1/prod(size(x, region)) applied by the pass
    that writes the result (AbstractFFTs ScaledPlan semantics): equals the
    oracle's bfft scaled by 1/N, and ifft(fft(x)) returns x."""
    from fftahip.plan import Plan, mul_, plan_ifft

    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(shape, dt, seed=sum(shape) + precision)
    print(f"case ifft shape={shape} region={region} prec={precision}", flush=True)
    p = plan_ifft(x, region)
    tx, dx = to_device(torch_cuda, x)
    ty, dy = empty_device(torch_cuda, x.shape, x.dtype)
    mul_(dy, p, dx)
    y = to_host(torch_cuda, ty, x.shape)
    reg = region if region is not None else tuple(range(1, len(shape) + 1))
    reg = (reg,) if isinstance(reg, int) else reg
    n = int(np.prod([shape[r - 1] for r in reg]))
    ref = _oracle(x, 1, region) / n
    assert oracle.relerr(y, ref) <= tol(n, precision)
    # round trip through the forward plan
    f = Plan(x.shape, precision, region, -1)
    tz, dz = empty_device(torch_cuda, x.shape, x.dtype)
    mul_(dz, f, dx)
    mul_(dy, p, dz)
    back = to_host(torch_cuda, ty, x.shape)
    assert oracle.relerr(back, x.astype(np.complex128)) <= 2 * tol(n, precision)


def test_set_scale_changes_result(torch_cuda):
    """A scale factor set after a shape has run rebuilds that shape's schedule."""
    from fftahip.plan import Plan, mul_
    from fftahip import _lib

    x = _rand((4099,), np.complex64, seed=3)
    p = Plan(x.shape, 32, None, -1)
    tx, dx = to_device(torch_cuda, x)
    ty, dy = empty_device(torch_cuda, x.shape, x.dtype)
    mul_(dy, p, dx)
    y1 = to_host(torch_cuda, ty, x.shape)
    _lib.load().fftahip_plan_set_scale(p.handle, 0.25)
    mul_(dy, p, dx)
    y2 = to_host(torch_cuda, ty, x.shape)
    assert oracle.relerr(y2, 0.25 * y1.astype(np.complex128)) <= 1e-7


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n", [20011, 62539])
@pytest.mark.parametrize("setting", ["FFTAHIP_CHIRP_CALC=0", "FFTAHIP_BLUE_COST=0", "FFTAHIP_ROWSTEP=0"])
def test_planner_settings_vs_oracle(torch_cuda, monkeypatch, n, precision, setting):
    """The planner's alternative paths kept behind environment settings (chirp
    table reads in the register passes, the first splittable Bluestein
    length, the flat element mapping of row phases) still match the oracle."""
    name, value = setting.split("=")
    monkeypatch.setenv(name, value)
    dtype = np.complex64 if precision == 32 else np.complex128
    for direction in (-1, 1):
        x = _rand(n, dtype, seed=n + direction)
        y = _gpu_transform(torch_cuda, x, direction)
        assert oracle.relerr(y, _oracle(x, direction)) <= tol(n, precision)


CUTOFF_CASES = [
    # (length, cutoff): primes below the cutoff run as direct DFT stages,
    # primes at or above it through a convolution (src/callgraph.jl:95-100)
    (5, 2), (5, 7), (13, 11), (13, 73), (17, 17), (17, 18), (71, 73), (71, 71), (101, 73), (101, 200),
    (199 * 3, 200), (199 * 3, 199), (19 * 16 * 7, 19), (19 * 16 * 7, 20), (997, 1000), (1009 * 2, 2000), (743, 800), (743 * 4, 744),
    (73 * 73, 73), (73 * 73, 74), (4096, 2),
    # a direct prime above the strided-pass cap, and in FP64 a line too long
    # for one workgroup with its staged roots: a pass of its own
    (743 * 12, 752), (751 * 8, 752),
]


def _prime_factors(n):
    f, p = [], 2
    while p * p <= n:
        while n % p == 0:
            f.append(p)
            n //= p
        p += 1
    if n > 1:
        f.append(n)
    return f


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("n,cutoff", CUTOFF_CASES)
def test_bluestein_cutoff_on_device(torch_cuda, n, cutoff, precision):
    """BLUESTEIN_CUTOFF (src/plan.jl:81, src/callgraph.jl:44) decides on the
    device as in FFTA: a length whose primes are all below the cutoff (or 2,
    3) runs as plain Stockham passes with direct DFT stages; any other prime
    goes through a convolution.  The result matches the oracle run with the
    same cutoff, and FFTA's own tree for that cutoff is reported."""
    from fftahip.plan import Plan, mul_

    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(n, dt, seed=n + cutoff)
    print(f"case cutoff n={n} cutoff={cutoff} prec={precision}", flush=True)
    for direction in (-1, 1):
        p = Plan((n,), precision, None, direction, cutoff=cutoff)
        d = p.describe()["dims"][0]
        # the direct-DFT stage takes primes up to 751
        direct = all(q <= 3 or (q < cutoff and q <= 751) for q in _prime_factors(n))
        kind = d["device"]["contiguous"]["kind"]
        assert kind.startswith("stockham") == direct, (n, cutoff, kind)
        # every prime factor >= 5 ends as a prime leaf of FFTA's tree
        tree_types = {q["type"] for q in d["ffta"]}
        assert ("BLUESTEIN" in tree_types) == any(q >= 5 and q >= cutoff for q in _prime_factors(n))
        tx, dx = to_device(torch_cuda, x)
        ty, dy = empty_device(torch_cuda, x.shape, dt)
        mul_(dy, p, dx)
        y = to_host(torch_cuda, ty, x.shape)
        ref = oracle.fft(x.astype(np.complex128), direction, cutoff)
        err = oracle.relerr(y, ref)
        if err > tol(n, precision):
            # FFTA's DFT leaf steps its twiddles with Singleton's recurrence
            # (src/algos.jl:150-159), whose FP64 error grows with the prime:
            # near P = 1000 the oracle itself is ~1.1e-13 from the truth, above
            # the tolerance.  The bar is then the long-double truth, and the
            # gap to the oracle is bounded by the oracle's own error.
            truth = oracle.truth(x.astype(np.complex128), direction)
            own = oracle.relerr(ref, truth)
            print(f"  vs oracle {err:.3e} > tol; oracle vs truth {own:.3e}", flush=True)
            assert oracle.relerr(y, truth) <= tol(n, precision), (n, cutoff, direction)
            assert err <= tol(n, precision) + own, (n, cutoff, direction, err, own)


@pytest.mark.parametrize("cutoff", [2, 30, 200])
def test_bluestein_cutoff_batch(torch_cuda, cutoff):
    """The batch API takes the same cutoff."""
    from fftahip.plan import Batch

    lengths = np.array([5, 17, 29 * 8, 31, 101, 199, 4096, 6561, 1009, 13 * 17 * 19])
    dirs = np.where(np.arange(lengths.size) % 2 == 0, -1, 1)
    b = Batch(lengths, dirs, 32, cutoff=cutoff)
    rng = np.random.default_rng(cutoff)
    x = (rng.standard_normal(b.total) + 1j * rng.standard_normal(b.total)).astype(np.complex64)
    tx, _ = to_device(torch_cuda, x)
    ty, _ = empty_device(torch_cuda, (b.total,), np.complex64)
    print(f"case cutoff batch {cutoff}", flush=True)
    b.execute_device(tx.data_ptr(), ty.data_ptr(), torch_cuda.cuda.current_stream().cuda_stream)
    y = to_host(torch_cuda, ty, (b.total,))
    for n, d, o in zip(b.lengths, dirs, b.offsets):
        ref = oracle.fft(x[o:o + n].astype(np.complex128), int(d), cutoff)
        assert oracle.relerr(y[o:o + n], ref) <= tol(int(n), 32), (int(n), cutoff)


RADER_SIZES = [79, 101, 103, 139, 257, 641, 769, 1009, 1153, 3457, 4481, 7681, 8191, 12289, 97 * 8, 101 * 27,
               257 * 64, 3457 * 9, 12289 * 4, 139 * 35]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("direction", [-1, 1])
@pytest.mark.parametrize("n", RADER_SIZES)
def test_rader_vs_oracle(torch_cuda, n, direction, precision):
    """Lengths whose convolution prime P has a smooth P - 1 run Rader's
    algorithm on the device (FFTA runs Bluestein there, src/algos.jl:389-424;
    both compute the same DFT)."""
    from fftahip.plan import Plan

    dt = np.complex64 if precision == 32 else np.complex128
    x = _rand(n, dt, seed=n * 3 + (direction > 0))
    p = Plan((n,), precision, None, direction)
    kind = p.describe()["dims"][0]["device"]["contiguous"]["kind"]
    print(f"case rader n={n} dir={direction} prec={precision} kind={kind}", flush=True)
    y = _gpu_transform(torch_cuda, x, direction)
    err = oracle.relerr(y, _oracle(x, direction))
    assert err <= tol(n, precision), f"n={n} {kind}: relerr {err:.3e}"
