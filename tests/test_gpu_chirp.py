"""Device read-back of the Bluestein chirp (src/algos.jl:353-359): the
values the plans' chirp tables and register-pass chirp steps hold, computed
on the device, against independent truth.

* Small lengths against the 50-digit mpmath fixtures (tests/golden/chirp.npz):
  FP64 within 2 ulp, FP32 equal to the fixture rounded to ComplexF32 (1 ulp).
* Large lengths (up to 2^31 - 1) at the top of the index range, where
  n^2 reaches 2^62: the index p = n^2 mod 2N folded into (-N, N] is computed
  here with Python integers (exact) and the value with numpy's FP64
  cos/sin; a wrong reduction shows as an O(1) error, the evaluation error of
  both sides is a few ulp.
* Register-pass chirp steps d[t] = exp(i pi ((t s mod 2N)^2 mod 2N) / N) at
  N = 2^31 - 1 with large steps s, the same way.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chirp.npz")


def _exact(N: int, idx) -> np.ndarray:
    """exp(i pi p / N) with p = n^2 mod 2N folded into (-N, N], exact integer reduction."""
    p = np.array([(int(n) * int(n)) % (2 * N) for n in idx], dtype=object)
    p = np.array([int(v) - 2 * N if int(v) > N else int(v) for v in p], dtype=np.float64)
    return np.cos(np.pi * p / N) + 1j * np.sin(np.pi * p / N)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("N", [73, 103, 1009])
def test_chirp_vs_golden(torch_cuda, N, precision):
    from fftahip.plan import debug_chirp

    gold = np.load(GOLDEN)[f"chirp_{N}_-1"]
    dev = debug_chirp(N, precision)
    print(f"case chirp N={N} prec={precision}", flush=True)
    if precision == 64:
        assert np.max(np.abs(dev - gold)) <= 2 * 2.0 ** -52
    else:
        ref = gold.astype(np.complex64)
        assert np.max(np.abs(dev.astype(np.complex128) - ref.astype(np.complex128))) <= 2.0 ** -24


@pytest.mark.parametrize("N", [1000003, 20011, (1 << 31) - 1, (1 << 31) - 19])
def test_chirp_large_index_exact(torch_cuda, N):
    from fftahip.plan import debug_chirp

    count = 2048
    n0 = N - count
    dev = debug_chirp(N, 64, n0, count)
    ref = _exact(N, range(n0, N))
    err = float(np.max(np.abs(dev - ref)))
    print(f"case chirp N={N} n0={n0}: max abs err {err:.2e}", flush=True)
    # both sides evaluate cos/sin of the same exactly reduced angle in FP64;
    # the argument p/N itself is rounded once on each side
    assert err <= 64 * 2.0 ** -52
    # the middle of the range as well (p folds to negative values there)
    mid = N // 2
    dev2 = debug_chirp(N, 32, mid, 512)
    ref2 = _exact(N, range(mid, mid + 512))
    assert float(np.max(np.abs(dev2.astype(np.complex128) - ref2))) <= 4 * 2.0 ** -24


@pytest.mark.parametrize("N,s", [((1 << 31) - 1, 123456789), ((1 << 31) - 1, (1 << 32) - 7), (1000003, 999983),
                                 (65521, 8192)])
def test_chirp_steps_exact(torch_cuda, N, s):
    from fftahip.plan import debug_chirp_steps

    L = 48
    dev = debug_chirp_steps(N, s, L)
    u = [((t * s) % (2 * N)) for t in range(L)]
    p = [(v * v) % (2 * N) for v in u]
    p = np.array([v - 2 * N if v > N else v for v in p], dtype=np.float64)
    ref = np.cos(np.pi * p / N) + 1j * np.sin(np.pi * p / N)
    err = float(np.max(np.abs(dev - ref)))
    print(f"case chirp steps N={N} s={s}: max abs err {err:.2e}", flush=True)
    assert err <= 64 * 2.0 ** -52
