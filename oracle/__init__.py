"""CPU oracle for the FFTA.jl complex transform path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``liboracle.so`` (built from ``ffta_oracle.cpp`` by the
Makefile beside it), a C++ restatement of FFTA.jl's planner, Singleton
twiddles, radix-4 / radix-3 / DFT / composite / Bluestein kernels and N-D
ordering.  Only the tests, the smoke check and the benchmark's CPU-baseline
leg may import it, and only as the checker.  The product library never calls
it.

Parity status of each function (how it is pinned; see tests/test_oracle.py):
  plan_tree      planner trees at the reference's test and config sizes,
                 hand-derived; invariants (product of leaves, strides).
  fft / fftn     mpmath (50-digit) fixtures in tests/golden, closed forms
                 (ones -> N*delta, delta -> DFT row), the literal 3x3x3 table of
                 the reference's N-D test, long-double naive DFT and numpy's
                 long-double FFT; FP32 error envelope against the reference's
                 accuracy bounds (Singleton recurrence precision).
  chirp_index    exact (n*n) mod 2N identity.
  chirp          50-digit mpmath exp(i*pi*n^2/N).
  naive_dft_ld   mpmath fixtures.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

COMPOSITE_FFT, DFT, POW3_FFT, POW2RADIX4_FFT, BLUESTEIN = range(5)
TYPE_NAMES = {0: "COMPOSITE", 1: "DFT", 2: "POW3", 3: "POW2RADIX4", 4: "BLUESTEIN"}
FFT_FORWARD, FFT_BACKWARD = -1, 1
DEFAULT_BLUESTEIN_CUTOFF = 73

_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_plan_tree.restype = ctypes.c_int64
        L.oracle_plan_tree.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64]
        L.oracle_chirp_index.argtypes = [ctypes.c_int64, _i64p]
        L.oracle_chirp.argtypes = [ctypes.c_int64, ctypes.c_int, _vp]
        for nm in ("oracle_fft_c64", "oracle_fft_c32"):
            f = getattr(L, nm)
            f.restype = ctypes.c_int
            f.argtypes = [_vp, _vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int64]
        for nm in ("oracle_fftn_c64", "oracle_fftn_c32"):
            f = getattr(L, nm)
            f.restype = ctypes.c_int
            f.argtypes = [_vp, _vp, ctypes.c_int, _i64p, ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int64]
        for nm in ("oracle_fft_batch_c32", "oracle_fft_batch_c64"):
            f = getattr(L, nm)
            f.restype = None
            f.argtypes = [ctypes.c_int64, _i64p, _i64p, ctypes.POINTER(ctypes.c_int), _vp, _vp, ctypes.c_int,
                          ctypes.c_int64]
        L.oracle_naive_dft_ld.argtypes = [_vp, _vp, ctypes.c_int64, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def plan_tree(n: int, cutoff: int = DEFAULT_BLUESTEIN_CUTOFF):
    """FFTA's CallGraph node list for length n: list of (left, right, type, sz, s_in, s_out)."""
    cap = 256
    buf = np.zeros(6 * cap, dtype=np.int64)
    cnt = lib().oracle_plan_tree(n, cutoff, buf.ctypes.data_as(_i64p), cap)
    if cnt < 0:
        raise ValueError("Array length must be strictly positive")
    return [tuple(int(v) for v in buf[6 * i: 6 * i + 6]) for i in range(cnt)]


def chirp_index(n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.int64)
    lib().oracle_chirp_index(n, out.ctypes.data_as(_i64p))
    return out


def chirp(n: int, direction: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.complex128)
    lib().oracle_chirp(n, direction, _ptr(out))
    return out


def _as(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def fft(x, direction: int = FFT_FORWARD, cutoff: int = DEFAULT_BLUESTEIN_CUTOFF) -> np.ndarray:
    """1-D transform of a complex64/complex128 vector, in that precision."""
    x = np.asarray(x)
    dt = np.complex64 if x.dtype == np.complex64 else np.complex128
    xi = _as(x, dt).ravel()
    out = np.empty_like(xi)
    f = lib().oracle_fft_c32 if dt == np.complex64 else lib().oracle_fft_c64
    if f(_ptr(xi), _ptr(out), xi.size, direction, cutoff) != 0:
        raise ValueError("Array length must be strictly positive")
    return out


def bfft(x, cutoff: int = DEFAULT_BLUESTEIN_CUTOFF) -> np.ndarray:
    return fft(x, FFT_BACKWARD, cutoff)


def fftn(x, region=None, direction: int = FFT_FORWARD, cutoff: int = DEFAULT_BLUESTEIN_CUTOFF) -> np.ndarray:
    """Transform a column-major (Fortran-order) array over a 1-based region."""
    x = np.asarray(x)
    dt = np.complex64 if x.dtype == np.complex64 else np.complex128
    xf = np.asfortranarray(x, dtype=dt)
    if region is None:
        region = tuple(range(1, x.ndim + 1))
    if isinstance(region, int):
        region = (region,)
    dims = np.array(x.shape, dtype=np.int64)
    reg = np.array(region, dtype=np.int64)
    out = np.empty_like(xf, order="F")
    f = lib().oracle_fftn_c32 if dt == np.complex64 else lib().oracle_fftn_c64
    rc = f(_ptr(xf), _ptr(out), x.ndim, dims.ctypes.data_as(_i64p), reg.size, reg.ctypes.data_as(_i64p), direction,
           cutoff)
    if rc != 0:
        raise ValueError("Array length must be strictly positive")
    return out


def fft_batch(x: np.ndarray, lengths, offsets, directions, nthreads: int = 1,
              cutoff: int = DEFAULT_BLUESTEIN_CUTOFF) -> np.ndarray:
    """Independent 1-D transforms of packed data x[offsets[i]: offsets[i]+lengths[i]]."""
    x = np.ascontiguousarray(x)
    lengths = np.ascontiguousarray(lengths, dtype=np.int64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    dirs = np.ascontiguousarray(directions, dtype=np.int32)
    out = np.empty_like(x)
    f = lib().oracle_fft_batch_c32 if x.dtype == np.complex64 else lib().oracle_fft_batch_c64
    f(lengths.size, lengths.ctypes.data_as(_i64p), offsets.ctypes.data_as(_i64p),
      dirs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(x), _ptr(out), nthreads, cutoff)
    return out


def naive_dft_ld(x, direction: int = FFT_FORWARD) -> np.ndarray:
    """Naive O(n^2) DFT accumulated in long double, exact angle reduction; rounded to complex128."""
    xi = _as(x, np.complex128).ravel()
    out = np.empty_like(xi)
    lib().oracle_naive_dft_ld(_ptr(xi), _ptr(out), xi.size, direction)
    return out


def truth(x, direction: int = FFT_FORWARD) -> np.ndarray:
    """Long-double FFT of x (numpy pocketfft in 80-bit), as complex256."""
    xl = np.asarray(x).astype(np.clongdouble)
    if direction == FFT_FORWARD:
        return np.fft.fft(xl)
    return np.fft.ifft(xl) * xl.size


def truth_n(x, axes, direction: int = FFT_FORWARD) -> np.ndarray:
    xl = np.asarray(x).astype(np.clongdouble)
    if direction == FFT_FORWARD:
        return np.fft.fftn(xl, axes=axes)
    n = 1
    for a in axes:
        n *= xl.shape[a]
    return np.fft.ifftn(xl, axes=axes) * n


def relerr(a, b) -> float:
    """Relative L2 error ||a-b|| / ||b||, accumulated in long double."""
    a = np.asarray(a).astype(np.clongdouble).ravel()
    b = np.asarray(b).astype(np.clongdouble).ravel()
    nb = np.sqrt(np.sum(np.abs(b) ** 2))
    na = np.sqrt(np.sum(np.abs(a - b) ** 2))
    if nb == 0:
        return float(na)
    return float(na / nb)
