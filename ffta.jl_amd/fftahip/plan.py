"""Plan objects mirroring FFTA.jl's FFTAPlan_cx (src/plan.jl:3-20) over libfftahip.so."""
from __future__ import annotations

import ctypes
import json
import math
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import i64_array, load

FFT_FORWARD = -1   # src/callgraph.jl:1
FFT_BACKWARD = 1
DEFAULT_BLUESTEIN_CUTOFF = 73   # src/callgraph.jl:44


class FFTAHipError(RuntimeError):
    """A failure of the library itself (HIP runtime, allocation, capacity)."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DomainError(ValueError):
    """Julia's DomainError."""


_EXC = {
    _lib.ERR_DIMENSION_MISMATCH: DimensionMismatch,
    _lib.ERR_ARGUMENT: ArgumentError,
    _lib.ERR_DOMAIN: DomainError,
}


def _check(st: int):
    if st == _lib.SUCCESS:
        return
    msg = _lib.last_error()
    raise _EXC.get(st, FFTAHipError)(msg)


def version() -> str:
    return load().fftahip_version().decode()


def device_info() -> dict:
    lds, cus, warp = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(load().fftahip_device_info(ctypes.byref(lds), ctypes.byref(cus), ctypes.byref(warp)))
    return {"lds_per_block": lds.value, "cu_count": cus.value, "warp_size": warp.value}


def _precision_of(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.complex64:
        return 32
    if dt == np.complex128:
        return 64
    raise ArgumentError(f"unsupported element type {dt}; ComplexF32 or ComplexF64 expected")


def _dtype_of(precision: int):
    return np.complex64 if precision == 32 else np.complex128


def _region_tuple(region, ndims: int):
    """Normalise a Julia region (Int, vector, tuple, range) to a sorted tuple (src/plan.jl:51-69)."""
    if region is None:
        return tuple(range(1, ndims + 1))
    if isinstance(region, (int, np.integer)):
        return (int(region),)
    return tuple(sorted(int(r) for r in region))


class DeviceArray:
    """A column-major device array: a raw device pointer plus Julia dims and precision."""

    def __init__(self, ptr: int, dims: Sequence[int], precision: int):
        self.ptr = int(ptr)
        self.dims = tuple(int(d) for d in dims)
        self.precision = int(precision)

    @property
    def shape(self):
        return self.dims


class Plan:
    """FFTAPlan_cx: one planned call graph per region dimension, plus direction.

    `region` is held as FFTA holds it (an int for 1-D plans, else a mutable
    list), and is what mul_ validates against (src/plan.jl:220, 268-271)."""

    def __init__(self, shape, precision: int, region, direction: int, cutoff: int = DEFAULT_BLUESTEIN_CUTOFF,
                 scale: float = 1.0):
        self._h = ctypes.c_void_p()
        dims = tuple(int(d) for d in shape)
        reg = _region_tuple(region, len(dims))
        self.precision = precision
        self.direction = direction
        self.cutoff = cutoff
        self.scale = scale
        self._M = len(reg)
        self.region = reg[0] if self._M == 1 else list(reg)
        L = load()
        _check(L.fftahip_plan_create(ctypes.byref(self._h), precision, direction, len(dims), i64_array(dims),
                                     len(reg), i64_array(reg), cutoff))
        if scale != 1.0:
            _check(L.fftahip_plan_set_scale(self._h, float(scale)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.fftahip_plan_destroy(h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    @property
    def ndims(self) -> int:
        return self._M

    def size(self, i: int | None = None):
        """Base.size(p) / size(p, i) (src/plan.jl:36-47)."""
        L = load()
        if i is None:
            return tuple(self.size(k) for k in range(1, self._M + 1))
        out = ctypes.c_int64()
        _check(L.fftahip_plan_size(self._h, int(i), ctypes.byref(out)))
        return out.value

    def _region_list(self):
        if isinstance(self.region, (int, np.integer)):
            return [int(self.region)]
        return [int(r) for r in self.region]

    def check(self, xdims, ydims):
        reg = self._region_list()
        _check(load().fftahip_plan_check(self._h, len(reg), i64_array(reg), len(xdims), i64_array(xdims),
                                         len(ydims), i64_array(ydims)))

    def describe(self) -> dict:
        st, txt = _lib.read_json(load().fftahip_plan_describe, self._h)
        _check(st)
        return json.loads(txt)

    def execute_device(self, dims, in_ptr: int, out_ptr: int, stream: int = 0):
        """mul!(y, p, x) on device memory (asynchronous on `stream`)."""
        _check(load().fftahip_execute(self._h, len(dims), i64_array(dims), ctypes.c_void_p(int(in_ptr)),
                                      ctypes.c_void_p(int(out_ptr)), ctypes.c_void_p(int(stream))))

    def __mul__(self, x):
        """p * x (src/plan.jl:374-384): allocate y like x, then mul!."""
        if isinstance(x, DeviceArray):
            raise ArgumentError("p * x needs a host array; use mul_(y, p, x) for device arrays")
        xa = _as_julia_array(x, self.precision)
        y = np.empty(xa.shape, dtype=xa.dtype, order="F")
        mul_(y, self, xa)
        return y


def _as_julia_array(x, precision: int | None = None) -> np.ndarray:
    a = np.asarray(x)
    if precision is not None:
        want = _dtype_of(precision)
        if a.dtype != want:
            raise ArgumentError(f"plan element type {np.dtype(want)} does not match array element type {a.dtype}")
    return np.asfortranarray(a)


def mul_(y, p: Plan, x):
    """LinearAlgebra.mul!(y, p, x): argument checks of src/plan.jl, then the device transform."""
    if isinstance(x, DeviceArray) or isinstance(y, DeviceArray):
        if not (isinstance(x, DeviceArray) and isinstance(y, DeviceArray)):
            raise ArgumentError("x and y must both be device arrays or both host arrays")
        if x.precision != p.precision or y.precision != p.precision:
            raise ArgumentError("array precision does not match the plan")
        p.check(x.dims, y.dims)
        p.execute_device(x.dims, x.ptr, y.ptr)
        return y
    if not isinstance(y, np.ndarray) or not y.flags.f_contiguous:
        raise ArgumentError("y must be a column-major (Fortran-order) numpy array")
    xa = _as_julia_array(x, p.precision)
    if y.dtype != xa.dtype:
        raise ArgumentError("output element type must match the plan")
    p.check(xa.shape, y.shape)
    L = load()
    _check(L.fftahip_execute_host(p.handle, xa.ndim, i64_array(xa.shape), xa.ctypes.data_as(ctypes.c_void_p),
                                  y.ctypes.data_as(ctypes.c_void_p)))
    return y


def _complexify(x) -> np.ndarray:
    a = np.asarray(x)
    if a.dtype in (np.complex64, np.complex128):
        return a
    if a.dtype == np.float32:
        return a.astype(np.complex64)
    return a.astype(np.complex128)


def plan_fft(x, region=None, BLUESTEIN_CUTOFF: int = DEFAULT_BLUESTEIN_CUTOFF, **_kw) -> Plan:
    """AbstractFFTs.plan_fft(x, region) -> FFTA _plan_fft(..., FFT_FORWARD) (src/plan.jl:71-72)."""
    a = _complexify(x) if not isinstance(x, DeviceArray) else x
    prec = a.precision if isinstance(a, DeviceArray) else _precision_of(a.dtype)
    return Plan(a.shape, prec, region, FFT_FORWARD, BLUESTEIN_CUTOFF)


def plan_bfft(x, region=None, BLUESTEIN_CUTOFF: int = DEFAULT_BLUESTEIN_CUTOFF, **_kw) -> Plan:
    """AbstractFFTs.plan_bfft(x, region) -> FFTA _plan_fft(..., FFT_BACKWARD) (src/plan.jl:74-75)."""
    a = _complexify(x) if not isinstance(x, DeviceArray) else x
    prec = a.precision if isinstance(a, DeviceArray) else _precision_of(a.dtype)
    return Plan(a.shape, prec, region, FFT_BACKWARD, BLUESTEIN_CUTOFF)


def fft(x, region=None, **kw) -> np.ndarray:
    """fft(x[, region]) = plan_fft(x, region) * x."""
    a = _complexify(x)
    return plan_fft(a, region, **kw) * a


def bfft(x, region=None, **kw) -> np.ndarray:
    a = _complexify(x)
    return plan_bfft(a, region, **kw) * a


def plan_ifft(x, region=None, BLUESTEIN_CUTOFF: int = DEFAULT_BLUESTEIN_CUTOFF, **_kw) -> Plan:
    """AbstractFFTs.plan_ifft: plan_bfft scaled by 1/prod(size(x, region)), the
    factor applied by the pass that writes the result."""
    a = _complexify(x) if not isinstance(x, DeviceArray) else x
    prec = a.precision if isinstance(a, DeviceArray) else _precision_of(a.dtype)
    shape = a.shape
    reg = _region_tuple(region, len(shape))
    n = 1
    for r in reg:
        n *= shape[r - 1] if r <= len(shape) else 1
    return Plan(shape, prec, region, FFT_BACKWARD, BLUESTEIN_CUTOFF, scale=1.0 / n)


def ifft(x, region=None, **kw) -> np.ndarray:
    """AbstractFFTs ifft = plan_ifft(x, region) * x."""
    a = _complexify(x)
    return plan_ifft(a, region, **kw) * a


class RPlan:
    """FFTAPlan_re: plan_rfft (direction -1, real input of `shape`) or
    plan_brfft (direction +1, complex input of `shape`, real length `length`
    along the first region dimension), src/plan.jl:102-150."""

    def __init__(self, shape, precision: int, region, direction: int, length: int | None = None,
                 cutoff: int = DEFAULT_BLUESTEIN_CUTOFF):
        self._h = ctypes.c_void_p()
        self.shape = tuple(int(d) for d in shape)
        self.precision = precision
        self.direction = direction
        reg = _region_tuple(region, len(self.shape))
        self.region = list(reg)
        length = 0 if length is None else int(length)
        L = load()
        _check(L.fftahip_rplan_create(ctypes.byref(self._h), precision, direction, len(self.shape),
                                      i64_array(self.shape), length, len(reg), i64_array(reg), cutoff))
        nd = ctypes.c_int()
        _check(L.fftahip_rplan_output_dims(self._h, ctypes.byref(nd), None))
        dims = (ctypes.c_int64 * max(1, nd.value))()
        _check(L.fftahip_rplan_output_dims(self._h, ctypes.byref(nd), dims))
        self.output_shape = tuple(int(dims[i]) for i in range(nd.value))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.fftahip_rplan_destroy(h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def execute_device(self, in_ptr: int, out_ptr: int, stream: int = 0):
        _check(load().fftahip_rplan_execute(self._h, ctypes.c_void_p(int(in_ptr)), ctypes.c_void_p(int(out_ptr)),
                                            ctypes.c_void_p(int(stream))))

    def __mul__(self, x):
        """p * x on a host array (staged through device memory)."""
        real_t = np.float32 if self.precision == 32 else np.float64
        cplx_t = _dtype_of(self.precision)
        want = real_t if self.direction == -1 else cplx_t
        a = np.asarray(x)
        if a.dtype != want:
            raise ArgumentError(f"plan element type {np.dtype(want)} does not match array element type {a.dtype}")
        if tuple(a.shape) + (1,) * (len(self.shape) - a.ndim) != self.shape:
            raise DimensionMismatch("plan and input array axes do not match")
        a = np.asfortranarray(a)
        out_t = cplx_t if self.direction == -1 else real_t
        y = np.empty(self.output_shape, dtype=out_t, order="F")
        _check(load().fftahip_rplan_execute_host(self._h, a.ctypes.data_as(ctypes.c_void_p),
                                                 y.ctypes.data_as(ctypes.c_void_p)))
        return y


def plan_rfft(x, region=None, BLUESTEIN_CUTOFF: int = DEFAULT_BLUESTEIN_CUTOFF, **_kw) -> RPlan:
    """AbstractFFTs.plan_rfft(x, region) (src/plan.jl:102-125): x real (Float32/Float64)."""
    a = np.asarray(x)
    if a.dtype not in (np.float32, np.float64):
        raise ArgumentError(f"rfft needs a real array, got {a.dtype}")
    return RPlan(a.shape, 32 if a.dtype == np.float32 else 64, region, FFT_FORWARD, None, BLUESTEIN_CUTOFF)


def plan_brfft(x, length: int, region=None, BLUESTEIN_CUTOFF: int = DEFAULT_BLUESTEIN_CUTOFF, **_kw) -> RPlan:
    """AbstractFFTs.plan_brfft(x, len, region) (src/plan.jl:127-150): x complex, result real with
    `length` entries along the first region dimension."""
    a = np.asarray(x)
    return RPlan(a.shape, _precision_of(a.dtype), region, FFT_BACKWARD, length, BLUESTEIN_CUTOFF)


def rfft(x, region=None, **kw) -> np.ndarray:
    a = np.asarray(x)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    return plan_rfft(a, region, **kw) * a


def brfft(x, length: int, region=None, **kw) -> np.ndarray:
    a = np.asarray(x)
    return plan_brfft(a, length, region, **kw) * a


def irfft(x, length: int, region=None, **kw) -> np.ndarray:
    """AbstractFFTs irfft = brfft scaled by 1/(length * prod(other region sizes))."""
    a = np.asarray(x)
    reg = _region_tuple(region, a.ndim)
    n = int(length)
    for r in reg[1:]:
        n *= a.shape[r - 1] if r <= a.ndim else 1
    y = brfft(a, length, region, **kw)
    y *= 1.0 / n
    return y


class Batch:
    """Many 1-D transforms of uneven lengths in one work-list launch per pass."""

    def __init__(self, lengths: Iterable[int], directions: Iterable[int], precision: int = 32,
                 offsets: Iterable[int] | None = None, cutoff: int = DEFAULT_BLUESTEIN_CUTOFF):
        lengths = [int(v) for v in lengths]
        directions = [int(v) for v in directions]
        self.count = len(lengths)
        self.precision = precision
        self.lengths = np.array(lengths, dtype=np.int64)
        if offsets is None:
            self.offsets = np.concatenate([[0], np.cumsum(self.lengths)[:-1]]).astype(np.int64) if self.count else \
                np.zeros(0, np.int64)
        else:
            self.offsets = np.array(list(offsets), dtype=np.int64)
        self.total = int((self.offsets + self.lengths).max()) if self.count else 0
        self._h = ctypes.c_void_p()
        dirs = (ctypes.c_int * max(1, self.count))(*directions)
        offs = None if offsets is None else i64_array(self.offsets)
        _check(load().fftahip_batch_create(ctypes.byref(self._h), precision, self.count, i64_array(lengths), offs,
                                           dirs, cutoff))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.fftahip_batch_destroy(h)
            self._h = ctypes.c_void_p()

    def execute_device(self, in_ptr: int, out_ptr: int, stream: int = 0, graph: bool = False):
        fn = load().fftahip_batch_execute_graph if graph else load().fftahip_batch_execute
        _check(fn(self._h, ctypes.c_void_p(int(in_ptr)), ctypes.c_void_p(int(out_ptr)), ctypes.c_void_p(int(stream))))

    def describe(self) -> dict:
        st, txt = _lib.read_json(load().fftahip_batch_describe, self._h)
        _check(st)
        return json.loads(txt)


def debug_schedule(shape, precision: int, direction: int, region=None, lds_per_block: int = 160 * 1024,
                   cutoff: int = 73) -> dict:
    """Dry-run schedule of mul!(y, plan, x) for an array of `shape` (no device
    needed), planned with BLUESTEIN_CUTOFF `cutoff`."""
    dims = tuple(int(d) for d in shape)
    reg = _region_tuple(region, len(dims))
    st, txt = _lib.read_json(load().fftahip_debug_schedule, precision, direction, len(dims), i64_array(dims),
                             len(reg), i64_array(reg), int(cutoff), lds_per_block)
    _check(st)
    return json.loads(txt)


def debug_batch_schedule(lengths, directions, precision: int = 32, lds_per_block: int = 160 * 1024) -> dict:
    lengths = [int(v) for v in lengths]
    dirs = (ctypes.c_int * max(1, len(lengths)))(*[int(d) for d in directions])
    st, txt = _lib.read_json(load().fftahip_debug_batch_schedule, precision, len(lengths), i64_array(lengths), dirs,
                             lds_per_block)
    _check(st)
    return json.loads(txt)


def debug_chirp(N: int, precision: int = 64, n0: int = 0, count: int | None = None) -> np.ndarray:
    """The device's chirp entries n0 .. n0 + count - 1 of length N (needs a GPU)."""
    count = N - n0 if count is None else count
    out = np.empty(count, dtype=np.complex64 if precision == 32 else np.complex128)
    _check(load().fftahip_debug_chirp(precision, N, n0, count, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def debug_chirp_steps(N: int, s: int, L: int) -> np.ndarray:
    """A register pass's FP64 chirp steps d[t], t < L, computed on the device (needs a GPU)."""
    out = np.empty(L, dtype=np.complex128)
    _check(load().fftahip_debug_chirp_steps(N, s, L, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def debug_fastdiv(d: int, n: np.ndarray) -> np.ndarray:
    n = np.ascontiguousarray(n, dtype=np.uint32)
    q = np.empty_like(n)
    _check(load().fftahip_debug_fastdiv(d, n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                        q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n.size))
    return q


def flops(n: int) -> float:
    """5 N log2 N, the metric's flop count of one length-n transform."""
    return 5.0 * n * math.log2(n) if n > 1 else 0.0
