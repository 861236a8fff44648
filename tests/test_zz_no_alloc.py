"""GPU test of fftahip_execute_noalloc, the entry point that allocates
nothing (its scratch is a fixed per-device area inside the library).

Called through the C ABI, its results must equal fftahip_execute's bit for
bit (same schedules, same kernels) on consecutive calls, the second with a
smaller input; shapes without a schedule and transforms needing more scratch
than the area holds are refused with clear errors."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NOT_READY = 8
CAPACITY = 7


def _lib():
    from fftahip import _lib as L

    return L, L.load()


def _run(torch, L, lib, fn, plan, x):
    dims = x.shape
    tx = torch.from_numpy(np.ascontiguousarray(x.ravel(order="F"))).to("cuda")
    ty = torch.full_like(tx, complex(np.nan, np.nan))
    stream = torch.cuda.current_stream().cuda_stream
    st = fn(plan.handle, len(dims), L.i64_array(dims), ctypes.c_void_p(tx.data_ptr()), ctypes.c_void_p(ty.data_ptr()),
            ctypes.c_void_p(stream))
    torch.cuda.synchronize()
    return st, ty.cpu().numpy()


CASES = [
    # (shape, precision, direction): the first needs scratch of a three-pass
    # Bluestein, the second (smaller) a multi-pass Stockham, then 2-D
    ((1000003,), 32, -1),
    ((196608,), 32, 1),
    ((3000, 2187), 32, 1),
    ((65537,), 64, -1),
    ((256, 243), 64, -1),
]


def test_noalloc_matches_execute(torch_cuda):
    """fftahip_execute first on every case (the reference outputs), then the
    no-alloc calls back to back, each later one on a smaller input than the
    first, twice over; every output must equal fftahip_execute's."""
    from fftahip.plan import Plan

    L, lib = _lib()
    rng = np.random.default_rng(42)
    inputs, plans, refs = [], [], []
    for shape, prec, direction in CASES:
        dt = np.complex64 if prec == 32 else np.complex128
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)
        p = Plan(shape, prec, None, direction)
        print(f"case execute shape={shape} prec={prec} dir={direction}", flush=True)
        st, y0 = _run(torch_cuda, L, lib, lib.fftahip_execute, p, x)
        assert st == 0, L.last_error()
        inputs.append(x)
        plans.append(p)
        refs.append(y0)
    for rep in range(2):
        for (shape, prec, direction), x, p, y0 in zip(CASES, inputs, plans, refs):
            print(f"case noalloc rep={rep} shape={shape} prec={prec} dir={direction}", flush=True)
            st, y1 = _run(torch_cuda, L, lib, lib.fftahip_execute_noalloc, p, x)
            assert st == 0, L.last_error()
            assert not np.isnan(y1).any()
            assert np.array_equal(y0, y1)


def test_noalloc_refusals(torch_cuda):
    from fftahip.plan import Plan

    L, lib = _lib()
    rng = np.random.default_rng(7)
    # a 1-D plan made for (4096,) applied along dim 1 of a (4096, 3) array:
    # valid, but that shape has no schedule until fftahip_execute ran it once
    p = Plan((4096,), 32, None, -1)
    x = (rng.standard_normal((4096, 3)) + 1j * rng.standard_normal((4096, 3))).astype(np.complex64)
    print("case noalloc not-ready shape", flush=True)
    st, _ = _run(torch_cuda, L, lib, lib.fftahip_execute_noalloc, p, x)
    assert st == NOT_READY and "no schedule" in L.last_error()
    st, y0 = _run(torch_cuda, L, lib, lib.fftahip_execute, p, x)
    assert st == 0
    st, y1 = _run(torch_cuda, L, lib, lib.fftahip_execute_noalloc, p, x)
    assert st == 0 and np.array_equal(y0, y1)
    # more scratch than the fixed area holds: refused, nothing launched
    print("case noalloc capacity", flush=True)
    n = 31104000
    big = Plan((n,), 32, None, -1)
    st = lib.fftahip_execute_noalloc(big.handle, 1, L.i64_array([n]), ctypes.c_void_p(0x10), ctypes.c_void_p(0x10),
                                     None)
    assert st == CAPACITY and "scratch" in L.last_error()
    # a mismatched shape fails the mul! checks first
    st = lib.fftahip_execute_noalloc(p.handle, 1, L.i64_array([4095]), ctypes.c_void_p(0x10), ctypes.c_void_p(0x10),
                                     None)
    assert st == 1
