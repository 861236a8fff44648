"""Serving loop: continuous batching of fft/bfft requests (libfftahip's fftahip_server_*)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import i64_array, load
from .plan import _check


class Server:
    """A device-side request loop.  Requests are admitted in arrival order
    against a device pool of `pool_bytes` and run on `nstreams` streams."""

    def __init__(self, pool_bytes: int = 1 << 30, nstreams: int = 4):
        self._h = ctypes.c_void_p()
        _check(load().fftahip_server_create(ctypes.byref(self._h), int(pool_bytes), int(nstreams)))

    def close(self):
        if self._h is not None and self._h.value and _lib._lib is not None:
            _lib._lib.fftahip_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def submit(self, precision: int, direction: int, dims, in_ptr: int, out_ptr: int, host: bool = False) -> int:
        t = ctypes.c_uint64()
        dims = [int(d) for d in dims]
        _check(load().fftahip_server_submit(self._h, precision, direction, len(dims), i64_array(dims),
                                            ctypes.c_void_p(int(in_ptr)), ctypes.c_void_p(int(out_ptr)),
                                            1 if host else 0, ctypes.byref(t)))
        return t.value

    def replay(self, t_offset, precision, direction, dims3, in_ptrs, out_ptrs, host: bool = False):
        """Submit request i at t_offset[i] seconds after the call (native load generator)."""
        n = len(t_offset)
        T = (ctypes.c_double * n)(*[float(v) for v in t_offset])
        Pr = (ctypes.c_int * n)(*[int(v) for v in precision])
        Dr = (ctypes.c_int * n)(*[int(v) for v in direction])
        nd = [sum(1 for d in row if d > 0) for row in dims3]
        Nd = (ctypes.c_int * n)(*nd)
        flat = []
        for row in dims3:
            flat += [int(d) for d in row] + [1] * (3 - len(row))
        Dm = i64_array(flat)
        Ip = (ctypes.c_void_p * n)(*[int(p) for p in in_ptrs])
        Op = (ctypes.c_void_p * n)(*[int(p) for p in out_ptrs])
        tickets = (ctypes.c_uint64 * n)()
        _check(load().fftahip_server_replay(self._h, n, T, Pr, Dr, Nd, Dm, Ip, Op, 1 if host else 0, tickets))
        return list(tickets)

    def wait(self, ticket: int) -> float:
        lat = ctypes.c_double()
        _check(load().fftahip_server_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(lat)))
        return lat.value

    def done(self, ticket: int) -> bool:
        d = ctypes.c_int()
        _check(load().fftahip_server_test(self._h, ctypes.c_uint64(ticket), ctypes.byref(d)))
        return bool(d.value)

    def drain(self):
        _check(load().fftahip_server_drain(self._h))

    def stats(self) -> dict:
        st = _lib.ServerStats()
        _check(load().fftahip_server_stats_get(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def reset_stats(self):
        _check(load().fftahip_server_reset_stats(self._h))
