"""Autotuner of the in-LDS radix schedules and the strided tile shapes
(include/fftahip.h, autotuner).

FFTA picks radices by fixed rules (src/callgraph.jl:76-126); here the library
times the candidate schedules of each line length on the device and keeps the
fastest.  `tune_batch` tunes the line lengths a ragged batch runs, heaviest
first; `tune_shape` times the strided tile shapes (LDS bytes per workgroup,
lines per tile) of one array shape.  The resulting table can be saved and
loaded (text, one line per entry), and applies to every plan created after it
is set.
"""
from __future__ import annotations

import collections
import ctypes
import os

from . import _lib
from ._lib import i64_array, load
from .plan import _check, debug_batch_schedule

TUNING_DIR = os.path.join(_lib.PKG_ROOT, "tuning")


def line_lengths(lengths, directions, precision: int = 32) -> list[tuple[int, int]]:
    """(L, elements) of the in-LDS line passes a batch runs, heaviest first."""
    s = debug_batch_schedule(lengths, directions, precision)
    mass = collections.Counter()
    for g in s["groups"]:
        for la in g:
            if la["op"] == 0 and not la.get("reg"):
                mass[int(la["L"])] += int(la["nlines"]) * int(la["L"])
    return sorted(mass.items(), key=lambda kv: -kv[1])


def tune_lines(lengths, precision: int = 32) -> int:
    """Time the candidate schedules of each line length; returns how many differ from the default rule."""
    lengths = [int(v) for v in lengths]
    changed = ctypes.c_int64(0)
    _check(load().fftahip_autotune_lines(precision, len(lengths), i64_array(lengths), ctypes.byref(changed)))
    return changed.value


def tune_batch(lengths, directions, precision: int = 32, top: int | None = None) -> dict:
    ll = line_lengths(lengths, directions, precision)
    if top is not None:
        ll = ll[:top]
    changed = tune_lines([L for L, _ in ll], precision)
    return {"line_lengths": len(ll), "changed": changed}


def tune_shape(shape, precision: int = 32, region=None) -> int:
    """Time the strided tile shapes for the plan of `shape` over `region`
    (default: every dimension) and keep the fastest for its transform lengths;
    returns how many lengths now use a non-default tile."""
    dims = [int(v) for v in shape]
    reg = list(range(1, len(dims) + 1)) if region is None else ([int(region)] if isinstance(region, int)
                                                                  else [int(r) for r in region])
    changed = ctypes.c_int64(0)
    _check(load().fftahip_autotune_shape(precision, len(dims), i64_array(dims), len(reg), i64_array(reg),
                                         ctypes.byref(changed)))
    return changed.value


def tile(precision: int, length: int) -> tuple[int, int]:
    """(LDS bytes per tile, fewest lines) of the strided passes of a transform of
    `length`; (0, 0) is the default rule."""
    b, c = ctypes.c_int64(0), ctypes.c_int64(0)
    _check(load().fftahip_autotune_get_tile(precision, int(length), ctypes.byref(b), ctypes.byref(c)))
    return b.value, c.value


def save(path: str) -> None:
    _check(load().fftahip_autotune_save(path.encode()))


def load_table(path: str) -> int:
    n = ctypes.c_int64(0)
    _check(load().fftahip_autotune_load(path.encode(), ctypes.byref(n)))
    return n.value


def clear() -> None:
    _check(load().fftahip_autotune_clear())


def schedule(precision: int, length: int) -> list[int]:
    """Radix schedule (stage order) plans use for an in-LDS line of `length`."""
    n = ctypes.c_int(0)
    rad = (ctypes.c_int * 16)()
    _check(load().fftahip_autotune_get(precision, int(length), ctypes.byref(n), rad))
    return [rad[i] for i in range(n.value)]
