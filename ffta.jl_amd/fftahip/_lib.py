"""ctypes binding of libfftahip.so (include/fftahip.h).

The library is the product: every transform runs in its gfx950 kernels.  If
the shared library is missing this module raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)                 # ffta.jl_amd/
# FFTAHIP_LIB selects another build of the same library (the host
# AddressSanitizer build, lib/libfftahip_asan.so, in the sanitizer test)
LIB_PATH = os.environ.get("FFTAHIP_LIB") or os.path.join(PKG_ROOT, "lib", "libfftahip.so")

SUCCESS = 0
ERR_DIMENSION_MISMATCH = 1
ERR_ARGUMENT = 2
ERR_DOMAIN = 3
ERR_HIP = 4
ERR_ALLOC = 5
ERR_NOT_IMPLEMENTED = 6
ERR_CAPACITY = 7
ERR_NOT_READY = 8

_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_st = ctypes.c_int

# name -> (restype, argtypes); every symbol declared in include/fftahip.h
SIGNATURES = {
    "fftahip_version": (ctypes.c_char_p, []),
    "fftahip_status_string": (ctypes.c_char_p, [_st]),
    "fftahip_last_error": (ctypes.c_char_p, []),
    "fftahip_plan_create": (_st, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int,
                                  _i64p, ctypes.c_int64]),
    "fftahip_plan_size": (_st, [_vp, ctypes.c_int, _i64p]),
    "fftahip_plan_rank": (_st, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "fftahip_plan_precision": (_st, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "fftahip_plan_direction": (_st, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "fftahip_plan_set_scale": (_st, [_vp, ctypes.c_double]),
    "fftahip_plan_check": (_st, [_vp, ctypes.c_int, _i64p, ctypes.c_int, _i64p, ctypes.c_int, _i64p]),
    "fftahip_execute": (_st, [_vp, ctypes.c_int, _i64p, _vp, _vp, _vp]),
    "fftahip_execute_host": (_st, [_vp, ctypes.c_int, _i64p, _vp, _vp]),
    "fftahip_execute_noalloc": (_st, [_vp, ctypes.c_int, _i64p, _vp, _vp, _vp]),
    "fftahip_plan_describe": (_st, [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "fftahip_plan_destroy": (_st, [_vp]),
    "fftahip_rplan_create": (_st, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p,
                                   ctypes.c_int64, ctypes.c_int, _i64p, ctypes.c_int64]),
    "fftahip_rplan_output_dims": (_st, [_vp, ctypes.POINTER(ctypes.c_int), _i64p]),
    "fftahip_rplan_execute": (_st, [_vp, _vp, _vp, _vp]),
    "fftahip_rplan_execute_host": (_st, [_vp, _vp, _vp]),
    "fftahip_rplan_destroy": (_st, [_vp]),
    "fftahip_batch_create": (_st, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int64, _i64p, _i64p,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int64]),
    "fftahip_batch_execute": (_st, [_vp, _vp, _vp, _vp]),
    "fftahip_batch_execute_graph": (_st, [_vp, _vp, _vp, _vp]),
    "fftahip_batch_describe": (_st, [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "fftahip_batch_destroy": (_st, [_vp]),
    "fftahip_server_create": (_st, [ctypes.POINTER(_vp), ctypes.c_int64, ctypes.c_int]),
    "fftahip_server_submit": (_st, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, _vp, _vp, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_uint64)]),
    "fftahip_server_replay": (_st, [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), _i64p, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "fftahip_server_test": (_st, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "fftahip_server_wait": (_st, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    "fftahip_server_drain": (_st, [_vp]),
    "fftahip_server_stats_get": (_st, [_vp, _vp]),
    "fftahip_server_reset_stats": (_st, [_vp]),
    "fftahip_server_destroy": (_st, [_vp]),
    "fftahip_autotune_lines": (_st, [ctypes.c_int, ctypes.c_int64, _i64p, _i64p]),
    "fftahip_autotune_shape": (_st, [ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int, _i64p, _i64p]),
    "fftahip_autotune_get_tile": (_st, [ctypes.c_int, ctypes.c_int64, _i64p, _i64p]),
    "fftahip_autotune_save": (_st, [ctypes.c_char_p]),
    "fftahip_autotune_load": (_st, [ctypes.c_char_p, _i64p]),
    "fftahip_autotune_clear": (_st, []),
    "fftahip_autotune_get": (_st, [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]),
    "fftahip_malloc": (_st, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "fftahip_free": (_st, [_vp]),
    "fftahip_memcpy_htod": (_st, [_vp, _vp, ctypes.c_size_t]),
    "fftahip_memcpy_dtoh": (_st, [_vp, _vp, ctypes.c_size_t]),
    "fftahip_synchronize": (_st, []),
    "fftahip_device_info": (_st, [_i64p, _i64p, _i64p]),
    "fftahip_debug_schedule": (_st, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int, _i64p,
                                     ctypes.c_int64, ctypes.c_int64, ctypes.c_char_p, ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_size_t)]),
    "fftahip_debug_batch_schedule": (_st, [ctypes.c_int, ctypes.c_int64, _i64p, ctypes.POINTER(ctypes.c_int),
                                           ctypes.c_int64, ctypes.c_char_p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "fftahip_debug_chirp": (_st, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]),
    "fftahip_debug_chirp_steps": (_st, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]),
    "fftahip_debug_mod_q31": (_st, [ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64]),
    "fftahip_debug_fastdiv": (_st, [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_uint32), ctypes.c_int64]),
}

class ServerStats(ctypes.Structure):
    """fftahip_server_stats (include/fftahip.h)."""
    _fields_ = [("completed", ctypes.c_int64), ("rejected", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("pending", ctypes.c_int64), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("seconds", ctypes.c_double),
                ("requests_per_s", ctypes.c_double), ("payload_gbps", ctypes.c_double),
                ("pool_peak_bytes", ctypes.c_int64), ("plans", ctypes.c_int64), ("coalesced", ctypes.c_int64),
                ("groups", ctypes.c_int64)]


_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libfftahip.so (built by ffta.jl_amd/Makefile) and bind every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `make -C {PKG_ROOT}` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def i64_array(vals):
    vals = [int(v) for v in vals]
    arr = (ctypes.c_int64 * max(1, len(vals)))(*vals)
    return arr


def last_error() -> str:
    return load().fftahip_last_error().decode()


def read_json(fn, *args):
    """Call a (buf, cap, needed) JSON-emitting entry point and return the text."""
    need = ctypes.c_size_t(0)
    st = fn(*args, None, 0, ctypes.byref(need))
    if st != SUCCESS:
        return st, None
    buf = ctypes.create_string_buffer(need.value)
    st = fn(*args, buf, need.value, ctypes.byref(need))
    return st, buf.value.decode()
