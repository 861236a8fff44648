"""The host code of the library (planner, schedule builder, batches, request
loop) under AddressSanitizer: the CPU tests of the C ABI (tests/test_host.py:
dry-run schedules, argument and capacity checks, the emulator's replay of
every emitted pass) run in a child process against lib/libfftahip_asan.so
(make -C ffta.jl_amd asan; host code built with -fsanitize=address, the
kernels' objects unchanged) with the clang ASan runtime preloaded.  Any
AddressSanitizer report fails the test."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ffta.jl_amd")
ASAN_LIB = os.path.join(PKG, "lib", "libfftahip_asan.so")


def _asan_runtime():
    try:
        out = subprocess.run(["/opt/rocm/bin/hipcc", "-print-file-name=libclang_rt.asan-x86_64.so"],
                             capture_output=True, text=True, timeout=60).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return out if out and os.path.isabs(out) and os.path.exists(out) else None


def test_host_abi_under_asan():
    if not os.path.exists(ASAN_LIB):
        pytest.skip("lib/libfftahip_asan.so not built (make -C ffta.jl_amd asan)")
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("clang ASan runtime not found")
    env = dict(os.environ)
    env.update({"LD_PRELOAD": rt, "FFTAHIP_LIB": ASAN_LIB,
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1:allocator_may_return_null=1"})
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_host.py"), "-q", "-x",
                        "-m", "not gpu", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert "AddressSanitizer" not in r.stdout + r.stderr, tail
    assert r.returncode == 0, tail
    assert " passed" in r.stdout, tail
