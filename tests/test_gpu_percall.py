"""The native per-call driver (ffta.jl_amd/bin/fftahip_percall) that bench.py
times the per-call FP64 path with: it runs, executes every call and reports
consistent counts."""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ffta.jl_amd", "bin", "fftahip_percall")


@pytest.mark.parametrize("precision", [32, 64])
def test_percall_driver(torch_cuda, precision):
    assert os.path.exists(BIN), "ffta.jl_amd/bin/fftahip_percall is not built"
    L = np.array([16, 1000, 4096, 73, 1000003, 243, 16], dtype=np.int64)
    D = np.array([-1, 1, -1, -1, 1, 1, -1], dtype=np.int32)
    with tempfile.NamedTemporaryFile(suffix=".calls", delete=False) as f:
        f.write(np.int64(L.size).tobytes() + L.tobytes() + D.tobytes())
        path = f.name
    try:
        print(f"case percall driver precision={precision}", flush=True)
        r = subprocess.run([BIN, path, str(precision), "2"], capture_output=True, text=True, timeout=120)
    finally:
        os.unlink(path)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["transforms"] == L.size and d["repeats"] == 2 and d["distinct_plans"] == 6
    assert d["device_ms"] > 0 and d["transforms_per_s"] > 0
