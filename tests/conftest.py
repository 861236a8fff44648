"""Shared test configuration: the `gpu` marker, import paths, and device helpers."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ffta.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def tol(n: int, precision: int) -> float:
    """North-star tolerance on relative L2 error: 1e-14*log2 N (CF64), 2e-6*log2 N (CF32); N=1 is exact."""
    if n <= 1:
        return 0.0
    return (1e-14 if precision == 64 else 2e-6) * np.log2(n)


@pytest.fixture(scope="session")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch
