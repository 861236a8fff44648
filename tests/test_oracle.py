"""Pins the CPU oracle (oracle/ffta_oracle.cpp, FFTA's algorithm restated) -- CPU only.

Each check is independent of the oracle's own formulas:
  * 50-digit mpmath DFTs of seeded inputs (tests/golden, written by
    make_golden.py) at the reference's test lengths,
  * the reference's closed forms (ones -> N*delta; the 3x3x3 table of
    test/ndim/minimal_complex.jl:9-26),
  * delta probes (e_j -> row j of the DFT matrix, which fixes every index map
    and the sign),
  * numpy's long-double FFT at config-scale lengths,
  * FFTA's own FP32 accuracy bounds (test/onedim/accuracy.jl:15-43), which a
    naive twiddle recurrence fails,
  * planner trees derived by hand from src/callgraph.jl:76-126,
  * the Bluestein chirp index identity p_n = n^2 mod 2N (src/algos.jl:354-358).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from conftest import tol

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gold(name):
    return np.load(os.path.join(GOLD, name))


def test_golden_1d_mpmath():
    d = _gold("dft_small.npz")
    for n in d["lengths"]:
        n = int(n)
        x = d[f"x_{n}"]
        for direction, key in ((-1, "fwd"), (1, "bwd")):
            ref = d[f"{key}_{n}"]
            y = oracle.fft(x, direction)
            err = oracle.relerr(y, ref)
            bound = max(tol(n, 64), 0.0)
            assert err <= bound, f"n={n} dir={direction}: {err:.3e} > {bound:.3e}"
            if n == 1:
                assert np.array_equal(y, x)


def test_golden_naive_long_double_truth():
    d = _gold("dft_small.npz")
    for n in (1, 2, 7, 16, 64, 103, 275):
        x = d[f"x_{n}"]
        assert oracle.relerr(oracle.naive_dft_ld(x, -1), d[f"fwd_{n}"]) < 1e-16
        assert oracle.relerr(oracle.naive_dft_ld(x, 1), d[f"bwd_{n}"]) < 1e-16


def test_golden_2d_mpmath():
    d = _gold("dft2_small.npz")
    for m, n in d["shapes"]:
        m, n = int(m), int(n)
        X = np.asfortranarray(d[f"x_{m}_{n}"])
        for direction, key in ((-1, "fwd"), (1, "bwd")):
            y = oracle.fftn(X, None, direction)
            err = oracle.relerr(y, d[f"{key}_{m}_{n}"])
            assert err <= max(tol(m * n, 64), 1e-300 if m * n > 1 else 0.0), (m, n, direction, err)


@pytest.mark.parametrize("n", [8, 11, 15, 16, 27, 100, 275])
def test_ones_forward(n):
    # test/onedim/complex_forward.jl:3-12
    y = oracle.fft(np.ones(n, dtype=np.complex128), -1)
    ref = np.zeros(n, dtype=np.complex128)
    ref[0] = n
    assert np.max(np.abs(y - ref)) <= 1e-12


@pytest.mark.parametrize("n", [8, 11, 15, 16, 27, 100, 103])
def test_ones_backward(n):
    # test/onedim/complex_backward.jl:4-11 (103 takes the Bluestein path)
    y = oracle.fft(np.ones(n, dtype=np.complex128), 1)
    ref = np.zeros(n, dtype=np.complex128)
    ref[0] = n
    assert np.max(np.abs(y - ref)) <= 1e-12


@pytest.mark.parametrize("n", [12, 15, 16, 27, 64, 73, 100, 275, 5329])
def test_delta_probe_rows(n):
    # e_j -> X[k] = exp(dir * 2 pi i j k / n): fixes index maps and sign
    for j in sorted({0, 1, 2, n // 3, n - 1}):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1
        k = np.arange(n)
        for direction in (-1, 1):
            y = oracle.fft(e, direction)
            ref = np.exp(direction * 2j * np.pi * ((j * k) % n) / n)
            assert np.max(np.abs(y - ref)) <= 1e-13, (n, j, direction)


def test_ndim_333_table():
    # test/ndim/minimal_complex.jl:9-26
    ref = _gold("ndim_333.npz")["y_ref"]
    y = np.zeros((3, 3, 3), dtype=np.complex128, order="F")
    y[1, 1, 1] = 1
    assert np.max(np.abs(oracle.fftn(y) - ref)) <= 1e-14


@pytest.mark.parametrize("sz", [(3, 5, 7), (4, 14, 9), (103, 5, 13), (26, 33, 35, 4), (3, 3, 3, 3, 3)])
def test_ndim_ones(sz):
    # test/ndim/minimal_complex.jl:4-7
    x = np.ones(sz, dtype=np.complex128, order="F")
    ref = np.zeros(sz, dtype=np.complex128)
    ref.flat[0] = np.prod(sz)
    assert np.max(np.abs(oracle.fftn(x) - ref)) <= 1e-8 * np.prod(sz)


@pytest.mark.parametrize("n", [177, 4096, 3000, 2187, 5329, 31104, 65536, 177147])
def test_long_double_truth(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for direction in (-1, 1):
        err = oracle.relerr(oracle.fft(x, direction), oracle.truth(x, direction))
        assert err <= tol(n, 64), (n, direction, err)


def test_mapslices_bitwise():
    # test/onedim/complex_forward.jl:27-33: fft(x, r) == mapslices(fft, x; dims=r)
    rng = np.random.default_rng(3)
    n = 7
    x = np.asfortranarray(rng.standard_normal((n, n + 1, n + 2)) + 1j * rng.standard_normal((n, n + 1, n + 2)))
    for r in (1, 2, 3):
        y = oracle.fftn(x, r)
        ref = np.apply_along_axis(lambda v: oracle.fft(v), r - 1, x)
        assert np.array_equal(y, ref)


POW2_BOUNDS = [(1 << 4, 3.0), (1 << 5, 3.0), (1 << 6, 3.0), (1 << 7, 3.0), (1 << 8, 4.0), (1 << 9, 4.0),
               (1 << 10, 8.0), (1 << 11, 8.0), (1 << 12, 10.0), (1 << 13, 12.0), (1 << 14, 14.0),
               (1 << 15, 20.0), (1 << 16, 22.0), (1 << 17, 28.0), (1 << 18, 28.0)]
POW3_BOUNDS = [(3, 3.0), (9, 3.0), (27, 3.0), (81, 4.0), (243, 5.0), (729, 7.0), (2187, 10.0), (6561, 13.0),
               (19683, 18.0)]


def _worst_eps(n):
    # test/onedim/accuracy.jl:45-56 (worst over 5 seeds, in units of eps(Float32))
    worst = 0.0
    for seed in range(1, 6):
        rng = np.random.default_rng(seed)
        x32 = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        y32 = oracle.fft(x32, -1)
        yref = oracle.fft(x32.astype(np.complex128), -1).astype(np.complex64)
        err = np.linalg.norm((y32 - yref).astype(np.complex128)) / np.linalg.norm(yref.astype(np.complex128))
        worst = max(worst, err / np.finfo(np.float32).eps)
    return worst


@pytest.mark.parametrize("n,bound", POW2_BOUNDS + POW3_BOUNDS)
def test_fp32_accuracy_envelope(n, bound):
    """FFTA's Singleton-recurrence error bounds hold for the restatement's FP32 path."""
    assert _worst_eps(n) <= bound


# Planner trees: (type, sz, s_in, s_out) in FFTA's pre-order, derived by hand
# from src/callgraph.jl:76-126.
TREES = {
    4096: [("POW2RADIX4", 4096, 1, 1)],
    177147: [("POW3", 177147, 1, 1)],
    1000003: [("BLUESTEIN", 1000003, 1, 1)],
    71: [("DFT", 71, 1, 1)],
    1: [("DFT", 1, 1, 1)],
    2187: [("POW3", 2187, 1, 1)],
    275: [("COMPOSITE", 275, 1, 1), ("COMPOSITE", 25, 11, 11), ("DFT", 5, 5, 55), ("DFT", 5, 55, 1),
          ("DFT", 11, 25, 1)],
    5329: [("COMPOSITE", 5329, 1, 1), ("BLUESTEIN", 73, 73, 73), ("BLUESTEIN", 73, 73, 1)],
    3000: [("COMPOSITE", 3000, 1, 1), ("POW2RADIX4", 8, 375, 375), ("COMPOSITE", 375, 8, 1), ("POW3", 3, 125, 125),
           ("COMPOSITE", 125, 24, 1), ("DFT", 5, 25, 25), ("COMPOSITE", 25, 120, 1), ("DFT", 5, 5, 5),
           ("DFT", 5, 600, 1)],
    31104000: [("COMPOSITE", 31104000, 1, 1), ("POW2RADIX4", 1024, 30375, 30375), ("COMPOSITE", 30375, 1024, 1),
               ("POW3", 243, 125, 125), ("COMPOSITE", 125, 248832, 1), ("DFT", 5, 25, 25),
               ("COMPOSITE", 25, 1244160, 1), ("DFT", 5, 5, 5), ("DFT", 5, 6220800, 1)],
    2000006: [("COMPOSITE", 2000006, 1, 1), ("POW2RADIX4", 2, 1000003, 1000003), ("BLUESTEIN", 1000003, 2, 1)],
}


@pytest.mark.parametrize("n", sorted(TREES))
def test_planner_trees(n):
    got = [(oracle.TYPE_NAMES[t], sz, si, so) for (_, _, t, sz, si, so) in oracle.plan_tree(n)]
    assert got == TREES[n]


def test_planner_invariants():
    for n in list(range(1, 400)) + [5329, 31104000, 1000003 * 2]:
        nodes = oracle.plan_tree(n)
        # the leaves multiply to n, and each composite node's children sizes multiply to it
        for i, (left, right, t, sz, _, _) in enumerate(nodes):
            if t == oracle.COMPOSITE_FFT:
                assert nodes[i + left][3] * nodes[i + right][3] == sz
        assert nodes[0][3] == n


def test_planner_cutoff():
    assert [t for (_, _, t, *_rest) in oracle.plan_tree(71, cutoff=73)] == [oracle.DFT]
    assert [t for (_, _, t, *_rest) in oracle.plan_tree(73, cutoff=73)] == [oracle.BLUESTEIN]
    assert [t for (_, _, t, *_rest) in oracle.plan_tree(71, cutoff=50)] == [oracle.BLUESTEIN]


def test_planner_rejects_nonpositive():
    with pytest.raises(ValueError):
        oracle.plan_tree(0)


@pytest.mark.parametrize("n", [73, 103, 1009, 1000003])
def test_chirp_index_exact(n):
    p = oracle.chirp_index(n)
    k = np.arange(n, dtype=np.int64)
    q = (k * k) % (2 * n)
    q = np.where(q > n, q - 2 * n, q)
    assert np.array_equal(p, q)


def test_chirp_values_mpmath():
    d = _gold("chirp.npz")
    for n in (73, 103, 1009):
        for direction in (-1, 1):
            c = oracle.chirp(n, direction)
            assert np.max(np.abs(c - d[f"chirp_{n}_{direction}"])) <= 2e-16 * 4


def test_planted_twiddle_sign_is_caught():
    """A sign slip in a transform (here: conjugating the output) fails the fixtures."""
    d = _gold("dft_small.npz")
    x = d["x_16"]
    wrong = np.conj(oracle.fft(np.conj(x), -1))  # = backward transform
    assert oracle.relerr(wrong, d["fwd_16"]) > 0.1
