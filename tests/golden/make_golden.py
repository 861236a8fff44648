"""Writes the golden fixtures under tests/golden/ (run once; outputs are committed).

Truth is computed with mpmath at 50 significant digits, independently of the
oracle and of the reference's code:
  dft_small.npz  1-D forward/backward DFTs of seeded complex inputs for the
                 lengths the reference's own tests use (test/onedim/*.jl:
                 1..64, 73, 100, 103, 275) plus pow2/pow3/composite extras.
  dft2_small.npz 2-D forward/backward DFTs of seeded column-major matrices
                 (test/twodim/complex_forward.jl:15-27 uses (m, n), m in n..n+1).
  chirp.npz      Bluestein chirps exp(i*pi*sgn*(n^2 mod 2N)/N) for N in
                 {73, 103, 1009} (src/algos.jl:353-359 defines the sequence).
  ndim_333.npz   the literal 3x3x3 table of test/ndim/minimal_complex.jl:9-26
                 (fft of a delta at (2,2,2), 1-based), transcribed as data.

Inputs: numpy default_rng(seed) standard normal real/imag parts.
"""
from __future__ import annotations

import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))

LENGTHS_1D = list(range(1, 65)) + [73, 81, 96, 100, 103, 121, 125, 128, 243, 256, 275]
SHAPES_2D = [(1, 1), (2, 1), (3, 5), (4, 4), (5, 4), (8, 11), (9, 8), (16, 17), (12, 15)]


def mp_dft(x, direction):
    n = len(x)
    xs = [mp.mpc(complex(v)) for v in x]
    out = np.zeros(n, dtype=np.complex128)
    for u in range(n):
        acc = mp.mpc(0)
        for v in range(n):
            m = (u * v) % n
            acc += xs[v] * mp.expjpi(mp.mpf(2 * direction * m) / n)
        out[u] = complex(acc)
    return out


def mp_dft2(X, direction):
    m, n = X.shape
    Xs = [[mp.mpc(complex(X[i, j])) for j in range(n)] for i in range(m)]
    # columns first (sum over i), then rows (sum over j), all in mpmath
    cols = [[mp.mpc(0)] * n for _ in range(m)]
    for j in range(n):
        for u in range(m):
            acc = mp.mpc(0)
            for i in range(m):
                acc += Xs[i][j] * mp.expjpi(mp.mpf(2 * direction * ((u * i) % m)) / m)
            cols[u][j] = acc
    out = np.zeros((m, n), dtype=np.complex128)
    for u in range(m):
        for v in range(n):
            acc = mp.mpc(0)
            for j in range(n):
                acc += cols[u][j] * mp.expjpi(mp.mpf(2 * direction * ((v * j) % n)) / n)
            out[u, v] = complex(acc)
    return out


def main():
    rng = np.random.default_rng(20240601)
    d = {}
    for n in LENGTHS_1D:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        d[f"x_{n}"] = x
        d[f"fwd_{n}"] = mp_dft(x, -1)
        d[f"bwd_{n}"] = mp_dft(x, +1)
    np.savez_compressed(os.path.join(HERE, "dft_small.npz"), lengths=np.array(LENGTHS_1D), **d)

    d2 = {}
    for (m, n) in SHAPES_2D:
        X = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        d2[f"x_{m}_{n}"] = X
        d2[f"fwd_{m}_{n}"] = mp_dft2(X, -1)
        d2[f"bwd_{m}_{n}"] = mp_dft2(X, +1)
    np.savez_compressed(os.path.join(HERE, "dft2_small.npz"), shapes=np.array(SHAPES_2D), **d2)

    dc = {}
    for n in (73, 103, 1009):
        for direction in (-1, 1):
            sgn = -direction
            vals = np.zeros(n, dtype=np.complex128)
            for k in range(n):
                p = (k * k) % (2 * n)
                vals[k] = complex(mp.expjpi(mp.mpf(sgn * p) / n))
            dc[f"chirp_{n}_{direction}"] = vals
    np.savez_compressed(os.path.join(HERE, "chirp.npz"), **dc)

    # test/ndim/minimal_complex.jl:9-26: y[2,2,2] = 1; fft(y) == this table.
    w1 = -0.5 - np.sqrt(3) * 1j / 2
    w2 = np.conj(w1)
    slabs = [
        [[1, w1, w2], [w1, w2, 1], [w2, 1, w1]],
        [[w1, w2, 1], [w2, 1, w1], [1, w1, w2]],
        [[w2, 1, w1], [1, w1, w2], [w1, w2, 1]],
    ]
    ref = np.zeros((3, 3, 3), dtype=np.complex128)
    for k in range(3):
        for i in range(3):
            for j in range(3):
                ref[i, j, k] = slabs[k][i][j]
    np.savez_compressed(os.path.join(HERE, "ndim_333.npz"), y_ref=ref)


if __name__ == "__main__":
    main()
