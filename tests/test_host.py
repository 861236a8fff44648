"""Host-side tests of the library (no GPU): C ABI exports, planner, the error
contract of src/plan.jl / test/argument_checking.jl, and CPU emulation of the
launch schedules the planner emits."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import emulator as E
import fftahip
from fftahip import _lib
from fftahip import plan as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "fftahip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fftahip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = _lib.load()
    names = _header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/fftahip.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version_and_status_strings():
    assert "fftahip" in fftahip.version()
    L = _lib.load()
    assert L.fftahip_status_string(1).decode() == "DimensionMismatch"


@pytest.mark.parametrize("d", [1, 2, 3, 5, 7, 9, 15, 16, 17, 243, 255, 256, 1000, 3000, 8192, 65535, 131072,
                               1000003, 2**20 + 7, 31104000])
def test_fastdiv_exact(d):
    rng = np.random.default_rng(d)
    n = np.concatenate([np.arange(0, 70000, dtype=np.uint64),
                        rng.integers(0, 2**31, 200000, dtype=np.uint64),
                        np.array([2**31 - 1, 2**32 - 1, d - 1, d, d + 1, 2 * d - 1, 2 * d], dtype=np.uint64)])
    n = n[n < 2**32].astype(np.uint32)
    q = P.debug_fastdiv(d, n)
    assert np.array_equal(q.astype(np.uint64), n.astype(np.uint64) // d)


# ---------------------------------------------------------------- plan API / error contract


def test_plan_size_and_domain_error():
    # test/argument_checking.jl:229-237 (plan_fft sizes; non-Bluestein lengths plan on the host)
    for sz in [(8, 11), (6, 24, 25), (100,)]:
        x = np.zeros(sz, dtype=np.complex128)
        p = P.plan_fft(x)
        assert p.size() == sz
        assert p.size(1 + len(sz)) == 1
        with pytest.raises(P.DomainError):
            p.size(0)
        with pytest.raises(P.DomainError):
            p.size(-1)


def test_empty_array_dimension_mismatch():
    # test/onedim/complex_forward.jl:35-37, twodim/complex_forward.jl:45-47
    with pytest.raises(P.DimensionMismatch, match="strictly positive"):
        P.plan_fft(np.zeros(0, dtype=np.complex128))
    with pytest.raises(P.DimensionMismatch):
        P.plan_bfft(np.zeros((0, 0), dtype=np.complex128))


def test_mismatch_plan_and_array():
    # test/argument_checking.jl:131-188 (complex parts)
    xc1 = np.zeros(3, dtype=np.complex128)
    xc1p = np.zeros(4, dtype=np.complex128)
    with pytest.raises(P.DimensionMismatch):
        P.plan_fft(xc1) * xc1p
    with pytest.raises(P.DimensionMismatch):
        P.plan_bfft(xc1) * xc1p
    xc2 = np.zeros((3, 3), dtype=np.complex128)
    xc2p = np.zeros((4, 4), dtype=np.complex128)
    for region in (1, 2):
        with pytest.raises(P.DimensionMismatch):
            P.plan_fft(xc2, region) * xc2p
        with pytest.raises(P.DimensionMismatch):
            P.plan_bfft(xc2, region) * xc2p
    with pytest.raises(P.DimensionMismatch):
        P.plan_fft(xc2) * xc2p


def test_region_outside_dimensions():
    # test/argument_checking.jl:189-203
    xc3 = np.zeros((3, 3, 3), dtype=np.complex128)
    yc3 = np.zeros((5, 5, 5), dtype=np.complex128)
    with pytest.raises(P.DimensionMismatch):
        P.plan_fft(xc3) * yc3
    invalid_p = P.plan_fft(np.zeros((3,) * 5, dtype=np.complex128), range(3, 6))
    with pytest.raises(P.DimensionMismatch, match=r"^Plan region is outside array dimensions\.$"):
        invalid_p * xc3
    xc4 = np.zeros((1, 5, 5, 5), dtype=np.complex128, order="F")
    with pytest.raises(P.DimensionMismatch, match=r"^Plan region is outside array dimensions\.$"):
        P.mul_(xc4, invalid_p, xc4)


def test_mismatch_input_output():
    # test/argument_checking.jl:206-227
    x1 = np.zeros(3, dtype=np.complex128)
    y1 = np.zeros(4, dtype=np.complex128)
    with pytest.raises(P.DimensionMismatch):
        P.mul_(y1, P.plan_fft(x1), x1)
    for N in (2, 3, 4):
        xN = np.zeros((3,) * N, dtype=np.complex128, order="F")
        yN = np.zeros((4,) * N, dtype=np.complex128, order="F")
        for region in range(1, N + 1):
            with pytest.raises(P.DimensionMismatch):
                P.mul_(yN, P.plan_fft(xN, region), xN)
        with pytest.raises(P.DimensionMismatch):
            P.mul_(yN, P.plan_fft(xN), xN)
        with pytest.raises(P.DimensionMismatch):
            P.mul_(yN, P.plan_fft(xN, range(1, N)), xN)


def test_mutated_region_is_invalid():
    # test/argument_checking.jl:249-266
    for n in (3, 4, 5):
        x = np.zeros((2,) * n, dtype=np.complex128, order="F")
        p1 = P.plan_fft(x, list(range(1, n)))
        p1.region.append(n)
        with pytest.raises(P.DimensionMismatch, match=r"^Region is invalid\.$"):
            p1 * x
        p2 = P.plan_fft(x, list(range(1, n)))
        p2.region[0:2] = [2, 1]
        with pytest.raises(P.DimensionMismatch, match=r"^Region is invalid\.$"):
            p2 * x


@pytest.mark.parametrize("shape,region,dims", [((64,), None, (65,)), ((64,), None, (32, 2)),
                                                ((16, 8), None, (16, 9)), ((16, 8), 2, (16, 9)),
                                                ((4, 4, 4), (1, 3), (4, 4)), ((4, 4), None, (5, 4, 4)),
                                                ((8, 8), (1, 2), (8,))])
def test_execute_checks_shape_before_any_launch(shape, region, dims):
    """fftahip_execute and fftahip_execute_host apply mul!'s argument checks
    themselves (src/plan.jl:157-164, 170-182, 210-222, 260-272): a shape that
    does not fit the plan returns DimensionMismatch (status 1) before anything
    is launched or copied.  The data pointers are deliberately bogus: they are
    never touched."""
    import ctypes

    L = _lib.load()
    p = P.Plan(shape, 64, region, -1)
    bogus = ctypes.c_void_p(0x10)
    d = _lib.i64_array(dims)
    assert L.fftahip_execute(p.handle, len(dims), d, bogus, bogus, None) == _lib.ERR_DIMENSION_MISMATCH
    msg = _lib.last_error()
    assert msg.startswith(("plan has", "Plan region is outside", "Region is invalid", "input array")), msg
    assert L.fftahip_execute_host(p.handle, len(dims), d, bogus, bogus) == _lib.ERR_DIMENSION_MISMATCH
    with pytest.raises(P.DimensionMismatch):
        p.execute_device(dims, 0x10, 0x10)


def test_execute_rejects_negative_dims():
    import ctypes

    L = _lib.load()
    p = P.Plan((8,), 32, None, -1)
    bogus = ctypes.c_void_p(0x10)
    assert L.fftahip_execute(p.handle, 1, _lib.i64_array([-8]), bogus, bogus, None) == _lib.ERR_ARGUMENT


@pytest.mark.parametrize("shape,region", [((2, 2**32), 1), ((4, 2**31), 1), ((2**31, 2), 2), ((2, 2**31, 2), 3)])
def test_line_count_capacity(shape, region):
    """Line indices are 32-bit on the device: a pass over 2^31 or more lines is
    refused with a capacity error at planning time instead of wrapping."""
    with pytest.raises(fftahip.FFTAHipError, match="2\\^31"):
        P.debug_schedule(shape, 32, -1, region)


def test_line_count_below_capacity_plans():
    s = P.debug_schedule((2, 2**30), 32, -1, 1)
    assert s["launches"] and all(la["nlines"] < 2**31 for la in s["launches"])


def test_describe_reports_ffta_tree():
    p = P.plan_fft(np.zeros(275, dtype=np.complex128))
    d = p.describe()
    tree = [(q["type"], q["sz"], q["s_in"], q["s_out"]) for q in d["dims"][0]["ffta"]]
    assert tree == [("COMPOSITE_FFT", 275, 1, 1), ("COMPOSITE_FFT", 25, 11, 11), ("DFT", 5, 5, 55),
                    ("DFT", 5, 55, 1), ("DFT", 11, 25, 1)]


# ---------------------------------------------------------------- schedule emulation

SHAPES_1D = [1, 2, 3, 16, 17, 64, 100, 103, 243, 1000, 1009, 2187, 3000, 4096, 5329, 8192, 10007, 16384,
             30000, 65536, 131072, 177147, 262144]


def _check_schedule(shape, prec, direction, region=None, lds=160 * 1024, cutoff=73):
    s = P.debug_schedule(shape, prec, direction, region, lds, cutoff)
    rng = np.random.default_rng(sum(shape))
    x = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    y = E.run_plan_schedule(s, x, 8 if prec == 32 else 16, lds)
    reg = P._region_tuple(region, len(shape))
    axes = [r - 1 for r in reg if r <= len(shape)]
    if direction == -1:
        ref = np.fft.fftn(x, axes=axes)
    else:
        ref = np.fft.ifftn(x, axes=axes) * np.prod([shape[a] for a in axes])
    err = np.linalg.norm(y - ref) / max(np.linalg.norm(ref), 1e-300)
    assert err < 1e-12, (shape, prec, direction, region, lds, err)


@pytest.mark.parametrize("lds", [160 * 1024, 64 * 1024])
@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("n", SHAPES_1D)
def test_schedule_1d(n, prec, lds):
    _check_schedule((n,), prec, -1 if n % 2 else 1, None, lds)


@pytest.mark.parametrize("shape,region", [((8, 11), None), ((64, 64), None), ((30, 21), 2), ((5, 6, 7), 2),
                                          ((5, 6, 7), (1, 3)), ((5, 6, 7), None), ((4096, 8), 2),
                                          ((3000, 9), None), ((103, 5, 13), None), ((1024, 1024), None),
                                          ((17, 4099), None), ((26, 33, 35, 4), (2, 4)), ((3, 5), 3)])
@pytest.mark.parametrize("prec", [32, 64])
def test_schedule_nd(shape, region, prec):
    _check_schedule(shape, prec, -1, region)
    _check_schedule(shape, prec, 1, region, 64 * 1024)


def test_schedule_batch_uneven():
    rng = np.random.default_rng(5)
    lengths = np.concatenate([rng.integers(1, 300, 40), [4096, 4099, 8192, 16384, 20011, 30000, 65536, 65521, 59049],
                              rng.integers(16, 70000, 12)])
    dirs = np.where(rng.uniform(size=lengths.size) < 0.5, -1, 1)
    s = P.debug_batch_schedule(lengths, dirs, 32)
    x = rng.standard_normal(lengths.sum()) + 1j * rng.standard_normal(lengths.sum())
    y = E.run_batch_schedule(s, x, 8, 160 * 1024)
    off = 0
    for n, d in zip(lengths, dirs):
        xi = x[off:off + n]
        ref = np.fft.fft(xi) if d == -1 else np.fft.ifft(xi) * n
        assert np.linalg.norm(y[off:off + n] - ref) <= 1e-12 * max(np.linalg.norm(ref), 1e-300)
        off += n


def test_three_pass_rows_avoid_wide_workgroups():
    """The three-pass Bluestein row cost (plan.cpp row_cost) weights rows
    that need 1024-thread workgroups: N = 27779 runs 15 x 4096 (power-of-two
    rows at 256 threads), not 9 x 6561, and the rows stay within 512 threads
    for lengths across the three-pass range."""
    s = P.debug_schedule((27779,), 32, -1)
    assert [(la["L"], la["mid_mode"]) for la in s["launches"]] == [(15, 0), (4096, 1), (15, 0)]
    for n in (8209, 12011, 20011, 27779, 40009, 49195, 65521):
        rows = [la for la in P.debug_schedule((n,), 32, -1)["launches"] if la["mid_mode"] == 1 and not la["reg"]]
        assert rows and all(la["block"] <= 512 for la in rows), (n, [(la["L"], la["block"]) for la in rows])


# Bluestein lines of 2N-1 <= L points: the first sweep skips the zero tail,
# the second computes only the positions the chirp store reads
ZERO_PAD_LENGTHS = [457, 4099, 2 * 7919, 3 * 4099, 1031, 2053]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("n", ZERO_PAD_LENGTHS)
def test_schedule_zero_padded_bluestein(n, prec):
    """The planner marks one-kernel Bluestein lines zero-padded (plan.cpp
    set_zero_pad); the emulator replays both sweeps with the tail slots left
    as NaN and the last stage's outputs t >= (r + 1) // 2 unwritten, and the
    whole schedule against numpy's FFT."""
    s = P.debug_schedule((n,), prec, -1)
    zp = [la for la in s["launches"] if la.get("zero_pad")]
    # FP64 runs 2 * 7919 as three fused passes (no one-kernel line)
    assert zp or (prec == 64 and n == 2 * 7919), [(la["L"], la["rad"], la["load_mode"]) for la in s["launches"]]
    for la in zp:
        assert la["chirp_n"] <= la["zero_from"] < la["L"]
    _check_schedule((n,), prec, -1)
    _check_schedule((n,), prec, 1)


@pytest.mark.parametrize("rad", [[8, 16, 9, 8], [16, 8, 8, 16], [3, 5, 3], [5, 4, 5], [9, 4, 9], [2, 3, 2], [16],
                                 [4, 5, 4], [3, 16, 3]])
@pytest.mark.parametrize("Tl", [1, 4, 16, 64])
def test_emulate_sweep_zero_pad(rad, Tl):
    L = int(np.prod(rad))
    # at least the threads every stage's butterflies need
    Tl = max([Tl] + [-(-(L // r) // max(1, E.EPT // r)) for r in rad])
    E.emulate_sweep(L, rad, Tl, 0, zero_first=True)
    E.emulate_sweep(L, rad, Tl, 0, prune_last=True)
    E.emulate_sweep(L, rad, Tl, 5, zero_first=True, prune_last=True)


def test_zero_pad_checker_rejects_wrong_bounds():
    """A zero tail that starts one row of the first stage too early, or more
    inputs than the bound covers, or a radix without the variant, is refused."""
    s = P.debug_schedule((4099,), 32, -1)
    la = next(la for la in s["launches"] if la.get("zero_pad"))
    E.Emulator._check_zero_pad(la, la["zero_pad"])
    r0 = la["rad"][0]
    bad = dict(la, zero_from=la["zero_from"] - la["L"] // r0)
    with pytest.raises(E.ScheduleError):
        E.Emulator._check_zero_pad(bad, 1)
    bad = dict(la, chirp_n=la["zero_from"] + 1)
    with pytest.raises(E.ScheduleError):
        E.Emulator._check_zero_pad(bad, 1)
    with pytest.raises(E.ScheduleError):
        E.Emulator._check_zero_pad(bad, 2)
    bad = dict(la, rad=[7] + la["rad"][1:], zero_from=4 * (la["L"] // 7))
    with pytest.raises(E.ScheduleError):
        E.Emulator._check_zero_pad(bad, 1)
    bad = dict(la, store_mode=1)
    with pytest.raises(E.ScheduleError):
        E.Emulator._check_zero_pad(bad, 3)


# primes up to 751 as direct-DFT stages (BLUESTEIN_CUTOFF 752): long lines
# whose staged roots (four copies per stage) leave no room for one workgroup
# split into more passes; every schedule still emulates exactly
LARGE_DIRECT_CASES = [743, 743 * 4, 743 * 12, 751 * 8, 701 * 16, 509 * 15, 653 * 9, 599 * 2 * 3, 743 * 743]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("n", LARGE_DIRECT_CASES)
def test_schedule_large_direct_primes(n, prec):
    s = P.debug_schedule((n,), prec, -1, None, 160 * 1024, 752)
    assert all(la["load_mode"] == 0 for la in s["launches"]), [(la["L"], la["rad"]) for la in s["launches"]]
    for la in s["launches"]:
        assert la["lds"] <= 160 * 1024
    _check_schedule((n,), prec, -1 if n % 2 else 1, None, 160 * 1024, 752)


DIRECT_PRIME_LENGTHS = [17, 19, 23 * 16, 31 * 27, 37 * 64, 41 * 43, 47 * 128, 53 * 35, 59 * 3, 61 * 61, 67 * 2,
                        71 * 256, 71 * 71 * 4, 17 * 4096, 29 * 31 * 37, 43 * 1024, 67 * 67]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("n", DIRECT_PRIME_LENGTHS)
def test_schedule_direct_primes(n, prec):
    """Primes 17..71 (below the default BLUESTEIN_CUTOFF of 73) run as direct
    DFT stages (FFTA's DFT leaf, src/algos.jl:135-160), not as convolutions:
    every launch is a plain Stockham pass, and the emulator replays the
    stage's work items, write map and root table against numpy's FFT."""
    s = P.debug_schedule((n,), prec, -1)
    assert all(la["load_mode"] == 0 and la["mid_mode"] == 0 and la["store_mode"] == 0 for la in s["launches"]), n
    assert any(r >= 17 for la in s["launches"] for r in la["rad"]), n
    _check_schedule((n,), prec, -1)
    _check_schedule((n,), prec, 1, None, 64 * 1024)


@pytest.mark.parametrize("shape,region", [((51, 71), None), ((19, 5, 23), (1, 3)), ((64, 47), 2), ((1000, 59), 2)])
def test_schedule_direct_primes_nd(shape, region):
    for prec in (32, 64):
        _check_schedule(shape, prec, -1, region)


RADER_LENGTHS = [79, 101, 257, 641, 769, 1009, 1153, 3457, 4481, 7681, 8191, 12289, 103, 139, 283,
                 97 * 8, 101 * 27, 257 * 64, 3457 * 9, 12289 * 4, 139 * 5 * 7]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("n", RADER_LENGTHS)
def test_schedule_rader(n, prec):
    """Convolution primes P whose P - 1 has an in-LDS schedule run Rader's
    algorithm (a cyclic convolution of P - 1 points in one workgroup) instead
    of Bluestein's (>= 2P - 1 points), on their own or as the prime part of a
    mixed length; the emulator recomputes the slot table and spectrum from
    the smallest primitive root and checks the result against numpy's FFT."""
    s = P.debug_schedule((n,), prec, -1)
    rad = [la for la in s["launches"] if la["load_mode"] == 2]
    big = max(q for q in range(2, n + 1) if n % q == 0 and all(q % d for d in range(2, int(q ** 0.5) + 1)))
    # a Rader line of P - 1 points plus two side slots must fit one workgroup's
    # LDS; with a direct-DFT stage in P - 1 the cost model may keep Bluestein
    smooth13 = all(q <= 13 for q in range(2, big) if (big - 1) % q == 0 and all(q % d for d in range(2, q)))
    if big - 1 <= (16384 if prec == 32 else 8000) and smooth13:
        assert rad and rad[0]["L"] == big - 1, [(la["L"], la["load_mode"]) for la in s["launches"]]
    _check_schedule((n,), prec, -1)
    _check_schedule((n,), prec, 1, None, 64 * 1024)


def test_rader_not_for_rough_p_minus_1():
    """P - 1 with a prime factor above 71 (4099: 4098 = 2 3 683) stays Bluestein."""
    s = P.debug_schedule((4099,), 32, -1)
    assert all(la["load_mode"] != 2 for la in s["launches"])
    assert any(la["load_mode"] == 1 for la in s["launches"])


@pytest.mark.parametrize("shape,region", [((101, 12), 1), ((12, 101), 2), ((97 * 8, 3), None), ((5, 257, 2), 2)])
def test_schedule_rader_nd(shape, region):
    for prec in (32, 64):
        _check_schedule(shape, prec, -1, region)
        _check_schedule(shape, prec, 1, region)


@pytest.mark.parametrize("n,prec", [(20011, 32), (3199, 32), (28693, 32), (65521, 32), (10007, 64), (4036, 64),
                                    (3 * 4099, 32)])
def test_register_passes(n, prec):
    """Short column passes run as register line passes (one line per thread);
    the emulator checks their geometry and result like any other pass."""
    s = P.debug_schedule((n,), prec, -1)
    assert any(la.get("reg") for la in s["launches"]), [(la["L"], la["col_in"], la["col_out"]) for la in s["launches"]]
    _check_schedule((n,), prec, -1)
    _check_schedule((n,), prec, 1)


def test_autotune_table(tmp_path):
    """The autotuner's table: loaded entries replace the default radix rule for
    their (precision, length), malformed ones are skipped, plans built
    afterwards use them (and still transform exactly), save/load round-trips."""
    from fftahip import autotune

    autotune.clear()
    try:
        assert autotune.schedule(32, 2025) == [9, 9, 5, 5]
        src = tmp_path / "t.txt"
        src.write_text("# precision length radices\n32 2025 5 5 9 9\n32 4096 8 8 8 8\n64 100 7 7\n32 12 16\n"
                       "32 4096 32 128\n")
        assert autotune.load_table(str(src)) == 2
        assert autotune.schedule(32, 2025) == [5, 5, 9, 9]
        assert autotune.schedule(64, 2025) == [9, 9, 5, 5]
        assert autotune.schedule(32, 4096) == [8, 8, 8, 8]
        s = P.debug_schedule((2025,), 32, -1)
        assert s["launches"][0]["rad"] == [5, 5, 9, 9]
        _check_schedule((2025,), 32, -1)
        _check_schedule((4096, 3), 32, 1)
        out = tmp_path / "o.txt"
        autotune.save(str(out))
        autotune.clear()
        assert autotune.schedule(32, 4096) == [16, 16, 16]
        assert autotune.load_table(str(out)) == 2
        assert autotune.schedule(32, 4096) == [8, 8, 8, 8]
    finally:
        autotune.clear()


def test_autotune_tile_table(tmp_path):
    """Strided tile shapes in the tuning table ("S precision length bytes
    lines"): a long strided line of a tuned length runs in one pass of few
    lines per tile instead of two 40 KB passes, the schedule still transforms
    exactly and stays in bounds (emulator), malformed entries are skipped and
    save/load/clear round-trip them."""
    from fftahip import autotune

    autotune.clear()
    try:
        dflt = P.debug_schedule((8, 4096), 32, -1, 2)["launches"]
        assert len(dflt) == 2 and all(la["L"] == 64 for la in dflt)
        src = tmp_path / "t.txt"
        src.write_text("S 32 4096 143360 4\nS 32 2187 71680 8\nS 64 12 100 4\nS 32 99 143360 0\nS 16 5 143360 4\n")
        assert autotune.load_table(str(src)) == 2
        assert autotune.tile(32, 4096) == (143360, 4)
        assert autotune.tile(64, 4096) == (0, 0)
        la = P.debug_schedule((8, 4096), 32, -1, 2)["launches"]
        assert len(la) == 1 and la[0]["L"] == 4096 and la[0]["C"] == 4 and la[0]["col_in"] == 1
        assert la[0]["lds"] <= 160 * 1024
        _check_schedule((8, 4096), 32, -1, 2)
        _check_schedule((8, 4096), 32, 1, None)
        # the contiguous dimension keeps its one-pass schedule: same radices as the strided one
        la2 = P.debug_schedule((4096, 4096), 32, -1)["launches"]
        assert [x["L"] for x in la2] == [4096, 4096] and la2[0]["rad"] == la2[1]["rad"]
        out = tmp_path / "o.txt"
        autotune.save(str(out))
        assert "S 32 4096 143360 4" in out.read_text()
        autotune.clear()
        assert autotune.tile(32, 4096) == (0, 0)
        assert autotune.load_table(str(out)) == 2
        assert autotune.tile(32, 2187) == (71680, 8)
    finally:
        autotune.clear()


def test_mod_q31_exact():
    """The register passes reduce a^2 mod 2N and a*s mod N (a, s < N < 2^31)
    through an FP64 quotient (mod_q31); it must equal the exact remainder,
    including next to multiples of d where the quotient is off by one."""
    import ctypes

    lib = _lib.load()
    rng = np.random.default_rng(7)
    cases = []
    for N in [1, 2, 3, 73, 1009, 65521, 1000003, (1 << 26) - 5, (1 << 30) + 3, (1 << 31) - 19, (1 << 31) - 1]:
        a = np.concatenate([rng.integers(0, N, 200), [0, 1, N - 1, max(0, N - 2), N // 2]]).astype(object)
        s = np.concatenate([rng.integers(0, N, 200), [N - 1, 1, 0, N // 3, max(1, N - 7)]]).astype(object)
        cases.append((2 * N, [int(v) * int(v) for v in a]))
        cases.append((N, [int(x) * int(y) for x, y in zip(a, s)]))
        # multiples of d and their neighbours, quotients up to 2^31 - 1
        d = 2 * N
        k = np.concatenate([rng.integers(1, max(2, min(N, 1 << 31)), 100), [1, max(1, min(N, (1 << 31) - 1) - 1)]])
        near = []
        for q in k:
            for e in (-1, 0, 1):
                m = int(q) * d + e
                if 0 <= m < (1 << 62) and m // d < (1 << 31):
                    near.append(m)
        cases.append((d, near))
    for d, ms in cases:
        m = np.array(ms, dtype=np.uint64)
        r = np.zeros_like(m)
        st = lib.fftahip_debug_mod_q31(m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), d,
                                       r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), m.size)
        assert st == 0
        exact = np.array([int(v) % d for v in ms], dtype=np.uint64)
        assert np.array_equal(r, exact), (d, [(int(a), int(b), int(c)) for a, b, c in zip(m, r, exact) if b != c][:5])


def test_debug_chirp_argument_checks():
    """The chirp read-back entry points refuse bad arguments before touching a device."""
    import ctypes

    lib = _lib.load()
    out = np.zeros(8, dtype=np.complex128)
    ptr = out.ctypes.data_as(ctypes.c_void_p)
    assert lib.fftahip_debug_chirp(16, 10, 0, 1, ptr) == 2
    assert lib.fftahip_debug_chirp(64, 1 << 31, 0, 1, ptr) == 2
    assert lib.fftahip_debug_chirp(64, 10, 8, 3, ptr) == 2
    assert lib.fftahip_debug_chirp(64, 10, -1, 1, ptr) == 2
    assert lib.fftahip_debug_chirp(64, 10, 0, 0, None) == 0
    dptr = out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.fftahip_debug_chirp_steps(10, 20, 4, dptr) == 2
    assert lib.fftahip_debug_chirp_steps(10, 3, 0, dptr) == 2


def _chirp_direct(n: int, N: int) -> complex:
    # the chirp table's definition: exp(i pi r / N), r = n^2 mod 2N folded into (-N, N]
    r = (n * n) % (2 * N)
    if r > N:
        r -= 2 * N
    return complex(np.cos(np.pi * r / N), np.sin(np.pi * r / N))


def _chirp_root(m: int, N: int) -> complex:
    # kernels.hip chirp_root: (m mod 2N) folded, with m < 2^64
    assert 0 <= m < 1 << 64
    return _chirp_direct_r(m % (2 * N), N)


def _chirp_direct_r(r: int, N: int) -> complex:
    if r > N:
        r -= 2 * N
    return complex(np.cos(np.pi * r / N), np.sin(np.pi * r / N))


@pytest.mark.parametrize("N", [73, 1009, 20011, 1000003, (1 << 31) - 1])
def test_register_pass_chirp_factorisation(N):
    """The register passes form the chirp of positions n = a + t*s as
    c[a] * w^t * d[t] with the kernel's unsigned 64-bit reductions: a*a and
    2*((a*s) mod N) for c[a] and w, ((t*s) mod 2N)^2 for the step table d[t].
    Each intermediate stays below 2^64 for N < 2^31, and the product equals
    the table's chirp at every position (incl. positions a + t*s >= 2N)."""
    rng = np.random.default_rng(N)
    for _ in range(40):
        a = int(rng.integers(0, N))
        s = int(rng.integers(1, N))
        L = int(rng.integers(2, 49))
        c = _chirp_root(a * a, N)
        w = _chirp_root(2 * ((a * s) % N), N)
        for t in range(L):
            u = (t * s) % (2 * N)
            d = _chirp_root(u * u, N)
            got = c * (w ** t) * d
            want = _chirp_direct(a + t * s, N)
            assert abs(got - want) < 1e-9, (N, a, s, t)
