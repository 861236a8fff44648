"""CPU emulation of the library's launch schedules (test helper).

Runs the line passes that the planner emits (fftahip_debug_schedule /
fftahip_debug_batch_schedule dry runs) in exact complex128 arithmetic with
numpy, so that the pass geometry -- line/element addressing, multi-pass
twiddle exponents, Bluestein chirp and spectrum indices, scratch layout --
can be checked on the host against a reference FFT.  Every global access is
bounds-checked and every pass is checked for two lines writing one address.

`emulate_sweep` follows the kernel's in-LDS Stockham stage loop (per-thread
butterfly coverage, stage index maps, stage twiddle indices) for one line.
"""
from __future__ import annotations

import numpy as np

TK_STAGE, TK_TWLO, TK_TWHI, TK_CHIRP, TK_BHAT, TK_IN, TK_OUT, TK_SCRATCH, TK_BHATROWS, TK_OFFTAB, TK_TW16, \
    TK_CHIRPD, TK_RTAB = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13
LOAD_RADER, STORE_RADER, RADER_SIDE = 2, 3, 2
EPT = 16
# in-register codelets (codelets.h); the primes 17, 19 and 23 only in the
# FP32 direct-DFT kernel variants, one butterfly per thread
CODELETS = {2, 3, 4, 5, 7, 8, 9, 11, 13, 16}
PRIME_CODELETS = {17, 19, 23}
# codelet radices with zero-padded first / last stage variants (common.h
# zero_pad_radix): inputs / outputs t >= (r + 1) // 2 dropped
ZERO_PAD_RADICES = {2, 3, 4, 5, 8, 9, 16}


def codelets(esize: int) -> set:
    """Radices that run as codelets in a pass of element size esize (8: FP32)."""
    return CODELETS | PRIME_CODELETS if esize == 8 else CODELETS
# direct-DFT stage of a prime radix P >= 17 (device.h stage_big): work items
# of BIG_KP output pairs, BIG_HOLD per thread
BIG_KP, BIG_HOLD = 4, 2
LINE_TAB_BYTES = 6 * 8  # per-line offset tables after a tile's lines
# register line passes: lines per workgroup, and the lengths with a codelet
REG_BLOCK = 256
REG_LENGTHS = {n for n in range(2, 49) if all(q in (2, 3, 5, 7, 11, 13) for q in
               [f for f in range(2, n + 1) if n % f == 0 and all(f % d for d in range(2, f))])}


class ScheduleError(AssertionError):
    pass


def chirp_fwd(n: int) -> np.ndarray:
    k = np.arange(n, dtype=np.int64)
    p = (k * k) % (2 * n)
    return np.exp(1j * np.pi * p / n)


def bhat_of(n: int, m: int) -> np.ndarray:
    b = chirp_fwd(n)
    ext = np.zeros(m, dtype=np.complex128)
    ext[:n] = b
    ext[m - np.arange(1, n)] = b[1:]
    return np.fft.fft(ext) / m


def primitive_root(P: int) -> int:
    qs = {q for q in range(2, P) if (P - 1) % q == 0 and all(q % d for d in range(2, int(q ** 0.5) + 1))}
    return next(g for g in range(2, P) if all(pow(g, (P - 1) // q, P) != 1 for q in qs))


def rader_tables(P: int):
    """Slot table rtab[t] = (-log_g t) mod (P-1) and spectrum FFT(b)/(P-1),
    b_r = exp(-2 pi i g^r / P), g the smallest primitive root."""
    L = P - 1
    g = primitive_root(P)
    rtab = np.zeros(P, dtype=np.int64)
    cur = 1
    for e in range(L):
        rtab[cur] = (L - e) % L
        cur = cur * g % P
    b = np.exp(-2j * np.pi * np.array([pow(g, r, P) for r in range(L)]) / P)
    return rtab, np.fft.fft(b) / L


def stage_table(L: int, rad, cod=CODELETS) -> np.ndarray:
    """Per-stage twiddles as the device table holds them: stage s (Ns = product
    of earlier radices, radix R) stores w_{Ns*R}^{t*k}, k < Ns, t = 1..R-1, at
    (Ns-1) + k*(R-1) + t-1."""
    extra = sum(R for R in rad if R not in cod)
    tab = np.zeros(max(1, L - 1 + extra), dtype=np.complex128)
    Ns = 1
    for R in rad:
        for k in range(Ns):
            t = np.arange(1, R)
            tab[(Ns - 1) + k * (R - 1) + t - 1] = np.exp(-2j * np.pi * ((t * k) % (Ns * R)) / (Ns * R))
        Ns *= R
    # then the P roots w_P^m of each direct-DFT stage
    off = L - 1
    for R in rad:
        if R not in cod:
            tab[off:off + R] = np.exp(-2j * np.pi * np.arange(R) / R)
            off += R
    return tab


def _big_stage(lds, tw, L, R, Ns, Tl, sh, roots_off, P):
    """The direct-DFT stage: work item w = jl + h*Tl (h < BIG_HOLD) covers
    butterfly j = w % Lr and output pairs k = g*BIG_KP+1 .. of group g = w // Lr
    (plus X_0 for g = 0); reads unstepped, writes base + k*Ns."""
    Lr = L // R
    H = (R - 1) // 2
    G = (H + BIG_KP - 1) // BIG_KP
    if Tl * BIG_HOLD < Lr * G:
        raise ScheduleError(f"L={L} direct radix {R}: {Tl} threads x {BIG_HOLD} items < {Lr * G}")
    if roots_off + R > tw.size:
        raise ScheduleError("direct-DFT roots beyond the stage table")
    roots = tw[roots_off:roots_off + R]
    if not np.allclose(roots, np.exp(-2j * np.pi * np.arange(R) / R), rtol=0, atol=1e-15):
        raise ScheduleError("direct-DFT roots wrong")
    # the kernel reads root (i*k0 mod R) + u*i for k = k0 + u from the R roots
    # repeated BIG_KP times in LDS: in range, and the root of i*k mod R
    i = np.arange(1, H + 1)[:, None, None]
    k0 = (np.arange(G) * BIG_KP + 1)[None, :, None]
    u = np.arange(BIG_KP)[None, None, :]
    ridx = (i * k0) % R + u * i
    if ridx.max() >= BIG_KP * R or np.any(ridx % R != (i * (k0 + u)) % R):
        raise ScheduleError(f"direct radix {R}: root index beyond the staged table")
    new = np.full(P, np.nan + 0j)
    hits = np.zeros(P, dtype=np.int64)
    t = np.arange(R)
    w = np.arange(Tl * BIG_HOLD)
    w = w[w < Lr * G]
    g, j = w // Lr, w % Lr
    v = lds[_pad(j[:, None] + t[None, :] * Lr, sh)]
    k = j % Ns
    if Ns > 1:
        widx = (Ns - 1) + k[:, None] * (R - 1) + t[None, 1:] - 1
        if widx.max() >= L - 1:
            raise ScheduleError("stage twiddle index out of range")
        v[:, 1:] = v[:, 1:] * tw[widx]
    V = np.fft.fft(v, axis=1)
    base = (j // Ns) * Ns * R + k
    for row in range(w.size):
        ks = [kk for kk in range(g[row] * BIG_KP + 1, g[row] * BIG_KP + BIG_KP + 1) if kk <= H]
        outs = ks + [R - kk for kk in ks] + ([0] if g[row] == 0 else [])
        for kk in outs:
            dst = _pad(base[row] + kk * Ns, sh)
            if dst >= P:
                raise ScheduleError("stage write out of range")
            new[dst] = V[row, kk]
            hits[dst] += 1
    return new, hits


def _pad(i, sh):
    return i + (i >> sh) if sh else i


def emulate_sweep(L: int, rad, Tl: int, sh: int = 0, rng=None, cod=CODELETS, zero_first: bool = False,
                  prune_last: bool = False) -> None:
    """Replay the kernel's stage loop for one line with its exact LDS
    positions (padding, stepped addressing) and table indexing; raise if an
    output slot is missed or written twice, or if the result is not the DFT.

    zero_first: the line is zero from ((r + 1) // 2) * (L // r) on (r the
    first radix), those LDS slots hold garbage (NaN) and the first stage takes
    its inputs t >= (r + 1) // 2 as zero without reading them.  prune_last:
    the last stage writes its outputs t < (r + 1) // 2 only, and only the
    positions below ((r + 1) // 2) * (L // r) are checked."""
    rng = rng or np.random.default_rng(0)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    tw = stage_table(L, rad, cod)
    P = _pad(L - 1, sh) + 1
    lds = np.full(P, np.nan + 0j)
    zf = L
    if zero_first:
        if rad[0] not in ZERO_PAD_RADICES or rad[0] not in cod:
            raise ScheduleError(f"zero-padded first stage of radix {rad[0]}")
        zf = (rad[0] + 1) // 2 * (L // rad[0])
        x[zf:] = 0
    lds[_pad(np.arange(zf), sh)] = x[:zf]
    keep = L
    if prune_last:
        if rad[-1] not in ZERO_PAD_RADICES or rad[-1] not in cod:
            raise ScheduleError(f"pruned last stage of radix {rad[-1]}")
        keep = (rad[-1] + 1) // 2 * (L // rad[-1])
    pmask = (1 << sh) - 1 if sh else 0
    Ns = 1
    roots_off = L - 1
    for s_i, R in enumerate(rad):
        Lr = L // R
        if R not in cod:
            new, hits = _big_stage(lds, tw, L, R, Ns, Tl, sh, roots_off, P)
            roots_off += R
            if not np.all(hits[_pad(np.arange(L), sh)] == 1) or hits.sum() != L:
                raise ScheduleError(f"L={L} direct radix {R}: slots written {np.unique(hits)} times")
            lds = new
            Ns *= R
            continue
        NB = max(1, EPT // R)
        if NB < 1:
            raise ScheduleError(f"radix {R} exceeds {EPT} registers")
        if Tl * NB < Lr:
            raise ScheduleError(f"L={L} radix {R}: {Tl} threads x {NB} butterflies < {Lr}")
        rstep = (Lr & pmask) == 0
        wstep = (Ns & pmask) == 0 or (Ns == 1 and (pmask + 1) % R == 0)
        rstride = Lr + ((Lr >> sh) if sh else 0)
        wstride = Ns + ((Ns >> sh) if sh else 0)
        new = np.full(P, np.nan + 0j)
        hits = np.zeros(P, dtype=np.int64)
        t = np.arange(R)
        for b in range(NB):
            j = np.arange(Tl) + b * Tl
            j = j[j < Lr]
            if j.size == 0:
                continue
            if rstep:
                src = _pad(j, sh)[:, None] + t[None, :] * rstride
            else:
                src = _pad(j[:, None] + t[None, :] * Lr, sh)
            kin = (R + 1) // 2 if (zero_first and s_i == 0) else R
            v = lds[src]
            v[:, kin:] = 0
            k = j % Ns
            if Ns > 1:
                widx = (Ns - 1) + k[:, None] * (R - 1) + t[None, 1:] - 1
                if widx.max() >= tw.size:
                    raise ScheduleError("stage twiddle index out of range")
                v[:, 1:] = v[:, 1:] * tw[widx]
            V = np.fft.fft(v, axis=1)
            base = (j // Ns) * Ns * R + k
            if wstep:
                dst = _pad(base, sh)[:, None] + t[None, :] * wstride
            else:
                dst = _pad(base[:, None] + t[None, :] * Ns, sh)
            if dst.max() >= P:
                raise ScheduleError("stage write out of range")
            kout = (R + 1) // 2 if (prune_last and s_i == len(rad) - 1) else R
            new[dst[:, :kout]] = V[:, :kout]
            np.add.at(hits, dst[:, :kout].ravel(), 1)
        nw = keep if (prune_last and s_i == len(rad) - 1) else L
        if not np.all(hits[_pad(np.arange(nw), sh)] == 1) or hits.sum() != nw:
            raise ScheduleError(f"L={L} stage radix {R}: slots written {np.unique(hits)} times")
        lds = new
        Ns *= R
    if Ns != L:
        raise ScheduleError(f"radices {rad} multiply to {Ns}, not {L}")
    line = lds[_pad(np.arange(keep), sh)]
    ref = np.fft.fft(x)[:keep]
    err = np.linalg.norm(line - ref) / max(np.linalg.norm(ref), 1e-300)
    if not err <= 1e-12:  # NaN: a slot read before it was written
        raise ScheduleError(f"sweep of L={L} radices {rad} wrong (rel err {err})")


class Emulator:
    def __init__(self, x: np.ndarray, out_elems: int, scratch_bytes: int, esize: int, offtab=None):
        self.esize = esize
        self.bufs = {
            TK_IN: np.asarray(x, dtype=np.complex128).ravel(order="F").copy(),
            TK_OUT: np.zeros(out_elems, dtype=np.complex128),
            TK_SCRATCH: np.full(scratch_bytes // esize, np.nan + 1j * np.nan, dtype=np.complex128),
        }
        self.offtab = np.asarray(offtab if offtab is not None else [], dtype=np.int64)
        self.sweeps_checked = set()

    def _buf(self, tok):
        kind, off = tok
        if kind not in self.bufs:
            raise ScheduleError(f"unexpected buffer token {tok}")
        if off % self.esize:
            raise ScheduleError("misaligned buffer offset")
        return self.bufs[kind], off // self.esize

    def run(self, launches, lds_per_block: int) -> None:
        for la in launches:
            if la["op"] == 1:
                src, so = self._buf(la["in"])
                dst, do = self._buf(la["out"])
                n = la["bytes"] // self.esize
                dst[do:do + n] = src[so:so + n]
                continue
            self._pass(la, lds_per_block)

    @staticmethod
    def _check_tw16(p, k, L):
        """The second twiddle factor is tw16[a*16 + j] with a = t (split 0) or
        a = k (split 1): the table must cover every a the pass uses."""
        tok = p["tw16"]
        if tok is None or tok[0] != TK_TW16:
            raise ScheduleError("pass twiddle without its second-factor table")
        amax = L if p["tw_split"] == 0 else int(k.max()) + 1
        if amax * 16 > tok[1]:
            raise ScheduleError(f"pass twiddle second factor: needs {amax * 16} entries, table has {tok[1]}")

    @staticmethod
    def _check_zero_pad(p, zp):
        """A zero-padded line (plan.cpp set_zero_pad) must be a Bluestein line
        whose inputs and read outputs all lie below the stage bounds."""
        L, rad = p["L"], p["rad"]
        if not (p["load_mode"] == 1 and p["mid_mode"] == 1 and p["store_mode"] == 2 and p["cae"] == 1
                and p["cbe"] == 1 and not p["pretw"] and not any(p["ca"]) and not any(p["cb"])
                and not p.get("mirror_rows", 0)):
            raise ScheduleError("zero-padded pass that is not a chirp-loaded Bluestein line")
        if p.get("mid_fused") and len(rad) < 2:
            raise ScheduleError("zero-padded pass with a one-stage fused boundary")
        if zp & 1:
            f0 = (rad[0] + 1) // 2 * (L // rad[0])
            if rad[0] not in ZERO_PAD_RADICES or p.get("zero_from") != f0 or p["chirp_n"] > f0:
                raise ScheduleError(f"zero tail from {p.get('zero_from')} (radix {rad[0]}, {p['chirp_n']} inputs)")
        if zp & 2:
            fl = (rad[-1] + 1) // 2 * (L // rad[-1])
            if rad[-1] not in ZERO_PAD_RADICES or p["chirp_n"] > fl:
                raise ScheduleError(f"pruned last stage of radix {rad[-1]} below {p['chirp_n']} outputs")

    def _pass(self, p, lds_per_block):
        L, C, Tl = p["L"], p["C"], p["Tl"]
        if p.get("reg"):
            # register line pass: one line per thread, blocks of REG_BLOCK lines
            mode = 2 if (p["load_mode"] or p["mid_mode"] or p["store_mode"] or p["pretw"] == 2) else int(p["pretw"] == 1)
            lmax = ({0: 48, 1: 48, 2: 24} if self.esize == 8 else {0: 24, 1: 16, 2: 12})[mode]
            if L not in REG_LENGTHS or L > lmax:
                raise ScheduleError(f"register pass of unsupported length {L}")
            if p["block"] != REG_BLOCK or p["lds"] != 0 or p["grid"] * REG_BLOCK < p["nlines"]:
                raise ScheduleError("register pass geometry")
            if p["mid_mode"] or not (p["col_in"] and p["col_out"]):
                raise ScheduleError("register pass needs column layout on both sides and no mid product")
            # computed chirp: the FP64 step table d[t] covers the line
            if p["load_mode"] or p["store_mode"] == 2:
                d = p.get("chirp_d")
                if d is not None and (d[0] != TK_CHIRPD or d[1] < L):
                    raise ScheduleError(f"chirp step table {d} for a line of {L}")
        else:
            if C < 1 or C > 256 or C * Tl > p["block"]:
                raise ScheduleError(f"bad tile C={C} Tl={Tl} block={p['block']}")
            if p.get("wave_sync") and (Tl > 64 or 64 % Tl):
                # wave-synchronised sweeps need every line within one wavefront
                raise ScheduleError(f"wave_sync with {Tl} threads per line")
            sh = p["lds_sh"]
            need = L + ((L - 1) >> sh if sh else 0) + (RADER_SIDE if p["load_mode"] == LOAD_RADER else 0)
            # each direct-DFT stage stages its P roots repeated BIG_KP times
            roots = sum(BIG_KP * r for r in p["rad"] if r not in codelets(self.esize)) * self.esize
            if p["ldl"] < need or p["lds"] != ((C * p["ldl"] * self.esize + 15) & ~15) + C * LINE_TAB_BYTES + roots:
                raise ScheduleError("LDS size mismatch")
            if p["lds"] > lds_per_block:
                raise ScheduleError(f"LDS {p['lds']} exceeds {lds_per_block}")
            if p["grid"] * C < p["nlines"]:
                raise ScheduleError("grid does not cover all lines")
            if p["xcd_remap"] and p["grid"] % 8:
                raise ScheduleError("XCD remap needs grid % 8 == 0")
            zp = p.get("zero_pad", 0)
            if zp:
                self._check_zero_pad(p, zp)
            key = (L, tuple(p["rad"]), Tl, p["lds_sh"], self.esize, zp)
            if key not in self.sweeps_checked:
                cod = codelets(self.esize)
                emulate_sweep(L, p["rad"], Tl, p["lds_sh"], cod=cod)
                # the first sweep skips the zero tail, the second prunes its outputs
                if zp & 1:
                    emulate_sweep(L, p["rad"], Tl, p["lds_sh"], cod=cod, zero_first=True)
                if zp & 2:
                    emulate_sweep(L, p["rad"], Tl, p["lds_sh"], cod=cod, prune_last=True)
                self.sweeps_checked.add(key)
        if p["tw"] is None or p["tw"][0] != TK_STAGE or p["tw"][1] != L:
            raise ScheduleError("stage twiddle table does not match L")
        nl = p["nlines"]
        n0, n1, n2 = p["n"]
        line = np.arange(nl, dtype=np.int64)
        idx = [line % n0]
        r = line // n0
        idx.append(r % n1)
        r = r // n1
        idx.append(r % n2)
        idx.append(r // n2)
        a, b, ca, cb = p["a"], p["b"], p["ca"], p["cb"]
        oin = sum(idx[d] * a[d] for d in range(4))
        oout = sum(idx[d] * b[d] for d in range(4))
        if p["in_tab"] is not None:
            base = p["in_tab"][1] // 8
            oin = oin + self.offtab[base + idx[3]]
        if p["out_tab"] is not None:
            base = p["out_tab"][1] // 8
            oout = oout + self.offtab[base + idx[3]]
        cin = sum(idx[d] * ca[d] for d in range(4))
        cout = sum(idx[d] * cb[d] for d in range(4))
        cmid = sum(idx[d] * p["cm"][d] for d in range(4))
        cms = np.full_like(cmid, p["cme"])
        mr = p.get("mirror_rows", 0)
        if mr:
            # rows k1 > M1/2 read row M1 - k1 backwards (device.h line_offsets)
            mir = 2 * idx[0] > mr
            cmid = np.where(mir, cmid + (mr - 2 * idx[0]) * p["cm"][0] + (L - 1) * p["cme"], cmid)
            cms = np.where(mir, -p["cme"], p["cme"])
        rader = p["load_mode"] == LOAD_RADER
        if rader:
            if p["rtab"] is None or p["rtab"][0] != TK_RTAB or p["rtab"][1] != L + 1 or p["chirp_n"] != L + 1:
                raise ScheduleError("Rader pass without its slot table")
            if p["store_mode"] != STORE_RADER or p["mid_mode"] != 1 or p["pretw"]:
                raise ScheduleError("Rader pass modes")
            if not p.get("reg") and not (p["col_in"] or p["row_in"]) or not (p["col_out"] or p["row_out"]):
                raise ScheduleError("Rader pass without a stepped layout")
        t = np.arange(L + 1 if rader else L, dtype=np.int64)
        src, so = self._buf(p["in"])
        dst, do = self._buf(p["out"])
        addr_in = so + oin[:, None] + t[None, :] * p["ae"]
        valid_in = np.ones_like(addr_in, dtype=bool)
        chirp = bhat = None
        if p["chirp"] is not None:
            chirp = chirp_fwd(p["chirp"][1])
        if rader:
            if p["bhat"] is None or p["bhat"][1] != L:
                raise ScheduleError("Rader spectrum table")
        elif p["bhat"] is not None:
            M = p["bhat"][1]
            bhat = bhat_of(p["chirp_n"], M)
            if p["bhat"][0] == TK_BHATROWS:
                # rows 0 .. M1/2 of the fused Bluestein row pass: [k1*M2 + k2] =
                # bhat[k1 + M1*k2], M2 = L (the others are read mirrored)
                M2 = L
                M1 = M // M2
                if p.get("mirror_rows", 0) != M1:
                    raise ScheduleError("row spectrum without mirrored rows")
                k1, k2 = np.divmod(np.arange((M1 // 2 + 1) * M2), M2)
                bhat = bhat[k1 + M1 * k2]
        if p["load_mode"] == 1:
            n = cin[:, None] + t[None, :] * p["cae"]
            valid_in = n < p["chirp_n"]
        sel = addr_in[valid_in]
        if sel.size and (sel.min() < 0 or sel.max() >= src.size):
            raise ScheduleError(f"load out of bounds [{sel.min()}, {sel.max()}] of {src.size}")
        if p.get("reg") and p["load_mode"] == 1:
            # dead chirp positions read the line's first element instead
            first = addr_in[:, 0]
            if first.min() < 0 or first.max() >= src.size:
                raise ScheduleError("register pass: first element of a line out of bounds")
        v = np.zeros(addr_in.shape, dtype=np.complex128)
        v[valid_in] = src[sel]
        if np.isnan(v[valid_in]).any():
            raise ScheduleError("pass reads scratch that no earlier pass wrote")
        if p["conj_in"]:
            v = np.conj(v)
        if p["load_mode"] == 1:
            n = cin[:, None] + t[None, :] * p["cae"]
            nn = np.where(valid_in, n, 0)
            v = np.where(valid_in, v * np.conj(chirp[nn]), 0)
        if rader:
            self._rader(p, v, L, cmid, do, oout, dst)
            return
        if p["pretw"] == 1:
            M = p["tw_lo"][1]
            m = idx[p["twdim"]][:, None] * t[None, :] * p["tw_scale"]
            if m.max() >= M:
                raise ScheduleError("pass twiddle exponent out of range")
            self._check_tw16(p, idx[p["twdim"]], L)
            v = v * np.exp(-2j * np.pi * m / M)
        v = np.fft.fft(v, axis=1)
        if p["mid_mode"] == 1:
            n = cmid[:, None] + t[None, :] * cms[:, None]
            if bhat is None or n.max() >= bhat.size or n.min() < 0:
                raise ScheduleError("mid spectrum index out of range")
            v = np.conj(v * bhat[n])
            v = np.fft.fft(v, axis=1)
        if p["pretw"] == 2:
            M = p["tw_lo"][1]
            m = idx[p["twdim"]][:, None] * t[None, :] * p["tw_scale"]
            if m.max() >= M:
                raise ScheduleError("pass twiddle exponent out of range")
            self._check_tw16(p, idx[p["twdim"]], L)
            v = v * np.exp(-2j * np.pi * m / M)
        addr_out = do + oout[:, None] + t[None, :] * p["be"]
        valid_out = np.ones_like(addr_out, dtype=bool)
        if p["store_mode"] == 1:
            n = cout[:, None] + t[None, :] * p["cbe"]
            if n.max() >= bhat.size:
                raise ScheduleError("spectrum index out of range")
            v = np.conj(v * bhat[n])
        elif p["store_mode"] == 2:
            n = cout[:, None] + t[None, :] * p["cbe"]
            valid_out = n < p["chirp_n"]
            nn = np.where(valid_out, n, 0)
            v = np.conj(chirp[nn] * v)
        if p["conj_out"]:
            v = np.conj(v)
        v = v * p["scale"]
        sel = addr_out[valid_out]
        if sel.size and (sel.min() < 0 or sel.max() >= dst.size):
            raise ScheduleError(f"store out of bounds [{sel.min()}, {sel.max()}] of {dst.size}")
        if np.unique(sel).size != sel.size:
            raise ScheduleError("two lines store to one address")
        dst[sel] = v[valid_out]


def _rader_pass(p, v, L, cmid, do, oout, dst):
    P = L + 1
    rtab, bhat = rader_tables(P)
    x0 = v[:, 0]
    a = np.zeros((v.shape[0], L), dtype=np.complex128)
    a[:, rtab[1:]] = v[:, 1:]
    V = np.fft.fft(a, axis=1)
    v0 = V[:, 0]
    n = cmid[:, None] + np.arange(L)[None, :] * p["cme"]
    if n.max() >= bhat.size:
        raise ScheduleError("Rader spectrum index out of range")
    W = np.fft.fft(np.conj(V * bhat[n]), axis=1)
    k = np.arange(1, P)
    m = (L - rtab[k]) % L
    X = np.empty((v.shape[0], P), dtype=np.complex128)
    X[:, 0] = x0 + v0
    X[:, 1:] = x0[:, None] + np.conj(W[:, m])
    if p["conj_out"]:
        X = np.conj(X)
    X = X * p["scale"]
    addr_out = do + oout[:, None] + np.arange(P)[None, :] * p["be"]
    sel = addr_out.ravel()
    if sel.min() < 0 or sel.max() >= dst.size:
        raise ScheduleError(f"store out of bounds [{sel.min()}, {sel.max()}] of {dst.size}")
    if np.unique(sel).size != sel.size:
        raise ScheduleError("two lines store to one address")
    dst[sel] = X.ravel()


Emulator._rader = staticmethod(_rader_pass)


def run_plan_schedule(sched: dict, x: np.ndarray, esize: int, lds_per_block: int) -> np.ndarray:
    em = Emulator(x, x.size, sched["scratch_bytes"], esize)
    em.run(sched["launches"], lds_per_block)
    return em.bufs[TK_OUT].reshape(x.shape, order="F")


def run_batch_schedule(sched: dict, x: np.ndarray, esize: int, lds_per_block: int) -> np.ndarray:
    em = Emulator(x, x.size, sched["scratch_bytes"], esize, offtab=sched["offtab"])
    ngroups = len(sched["groups"])
    maxp = max((len(g) for g in sched["groups"]), default=0)
    for ph in range(maxp):
        for g in range(ngroups):
            if ph < len(sched["groups"][g]):
                em.run([sched["groups"][g][ph]], lds_per_block)
    return em.bufs[TK_OUT]
