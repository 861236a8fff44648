// device.h -- device-side templates of the line passes (in-LDS Stockham line
// FFTs and register line passes), shared by the translation units that
// instantiate them (kern_lds.hip per precision and workgroup size,
// kern_reg.hip per precision) and by kernels.hip (tables, dispatch).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.h"
#include "codelets.h"

namespace fftahip {

// ------------------------------------------------------------------ helpers

// 32 or 64: the precision argument of prime_codelet / big_stage
template <class Rt>
constexpr int kPrec = sizeof(Rt) == 4 ? 32 : 64;

template <class R>
__device__ __forceinline__ cx<R> from_d2(double2 w) {
  return {(R)w.x, (R)w.y};
}

// Global-memory accesses.  Buffer and table pointers reach the kernels
// through pass descriptors, so the compiler cannot prove that they point to
// global memory and would emit flat instructions, which share their wait
// counter with LDS traffic and so serialise table loads against the LDS
// sweeps.  These helpers issue global instructions explicitly; a complex
// value moves as one 2-vector (one dwordx2 / dwordx4 access per element).
#define FFTAHIP_GLOBAL __attribute__((address_space(1)))
template <class R>
__device__ __forceinline__ cx<R> gld(const cx<R>* p) {
  typedef R V __attribute__((ext_vector_type(2)));
  const V v = *(const FFTAHIP_GLOBAL V*)p;
  return {v.x, v.y};
}
__device__ __forceinline__ double2 gld(const double2* p) {
  typedef double V __attribute__((ext_vector_type(2)));
  const V v = *(const FFTAHIP_GLOBAL V*)p;
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ int64_t gld(const int64_t* p) { return *(const FFTAHIP_GLOBAL int64_t*)p; }
__device__ __forceinline__ int gld(const uint16_t* p) { return *(const FFTAHIP_GLOBAL uint16_t*)p; }
// A table entry that every lane of a wave reads at the same address, through
// the constant address space (a scalar load into SGPRs, no VGPRs held).
__device__ __forceinline__ double2 uld(const double2* p) {
  typedef double V __attribute__((ext_vector_type(2)));
  const V v = *(const __attribute__((address_space(4))) V*)p;
  return make_double2(v.x, v.y);
}
template <class R>
__device__ __forceinline__ void gst(cx<R>* p, cx<R> a) {
  typedef R V __attribute__((ext_vector_type(2)));
  V v;
  v.x = a.x;
  v.y = a.y;
  *(FFTAHIP_GLOBAL V*)p = v;
}

// Conjugate when sgn = 0x80000000, identity when sgn = 0 (sign-bit flip of
// the imaginary part, no branch).
__device__ __forceinline__ cx<float> cconj_if(cx<float> a, uint32_t sgn) {
  return {a.x, __uint_as_float(__float_as_uint(a.y) ^ sgn)};
}
__device__ __forceinline__ cx<double> cconj_if(cx<double> a, uint32_t sgn) {
  return {a.x, __longlong_as_double(__double_as_longlong(a.y) ^ ((long long)sgn << 32))};
}

// exp(-2 pi i m / n) in FP64 for 0 <= m < n: reduce m to (-n/2, n/2] first so
// that sincospi sees |x| <= 1.
__device__ __forceinline__ double2 unit_root(int64_t m, int64_t n) {
  int64_t mm = m % n;
  if (mm < 0) mm += n;
  if (2 * mm > n) mm -= n;
  double x = -2.0 * (double)mm / (double)n;
  double s, c;
  sincospi(x, &s, &c);
  return make_double2(c, s);
}

// exp(i pi r / N) for the chirp exponent r = (m mod 2N) folded into (-N, N]
// (the chirp table's own convention), m >= 0, N < 2^31.
__device__ __forceinline__ double2 chirp_root(uint64_t m, int64_t N) {
  const uint64_t twoN = 2 * (uint64_t)N;
  int64_t r = (int64_t)(m % twoN);
  if (r > N) r -= (int64_t)twoN;
  double sn, cs;
  sincospi((double)r / (double)N, &sn, &cs);
  return make_double2(cs, sn);
}

// exp(i pi r / N) for r in [0, 2N) (folded into (-N, N] as the chirp table).
__device__ __forceinline__ double2 chirp_root_r(int64_t r, int64_t N) {
  if (r > N) r -= 2 * N;
  double sn, cs;
  sincospi((double)r / (double)N, &sn, &cs);
  return make_double2(cs, sn);
}

// ------------------------------------------------------------------ stages

// Padded position of element i inside an LDS line (one spare slot every
// 2^sh elements breaks the power-of-two strides of the radix stages across
// the LDS banks).  Takes the effective shift she = sh ? sh : 31, so that no
// padding needs no select (i < 2^31).
__device__ __forceinline__ int pad_shift(int sh) { return sh ? sh : 31; }
__device__ __forceinline__ int pad_idx(int i, int she) { return i + (i >> she); }

// Barrier between the LDS phases of a radix sweep.  When every line of the
// tile lies within one wavefront (threads per line divide 64: the planner's
// wave_sync), only the lanes of one wave exchange data, and one wave's LDS
// instructions execute in program order, so ordering the wave's own LDS
// accesses (wavefront-scope fences around a wave barrier) replaces the
// workgroup barrier: the waves of a tile then run their sweeps independently.
__device__ __forceinline__ void line_sync(bool wave) {
  if (wave) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One radix-RAD Stockham stage over a line in LDS.  Stage twiddles come from
// a per-stage table: stage s (Ns = product of earlier radices) keeps
// w_{Ns*RAD}^{t*k} for k < Ns, t = 1..RAD-1 at tw[(Ns-1) + k*(RAD-1) + t-1],
// so one butterfly reads its RAD-1 factors contiguously.  LDS positions are
// padded (i + (i >> sh)); when a stride is a multiple of 2^sh the padded
// positions of a butterfly are an arithmetic sequence and are stepped
// instead of recomputed.
// Zero-padded convolution lines (ZM, compile time): ZM == ZM_IN is the first
// stage of a sweep over a line that is zero at positions >= KH * L / RAD,
// KH = (RAD + 1) / 2: its inputs t >= KH are taken as zero without an LDS
// read (and the butterfly folds them away); ZM == ZM_OUT is the last stage
// of a sweep whose outputs are read at positions < KH * L / RAD only: the
// outputs t >= KH are not computed or written.
enum ZeroMode { ZM_NONE = 0, ZM_IN = 1, ZM_OUT = 2 };

template <class Rt, int RAD, int EPT, bool FUSE, int ZM = ZM_NONE, class LdF, class StF>
__device__ __forceinline__ void stage(cx<Rt>* line, int jl, bool active, int Tl, int L, int Ns, FastDiv nsd,
                                      const cx<Rt>* __restrict__ tw, int sh, bool gload, bool gstore, const LdF& ld,
                                      const StF& st, const cx<Rt>* __restrict__ pm, int pms, bool ws) {
  constexpr int KH = (RAD + 1) / 2;
  // butterflies per thread (a prime codelet above EPT points: one)
  constexpr int NB = EPT / RAD > 0 ? EPT / RAD : 1;
  const int Lr = L / RAD;
  const int she = pad_shift(sh);
  const int pmask = sh ? (1 << sh) - 1 : 0;
  const bool rstep = (Lr & pmask) == 0;
  // the first stage (Ns = 1) writes RAD consecutive positions j*RAD + t,
  // which stay inside one padding block when RAD divides the block size
  const bool wstep = (Ns & pmask) == 0 || (Ns == 1 && ((pmask + 1) % RAD) == 0);
  const int rstride = Lr + (sh ? (Lr >> sh) : 0);
  const int wstride = Ns + (sh ? (Ns >> sh) : 0);
  cx<Rt> v[NB][RAD];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = jl + b * Tl;
    if (active && j < Lr && pm) {
      // a spectrum product applied on the way in (first stage of a sweep:
      // no twiddles): v = conj(v * pm[i * pms]).  FP32 requests its factors
      // before the LDS reads; FP64 loads each as it consumes it (register
      // budget)
      constexpr bool kPre = sizeof(Rt) == 4;
      cx<Rt> f[kPre ? RAD : 1];
      if (kPre) {
#pragma unroll
        for (int t = 0; t < RAD; ++t) f[kPre ? t : 0] = gld(pm + (j + t * Lr) * pms);
      }
      if (rstep) {
        const int p0 = pad_idx(j, she);
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = line[p0 + t * rstride];
      } else {
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = line[pad_idx(j + t * Lr, she)];
      }
#pragma unroll
      for (int t = 0; t < RAD; ++t)
        v[b][t] = cconj(cmul(v[b][t], kPre ? f[kPre ? t : 0] : gld(pm + (j + t * Lr) * pms)));
      dft_any<Rt, RAD>(v[b]);
    } else if (active && j < Lr) {
      // FP32 pre-twiddle and full passes: the stage twiddles are requested
      // first (global loads, L2 latency) and consumed in the same branch
      // region, after the LDS reads.  FP64 loads them as it consumes them (a
      // full set in flight would not fit the register budget), and so does
      // the FP32 plain variant, which is built for 7 resident waves per SIMD
      // (<= 72 VGPRs) and hides the latency with waves instead.
      constexpr bool kPre = sizeof(Rt) == 4 && !FUSE;
      const cx<Rt>* __restrict__ w = tw;
      cx<Rt> wv[RAD > 1 ? RAD - 1 : 1];
      if (Ns > 1) {
        const uint32_t k = (uint32_t)j - fdiv((uint32_t)j, nsd) * (uint32_t)Ns;
        w = tw + (Ns - 1) + k * (RAD - 1);
        if (kPre) {
#pragma unroll
          for (int t = 0; t < RAD - 1; ++t) wv[t] = gld(w + t);
        }
      }
      if (FUSE && gload) {
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = ld(j + t * Lr);
      } else if (ZM == ZM_IN) {
        const int p0 = pad_idx(j, she);
#pragma unroll
        for (int t = 0; t < RAD; ++t)
          v[b][t] = t < KH ? (rstep ? line[p0 + t * rstride] : line[pad_idx(j + t * Lr, she)]) : cx<Rt>{(Rt)0, (Rt)0};
      } else if (rstep) {
        const int p0 = pad_idx(j, she);
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = line[p0 + t * rstride];
      } else {
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = line[pad_idx(j + t * Lr, she)];
      }
      if (Ns > 1) {
#pragma unroll
        for (int t = 1; t < RAD; ++t) v[b][t] = cmul(v[b][t], kPre ? wv[t - 1] : gld(w + t - 1));
      }
      dft_any<Rt, RAD>(v[b]);
    }
  }
  line_sync(ws);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = jl + b * Tl;
    if (active && j < Lr) {
      const uint32_t q = fdiv((uint32_t)j, nsd);
      const uint32_t k = (uint32_t)j - q * (uint32_t)Ns;
      const int base = (int)(q * (uint32_t)Ns * RAD + k);
      if (FUSE && gstore) {
#pragma unroll
        for (int t = 0; t < RAD; ++t) st(base + t * Ns, v[b][t]);
      } else if (wstep) {
        const int p0 = pad_idx(base, she);
#pragma unroll
        for (int t = 0; t < RAD; ++t)
          if (ZM != ZM_OUT || t < KH) line[p0 + t * wstride] = v[b][t];
      } else {
#pragma unroll
        for (int t = 0; t < RAD; ++t)
          if (ZM != ZM_OUT || t < KH) line[pad_idx(base + t * Ns, she)] = v[b][t];
      }
    }
  }
  line_sync(ws);
}

// One radix-P stage for a prime P >= 17 over a line in LDS: the direct DFT
// of P points (FFTA runs primes below BLUESTEIN_CUTOFF as O(P^2) DFT leaves,
// src/algos.jl:135-160), in Stockham order like stage().  With
// s_j = x_j + x_{P-j} and d_j = x_j - x_{P-j} (j = 1..H, H = (P-1)/2):
//   X_0 = x_0 + sum_j s_j,
//   X_k = A_k - i B_k,  X_{P-k} = A_k + i B_k,
//   A_k = x_0 + sum_j cos(2 pi jk/P) s_j,  B_k = sum_j sin(2 pi jk/P) d_j.
// A work item is (butterfly j, group g of kBigKP consecutive k); butterflies
// run fastest across threads, and a thread holds at most kBigHold work
// items across the barrier (the planner sizes Tl for that).  `roots` holds
// w_P^(m mod P) = exp(-2 pi i m/P) for m < kBigKP * P (the P roots repeated),
// in the pass precision, staged in LDS (the tile copies them there before its
// first barrier).  For the work item's k = k0 + u, the exponent i*k mod P is
// (i*k0 mod P) + u*i: one reduction per step for all kBigKP roots, and the
// repeated table absorbs the rest.  Positions j + i*L/P whose padding is
// linear in i (L/P a multiple of the padding block) are stepped.  Stage
// twiddles (Ns > 1) are applied to the inputs as they are read, from the same
// per-stage table as stage().
constexpr int kBigKP = 4;
constexpr int kBigHold = 2;

template <class Rt>
__device__ __forceinline__ void stage_big(cx<Rt>* line, int jl, bool active, int Tl, int L, int P, int Ns, FastDiv nsd,
                                          const cx<Rt>* __restrict__ tw, const cx<Rt>* __restrict__ roots, int sh,
                                          bool ws) {
  const int Lr = L / P;
  const int H = (P - 1) >> 1;
  const int G = (H + kBigKP - 1) / kBigKP;
  const int nwork = Lr * G;
  const int she = pad_shift(sh);
  const int pmask = sh ? (1 << sh) - 1 : 0;
  const bool rstep = (Lr & pmask) == 0;
  const int rstride = Lr + (sh ? (Lr >> sh) : 0);
  cx<Rt> out[kBigHold][2 * kBigKP + 1];
  int obase[kBigHold], og[kBigHold];
#pragma unroll
  for (int h = 0; h < kBigHold; ++h) {
    const int w = jl + h * Tl;
    og[h] = -1;
    obase[h] = 0;
    if (!(active && w < nwork)) continue;
    const int g = w / Lr;
    const int j = w - g * Lr;
    const uint32_t q = fdiv((uint32_t)j, nsd);
    const int k = j - (int)q * Ns;
    const cx<Rt>* __restrict__ wt = tw + (Ns - 1) + k * (P - 1);
    const int pj = pad_idx(j, she);
    const cx<Rt> x0 = line[pj];
    cx<Rt> A[kBigKP], B[kBigKP];
#pragma unroll
    for (int u = 0; u < kBigKP; ++u) {
      A[u] = x0;
      B[u] = {(Rt)0, (Rt)0};
    }
    cx<Rt> sum = x0;
    const int k0 = g * kBigKP + 1;
    int mb = 0;  // i * k0 mod P
    // two steps in flight for FP32; FP64 has no registers for a second
    constexpr int kUnroll = sizeof(Rt) == 4 ? 2 : 1;
#pragma unroll kUnroll
    for (int i = 1; i <= H; ++i) {
      cx<Rt> a, b;
      if (rstep) {
        a = line[pj + i * rstride];
        b = line[pj + (P - i) * rstride];
      } else {
        a = line[pad_idx(j + i * Lr, she)];
        b = line[pad_idx(j + (P - i) * Lr, she)];
      }
      if (Ns > 1) {
        a = cmul(a, gld(wt + (i - 1)));
        b = cmul(b, gld(wt + (P - i - 1)));
      }
      const cx<Rt> s = cadd(a, b), d = csub(a, b);
      sum = cadd(sum, s);
      mb += k0;  // < 2P: the unsigned minimum subtracts P when mb >= P
      mb = (int)min((unsigned)mb, (unsigned)(mb - P));
      const cx<Rt>* __restrict__ rr = roots + mb;
#pragma unroll
      for (int u = 0; u < kBigKP; ++u) {
        const cx<Rt> r = rr[u * i];  // (cos, -sin) of 2 pi (i (k0 + u)) / P
        A[u].x += r.x * s.x;
        A[u].y += r.x * s.y;
        B[u].x -= r.y * d.x;
        B[u].y -= r.y * d.y;
      }
    }
#pragma unroll
    for (int u = 0; u < kBigKP; ++u) {
      out[h][2 * u] = {A[u].x + B[u].y, A[u].y - B[u].x};
      out[h][2 * u + 1] = {A[u].x - B[u].y, A[u].y + B[u].x};
    }
    out[h][2 * kBigKP] = sum;
    og[h] = g;
    obase[h] = (int)(q * (uint32_t)Ns * (uint32_t)P) + k;
  }
  line_sync(ws);
#pragma unroll
  for (int h = 0; h < kBigHold; ++h) {
    if (og[h] < 0) continue;
    const int base = obase[h];
#pragma unroll
    for (int u = 0; u < kBigKP; ++u) {
      const int kk = og[h] * kBigKP + 1 + u;
      if (kk > H) break;
      line[pad_idx(base + kk * Ns, she)] = out[h][2 * u];
      line[pad_idx(base + (P - kk) * Ns, she)] = out[h][2 * u + 1];
    }
    if (og[h] == 0) line[pad_idx(base, she)] = out[h][2 * kBigKP];
  }
  line_sync(ws);
}

// The boundary between the two sweeps of a convolution line, in one stage:
// the last radix-RAD stage of the first sweep (Ns = L / RAD: butterfly j
// reads positions j + t*L/RAD with its stage twiddles and produces the
// outputs at the same positions), the spectrum product v = conj(v * pm[i *
// pms]) at those positions, and the first stage of the second sweep (Ns = 1,
// no twiddles, the same radix over the same positions), which writes
// positions j*RAD + t.  One LDS round trip and two barriers fewer than
// running the two stages apart.  With rader_dc the first sweep's DC term
// (position 0) is kept in the slot rader_dc before the product.
template <class Rt, int RAD, int EPT>
__device__ __forceinline__ void stage_mid(cx<Rt>* line, int jl, bool active, int Tl, int L,
                                          const cx<Rt>* __restrict__ tw, int sh, const cx<Rt>* __restrict__ pm, int pms,
                                          int rader_dc, bool ws) {
  constexpr int NB = EPT / RAD;
  const int Lr = L / RAD;
  const int Ns = Lr;  // the first sweep's last stage
  const int she = pad_shift(sh);
  const int pmask = sh ? (1 << sh) - 1 : 0;
  const bool rstep = (Lr & pmask) == 0;
  const int rstride = Lr + (sh ? (Lr >> sh) : 0);
  cx<Rt> v[NB][RAD];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = jl + b * Tl;
    if (active && j < Lr) {
      // FP32 requests the spectrum factors first; FP64 loads each as it
      // consumes it (register budget).  The stage twiddles of butterfly j
      // (k = j, q = 0) are read as they are consumed.
      constexpr bool kPre = sizeof(Rt) == 4;
      cx<Rt> f[kPre ? RAD : 1];
      if (kPre) {
#pragma unroll
        for (int t = 0; t < RAD; ++t) f[kPre ? t : 0] = gld(pm + (j + t * Lr) * pms);
      }
      if (rstep) {
        const int p0 = pad_idx(j, she);
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = line[p0 + t * rstride];
      } else {
#pragma unroll
        for (int t = 0; t < RAD; ++t) v[b][t] = line[pad_idx(j + t * Lr, she)];
      }
      if (Ns > 1) {
        const cx<Rt>* __restrict__ w = tw + (Ns - 1) + j * (RAD - 1);
#pragma unroll
        for (int t = 1; t < RAD; ++t) v[b][t] = cmul(v[b][t], gld(w + t - 1));
      }
      dft_any<Rt, RAD>(v[b]);
      if (rader_dc >= 0 && j == 0) line[rader_dc] = v[b][0];
#pragma unroll
      for (int t = 0; t < RAD; ++t)
        v[b][t] = cconj(cmul(v[b][t], kPre ? f[kPre ? t : 0] : gld(pm + (j + t * Lr) * pms)));
      dft_any<Rt, RAD>(v[b]);
    }
  }
  line_sync(ws);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = jl + b * Tl;
    if (active && j < Lr) {
      const int base = j * RAD;
      if (sh == 0 || ((1 << sh) % RAD) == 0) {
        // the RAD positions lie within one padding block (or no padding)
        const int p0 = pad_idx(base, she);
#pragma unroll
        for (int t = 0; t < RAD; ++t) line[p0 + t] = v[b][t];
      } else {
#pragma unroll
        for (int t = 0; t < RAD; ++t) line[pad_idx(base + t, she)] = v[b][t];
      }
    }
  }
  line_sync(ws);
}

// All radix stages of a line.  With fuse_first the first stage reads its
// inputs through ld(element) from global memory instead of LDS; with
// fuse_last the last stage writes through st(element, value).  Stages
// [s_begin, s_end) run (default: all).  ZP (compile time) carries the
// zero-padded first / last stage variants, taken at run time by zero_first
// (stage 0) and prune_last (stage nst - 1).
template <class Rt, int EPT, bool FUSE, bool BIG, bool ZP = false, class LdF, class StF>
__device__ __forceinline__ void sweep(cx<Rt>* line, int jl, bool active, const LinePass& p,
                                      const cx<Rt>* __restrict__ tw, const cx<Rt>* roots, bool fuse_first,
                                      bool fuse_last, const LdF& ld, const StF& st,
                                      const cx<Rt>* __restrict__ pm = nullptr, int pms = 0, int s_begin = 0,
                                      int s_end = -1, bool zero_first = false, bool prune_last = false) {
  int Ns = 1;
  const int sh = p.lds_sh;
  const bool ws = p.wave_sync != 0;
  const int nst = s_end < 0 ? p.nst : s_end;
  // roots of the radix-P stages with P >= 17, one block of kBigKP * P per
  // stage (LDS)
  int roots_off = 0;
  for (int s = 0; s < s_begin; ++s) {
    if (BIG && big_stage(p.rad[s], kPrec<Rt>)) roots_off += kBigKP * p.rad[s];
    Ns *= p.rad[s];
  }
  for (int s = s_begin; s < nst; ++s) {
    const int r = p.rad[s];
    const FastDiv nsd = p.nsdiv[s];
    const bool gl = fuse_first && s == 0;
    const bool gs = fuse_last && s == p.nst - 1;
    if (BIG && big_stage(r, kPrec<Rt>)) {
      // the planner keeps fused global access and the in-stage spectrum
      // product off stages of this kind
      stage_big<Rt>(line, jl, active, p.Tl, p.L, r, Ns, nsd, tw, roots + roots_off, sh, ws);
      roots_off += kBigKP * r;
      Ns *= r;
      continue;
    }
    switch (r) {
#define FFTAHIP_STAGE(R) \
  case R: \
    if constexpr (ZP && zero_pad_radix(R)) { \
      if (zero_first && s == 0) { \
        stage<Rt, R, EPT, FUSE, ZM_IN>(line, jl, active, p.Tl, p.L, Ns, nsd, tw, sh, gl, gs, ld, st, nullptr, 0, ws); \
        break; \
      } \
      if (prune_last && s == p.nst - 1) { \
        stage<Rt, R, EPT, FUSE, ZM_OUT>(line, jl, active, p.Tl, p.L, Ns, nsd, tw, sh, gl, gs, ld, st, s == s_begin ? pm : nullptr, pms, ws); \
        break; \
      } \
    } \
    stage<Rt, R, EPT, FUSE>(line, jl, active, p.Tl, p.L, Ns, nsd, tw, sh, gl, gs, ld, st, s == s_begin ? pm : nullptr, pms, ws); \
    break;
      FFTAHIP_STAGE(2)
      FFTAHIP_STAGE(3)
      FFTAHIP_STAGE(4)
      FFTAHIP_STAGE(5)
      FFTAHIP_STAGE(7)
      FFTAHIP_STAGE(8)
      FFTAHIP_STAGE(9)
      FFTAHIP_STAGE(11)
      FFTAHIP_STAGE(13)
      FFTAHIP_STAGE(16)
#undef FFTAHIP_STAGE
      // prime codelets: direct-DFT variants only (their registers)
#define FFTAHIP_PSTAGE(R) \
  case R: \
    if constexpr (BIG && kPrec<Rt> == 32) \
      stage<Rt, R, EPT, FUSE>(line, jl, active, p.Tl, p.L, Ns, nsd, tw, sh, gl, gs, ld, st, s == s_begin ? pm : nullptr, pms, ws); \
    break;
      FFTAHIP_PSTAGE(17)
      FFTAHIP_PSTAGE(19)
      FFTAHIP_PSTAGE(23)
#undef FFTAHIP_PSTAGE
      default: break;
    }
    Ns *= r;
  }
}

// ------------------------------------------------------------------ line pass

// Positions t < live_bound(rem, step, L) of a line satisfy t*step < rem
// (rem = chirp_n minus the line's first chirp index): the live prefix of a
// Bluestein line, as a 32-bit bound computed once per line.
__device__ __forceinline__ int live_bound(int64_t rem, int64_t step, int L) {
  if (rem <= 0) return 0;
  if (step <= 0) return L;
  const int64_t b = (rem + step - 1) / step;
  return b < (int64_t)L ? (int)b : L;
}

// Per-line offsets of a tile (input, output, pass-twiddle index, chirp in /
// out, spectrum), kept in LDS right after the tile's lines and sized by its
// line count (tile_lds_bytes), so that a one-line tile does not reserve
// tables for kMaxLines.
struct LineShared {
  int64_t *in, *out, *twk, *cin, *cout, *cmid;
};

// Spectrum position of a line's first element and its step along the line,
// from the tile's table entry (complemented for a mirrored row).
__device__ __forceinline__ void mid_ref(int64_t v, int64_t cme, int64_t& base, int64_t& step) {
  const bool mir = v < 0;
  base = mir ? ~v : v;
  step = mir ? -cme : cme;
}

// Kernel variants (compile-time), so that the common passes carry no
// register cost for operations they do not use:
//   MODE_PLAIN  load, radix sweeps, store (optionally conjugating / scaling);
//   MODE_PRETW  as PLAIN with the multi-pass twiddle applied on load;
//   MODE_FULL   every fused operation, selected at run time (Bluestein).
enum PassMode { MODE_PLAIN = 0, MODE_PRETW = 1, MODE_FULL = 2 };

// Global loads are issued in batches per thread (kept in flight together).
template <int MODE>
struct Batch {
  static constexpr int U = MODE == MODE_PLAIN ? 8 : 4;
};

__device__ __forceinline__ double2 dmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Offsets of line `line` of a pass: input/output element offsets, the index
// the pass twiddle is taken over, and the chirp / spectrum positions.  The
// line index is split over n[0], n[1], n[2] and a last (table) dimension.
struct LineOffsets {
  int64_t in, out, twk, cin, cout, cmid;
  int64_t cms;  // direction of the spectrum row (+1, or -1 for a mirrored row)
};

template <int MODE>
__device__ __forceinline__ LineOffsets line_offsets(const LinePass& p, uint32_t line) {
  int64_t idx[4];
  const uint32_t q = fdiv(line, p.divn[0]);
  idx[0] = line - q * p.divn[0].d;
  const uint32_t q1 = fdiv(q, p.divn[1]);
  idx[1] = q - q1 * p.divn[1].d;
  const uint32_t q2 = fdiv(q1, p.divn[2]);
  idx[2] = q1 - q2 * p.divn[2].d;
  idx[3] = q2;
  LineOffsets o;
  o.in = idx[0] * p.a[0] + idx[1] * p.a[1] + idx[2] * p.a[2] + idx[3] * p.a[3];
  o.out = idx[0] * p.b[0] + idx[1] * p.b[1] + idx[2] * p.b[2] + idx[3] * p.b[3];
  if (p.in_tab) o.in += gld(p.in_tab + idx[3]);
  if (p.out_tab) o.out += gld(p.out_tab + idx[3]);
  o.twk = o.cin = o.cout = o.cmid = 0;
  o.cms = 1;
  if (MODE != MODE_PLAIN) {
    const int d = p.twdim & 3;
    o.twk = d == 0 ? idx[0] : (d == 1 ? idx[1] : (d == 2 ? idx[2] : idx[3]));
  }
  if (MODE == MODE_FULL) {
    o.cin = idx[0] * p.ca[0] + idx[1] * p.ca[1] + idx[2] * p.ca[2] + idx[3] * p.ca[3];
    o.cout = idx[0] * p.cb[0] + idx[1] * p.cb[1] + idx[2] * p.cb[2] + idx[3] * p.cb[3];
    o.cmid = idx[0] * p.cm[0] + idx[1] * p.cm[1] + idx[2] * p.cm[2] + idx[3] * p.cm[3];
    // mirrored spectrum rows (the three-pass row pass): row k1 > M1/2 reads
    // row M1 - k1 backwards, B[k1 + M1 k2] = B[M - k1 - M1 k2] = row (M1 - k1)
    // at M2 - 1 - k2 (the Bluestein spectrum is symmetric, B[M - k] = B[k])
    if (p.mirror_rows && 2 * idx[0] > p.mirror_rows) {
      o.cmid += (p.mirror_rows - 2 * idx[0]) * p.cm[0] + (int64_t)(p.L - 1) * p.cme;
      o.cms = -1;
    }
  }
  return o;
}

// Pass twiddle w_M^{k*t*s} (s = tw_scale; k*t*s < M < 2^31) from its two
// factors (see LinePass): requests, then the product.
template <class Rt>
struct PassTw {
  double2 lo, hi;
  cx<Rt> b;
};

template <class Rt>
__device__ __forceinline__ PassTw<Rt> pass_twiddle_load(const LinePass& p, int64_t twk, int t) {
  const uint32_t k = (uint32_t)twk, tt = (uint32_t)t;
  uint32_t a, j, m;
  if (p.tw_split == 0) {
    j = k & 15u;
    a = tt;
    m = (k - j) * tt;
  } else {
    j = tt & 15u;
    a = k;
    m = k * (tt - j);
  }
  m *= (uint32_t)p.tw_scale;
  const uint32_t mask = (1u << p.tw_sh) - 1u;
  PassTw<Rt> w;
  w.lo = gld(p.tw_lo + (m & mask));
  w.hi = gld(p.tw_hi + (m >> p.tw_sh));
  w.b = gld(reinterpret_cast<const cx<Rt>*>(p.tw16) + (a * 16u + j));
  return w;
}

template <class Rt>
__device__ __forceinline__ cx<Rt> pass_twiddle_value(const PassTw<Rt>& w) {
  return cmul(from_d2<Rt>(dmul2(w.lo, w.hi)), w.b);
}

template <class Rt>
__device__ __forceinline__ cx<Rt> pass_twiddle(const LinePass& p, int64_t twk, int t) {
  return pass_twiddle_value<Rt>(pass_twiddle_load<Rt>(p, twk, t));
}

// w_M^{k * s} in FP64 (two-level tables, FP64 product): the ratio of
// consecutive pass twiddles along a line.
__device__ __forceinline__ double2 step_twiddle(const LinePass& p, int64_t twk) {
  const uint32_t m = (uint32_t)twk * (uint32_t)p.tw_scale;
  const uint32_t mask = (1u << p.tw_sh) - 1u;
  return dmul2(gld(p.tw_lo + (m & mask)), gld(p.tw_hi + (m >> p.tw_sh)));
}

template <class Rt, int EPT, int BLOCK, int MODE, bool BIG>
__device__ __forceinline__ void run_line_pass(const LinePass& p, const BufPtrs& bp, uint32_t tile, cx<Rt>* lds,
                                              const LineShared& sh) {
  constexpr int U = Batch<MODE>::U;
  const int tid = threadIdx.x;
  const int C = p.C;
  const int L = p.L;
  const int ldl = p.ldl;
  const int psh = p.lds_sh;
  const int pshe = pad_shift(psh);
  const int64_t line0 = (int64_t)tile * C;
  const int64_t rem = p.nlines - line0;
  const int nvalid = rem < C ? (int)rem : C;

  // direct-DFT roots (BIG variants): the stage table's entries after its
  // L - 1 stage twiddles, P per direct-DFT stage, staged after the line
  // tables as kBigKP * P entries per stage (the P roots repeated)
  cx<Rt>* roots = nullptr;
  if (BIG) {
    roots = reinterpret_cast<cx<Rt>*>(reinterpret_cast<unsigned char*>(lds) + tile_lds_bytes(C, ldl, sizeof(cx<Rt>)));
    const cx<Rt>* __restrict__ src = reinterpret_cast<const cx<Rt>*>(p.tw) + (L - 1);
    int lo = 0;
    for (int s = 0; s < p.nst; ++s) {
      const int P = p.rad[s];
      if (!big_stage(P, kPrec<Rt>)) continue;
      for (int i = tid; i < kBigKP * P; i += BLOCK) roots[lo + i] = gld(src + i % P);
      lo += kBigKP * P;
      src += P;
    }
  }
  for (int c = tid; c < nvalid; c += BLOCK) {
    const LineOffsets o = line_offsets<MODE>(p, (uint32_t)(line0 + c));
    sh.in[c] = o.in;
    sh.out[c] = o.out;
    if (MODE != MODE_PLAIN) sh.twk[c] = o.twk;
    if (MODE == MODE_FULL) {
      sh.cin[c] = o.cin;
      sh.cout[c] = o.cout;
      // a mirrored row keeps its offset complemented (the offsets are >= 0)
      sh.cmid[c] = o.cms < 0 ? ~o.cmid : o.cmid;
    }
  }
  __syncthreads();

  const void* inb = p.in_buf == BUF_IN ? bp.in : (p.in_buf == BUF_OUT ? (const void*)bp.out : (const void*)bp.scratch);
  void* outb = p.out_buf == BUF_OUT ? bp.out : bp.scratch;
  const cx<Rt>* __restrict__ in = reinterpret_cast<const cx<Rt>*>(inb) + p.in_off;
  cx<Rt>* __restrict__ out = reinterpret_cast<cx<Rt>*>(outb) + p.out_off;
  const cx<Rt>* __restrict__ tw = reinterpret_cast<const cx<Rt>*>(p.tw);
  const cx<Rt>* __restrict__ chirp = reinterpret_cast<const cx<Rt>*>(p.chirp);
  const cx<Rt>* __restrict__ bhat = reinterpret_cast<const cx<Rt>*>(p.bhat);
  const int total = C * L;
  const uint32_t sgn_in = p.conj_in ? 0x80000000u : 0u;
  const uint32_t sgn_out = p.conj_out ? 0x80000000u : 0u;
  const bool chirp_in = MODE == MODE_FULL && p.load_mode == LOAD_CHIRP;
  const bool pretw_in = MODE != MODE_PLAIN && p.pretw == 1;
  const bool pretw_out = MODE == MODE_FULL && p.pretw == 2;
  const int store_mode = MODE == MODE_FULL ? p.store_mode : STORE_PLAIN;
  // Rader line: P = chirp_n elements on the global side, L = P - 1 in LDS,
  // x0 and the first sweep's DC term in the two side slots after the line
  const bool rader = MODE == MODE_FULL && p.load_mode == LOAD_RADER;

  const bool fuse_in = MODE == MODE_PLAIN && p.fuse_in != 0;
  const bool fuse_out = MODE == MODE_PLAIN && p.fuse_out != 0;
  const bool has_mid = MODE == MODE_FULL && p.mid_mode == MID_BHAT;
  // zero-padded Bluestein lines (5-smooth, so never the direct-DFT variant):
  // the line is zero from p.zero_from on, and the output is read below the
  // last stage's pruning bound
  constexpr bool kZero = MODE == MODE_FULL && !BIG;
  const bool zero_in = kZero && (p.zero_pad & 1) != 0;
  const bool prune_out = kZero && (p.zero_pad & 2) != 0;

  // FP32 passes take the pass twiddles along a thread's run of positions as
  // powers of one FP64 ratio (see reg_line); FP64 passes look each one up.
  constexpr bool kGeo = sizeof(Rt) == 4;
  // Stepped phases: thread tid < C * (BLOCK / C) stays on one line and steps
  // along it by tstep = BLOCK / C positions, so that offsets, LDS slots and
  // pass twiddles advance incrementally instead of being re-derived per
  // element; the remaining BLOCK % C threads sit out the phase.  Column layout
  // (lines adjacent in memory): c = tid % C, neighbouring lanes on
  // neighbouring lines.  Row layout (row_in / row_mid / row_out): c =
  // tid / tstep, neighbouring lanes on neighbouring positions of one line.
  const int tstep = BLOCK / C;
  const bool col_step = tstep >= 1;
  const int my_t0 = (int)fdiv((uint32_t)tid, p.divC);
  const int my_c = my_t0 < tstep ? (int)(tid - (uint32_t)my_t0 * (uint32_t)C) : C;
  const int row_c = (int)fdiv((uint32_t)tid, p.divR);  // >= C past the last line's threads
  const int row_t0 = tid - row_c * tstep;

  // ---- load (global -> LDS), coalesced across lines or along a line.  All
  // global loads of a batch (element, chirp factor, pass-twiddle halves) are
  // issued before any of them is consumed.
  if (rader) {
    // element t >= 1 of the prime-length line to slot rtab[t], element 0 to
    // the x0 side slot (the planner gives Rader passes a stepped layout)
    const int in_c = p.col_in ? my_c : row_c;
    const int in_t0 = p.col_in ? my_t0 : row_t0;
    if (in_c < nvalid) {
      const cx<Rt>* __restrict__ src = in + sh.in[in_c];
      cx<Rt>* line = lds + in_c * ldl;
      const int P = L + 1;
      for (int t0 = in_t0; t0 < P; t0 += tstep * U) {
        cx<Rt> v[U];
        int slot[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          slot[u] = -1;
          if (t < P) {
            v[u] = gld(src + (int64_t)t * p.ae);
            slot[u] = t ? pad_idx(gld(p.rtab + t), pshe) : ldl - kRaderSide;
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (slot[u] >= 0) line[slot[u]] = cconj_if(v[u], sgn_in);
      }
    }
  } else if (!fuse_in && (p.col_in || p.row_in) && col_step) {
    const int in_c = p.col_in ? my_c : row_c;
    const int in_t0 = p.col_in ? my_t0 : row_t0;
    if (in_c < nvalid) {
      const int c = in_c;
      // the element and chirp pointers advance by a fixed stride per step
      // (no per-element 64-bit products); positions t < tlive have a chirp
      // factor (n = cin0 + t*cae < chirp_n), the others load zero
      const int64_t sstep = (int64_t)tstep * p.ae;
      const cx<Rt>* __restrict__ sp = in + sh.in[c] + (int64_t)in_t0 * p.ae;
      const cx<Rt>* __restrict__ cp = chirp;
      int64_t cstep = 0;
      int tlive = L;
      if (chirp_in) {
        const int64_t cin0 = sh.cin[c];
        cp = chirp + cin0 + (int64_t)in_t0 * p.cae;
        cstep = (int64_t)tstep * p.cae;
        tlive = live_bound(p.chirp_n - cin0, p.cae, L);
      }
      double2 w = make_double2(1.0, 0.0), wr = w;
      // exponents k*t*s with t < L only (the tables end at M)
      if (kGeo && pretw_in && in_t0 < L) {
        w = step_twiddle(p, sh.twk[c] * (int64_t)in_t0);
        if (tstep < L) wr = step_twiddle(p, sh.twk[c] * (int64_t)tstep);
      }
      // a zero-padded line is written up to zero_from (the first stage reads
      // no position past it)
      const int lend = zero_in ? p.zero_from : L;
      for (int t0 = in_t0; t0 < lend; t0 += tstep * U) {
        cx<Rt> v[U], f[U];
        PassTw<Rt> pw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          v[u] = {(Rt)0, (Rt)0};
          f[u] = {(Rt)0, (Rt)0};
          if (t < tlive) {
            if (chirp_in) f[u] = gld(cp);
            v[u] = gld(sp);
          }
          if (t < L && !kGeo && pretw_in) pw[u] = pass_twiddle_load<Rt>(p, sh.twk[c], t);
          sp += sstep;
          cp += cstep;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          if (t >= lend) continue;
          cx<Rt> x = cconj_if(v[u], sgn_in);
          if (chirp_in) x = cmul(x, cconj(f[u]));
          if (pretw_in) {
            if (kGeo) {
              x = cmul(x, from_d2<Rt>(w));
              w = dmul2(w, wr);
            } else {
              x = cmul(x, pass_twiddle_value<Rt>(pw[u]));
            }
          }
          lds[c * ldl + pad_idx(t, pshe)] = x;
        }
      }
    }
  } else
  for (int e0 = 0; e0 < (fuse_in ? 0 : total); e0 += BLOCK * U) {
    cx<Rt> v[U], f[U];
    PassTw<Rt> w[U];
    int slot[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * BLOCK + tid;
      slot[u] = -1;
      v[u] = {(Rt)0, (Rt)0};
      f[u] = {(Rt)0, (Rt)0};
      if (e < total) {
        int c, t;
        if (p.col_in) {
          t = (int)fdiv((uint32_t)e, p.divC);
          c = e - t * C;
        } else if (C == 1) {
          c = 0;
          t = e;
        } else {
          c = (int)fdiv((uint32_t)e, p.divL);
          t = e - c * L;
        }
        if (c < nvalid) {
          slot[u] = c * ldl + pad_idx(t, pshe);
          bool live = true;
          if (chirp_in) {
            const int64_t n = sh.cin[c] + (int64_t)t * p.cae;
            live = n < p.chirp_n;
            if (live) f[u] = gld(chirp + n);
          }
          if (live) v[u] = gld(in + sh.in[c] + (int64_t)t * p.ae);
          if (pretw_in) w[u] = pass_twiddle_load<Rt>(p, sh.twk[c], t);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (slot[u] < 0) continue;
      cx<Rt> x = cconj_if(v[u], sgn_in);
      if (chirp_in) x = cmul(x, cconj(f[u]));
      if (pretw_in) x = cmul(x, pass_twiddle_value<Rt>(w[u]));
      lds[slot[u]] = x;
    }
  }
  __syncthreads();

  // ---- radix sweeps in LDS (first stage from / last stage to global memory
  // when fused: the element reads/writes of line cl then go straight to HBM,
  // coalesced across the threads of the line)
  const int cl = tid / p.Tl;
  const int jl = tid - cl * p.Tl;
  const bool active = (cl < nvalid) && (tid < C * p.Tl);
  cx<Rt>* myline = lds + (active ? cl : 0) * ldl;
  const int cs = active ? cl : 0;
  const Rt scale = (Rt)p.scale;
  const bool has_scale = p.scale != 1.0;
  auto ld = [&](int e) -> cx<Rt> {
    cx<Rt> x = {(Rt)0, (Rt)0};
    if (chirp_in) {
      const int64_t n = sh.cin[cs] + (int64_t)e * p.cae;
      if (n >= p.chirp_n) return x;
      x = cmul(cconj_if(gld(in + sh.in[cs] + (int64_t)e * p.ae), sgn_in), cconj(gld(chirp + n)));
    } else {
      x = cconj_if(gld(in + sh.in[cs] + (int64_t)e * p.ae), sgn_in);
    }
    if (pretw_in) x = cmul(x, pass_twiddle<Rt>(p, sh.twk[cs], e));
    return x;
  };
  auto st = [&](int e, cx<Rt> x) {
    if (pretw_out) x = cmul(x, pass_twiddle<Rt>(p, sh.twk[cs], e));
    if (store_mode != STORE_PLAIN) {
      const int64_t n = sh.cout[cs] + (int64_t)e * p.cbe;
      if (store_mode == STORE_BHAT) {
        x = cconj(cmul(x, gld(bhat + n)));
      } else {
        if (n >= p.chirp_n) return;
        x = cconj(cmul(gld(chirp + n), x));
      }
    }
    x = cconj_if(x, sgn_out);
    if (has_scale) x = cscale(x, scale);
    gst(out + sh.out[cs] + (int64_t)e * p.be, x);
  };
  // only the plain variant carries the fused global paths (register cost)
  constexpr bool kFuse = MODE == MODE_PLAIN;
  // FP32 only: in FP64 the fused stage's values and factors exceed the full
  // variant's register budget (the planner sets mid_fused for FP32 passes)
  constexpr bool kMidFuse = sizeof(Rt) == 4;
  const bool mid_fused = kMidFuse && has_mid && p.mid_fused;
  // wave-local sweeps (line_sync): the phases that map threads to lines
  // differently (the separate spectrum phase, the store) need the workgroup
  // barrier the sweeps' last stage no longer gives
  const bool ws = p.wave_sync != 0;
  sweep<Rt, EPT, kFuse, BIG, kZero>(myline, jl, active, p, tw, roots, fuse_in, fuse_out && !has_mid, ld, st, nullptr,
                                    0, 0, mid_fused ? p.nst - 1 : p.nst, zero_in, false);
  if (rader && !mid_fused) {
    // the DC term of the first sweep (sum of elements 1..P-1) before the
    // spectrum product overwrites it
    if (active && jl == 0) myline[ldl - kRaderSide + 1] = myline[0];
    line_sync(ws);
  }

  if (kMidFuse && has_mid && p.mid_fused) {
    // first sweep's last stage, spectrum product and second sweep's first
    // stage in one LDS round trip (the planner sets mid_fused only when the
    // schedule's first and last radix are the same codelet radix)
    const int rdc = rader ? ldl - kRaderSide + 1 : -1;
    int64_t mb, ms;
    mid_ref(sh.cmid[cs], p.cme, mb, ms);
    const cx<Rt>* __restrict__ pm = bhat + mb;
    switch (p.rad[p.nst - 1]) {
#define FFTAHIP_MIDSTAGE(R) \
  case R: stage_mid<Rt, R, EPT>(myline, jl, active, p.Tl, p.L, tw, p.lds_sh, pm, (int)ms, rdc, ws); break;
      FFTAHIP_MIDSTAGE(2)
      FFTAHIP_MIDSTAGE(3)
      FFTAHIP_MIDSTAGE(4)
      FFTAHIP_MIDSTAGE(5)
      FFTAHIP_MIDSTAGE(7)
      FFTAHIP_MIDSTAGE(8)
      FFTAHIP_MIDSTAGE(9)
      FFTAHIP_MIDSTAGE(11)
      FFTAHIP_MIDSTAGE(13)
      FFTAHIP_MIDSTAGE(16)
#undef FFTAHIP_MIDSTAGE
      default: break;
    }
    sweep<Rt, EPT, kFuse, BIG, kZero>(myline, jl, active, p, tw, roots, false, fuse_out, ld, st, nullptr, 0, 1, -1, false,
                                      prune_out);
  } else if (has_mid && p.mid_in_stage) {
    int64_t mb, ms;
    mid_ref(sh.cmid[cs], p.cme, mb, ms);
    sweep<Rt, EPT, kFuse, BIG, kZero>(myline, jl, active, p, tw, roots, false, fuse_out, ld, st, bhat + mb, (int)ms, 0, -1,
                                      false, prune_out);
  } else if (has_mid) {
    if (ws) __syncthreads();
    if (p.row_mid && col_step) {
      if (row_c < nvalid) {
        const int c = row_c;
        int64_t mb, ms;
        mid_ref(sh.cmid[c], p.cme, mb, ms);
        const cx<Rt>* __restrict__ fp = bhat + mb + (int64_t)row_t0 * ms;
        const int64_t fstep = (int64_t)tstep * ms;
        cx<Rt>* line = lds + c * ldl;
        for (int t0 = row_t0; t0 < L; t0 += tstep * U) {
          cx<Rt> f[U], x[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int t = t0 + u * tstep;
            if (t < L) {
              f[u] = gld(fp);
              x[u] = line[pad_idx(t, pshe)];
            }
            fp += fstep;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int t = t0 + u * tstep;
            if (t < L) line[pad_idx(t, pshe)] = cconj(cmul(x[u], f[u]));
          }
        }
      }
    } else
    for (int e0 = 0; e0 < total; e0 += BLOCK * U) {
      cx<Rt> f[U];
      int slot[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * BLOCK + tid;
        slot[u] = -1;
        if (e < total) {
          const int c = C == 1 ? 0 : (int)fdiv((uint32_t)e, p.divL);
          const int t = e - c * L;
          if (c < nvalid) {
            slot[u] = c * ldl + pad_idx(t, pshe);
            int64_t mb, ms;
            mid_ref(sh.cmid[c], p.cme, mb, ms);
            f[u] = gld(bhat + mb + (int64_t)t * ms);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (slot[u] >= 0) lds[slot[u]] = cconj(cmul(lds[slot[u]], f[u]));
    }
    __syncthreads();
    sweep<Rt, EPT, kFuse, BIG, kZero>(myline, jl, active, p, tw, roots, false, fuse_out, ld, st, nullptr, 0, 0, -1, false,
                                      prune_out);
  }

  if (ws) __syncthreads();

  // ---- store (LDS -> global).  As on the load side, the table factors of a
  // batch are all requested before the first store of the batch.
  if (MODE == MODE_PLAIN) {
    if (!fuse_out && (p.col_out || p.row_out) && col_step) {
      const int oc = p.col_out ? my_c : row_c;
      const int ot0 = p.col_out ? my_t0 : row_t0;
      if (oc < nvalid) {
        cx<Rt>* __restrict__ dst = out + sh.out[oc];
        const cx<Rt>* line = lds + oc * ldl;
        for (int t = ot0; t < L; t += tstep) {
          cx<Rt> y = cconj_if(line[pad_idx(t, pshe)], sgn_out);
          if (has_scale) y = cscale(y, scale);
          gst(dst + (int64_t)t * p.be, y);
        }
      }
      return;
    }
    for (int e0 = 0; e0 < (fuse_out ? 0 : total); e0 += BLOCK * U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * BLOCK + tid;
        if (e >= total) continue;
        int c, t;
        if (p.col_out) {
          t = (int)fdiv((uint32_t)e, p.divC);
          c = e - t * C;
        } else if (C == 1) {
          c = 0;
          t = e;
        } else {
          c = (int)fdiv((uint32_t)e, p.divL);
          t = e - c * L;
        }
        if (c >= nvalid) continue;
        cx<Rt> y = cconj_if(lds[c * ldl + pad_idx(t, pshe)], sgn_out);
        if (has_scale) y = cscale(y, scale);
        gst(out + sh.out[c] + (int64_t)t * p.be, y);
      }
    }
    return;
  }
  if (rader) {
    // X_0 = x0 + v0; X_k = x0 + conj(v[m]), m = (L - rtab[k]) mod L
    const int oc = p.col_out ? my_c : row_c;
    const int ot0 = p.col_out ? my_t0 : row_t0;
    if (oc < nvalid) {
      cx<Rt>* __restrict__ dst = out + sh.out[oc];
      const cx<Rt>* line = lds + oc * ldl;
      const cx<Rt> x0 = line[ldl - kRaderSide];
      const cx<Rt> v0 = line[ldl - kRaderSide + 1];
      const int P = L + 1;
      for (int t0 = ot0; t0 < P; t0 += tstep * U) {
        int r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          r[u] = (t > 0 && t < P) ? gld(p.rtab + t) : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          if (t >= P) continue;
          cx<Rt> y = t == 0 ? cadd(x0, v0) : cadd(x0, cconj(line[pad_idx(r[u] ? L - r[u] : 0, pshe)]));
          y = cconj_if(y, sgn_out);
          if (has_scale) y = cscale(y, scale);
          gst(dst + (int64_t)t * p.be, y);
        }
      }
    }
    return;
  }
  if ((p.col_out || p.row_out) && col_step) {
    const int oc = p.col_out ? my_c : row_c;
    const int ot0 = p.col_out ? my_t0 : row_t0;
    if (oc < nvalid) {
      const int c = oc;
      // output and table pointers advance by a fixed stride per step; a
      // chirp store writes positions t < tlive only (n < chirp_n)
      const int64_t dstep = (int64_t)tstep * p.be;
      cx<Rt>* __restrict__ dp = out + sh.out[c] + (int64_t)ot0 * p.be;
      const cx<Rt>* __restrict__ tab = store_mode == STORE_BHAT ? bhat : chirp;
      const int64_t cout0 = sh.cout[c];
      const cx<Rt>* __restrict__ fp = tab + cout0 + (int64_t)ot0 * p.cbe;
      const int64_t fstep = (int64_t)tstep * p.cbe;
      const int tlive = store_mode == STORE_CHIRP ? live_bound(p.chirp_n - cout0, p.cbe, L) : L;
      double2 w = make_double2(1.0, 0.0), wr = w;
      // exponents k*t*s with t < L only (the tables end at M)
      if (kGeo && pretw_out && ot0 < L) {
        w = step_twiddle(p, sh.twk[c] * (int64_t)ot0);
        if (tstep < L) wr = step_twiddle(p, sh.twk[c] * (int64_t)tstep);
      }
      for (int t0 = ot0; t0 < tlive; t0 += tstep * U) {
        cx<Rt> x[U], f[U];
        PassTw<Rt> pw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          x[u] = f[u] = {(Rt)0, (Rt)0};
          if (t < tlive) {
            if (store_mode != STORE_PLAIN) f[u] = gld(fp);
            if (!kGeo && pretw_out) pw[u] = pass_twiddle_load<Rt>(p, sh.twk[c], t);
            x[u] = lds[c * ldl + pad_idx(t, pshe)];
          }
          fp += fstep;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + u * tstep;
          if (t < L) {
            cx<Rt> y = x[u];
            if (pretw_out) {
              if (kGeo) {
                y = cmul(y, from_d2<Rt>(w));
                w = dmul2(w, wr);
              } else {
                y = cmul(y, pass_twiddle_value<Rt>(pw[u]));
              }
            }
            if (store_mode == STORE_BHAT) y = cconj(cmul(y, f[u]));
            else if (store_mode == STORE_CHIRP) y = cconj(cmul(f[u], y));
            y = cconj_if(y, sgn_out);
            if (has_scale) y = cscale(y, scale);
            if (t < tlive) gst(dp, y);
          }
          dp += dstep;
        }
      }
    }
    return;
  }
  for (int e0 = 0; e0 < total; e0 += BLOCK * U) {
    cx<Rt> x[U], f[U];
    PassTw<Rt> w[U];
    int64_t o[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * BLOCK + tid;
      ok[u] = false;
      if (e >= total) continue;
      int c, t;
      if (p.col_out) {
        t = (int)fdiv((uint32_t)e, p.divC);
        c = e - t * C;
      } else if (C == 1) {
        c = 0;
        t = e;
      } else {
        c = (int)fdiv((uint32_t)e, p.divL);
        t = e - c * L;
      }
      if (c >= nvalid) continue;
      if (store_mode != STORE_PLAIN) {
        const int64_t n = sh.cout[c] + (int64_t)t * p.cbe;
        if (store_mode == STORE_BHAT) {
          f[u] = gld(bhat + n);
        } else {
          if (n >= p.chirp_n) continue;
          f[u] = gld(chirp + n);
        }
      }
      if (pretw_out) w[u] = pass_twiddle_load<Rt>(p, sh.twk[c], t);
      x[u] = lds[c * ldl + pad_idx(t, pshe)];
      o[u] = sh.out[c] + (int64_t)t * p.be;
      ok[u] = true;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      cx<Rt> y = x[u];
      if (pretw_out) y = cmul(y, pass_twiddle_value<Rt>(w[u]));
      if (store_mode == STORE_BHAT) y = cconj(cmul(y, f[u]));
      else if (store_mode == STORE_CHIRP) y = cconj(cmul(f[u], y));
      y = cconj_if(y, sgn_out);
      if (has_scale) y = cscale(y, scale);
      gst(out + o[u], y);
    }
  }
}

// Work-list launch of ntiles tiles over ngroups groups: with several groups,
// tile_start holds two words per tile, its group and its tile index within
// the group (one pair of loads per workgroup instead of a search); with one
// group it is not read.  Every group is an independent LinePass of the same precision,
// block size and mode; groups may differ in length, radices and fused
// operations.  Tiles are visited in XCD-aware order (workgroups are dealt
// round-robin to the 8 XCDs, so tile = (bid % 8) * (grid / 8) + bid / 8 with
// grid a multiple of 8 gives each XCD a contiguous range of tiles and its
// L2 the neighbouring lines).  At least 4 waves per SIMD (<= 128 VGPRs), 7
// for the FP32 plain variant (<= 72 VGPRs): the passes are latency bound on
// global memory.
// BIG: the variant whose sweeps also run direct-DFT stages of prime radices
// >= 17 (stage_big); kept apart so that the other variants carry none of its
// register cost.
template <class Rt, int EPT, int BLOCK, int MODE, bool BIG>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(MODE == MODE_PLAIN && sizeof(Rt) == 4 && !BIG ? 7 : 4)))
line_pass_group_kernel(const LinePass* __restrict__ passes, const uint32_t* __restrict__ tile_start, int ngroups,
                       BufPtrs bp, uint32_t ntiles) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const uint32_t bid = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (bid >= ntiles) return;
  // several groups: tile_start holds (group, tile within the group) per tile
  int lo = 0;
  uint32_t local = bid;
  if (ngroups > 1) {
    lo = (int)tile_start[2 * bid];
    local = tile_start[2 * bid + 1];
  }
  const LinePass& p = passes[lo];
  const int C = p.C;
  int64_t* tab = reinterpret_cast<int64_t*>(smem_raw + (tile_lds_bytes(C, p.ldl, sizeof(cx<Rt>)) - (size_t)C * kLineTabBytes));
  const LineShared sh{tab, tab + C, tab + 2 * C, tab + 3 * C, tab + 4 * C, tab + 5 * C};
  run_line_pass<Rt, EPT, BLOCK, MODE, BIG>(p, bp, local, reinterpret_cast<cx<Rt>*>(smem_raw), sh);
}

// ------------------------------------------------------------------ register line passes

// A pass of short lines laid out across memory (line c+1 starts one element
// after line c on both sides) runs one line per thread, entirely in
// registers: each of the L loads and L stores of a wave covers 64
// consecutive elements, and there is no LDS staging and no barrier.  The
// line FFT is one compile-time codelet (dft_any<L>).  Same pass semantics as
// run_line_pass except the mid-pass spectrum product (not supported here).
constexpr int kRegChunk = 4;

template <class Rt, int L, int MODE>
__device__ __forceinline__ void reg_line(const LinePass& p, const BufPtrs& bp, uint32_t line) {
  const LineOffsets o = line_offsets<MODE>(p, line);
  const void* inb = p.in_buf == BUF_IN ? bp.in : (p.in_buf == BUF_OUT ? (const void*)bp.out : (const void*)bp.scratch);
  void* outb = p.out_buf == BUF_OUT ? bp.out : bp.scratch;
  const cx<Rt>* __restrict__ in = reinterpret_cast<const cx<Rt>*>(inb) + p.in_off + o.in;
  cx<Rt>* __restrict__ out = reinterpret_cast<cx<Rt>*>(outb) + p.out_off + o.out;
  const cx<Rt>* __restrict__ chirp = reinterpret_cast<const cx<Rt>*>(p.chirp);
  const cx<Rt>* __restrict__ bhat = reinterpret_cast<const cx<Rt>*>(p.bhat);
  const uint32_t sgn_in = p.conj_in ? 0x80000000u : 0u;
  const uint32_t sgn_out = p.conj_out ? 0x80000000u : 0u;
  const bool chirp_in = MODE == MODE_FULL && p.load_mode == LOAD_CHIRP;
  const bool pretw_in = MODE != MODE_PLAIN && p.pretw == 1;
  const bool pretw_out = MODE == MODE_FULL && p.pretw == 2;
  const int store_mode = MODE == MODE_FULL ? p.store_mode : STORE_PLAIN;

  cx<Rt> v[L];
  // Computed chirp (chirp_d set): c[a + t*s] = c[a] * w^t * d[t], where
  // (a + t s)^2 = a^2 + 2 a s t + s^2 t^2; c[a] and w = exp(i pi 2as / N)
  // come from sincospi once per line (exact integer reduction), w^t as an
  // FP64 progression (< 48 steps), d[t] from the pass's FP64 table.  No
  // table loads per element, so every element load of the line is in flight
  // at once.
  if (chirp_in && p.chirp_d) {
    const int64_t N = p.chirp_n;
    const uint64_t a = (uint64_t)o.cin, st = (uint64_t)p.cae;
    // a, st < N (the planner attaches chirp steps only then): exact reductions
    double2 c = chirp_root_r((int64_t)mod_q31(a * a, 2 * (uint64_t)N, 0.5 * p.chirp_inv), N);
    const double2 w = chirp_root_r((int64_t)(2 * mod_q31(a * st, (uint64_t)N, p.chirp_inv)), N);
    // element loads in chunks of kChirpChunk in flight (register budget);
    // the live positions are a prefix t < tl, and the dead ones (zero in a
    // padded Bluestein line, about half of them) load and compute nothing
    const int tl = live_bound(N - o.cin, p.cae, L);
    constexpr int kChirpChunk = 8;
#pragma unroll
    for (int t0 = 0; t0 < L; t0 += kChirpChunk) {
#pragma unroll
      for (int t = t0; t < t0 + kChirpChunk && t < L; ++t) {
        v[t] = {(Rt)0, (Rt)0};
        if (t < tl) v[t] = gld(in + (int64_t)t * p.ae);
      }
#pragma unroll
      for (int t = t0; t < t0 + kChirpChunk && t < L; ++t) {
        if (t < tl) {
          const cx<Rt> ct = from_d2<Rt>(dmul2(c, uld(p.chirp_d + t)));
          v[t] = cmul(cconj_if(v[t], sgn_in), cconj(ct));
          c = dmul2(c, w);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else
  // positions past the chirp length read the line's first element (always
  // in bounds) and are zeroed, so the loads stay unconditional
#pragma unroll
  for (int t = 0; t < L; ++t) {
    if (chirp_in) {
      const int64_t n = o.cin + (int64_t)t * p.cae;
      const bool live = n < p.chirp_n;
      const cx<Rt> x = gld(in + (live ? (int64_t)t * p.ae : 0));
      const cx<Rt> c = gld(chirp + (live ? n : 0));
      const cx<Rt> y = cmul(cconj_if(x, sgn_in), cconj(c));
      v[t] = live ? y : cx<Rt>{(Rt)0, (Rt)0};
      // keep at most kRegChunk table loads in flight (register budget)
      if (t % kRegChunk == kRegChunk - 1) __builtin_amdgcn_sched_barrier(0);
    } else {
      v[t] = cconj_if(gld(in + (int64_t)t * p.ae), sgn_in);
    }
  }
  // FP32: the pass twiddles of a line, w^{k*t*s} for t = 1..L-1, are the
  // powers of w^{k*s}, formed by FP64 products (< 48 steps: error below
  // 1e-14 relative, invisible after rounding to FP32); FP64 looks every
  // factor up.
  constexpr bool kGeo = sizeof(Rt) == 4;
  double2 wr = make_double2(1.0, 0.0);
  if (kGeo && (pretw_in || pretw_out)) wr = step_twiddle(p, o.twk);
  if (pretw_in) {
    double2 w = wr;
#pragma unroll
    for (int t = 1; t < L; ++t) {
      if (kGeo) {
        v[t] = cmul(v[t], from_d2<Rt>(w));
        w = dmul2(w, wr);
      } else {
        v[t] = cmul(v[t], pass_twiddle<Rt>(p, o.twk, t));
        if (t % kRegChunk == 0) __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  dft_any<Rt, L>(v);
  const Rt scale = (Rt)p.scale;
  const bool has_scale = p.scale != 1.0;
  double2 w = wr;
  // computed output chirp (as on the load side): c[a + k*s] = c[a] w^k d[k];
  // such passes carry no post-twiddle (the planner attaches chirp steps to
  // a chirp store only without one), so this loop holds no twiddle state
  if (store_mode == STORE_CHIRP && p.chirp_d) {
    const uint64_t a = (uint64_t)o.cout, st = (uint64_t)p.cbe;
    const uint64_t Nc = (uint64_t)p.chirp_n;
    double2 oc = chirp_root_r((int64_t)mod_q31(a * a, 2 * Nc, 0.5 * p.chirp_inv), p.chirp_n);
    const double2 ow = chirp_root_r((int64_t)(2 * mod_q31(a * st, Nc, p.chirp_inv)), p.chirp_n);
    // the stored positions are a prefix k < kl (the rest of a padded
    // Bluestein column is dropped): no chirp work past it
    const int kl = live_bound(p.chirp_n - o.cout, p.cbe, L);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      if (k < kl) {
        cx<Rt> y = cconj(cmul(from_d2<Rt>(dmul2(oc, uld(p.chirp_d + k))), v[k]));
        oc = dmul2(oc, ow);
        y = cconj_if(y, sgn_out);
        if (has_scale) y = cscale(y, scale);
        gst(out + (int64_t)k * p.be, y);
      }
      if (k % kRegChunk == kRegChunk - 1) __builtin_amdgcn_sched_barrier(0);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < L; ++k) {
    cx<Rt> y = v[k];
    if (pretw_out && k > 0) {
      if (kGeo) {
        y = cmul(y, from_d2<Rt>(w));
        w = dmul2(w, wr);
      } else {
        y = cmul(y, pass_twiddle<Rt>(p, o.twk, k));
      }
    }
    bool live = true;
    if (store_mode == STORE_BHAT) {
      y = cconj(cmul(y, gld(bhat + o.cout + (int64_t)k * p.cbe)));
    } else if (store_mode == STORE_CHIRP) {
      const int64_t n = o.cout + (int64_t)k * p.cbe;
      live = n < p.chirp_n;
      y = cconj(cmul(gld(chirp + (live ? n : 0)), y));
    }
    y = cconj_if(y, sgn_out);
    if (has_scale) y = cscale(y, scale);
    if (live) gst(out + (int64_t)k * p.be, y);
    if (MODE != MODE_PLAIN && k % kRegChunk == kRegChunk - 1) __builtin_amdgcn_sched_barrier(0);
  }
}

// Line lengths of the register passes: every 13-smooth length up to
// reg_max (one line's values, and the table factors in flight, must fit the
// 128 registers): FP32 48 (24 with the full set of fused operations), FP64
// 24 for plain passes, 16 with the load-side pass twiddle, 12 with the full
// set.
template <class Rt, int MODE>
constexpr int reg_max() {
  if (sizeof(Rt) == 4) return MODE == MODE_FULL ? 24 : 48;
  return MODE == MODE_PLAIN ? 24 : (MODE == MODE_PRETW ? 16 : 12);
}

template <class Rt, int MODE>
__device__ __forceinline__ void reg_dispatch(const LinePass& p, const BufPtrs& bp, uint32_t line) {
  switch (p.L) {
#define FFTAHIP_REG(N) \
  case N:              \
    if constexpr (N <= reg_max<Rt, MODE>()) reg_line<Rt, N, MODE>(p, bp, line); \
    break;
    FFTAHIP_REG(2) FFTAHIP_REG(3) FFTAHIP_REG(4) FFTAHIP_REG(5) FFTAHIP_REG(6) FFTAHIP_REG(7) FFTAHIP_REG(8)
    FFTAHIP_REG(9) FFTAHIP_REG(10) FFTAHIP_REG(11) FFTAHIP_REG(12) FFTAHIP_REG(13) FFTAHIP_REG(14) FFTAHIP_REG(15)
    FFTAHIP_REG(16) FFTAHIP_REG(18) FFTAHIP_REG(20) FFTAHIP_REG(21) FFTAHIP_REG(22) FFTAHIP_REG(24) FFTAHIP_REG(25)
    FFTAHIP_REG(26) FFTAHIP_REG(27) FFTAHIP_REG(28) FFTAHIP_REG(30) FFTAHIP_REG(32) FFTAHIP_REG(33) FFTAHIP_REG(35)
    FFTAHIP_REG(36) FFTAHIP_REG(39) FFTAHIP_REG(40) FFTAHIP_REG(42) FFTAHIP_REG(44) FFTAHIP_REG(45) FFTAHIP_REG(48)
#undef FFTAHIP_REG
    default: break;
  }
}

// Work-list launch of register passes: tile = 256 consecutive lines of one
// group (pass); tiles in the XCD-aware order of line_pass_group_kernel.  The
// FP32 full variant (the three-pass Bluestein column passes, latency-bound)
// is built for 6 resident waves per SIMD (80 VGPRs, 8 bytes of spill in its
// longest lines): more loads in flight per CU; the batched step ran 3.36 ms
// against 3.46 ms with 4 waves and 3.39 ms with 5, while 8 waves (64 VGPRs)
// spill 136 bytes and ran 3.51 ms.
template <class Rt, int MODE>
__global__ void __launch_bounds__(kRegBlock)
__attribute__((amdgpu_waves_per_eu(sizeof(Rt) == 4 && MODE == MODE_FULL ? 6 : 4)))
reg_pass_group_kernel(const LinePass* __restrict__ passes, const uint32_t* __restrict__ tile_start, int ngroups,
                      BufPtrs bp, uint32_t ntiles) {
  const uint32_t bid = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (bid >= ntiles) return;
  // several groups: tile_start holds (group, tile within the group) per tile
  int lo = 0;
  uint32_t local = bid;
  if (ngroups > 1) {
    lo = (int)tile_start[2 * bid];
    local = tile_start[2 * bid + 1];
  }
  const LinePass& p = passes[lo];
  const int64_t line = (int64_t)local * kRegBlock + threadIdx.x;
  if (line >= p.nlines) return;
  reg_dispatch<Rt, MODE>(p, bp, (uint32_t)line);
}

}  // namespace fftahip
