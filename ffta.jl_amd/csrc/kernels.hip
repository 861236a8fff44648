// kernels.hip -- gfx950 kernels of the transform path.
//
// line_pass: a batch of independent in-LDS line FFTs (self-sorting Stockham
// radix sweeps with in-register codelets), with fused load/store operations
// (conjugation for the backward sign, multi-pass pre-twiddles, Bluestein
// chirp and spectrum products).  All passes of every plan are instances of
// this one kernel; see plan.cpp for how plans are built from it.
//
// Table generators: stage twiddles, two-level FP64 twiddles, Bluestein chirp
// and spectrum, all evaluated in FP64 with the angle reduced exactly in
// integers before sincospi.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.h"
#include "codelets.h"
#include "kernels.h"
#include "device.h"

namespace fftahip {

// ------------------------------------------------------------------ tables

// Per-stage twiddles of a radix schedule: stage s (Ns = product of earlier
// radices, radix R) holds w_{Ns*R}^{t*k} for k < Ns, t = 1..R-1 at
// (Ns-1) + k*(R-1) + t-1; L-1 entries in all.  After them, every stage of a
// prime radix P >= 17 (stage_big) appends the P roots w_P^m, m < P.
template <class Rt>
__global__ void gen_stage_twiddles_kernel(cx<Rt>* tw, RadixList rl, int64_t n) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  int64_t Ns = 1;
  for (int s = 0; s < rl.nst; ++s) {
    const int64_t R = rl.rad[s];
    const int64_t lo = Ns - 1, hi = Ns * R - 1;
    if (m >= lo && m < hi) {
      const int64_t r = m - lo;
      const int64_t k = r / (R - 1), t = r - k * (R - 1) + 1;
      tw[m] = from_d2<Rt>(unit_root(t * k, Ns * R));
      return;
    }
    Ns *= R;
  }
  int64_t off = Ns - 1;  // Ns = L here
  for (int s = 0; s < rl.nst; ++s) {
    const int64_t R = rl.rad[s];
    if (!big_stage((int)R, sizeof(Rt) == 4 ? 32 : 64)) continue;
    if (m >= off && m < off + R) {
      tw[m] = from_d2<Rt>(unit_root(m - off, R));
      return;
    }
    off += R;
  }
}

// lo[m] = w_N^m for m < 2^sh; hi[h] = w_N^{h * 2^sh} for h < nhi
__global__ void gen_twiddle_2level_kernel(double2* lo, double2* hi, int64_t N, int sh, int64_t nhi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nlo = 1ll << sh;
  if (i < nlo) lo[i] = unit_root(i, N);
  if (i < nhi) hi[i] = unit_root(i << sh, N);
}

// tw16[a*16 + j] = w_N^{(a*j*s) mod N} for a < amax, j < 16 (second factor
// of the pass twiddles), FP64 evaluated, stored in the pass precision.
template <class Rt>
__global__ void gen_tw16_kernel(cx<Rt>* out, int64_t N, int64_t s, int64_t amax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= amax * 16) return;
  const int64_t a = i >> 4, j = i & 15;
  out[i] = from_d2<Rt>(unit_root(((a * j) % N) * s % N, N));
}

// ------------------------------------------------------------------ real transforms
// Boundary kernels of rfft / brfft (src/plan.jl:386-500): even lengths
// n = 2m run one complex FFT of m points on the packed pairs (x[2t],
// x[2t+1]) with a split step before (brfft) or after (rfft) it (Sorensen et
// al., eqs. 6-9, as FFTA cites); odd lengths run a complex FFT of n points
// on the complexified input / Hermitian extension.  One thread per element
// of the (p, t, q) space, p fastest, so neighbouring threads touch
// neighbouring elements on both sides.

__device__ __forceinline__ void real_geom_index(const RealGeom& g, uint64_t i, int64_t& p, int64_t& t, int64_t& q) {
  const uint32_t a = fdiv((uint32_t)i, g.divP);
  p = (int64_t)((uint32_t)i - a * (uint32_t)g.P);
  const uint32_t b = fdiv(a, g.divT);
  t = (int64_t)(a - b * (uint32_t)g.T);
  q = (int64_t)b;
}

// w_n^{m'} for 0 <= m' < n from the two-level FP64 tables of n.
__device__ __forceinline__ double2 twiddle_2l(const double2* lo, const double2* hi, int sh, int64_t m) {
  const int64_t mask = (1ll << sh) - 1;
  return dmul2(gld(lo + (m & mask)), gld(hi + (m >> sh)));
}

// A[p,t,q] = x[p, 2t, q] + i x[p, 2t+1, q]   (T = m pairs)
template <class Rt>
__global__ void real_pack_kernel(cx<Rt>* A, const Rt* x, RealGeom g) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, t, q;
  real_geom_index(g, i, p, t, q);
  const Rt* src = x + p + q * g.in_q + 2 * t * g.in_t;
  A[p + t * g.out_t + q * g.out_q] = {src[0], src[g.in_t]};
}

// y[p, 2t, q] = Re B[p,t,q], y[p, 2t+1, q] = Im B[p,t,q]
template <class Rt>
__global__ void real_unpack_kernel(Rt* y, const cx<Rt>* B, RealGeom g) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, t, q;
  real_geom_index(g, i, p, t, q);
  const cx<Rt> v = B[p + t * g.in_t + q * g.in_q];
  Rt* dst = y + p + q * g.out_q + 2 * t * g.out_t;
  dst[0] = v.x;
  dst[g.out_t] = v.y;
}

// A[p,t,q] = x[p,t,q] + 0i   (odd rfft)
template <class Rt>
__global__ void real_complexify_kernel(cx<Rt>* A, const Rt* x, RealGeom g) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, t, q;
  real_geom_index(g, i, p, t, q);
  A[p + t * g.out_t + q * g.out_q] = {x[p + t * g.in_t + q * g.in_q], (Rt)0};
}

// y[p,t,q] = Re B[p,t,q]   (odd brfft)
template <class Rt>
__global__ void real_part_kernel(Rt* y, const cx<Rt>* B, RealGeom g) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, t, q;
  real_geom_index(g, i, p, t, q);
  y[p + t * g.out_t + q * g.out_q] = B[p + t * g.in_t + q * g.in_q].x;
}

// Y[p,k,q] = B[p,k,q] for k <= n/2   (odd rfft: T = n/2 + 1)
template <class Rt>
__global__ void copy_lines_kernel(cx<Rt>* Y, const cx<Rt>* B, RealGeom g) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, t, q;
  real_geom_index(g, i, p, t, q);
  Y[p + t * g.out_t + q * g.out_q] = B[p + t * g.in_t + q * g.in_q];
}

// A[p,t,q] = X[p,t,q] for t <= n/2, conj(X[p,n-t,q]) above   (odd brfft: T = n)
template <class Rt>
__global__ void hermitian_extend_kernel(cx<Rt>* A, const cx<Rt>* X, RealGeom g, int64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, t, q;
  real_geom_index(g, i, p, t, q);
  const int64_t s = t <= n / 2 ? t : n - t;
  const cx<Rt> v = X[p + s * g.in_t + q * g.in_q];
  A[p + t * g.out_t + q * g.out_q] = t <= n / 2 ? v : cconj(v);
}

// rfft split step, in place on Y (m + 1 entries along t; Y[0..m-1] hold the
// m-point FFT Z of the packed pairs): for k = 0..m/2 (T = m/2 + 1)
//   Y[0] = Re Z0 + Im Z0, Y[m] = Re Z0 - Im Z0;
//   XX = (Z_k + conj Z_{m-k}) / 2, XY = i (conj Z_{m-k} - Z_k) / 2,
//   Y[k] = XX + w^k XY, Y[m-k] = conj(XX - w^k XY),  w = exp(-2 pi i / n).
template <class Rt>
__global__ void rfft_split_kernel(cx<Rt>* Y, RealGeom g, int64_t m, const double2* lo, const double2* hi, int sh) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, k, q;
  real_geom_index(g, i, p, k, q);
  cx<Rt>* line = Y + p + q * g.out_q;
  if (k == 0) {
    const cx<Rt> z = line[0];
    line[0] = {z.x + z.y, (Rt)0};
    line[m * g.out_t] = {z.x - z.y, (Rt)0};
    return;
  }
  const cx<Rt> a = line[k * g.out_t];
  const cx<Rt> b = cconj(line[(m - k) * g.out_t]);
  const Rt h = (Rt)0.5;
  const cx<Rt> xx = {h * (a.x + b.x), h * (a.y + b.y)};
  const cx<Rt> d = {h * (b.x - a.x), h * (b.y - a.y)};
  const cx<Rt> xy = {-d.y, d.x};  // i * d
  const cx<Rt> w = from_d2<Rt>(twiddle_2l(lo, hi, sh, k));
  const cx<Rt> wxy = cmul(w, xy);
  // k = m/2 pairs with itself: FFTA's second assignment wins (src/plan.jl:438-439)
  if (2 * k != m) line[k * g.out_t] = cadd(xx, wxy);
  line[(m - k) * g.out_t] = cconj(csub(xx, wxy));
}

// brfft split step (X has m + 1 entries along t; A gets m), for j = 0..m/2:
//   A[0] = (Re X0 + Re Xm) + i (Re X0 - Re Xm);
//   XX = X_j + conj X_{m-j}, XY = w^j (X_j - conj X_{m-j}),  w = exp(+2 pi i / n),
//   A[j] = XX + i XY, A[m-j] = conj(XX - i XY).
template <class Rt>
__global__ void brfft_split_kernel(cx<Rt>* A, const cx<Rt>* X, RealGeom g, int64_t m, const double2* lo,
                                   const double2* hi, int sh, int64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)(g.P * g.T * g.Q)) return;
  int64_t p, j, q;
  real_geom_index(g, i, p, j, q);
  const cx<Rt>* src = X + p + q * g.in_q;
  cx<Rt>* dst = A + p + q * g.out_q;
  if (j == 0) {
    const Rt x0 = src[0].x, xm = src[m * g.in_t].x;
    dst[0] = {x0 + xm, x0 - xm};
    return;
  }
  const cx<Rt> a = src[j * g.in_t];
  const cx<Rt> b = cconj(src[(m - j) * g.in_t]);
  const cx<Rt> xx = cadd(a, b);
  // w^j with w = exp(+2 pi i / n) is exp(-2 pi i (n - j) / n)
  const cx<Rt> w = from_d2<Rt>(twiddle_2l(lo, hi, sh, n - j));
  const cx<Rt> xy = cmul(w, csub(a, b));
  const cx<Rt> ixy = {-xy.y, xy.x};
  // j = m/2 pairs with itself: the second assignment wins (src/plan.jl:481-482)
  if (2 * j != m) dst[j * g.out_t] = cadd(xx, ixy);
  dst[(m - j) * g.out_t] = cconj(csub(xx, ixy));
}

// Forward Bluestein chirp b_n = exp(+i pi p / N), p = n^2 mod 2N taken in
// (-N, N] (src/algos.jl:353-359 with sgn = +1), FP64 evaluated (exact
// reduction: n < N < 2^31, so n^2 < 2^62).
__device__ __forceinline__ double2 chirp_value(int64_t n, int64_t N) {
  const int64_t twoN = 2 * N;
  int64_t p = (int64_t)(((uint64_t)n * (uint64_t)n) % (uint64_t)twoN);
  if (p > N) p -= twoN;
  double s, c;
  sincospi((double)p / (double)N, &s, &c);
  return make_double2(c, s);
}

// The table of the plan precision (and its FP64 copy for the spectrum).
template <class Rt>
__global__ void gen_chirp_kernel(cx<Rt>* chirp, double2* chirp64, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double2 v = chirp_value(n, N);
  chirp[n] = {(Rt)v.x, (Rt)v.y};
  if (chirp64) chirp64[n] = v;
}

// Entries n0 .. n0 + count - 1 of the chirp of length N (read-back of the
// table values at any N without building the whole table).
template <class Rt>
__global__ void gen_chirp_range_kernel(cx<Rt>* out, int64_t n0, int64_t count, int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double2 v = chirp_value(n0 + i, N);
  out[i] = {(Rt)v.x, (Rt)v.y};
}

// Chirp steps of a register pass: d[t] = exp(i pi (t^2 s^2 mod 2N) / N),
// t < L, FP64 (exact integer reduction: t*s mod 2N < 2^32, squared < 2^64).
__global__ void gen_chirp_steps_kernel(double2* d, int64_t s, int64_t N, int L) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= L) return;
  const uint64_t twoN = 2 * (uint64_t)N;
  const uint64_t u = ((uint64_t)t * (uint64_t)s) % twoN;
  d[t] = chirp_root(u * u, N);
}

// Rader's kernel b_r = w_P^{g^r}, r < P - 1 (g a primitive root of P), FP64;
// g^r mod P by binary powering (P < 2^31, products below 2^62).
__global__ void gen_rader_b_kernel(double2* b, int64_t P, int64_t g) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P - 1) return;
  uint64_t acc = 1, base = (uint64_t)g, e = (uint64_t)r;
  while (e) {
    if (e & 1) acc = acc * base % (uint64_t)P;
    base = base * base % (uint64_t)P;
    e >>= 1;
  }
  b[r] = unit_root((int64_t)acc, P);
}

// Periodic extension of the chirp into M slots (b[m], b[M-m] = b[m] for
// 1 <= m < N, zeros between), FP64.
__global__ void gen_chirp_ext_kernel(double2* bext, const double2* chirp64, int64_t N, int64_t M) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  double2 v = make_double2(0.0, 0.0);
  if (i < N) v = chirp64[i];
  else if (M - i < N) v = chirp64[M - i];
  bext[i] = v;
}

// bhat[i] = spec64[i] / M, rounded to the plan precision.
template <class Rt>
__global__ void finish_bhat_kernel(cx<Rt>* bhat, const double2* spec64, int64_t M) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double inv = 1.0 / (double)M;
  bhat[i] = {(Rt)(spec64[i].x * inv), (Rt)(spec64[i].y * inv)};
}

// dst[k1*M2 + k2] = src[k1 + M1*k2]: the spectrum in the order the fused
// Bluestein row pass reads it (row k1 contiguous).
template <class Rt>
__global__ void permute_rows_kernel(cx<Rt>* dst, const cx<Rt>* src, int64_t M1, int64_t M2, int64_t rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * M2) return;
  const int64_t k1 = i / M2, k2 = i - k1 * M2;
  dst[i] = src[k1 + M1 * k2];
}

// ------------------------------------------------------------------ launchers

int pass_mode(const LinePass& p) {
  const bool full = p.load_mode != LOAD_PLAIN || p.mid_mode != MID_NONE || p.store_mode != STORE_PLAIN ||
                    p.pretw == 2;
  if (full) return MODE_FULL;
  return p.pretw == 1 ? MODE_PRETW : MODE_PLAIN;
}

int pass_variant(const LinePass& p) {
  int big = 0;
  for (int s = 0; s < p.nst; ++s)
    if (p.rad[s] >= 17) big = kBigVariant;
  return pass_mode(p) + big;
}

KernelRef lds_kernel(int precision, int block, int mode) {
  KernelRef k{nullptr, block};
  if (precision == 32) {
    if (block == 256) k.fn = lds_fn_f32_256(mode);
    else if (block == 512) k.fn = lds_fn_f32_512(mode);
    else if (block == 1024) k.fn = lds_fn_f32_1024(mode);
  } else {
    if (block == 256) k.fn = lds_fn_f64_256(mode);
    else if (block == 512) k.fn = lds_fn_f64_512(mode);
    else if (block == 1024) k.fn = lds_fn_f64_1024(mode);
  }
  return k;
}

KernelRef reg_kernel(int precision, int mode) {
  return KernelRef{precision == 32 ? reg_fn_f32(mode) : reg_fn_f64(mode), kRegBlock};
}

void init_kernels(size_t lds_per_block) {
  for (int prec : {32, 64})
    for (int block : {256, 512, 1024})
      for (int variant = 0; variant < kLdsVariants; ++variant)
        (void)hipFuncSetAttribute(lds_kernel(prec, block, variant).fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_per_block);
  (void)hipGetLastError();
}

// Grids are padded to a multiple of 8 (the kernels' XCD-aware tile order);
// the kernels take the real tile count.
LaunchArgs::LaunchArgs(const LinePass* P, const uint32_t* T, int ng, const BufPtrs& b, int64_t grid)
    : passes(P), tiles(T), ngroups(ng), bp(b), ntiles((uint32_t)grid) {
  args[0] = &passes;
  args[1] = &tiles;
  args[2] = &ngroups;
  args[3] = &bp;
  args[4] = &ntiles;
}

uint32_t padded_grid(int64_t grid) { return (uint32_t)((grid + 7) & ~7ll); }

hipError_t launch_line_pass_groups(const LinePass* d_passes, const uint32_t* d_tile_start, int ngroups, int variant,
                                   int precision, int block, const BufPtrs& bp, int remap, size_t lds_bytes,
                                   int64_t grid, hipStream_t stream) {
  (void)remap;
  const KernelRef k = lds_kernel(precision, block, variant);
  if (!k.fn) return hipErrorInvalidValue;
  LaunchArgs a(d_passes, d_tile_start, ngroups, bp, grid);
  return hipLaunchKernel(k.fn, dim3(padded_grid(grid)), dim3(k.block), a.args, lds_bytes, stream);
}

hipError_t launch_reg_pass_groups(const LinePass* d_passes, const uint32_t* d_tile_start, int ngroups, int mode,
                                  int precision, const BufPtrs& bp, int64_t grid, hipStream_t stream) {
  const KernelRef k = reg_kernel(precision, mode);
  LaunchArgs a(d_passes, d_tile_start, ngroups, bp, grid);
  return hipLaunchKernel(k.fn, dim3(padded_grid(grid)), dim3(k.block), a.args, 0, stream);
}

int reg_pass_max(int precision, int mode) {
  if (precision == 32)
    return mode == MODE_PLAIN ? reg_max<float, MODE_PLAIN>()
                              : (mode == MODE_PRETW ? reg_max<float, MODE_PRETW>() : reg_max<float, MODE_FULL>());
  return mode == MODE_PLAIN ? reg_max<double, MODE_PLAIN>()
                            : (mode == MODE_PRETW ? reg_max<double, MODE_PRETW>() : reg_max<double, MODE_FULL>());
}


hipError_t launch_gen_stage_twiddles(void* tw, int precision, const RadixList& rl, int64_t n, hipStream_t stream) {
  const int64_t blocks = (n + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(gen_stage_twiddles_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(tw), rl, n);
  else
    hipLaunchKernelGGL(gen_stage_twiddles_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(tw), rl, n);
  return hipGetLastError();
}

hipError_t launch_gen_twiddle_2level(double2* lo, double2* hi, int64_t N, int sh, int64_t nhi, hipStream_t stream) {
  const int64_t n = (1ll << sh) > nhi ? (1ll << sh) : nhi;
  hipLaunchKernelGGL(gen_twiddle_2level_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, lo, hi, N,
                     sh, nhi);
  return hipGetLastError();
}

hipError_t launch_gen_tw16(void* out, int precision, int64_t N, int64_t s, int64_t amax, hipStream_t stream) {
  const int64_t blocks = (amax * 16 + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(gen_tw16_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(out), N, s, amax);
  else
    hipLaunchKernelGGL(gen_tw16_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(out), N, s, amax);
  return hipGetLastError();
}

hipError_t launch_real_boundary(int op, int precision, void* dst, const void* src, const RealGeom& g, int64_t m,
                                const double2* lo, const double2* hi, int sh, int64_t n, hipStream_t s) {
  const uint64_t total = (uint64_t)(g.P * g.T * g.Q);
  if (total == 0) return hipSuccess;
  const dim3 grid((uint32_t)((total + 255) / 256)), block(256);
#define FFTAHIP_REAL(Rt)                                                                                           \
  switch (op) {                                                                                                   \
    case REAL_PACK:                                                                                               \
      hipLaunchKernelGGL(real_pack_kernel<Rt>, grid, block, 0, s, (cx<Rt>*)dst, (const Rt*)src, g);               \
      break;                                                                                                      \
    case REAL_UNPACK:                                                                                             \
      hipLaunchKernelGGL(real_unpack_kernel<Rt>, grid, block, 0, s, (Rt*)dst, (const cx<Rt>*)src, g);             \
      break;                                                                                                      \
    case REAL_COMPLEXIFY:                                                                                         \
      hipLaunchKernelGGL(real_complexify_kernel<Rt>, grid, block, 0, s, (cx<Rt>*)dst, (const Rt*)src, g);         \
      break;                                                                                                      \
    case REAL_PART:                                                                                               \
      hipLaunchKernelGGL(real_part_kernel<Rt>, grid, block, 0, s, (Rt*)dst, (const cx<Rt>*)src, g);               \
      break;                                                                                                      \
    case REAL_COPY:                                                                                               \
      hipLaunchKernelGGL(copy_lines_kernel<Rt>, grid, block, 0, s, (cx<Rt>*)dst, (const cx<Rt>*)src, g);          \
      break;                                                                                                      \
    case REAL_HERMITIAN:                                                                                          \
      hipLaunchKernelGGL(hermitian_extend_kernel<Rt>, grid, block, 0, s, (cx<Rt>*)dst, (const cx<Rt>*)src, g, n); \
      break;                                                                                                      \
    case REAL_RFFT_SPLIT:                                                                                         \
      hipLaunchKernelGGL(rfft_split_kernel<Rt>, grid, block, 0, s, (cx<Rt>*)dst, g, m, lo, hi, sh);               \
      break;                                                                                                      \
    case REAL_BRFFT_SPLIT:                                                                                        \
      hipLaunchKernelGGL(brfft_split_kernel<Rt>, grid, block, 0, s, (cx<Rt>*)dst, (const cx<Rt>*)src, g, m, lo,  \
                         hi, sh, n);                                                                              \
      break;                                                                                                      \
    default:                                                                                                      \
      return hipErrorInvalidValue;                                                                                \
  }
  if (precision == 32) {
    FFTAHIP_REAL(float)
  } else {
    FFTAHIP_REAL(double)
  }
#undef FFTAHIP_REAL
  return hipGetLastError();
}

hipError_t launch_gen_chirp(void* chirp, double2* chirp64, int precision, int64_t N, hipStream_t stream) {
  const int64_t blocks = (N + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(gen_chirp_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(chirp), chirp64, N);
  else
    hipLaunchKernelGGL(gen_chirp_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(chirp), chirp64, N);
  return hipGetLastError();
}

hipError_t launch_gen_chirp_range(void* out, int precision, int64_t n0, int64_t count, int64_t N,
                                  hipStream_t stream) {
  const int64_t blocks = (count + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(gen_chirp_range_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(out), n0, count, N);
  else
    hipLaunchKernelGGL(gen_chirp_range_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(out), n0, count, N);
  return hipGetLastError();
}

hipError_t launch_gen_chirp_steps(double2* d, int64_t s, int64_t N, int L, hipStream_t stream) {
  hipLaunchKernelGGL(gen_chirp_steps_kernel, dim3((uint32_t)((L + 63) / 64)), dim3(64), 0, stream, d, s, N, L);
  return hipGetLastError();
}

hipError_t launch_gen_rader_b(double2* b, int64_t P, int64_t g, hipStream_t stream) {
  hipLaunchKernelGGL(gen_rader_b_kernel, dim3((uint32_t)((P - 1 + 255) / 256)), dim3(256), 0, stream, b, P, g);
  return hipGetLastError();
}

hipError_t launch_gen_chirp_ext(double2* bext, const double2* chirp64, int64_t N, int64_t M, hipStream_t stream) {
  hipLaunchKernelGGL(gen_chirp_ext_kernel, dim3((uint32_t)((M + 255) / 256)), dim3(256), 0, stream, bext, chirp64,
                     N, M);
  return hipGetLastError();
}

hipError_t launch_permute_rows(void* dst, const void* src, int precision, int64_t M1, int64_t M2, int64_t rows,
                               hipStream_t stream) {
  const int64_t blocks = (rows * M2 + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(permute_rows_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(dst), reinterpret_cast<const cx<float>*>(src), M1, M2, rows);
  else
    hipLaunchKernelGGL(permute_rows_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(dst), reinterpret_cast<const cx<double>*>(src), M1, M2, rows);
  return hipGetLastError();
}

hipError_t launch_finish_bhat(void* bhat, const double2* spec64, int precision, int64_t M, hipStream_t stream) {
  const int64_t blocks = (M + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(finish_bhat_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(bhat), spec64, M);
  else
    hipLaunchKernelGGL(finish_bhat_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(bhat), spec64, M);
  return hipGetLastError();
}

}  // namespace fftahip
