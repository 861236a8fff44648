// common.h -- types shared by the host planner and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fftahip {

constexpr int kMaxStages = 12;      // radix stages inside one in-LDS line FFT

// Prime radices above 16 that run as compile-time codelets (dft_prime, one
// butterfly per thread) in the FP32 direct-DFT kernel variants (in FP64 their
// values exceed the register budget); every other prime >= 17 runs the
// table-driven direct-DFT stage, whose roots follow the stage twiddles in the
// stage table.  prec: 32 or 64.
__host__ __device__ constexpr bool prime_codelet(int r, int prec) {
  return prec == 32 && (r == 17 || r == 19 || r == 23);
}
__host__ __device__ constexpr bool big_stage(int r, int prec) { return r >= 17 && !prime_codelet(r, prec); }
// Codelet radices of 5-smooth convolution lines whose first / last stage has
// a zero-padded variant (inputs or outputs t >= (r + 1) / 2 dropped).
__host__ __device__ constexpr bool zero_pad_radix(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 9 || r == 16;
}
constexpr int kMaxLines = 256;      // lines per workgroup, at most
constexpr int kRegBlock = 256;      // lines (threads) per workgroup of a register line pass
// per-line offset tables of a line-pass tile, kept in LDS after its lines:
// six int64 per line
constexpr int kLineTabBytes = 6 * 8;
// LDS a tile of C lines of pitch ldl (elements of esz bytes) takes: the lines,
// 16-byte aligned, then the per-line tables
__host__ __device__ inline size_t tile_lds_bytes(int64_t C, int64_t ldl, int64_t esz) {
  return (size_t)(((C * ldl * esz + 15) & ~(int64_t)15) + C * kLineTabBytes);
}

// Unsigned 32-bit division by a runtime constant: q = (umulhi(n, m) + n) >> s.
// Exact for every n < 2^32 when d < 2^31 (round-up method of Granlund and
// Montgomery, 1994).
struct FastDiv {
  uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  f.s = s;
  uint64_t m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
  f.m = (uint32_t)m;
  return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, FastDiv f) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t hi = __umulhi(n, f.m);
#else
  uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
  return (uint32_t)(((uint64_t)hi + n) >> f.s);
}

// m mod d for d < 2^32 and m < 2^62 whose quotient is below 2^31, exact:
// the FP64 product m * (1/d) is within 2^-12 of m/d (m rounds to FP64 with an
// error below 2^9 when m >= 2^53, and then d > 2^22; the two roundings add a
// relative 2^-52), so the truncated quotient is off by at most one, which one
// correction step removes.  About a third of a generic 64-bit remainder.
__host__ __device__ inline uint64_t mod_q31(uint64_t m, uint64_t d, double inv_d) {
  const uint64_t q = (uint64_t)((double)m * inv_d);
  int64_t r = (int64_t)(m - q * d);
  if (r < 0) r += (int64_t)d;
  else if (r >= (int64_t)d) r -= (int64_t)d;
  return (uint64_t)r;
}

// What a line pass does when it loads a line into LDS.
enum LoadMode : int32_t {
  LOAD_PLAIN = 0,  // v = x
  LOAD_CHIRP = 1,  // Bluestein: v = x[n] * conj(chirp[n]) for n < chirp_n, 0 beyond
  // Rader: the line holds the chirp_n = P elements of a prime-length DFT;
  // element t >= 1 goes to LDS slot rtab[t] (a' = x permuted by powers of a
  // primitive root), element 0 to the line's side slot x0
  LOAD_RADER = 2,
};

// What a line pass does between its first and second radix sweep.
enum MidMode : int32_t {
  MID_NONE = 0,   // one sweep only
  MID_BHAT = 1,   // Bluestein spectrum product: v = conj(v * bhat[n]), then sweep again
};

// What a line pass does when it stores a line.
enum StoreMode : int32_t {
  STORE_PLAIN = 0,  // y = v
  STORE_BHAT = 1,   // Bluestein spectrum product: y = conj(v * bhat[n])
  STORE_CHIRP = 2,  // Bluestein end: y = conj(chirp[n] * v) for n < chirp_n only
  // Rader end: y[0] = x0 + v0 (v0 = the first sweep's DC term, side slot),
  // y[k] = x0 + conj(v[m]) with m = (L - rtab[k]) mod L, L = chirp_n - 1
  STORE_RADER = 3,
};

// LDS side slots of a Rader line: the two complex entries after the line's
// padded positions (the planner adds them to the pitch).
constexpr int kRaderSide = 2;

// One batched pass of in-LDS line FFTs.  A "line" is L elements of a
// transform; every line is independently transformed (forward sign), with
// optional fused pre/post operations.  Lines are indexed (i0, i1, i2, i3)
// with i0 fastest (extents n[0..2], i3 takes the rest); element t of a line
// is read from
//   buf[in_buf] + in_off + sum_d i_d*a[d] + (in_tab ? in_tab[i3] : 0) + t*ae
// and output element k is written to
//   buf[out_buf] + out_off + sum_d i_d*b[d] + (out_tab ? out_tab[i3] : 0) + k*be.
// Buffers a pass reads or writes, chosen at launch time (so that a schedule
// uploaded once can run on any arrays).
enum BufId : int32_t { BUF_IN = 0, BUF_OUT = 1, BUF_SCRATCH = 2 };

struct LinePass {
  int32_t in_buf, out_buf;   // BufId
  int64_t in_off, out_off;   // element offsets into those buffers
  int32_t L;
  int32_t nst;
  int32_t rad[kMaxStages];
  FastDiv nsdiv[kMaxStages];      // by Ns (product of earlier radices) per stage
  int32_t C;    // lines per workgroup
  int32_t Tl;   // threads per line (C*Tl <= block, C <= kMaxLines)
  int32_t ldl;  // LDS row pitch in elements (>= padded L)
  int32_t lds_sh;  // LDS padding: element i of a line sits at i + (i >> lds_sh) (0: none)
  int32_t col_in, col_out;  // 1: consecutive lines are adjacent in memory (coalesce across lines)
  int32_t fuse_in, fuse_out; // 1: first stage loads from / last stage stores to global memory directly
  // 1: the load / spectrum / store phase of a row-layout side runs stepped,
  // one line per block/C threads (each access of a wave covers >= 128
  // contiguous bytes of the line, or the whole line)
  int32_t row_in, row_mid, row_out;
  // 1: the spectrum product is applied by the second sweep's first radix
  // stage as it reads its inputs (no separate spectrum phase)
  int32_t mid_in_stage;
  // 1: the first sweep's last stage, the spectrum product and the second
  // sweep's first stage run as one stage (first and last radix equal)
  int32_t mid_fused;
  // 1: every line of a tile lies within one wavefront during the sweeps
  // (threads per line divide 64), so their stages synchronise per wave
  int32_t wave_sync;
  // zero-padded convolution lines (one-kernel Bluestein): bit 0, the line is
  // zero at positions >= zero_from, where zero_from = ((r + 1) / 2) * (L / r)
  // for the first radix r, and the first sweep's first stage skips those
  // inputs (the load phase writes positions < zero_from only); bit 1, the
  // store reads positions < ((r + 1) / 2) * (L / r) only, r the last radix,
  // and the second sweep's last stage computes and writes no others
  int32_t zero_pad;
  int32_t zero_from;
  FastDiv divC, divL;       // by C and by L
  FastDiv divR;             // by block / C
  int64_t nlines;           // < 2^31
  int64_t n[3];
  FastDiv divn[3];          // by n[0], n[1], n[2]
  int64_t a[4], b[4];
  int64_t ae, be;
  const int64_t* in_tab;
  const int64_t* out_tab;
  const void* tw;  // per-stage twiddles of this radix schedule (kernels.hip, stage()), pass precision
  // Pass twiddle: element t of line index k = i_twdim is multiplied by
  // w_N^{k * t * tw_scale}, on load (pretw = 1, multi-pass Stockham) or on
  // store (pretw = 2, four-step column/row passes).  The exponent is split
  // so that neighbouring lanes read neighbouring table entries:
  //   tw_split 0 (lanes run over lines):     k*t = (k & ~15)*t + (k & 15)*t
  //   tw_split 1 (lanes run along the line): k*t = k*(t & ~15) + k*(t & 15)
  // The first part is w_N^m = lo[m & (2^sh-1)] * hi[m >> sh] (FP64 tables,
  // product in FP64), the same for 16 neighbouring lanes; the second is
  // tw16[a*16 + j] = w_N^{a*j*tw_scale} (a = t or k, j < 16; FP64-generated,
  // stored in the pass precision).  Their product is taken in the pass
  // precision.
  int32_t pretw;
  int32_t twdim;
  int32_t tw_sh;
  int32_t tw_split;
  int64_t tw_scale;
  const double2* tw_lo;
  const double2* tw_hi;
  const void* tw16;
  int32_t conj_in, conj_out;
  double scale;  // multiply on store when != 1
  int32_t load_mode, mid_mode, store_mode;
  int32_t xcd_remap;  // host hint: single-pass launch may use the XCD-aware tile order
  // Chirp / spectrum index of a loaded element: n = sum_d i_d*ca[d] + t*cae;
  // of a stored element: n = sum_d i_d*cb[d] + k*cbe.
  int64_t chirp_n;
  int64_t ca[4], cb[4], cm[4];
  int64_t cae, cbe, cme;  // spectrum index in the middle of a pass: n = sum_d i_d*cm[d] + k*cme
  // > 0: lines of dimension 0 are the rows k1 < mirror_rows of a row-permuted
  // spectrum, of which only rows k1 <= mirror_rows / 2 are stored; a row
  // k1 above that reads row mirror_rows - k1 backwards (B[M - k] = B[k])
  int32_t mirror_rows;
  const void* chirp;  // chirp_n entries, pass precision
  const void* bhat;   // spectrum factor (mid: L entries; store: indexed by n)
  // Register passes: the chirp of a line's positions n = a + t*s is formed as
  // c[a] * w^t * d[t] (w = exp(2 pi i a s / N)), with d[t] =
  // exp(i pi (t^2 s^2 mod 2N) / N) from this FP64 table (L entries; null:
  // read the chirp table instead).
  const double2* chirp_d;
  double chirp_inv;  // 1 / chirp_n (the register passes' index reductions)
  // Rader passes: LDS slot of each element t < chirp_n of the prime-length
  // line (rtab[0] unused)
  const uint16_t* rtab;
};

// Radix schedule of an in-LDS line (for table generation).
struct RadixList {
  int32_t nst;
  int32_t rad[kMaxStages];
};

// Device pointers of the buffers a launch works on.
struct BufPtrs {
  const void* in;
  void* out;
  void* scratch;
};

// Element space of the real-transform boundary kernels: lines along one
// array dimension, indexed (p, t, q) with p fastest.  p runs over the
// dimensions before the transformed one (stride 1 on both sides), t along
// the line, q over the dimensions after it.  Offsets are in elements of
// each side's own type (real or complex).
struct RealGeom {
  int64_t P, T, Q;          // extents
  int64_t in_t, in_q;       // input strides of t and q
  int64_t out_t, out_q;     // output strides of t and q
  FastDiv divP, divT;
};

}  // namespace fftahip
