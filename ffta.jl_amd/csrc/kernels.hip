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

namespace fftahip {

// ------------------------------------------------------------------ helpers

template <class R>
__device__ __forceinline__ cx<R> from_d2(double2 w) {
  return {(R)w.x, (R)w.y};
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

// ------------------------------------------------------------------ stages

template <class Rt, int RAD, int EPT>
__device__ __forceinline__ void stage(cx<Rt>* line, int jl, bool active, int Tl, int L, int Ns, FastDiv nsd,
                                      int twst, const cx<Rt>* __restrict__ tw) {
  constexpr int NB = EPT / RAD;
  const int Lr = L / RAD;
  cx<Rt> v[NB][RAD];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = jl + b * Tl;
    if (active && j < Lr) {
#pragma unroll
      for (int t = 0; t < RAD; ++t) v[b][t] = line[j + t * Lr];
      if (Ns > 1) {
        const uint32_t k = (uint32_t)j - fdiv((uint32_t)j, nsd) * (uint32_t)Ns;
        const uint32_t step = k * (uint32_t)twst;
#pragma unroll
        for (int t = 1; t < RAD; ++t) v[b][t] = cmul(v[b][t], tw[t * step]);
      }
      dft_any<Rt, RAD>(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = jl + b * Tl;
    if (active && j < Lr) {
      const uint32_t q = fdiv((uint32_t)j, nsd);
      const uint32_t k = (uint32_t)j - q * (uint32_t)Ns;
      const int base = (int)(q * (uint32_t)Ns * RAD + k);
#pragma unroll
      for (int t = 0; t < RAD; ++t) line[base + t * Ns] = v[b][t];
    }
  }
  __syncthreads();
}

template <class Rt, int EPT>
__device__ __forceinline__ void sweep(cx<Rt>* line, int jl, bool active, const LinePass& p,
                                      const cx<Rt>* __restrict__ tw) {
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int r = p.rad[s];
    const FastDiv nsd = p.nsdiv[s];
    const int twst = p.twst[s];
    switch (r) {
      case 2: stage<Rt, 2, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 3: stage<Rt, 3, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 4: stage<Rt, 4, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 5: stage<Rt, 5, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 7: stage<Rt, 7, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 8: stage<Rt, 8, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 9: stage<Rt, 9, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 11: stage<Rt, 11, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 13: stage<Rt, 13, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      case 16: stage<Rt, 16, EPT>(line, jl, active, p.Tl, p.L, Ns, nsd, twst, tw); break;
      default: break;
    }
    Ns *= r;
  }
}

// ------------------------------------------------------------------ line pass

struct LineShared {
  int64_t in[kMaxLines];
  int64_t out[kMaxLines];
  int64_t twk[kMaxLines];
  int64_t cin[kMaxLines];
  int64_t cout[kMaxLines];
};
static_assert(sizeof(LineShared) == kLineSharedBytes, "static LDS size");

// Global accesses are issued in batches of kU per thread so that each thread
// keeps kU loads in flight before it consumes any of them.
constexpr int kU = 8;

template <class Rt, int EPT, int BLOCK>
__device__ __forceinline__ void run_line_pass(const LinePass& p, uint32_t tile, cx<Rt>* lds, LineShared& sh) {
  const int tid = threadIdx.x;
  const int C = p.C;
  const int L = p.L;
  const int ldl = p.ldl;
  const int64_t line0 = (int64_t)tile * C;
  const int64_t rem = p.nlines - line0;
  const int nvalid = rem < C ? (int)rem : C;

  for (int c = tid; c < nvalid; c += BLOCK) {
    const uint32_t line = (uint32_t)(line0 + c);
    int64_t idx[4];
    uint32_t q = fdiv(line, p.divn[0]);
    idx[0] = line - q * p.divn[0].d;
    uint32_t q1 = fdiv(q, p.divn[1]);
    idx[1] = q - q1 * p.divn[1].d;
    uint32_t q2 = fdiv(q1, p.divn[2]);
    idx[2] = q1 - q2 * p.divn[2].d;
    idx[3] = q2;
    int64_t oi = idx[0] * p.a[0] + idx[1] * p.a[1] + idx[2] * p.a[2] + idx[3] * p.a[3];
    int64_t oo = idx[0] * p.b[0] + idx[1] * p.b[1] + idx[2] * p.b[2] + idx[3] * p.b[3];
    if (p.in_tab) oi += p.in_tab[idx[3]];
    if (p.out_tab) oo += p.out_tab[idx[3]];
    sh.in[c] = oi;
    sh.out[c] = oo;
    sh.twk[c] = idx[p.twdim & 3];
    sh.cin[c] = idx[0] * p.ca[0] + idx[1] * p.ca[1] + idx[2] * p.ca[2] + idx[3] * p.ca[3];
    sh.cout[c] = idx[0] * p.cb[0] + idx[1] * p.cb[1] + idx[2] * p.cb[2] + idx[3] * p.cb[3];
  }
  __syncthreads();

  const cx<Rt>* __restrict__ in = reinterpret_cast<const cx<Rt>*>(p.in);
  cx<Rt>* __restrict__ out = reinterpret_cast<cx<Rt>*>(p.out);
  const cx<Rt>* __restrict__ tw = reinterpret_cast<const cx<Rt>*>(p.tw);
  const cx<Rt>* __restrict__ chirp = reinterpret_cast<const cx<Rt>*>(p.chirp);
  const cx<Rt>* __restrict__ bhat = reinterpret_cast<const cx<Rt>*>(p.bhat);
  const int total = C * L;
  const bool chirp_in = p.load_mode == LOAD_CHIRP;
  const int64_t twmask = (1ll << p.tw_sh) - 1;

  // ---- load (global -> LDS), coalesced across lines or along a line
  for (int e0 = 0; e0 < total; e0 += BLOCK * kU) {
    cx<Rt> v[kU];
    cx<Rt> ch[kU];
    double2 wlo[kU], whi[kU];
    int slot[kU];
    bool live[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int e = e0 + u * BLOCK + tid;
      slot[u] = -1;
      live[u] = false;
      v[u] = {(Rt)0, (Rt)0};
      if (e < total) {
        int c, t;
        if (p.col_in) {
          t = (int)fdiv((uint32_t)e, p.divC);
          c = e - t * C;
        } else {
          c = (int)fdiv((uint32_t)e, p.divL);
          t = e - c * L;
        }
        if (c < nvalid) {
          slot[u] = c * ldl + t;
          live[u] = true;
          if (chirp_in) {
            const int64_t n = sh.cin[c] + (int64_t)t * p.cae;
            live[u] = n < p.chirp_n;
            if (live[u]) ch[u] = chirp[n];
          }
          if (live[u]) {
            v[u] = in[sh.in[c] + (int64_t)t * p.ae];
            if (p.pretw) {
              const int64_t m = sh.twk[c] * (int64_t)t * p.tw_scale;
              wlo[u] = p.tw_lo[m & twmask];
              whi[u] = p.tw_hi[m >> p.tw_sh];
            }
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (slot[u] < 0) continue;
      cx<Rt> x = v[u];
      if (live[u]) {
        if (p.conj_in) x = cconj(x);
        if (chirp_in) x = cmul(x, cconj(ch[u]));
        if (p.pretw) {
          const double2 w = make_double2(wlo[u].x * whi[u].x - wlo[u].y * whi[u].y,
                                         wlo[u].x * whi[u].y + wlo[u].y * whi[u].x);
          x = cmul(x, from_d2<Rt>(w));
        }
      }
      lds[slot[u]] = x;
    }
  }
  __syncthreads();

  // ---- radix sweeps in LDS
  const int cl = tid / p.Tl;
  const int jl = tid - cl * p.Tl;
  const bool active = (cl < nvalid) && (tid < C * p.Tl);
  cx<Rt>* myline = lds + (active ? cl : 0) * ldl;
  sweep<Rt, EPT>(myline, jl, active, p, tw);

  if (p.mid_mode == MID_BHAT) {
    for (int e0 = 0; e0 < total; e0 += BLOCK * kU) {
      cx<Rt> f[kU];
      int slot[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int e = e0 + u * BLOCK + tid;
        slot[u] = -1;
        if (e < total) {
          const int c = (int)fdiv((uint32_t)e, p.divL);
          const int t = e - c * L;
          if (c < nvalid) {
            slot[u] = c * ldl + t;
            f[u] = bhat[t];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (slot[u] >= 0) lds[slot[u]] = cconj(cmul(lds[slot[u]], f[u]));
    }
    __syncthreads();
    sweep<Rt, EPT>(myline, jl, active, p, tw);
  }

  // ---- store (LDS -> global)
  const Rt scale = (Rt)p.scale;
  const bool has_scale = p.scale != 1.0;
  for (int e0 = 0; e0 < total; e0 += BLOCK * kU) {
    cx<Rt> f[kU];
    int slot[kU];
    int64_t dst[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int e = e0 + u * BLOCK + tid;
      slot[u] = -1;
      if (e < total) {
        int c, t;
        if (p.col_out) {
          t = (int)fdiv((uint32_t)e, p.divC);
          c = e - t * C;
        } else {
          c = (int)fdiv((uint32_t)e, p.divL);
          t = e - c * L;
        }
        if (c < nvalid) {
          bool keep = true;
          if (p.store_mode != STORE_PLAIN) {
            const int64_t n = sh.cout[c] + (int64_t)t * p.cbe;
            if (p.store_mode == STORE_BHAT) {
              f[u] = bhat[n];
            } else {
              keep = n < p.chirp_n;
              if (keep) f[u] = chirp[n];
            }
          }
          if (keep) {
            slot[u] = c * ldl + t;
            dst[u] = sh.out[c] + (int64_t)t * p.be;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (slot[u] < 0) continue;
      cx<Rt> x = lds[slot[u]];
      if (p.store_mode == STORE_BHAT) x = cconj(cmul(x, f[u]));
      else if (p.store_mode == STORE_CHIRP) x = cconj(cmul(f[u], x));
      if (p.conj_out) x = cconj(x);
      if (has_scale) x = cscale(x, scale);
      out[dst[u]] = x;
    }
  }
}

template <class Rt, int EPT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) line_pass_kernel(LinePass p) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ LineShared sh;
  uint32_t tile = blockIdx.x;
  if (p.xcd_remap) {
    const uint32_t per = gridDim.x >> 3;
    tile = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  }
  run_line_pass<Rt, EPT, BLOCK>(p, tile, reinterpret_cast<cx<Rt>*>(smem_raw), sh);
}

// Work-list launch: group g owns tiles [tile_start[g], tile_start[g+1]).
// Every group is an independent LinePass of the same precision and block
// size; groups may differ in length, radices and fused operations.
template <class Rt, int EPT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) line_pass_group_kernel(const LinePass* __restrict__ passes,
                                                                const uint32_t* __restrict__ tile_start,
                                                                int ngroups) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ LineShared sh;
  const uint32_t bid = blockIdx.x;
  int lo = 0, hi = ngroups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= bid) lo = mid;
    else hi = mid - 1;
  }
  const LinePass& p = passes[lo];
  run_line_pass<Rt, EPT, BLOCK>(p, bid - tile_start[lo], reinterpret_cast<cx<Rt>*>(smem_raw), sh);
}

// ------------------------------------------------------------------ tables

// tw[m] = exp(-2 pi i m / L), m < L
template <class Rt>
__global__ void gen_stage_twiddles_kernel(cx<Rt>* tw, int64_t L) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < L) tw[m] = from_d2<Rt>(unit_root(m, L));
}

// lo[m] = w_N^m for m < 2^sh; hi[h] = w_N^{h * 2^sh} for h < nhi
__global__ void gen_twiddle_2level_kernel(double2* lo, double2* hi, int64_t N, int sh, int64_t nhi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nlo = 1ll << sh;
  if (i < nlo) lo[i] = unit_root(i, N);
  if (i < nhi) hi[i] = unit_root(i << sh, N);
}

// Forward Bluestein chirp b_n = exp(+i pi p / N), p = n^2 mod 2N taken in
// (-N, N] (src/algos.jl:353-359 with sgn = +1), FP64 evaluated.
template <class Rt>
__global__ void gen_chirp_kernel(cx<Rt>* chirp, double2* chirp64, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t twoN = 2 * N;
  int64_t p = (int64_t)(((uint64_t)n * (uint64_t)n) % (uint64_t)twoN);  // exact: N < 2^31
  if (p > N) p -= twoN;
  double s, c;
  sincospi((double)p / (double)N, &s, &c);
  chirp[n] = {(Rt)c, (Rt)s};
  if (chirp64) chirp64[n] = make_double2(c, s);
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

// ------------------------------------------------------------------ launchers

namespace {
template <class Rt, int EPT, int BLOCK>
hipError_t launch_line_pass_t(const LinePass& p, size_t lds_bytes, int64_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((line_pass_kernel<Rt, EPT, BLOCK>), dim3((uint32_t)grid), dim3(BLOCK), lds_bytes, stream, p);
  return hipGetLastError();
}
template <class Rt, int EPT, int BLOCK>
hipError_t launch_group_t(const LinePass* passes, const uint32_t* tile_start, int ngroups, size_t lds_bytes,
                          int64_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((line_pass_group_kernel<Rt, EPT, BLOCK>), dim3((uint32_t)grid), dim3(BLOCK), lds_bytes, stream,
                     passes, tile_start, ngroups);
  return hipGetLastError();
}
}  // namespace

template <class Rt, int EPT, int BLOCK>
void allow_dynamic_lds(size_t bytes) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&line_pass_kernel<Rt, EPT, BLOCK>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&line_pass_group_kernel<Rt, EPT, BLOCK>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)hipGetLastError();
}

void init_kernels(size_t lds_per_block) {
  const size_t dyn = lds_per_block > (size_t)kLineSharedBytes ? lds_per_block - kLineSharedBytes : 0;
  allow_dynamic_lds<float, 16, 256>(dyn);
  allow_dynamic_lds<float, 16, 512>(dyn);
  allow_dynamic_lds<float, 16, 1024>(dyn);
  allow_dynamic_lds<double, 16, 256>(dyn);
  allow_dynamic_lds<double, 16, 512>(dyn);
  allow_dynamic_lds<double, 16, 1024>(dyn);
}

hipError_t launch_line_pass(const LinePass& p, int precision, int block, size_t lds_bytes, int64_t grid,
                            hipStream_t stream) {
  if (precision == 32) {
    if (block == 256) return launch_line_pass_t<float, 16, 256>(p, lds_bytes, grid, stream);
    if (block == 512) return launch_line_pass_t<float, 16, 512>(p, lds_bytes, grid, stream);
    if (block == 1024) return launch_line_pass_t<float, 16, 1024>(p, lds_bytes, grid, stream);
  } else {
    if (block == 256) return launch_line_pass_t<double, 16, 256>(p, lds_bytes, grid, stream);
    if (block == 512) return launch_line_pass_t<double, 16, 512>(p, lds_bytes, grid, stream);
    if (block == 1024) return launch_line_pass_t<double, 16, 1024>(p, lds_bytes, grid, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_line_pass_groups(const LinePass* d_passes, const uint32_t* d_tile_start, int ngroups,
                                   int precision, int block, size_t lds_bytes, int64_t grid, hipStream_t stream) {
  if (precision == 32) {
    if (block == 256) return launch_group_t<float, 16, 256>(d_passes, d_tile_start, ngroups, lds_bytes, grid, stream);
    if (block == 512) return launch_group_t<float, 16, 512>(d_passes, d_tile_start, ngroups, lds_bytes, grid, stream);
    if (block == 1024)
      return launch_group_t<float, 16, 1024>(d_passes, d_tile_start, ngroups, lds_bytes, grid, stream);
  } else {
    if (block == 256)
      return launch_group_t<double, 16, 256>(d_passes, d_tile_start, ngroups, lds_bytes, grid, stream);
    if (block == 512)
      return launch_group_t<double, 16, 512>(d_passes, d_tile_start, ngroups, lds_bytes, grid, stream);
    if (block == 1024)
      return launch_group_t<double, 16, 1024>(d_passes, d_tile_start, ngroups, lds_bytes, grid, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_gen_stage_twiddles(void* tw, int precision, int64_t L, hipStream_t stream) {
  const int64_t blocks = (L + 255) / 256;
  if (precision == 32)
    hipLaunchKernelGGL(gen_stage_twiddles_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<float>*>(tw), L);
  else
    hipLaunchKernelGGL(gen_stage_twiddles_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<cx<double>*>(tw), L);
  return hipGetLastError();
}

hipError_t launch_gen_twiddle_2level(double2* lo, double2* hi, int64_t N, int sh, int64_t nhi, hipStream_t stream) {
  const int64_t n = (1ll << sh) > nhi ? (1ll << sh) : nhi;
  hipLaunchKernelGGL(gen_twiddle_2level_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, lo, hi, N,
                     sh, nhi);
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

hipError_t launch_gen_chirp_ext(double2* bext, const double2* chirp64, int64_t N, int64_t M, hipStream_t stream) {
  hipLaunchKernelGGL(gen_chirp_ext_kernel, dim3((uint32_t)((M + 255) / 256)), dim3(256), 0, stream, bext, chirp64,
                     N, M);
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
