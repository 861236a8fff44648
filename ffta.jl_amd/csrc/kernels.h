// kernels.h -- host-side launchers for the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.h"

namespace fftahip {

void init_kernels(size_t lds_per_block);

// A line-pass kernel variant: its host entry (for hipLaunchKernel and graph
// kernel nodes) and workgroup size.
struct KernelRef {
  const void* fn;
  int block;
};
KernelRef lds_kernel(int precision, int block, int variant);
KernelRef reg_kernel(int precision, int mode);
// Kernel arguments of a work-list launch (passes, tile starts, group count,
// buffers, tile count), laid out for hipLaunchKernel / kernel nodes.
struct LaunchArgs {
  const LinePass* passes;
  const uint32_t* tiles;
  int ngroups;
  BufPtrs bp;
  uint32_t ntiles;
  void* args[5];
  LaunchArgs(const LinePass* P, const uint32_t* T, int ng, const BufPtrs& b, int64_t grid);
  LaunchArgs(const LaunchArgs&) = delete;
  LaunchArgs& operator=(const LaunchArgs&) = delete;
};
// Workgroups a launch of `grid` tiles takes (a multiple of 8).
uint32_t padded_grid(int64_t grid);
// Per-variant entries (kern_lds.hip / kern_reg.hip translation units).
const void* lds_fn_f32_256(int mode);
const void* lds_fn_f32_512(int mode);
const void* lds_fn_f32_1024(int mode);
const void* lds_fn_f64_256(int mode);
const void* lds_fn_f64_512(int mode);
const void* lds_fn_f64_1024(int mode);
const void* reg_fn_f32(int mode);
const void* reg_fn_f64(int mode);
// Kernel mode a pass needs (0 plain, 1 load-side pass twiddle, 2 full).
int pass_mode(const LinePass& p);
// In-LDS kernel variant of a pass: its mode, plus kBigVariant when a stage
// of its radix schedule is a prime >= 17 (the direct-DFT stage).
constexpr int kBigVariant = 3;
constexpr int kLdsVariants = 6;
int pass_variant(const LinePass& p);
// Launch ngroups independent passes (device descriptors) as one grid.
// `variant`: pass_variant of every group of the launch.  d_tile_start: with
// ngroups > 1, (group, tile within the group) for each of the grid's tiles;
// unused with one group.
hipError_t launch_line_pass_groups(const LinePass* d_passes, const uint32_t* d_tile_start, int ngroups, int variant,
                                   int precision, int block, const BufPtrs& bp, int remap, size_t lds_bytes,
                                   int64_t grid, hipStream_t stream);
// Register line passes (one short line per thread; lengths are the 13-smooth
// L <= reg_pass_max(precision, mode)); tiles of kRegBlock lines.
int reg_pass_max(int precision, int mode);
hipError_t launch_reg_pass_groups(const LinePass* d_passes, const uint32_t* d_tile_start, int ngroups, int mode,
                                  int precision, const BufPtrs& bp, int64_t grid, hipStream_t stream);
// n = table entries: L - 1 stage twiddles, then P roots per stage of a prime radix P >= 17.
hipError_t launch_gen_stage_twiddles(void* tw, int precision, const RadixList& rl, int64_t n, hipStream_t stream);
hipError_t launch_gen_twiddle_2level(double2* lo, double2* hi, int64_t N, int sh, int64_t nhi, hipStream_t stream);
hipError_t launch_gen_tw16(void* out, int precision, int64_t N, int64_t s, int64_t amax, hipStream_t stream);
// Real-transform boundary kernels (see kernels.hip); lo/hi/sh are the
// two-level FP64 twiddle tables of n for the split steps.
enum RealOp { REAL_PACK = 0, REAL_UNPACK, REAL_COMPLEXIFY, REAL_PART, REAL_COPY, REAL_HERMITIAN, REAL_RFFT_SPLIT,
              REAL_BRFFT_SPLIT };
hipError_t launch_real_boundary(int op, int precision, void* dst, const void* src, const RealGeom& g, int64_t m,
                                const double2* lo, const double2* hi, int sh, int64_t n, hipStream_t s);
hipError_t launch_gen_chirp(void* chirp, double2* chirp64, int precision, int64_t N, hipStream_t stream);
// Chirp steps d[t] = exp(i pi (t^2 s^2 mod 2N) / N), t < L (FP64), of a register pass.
hipError_t launch_gen_chirp_steps(double2* d, int64_t s, int64_t N, int L, hipStream_t stream);
hipError_t launch_gen_chirp_range(void* out, int precision, int64_t n0, int64_t count, int64_t N,
                                  hipStream_t stream);
hipError_t launch_gen_chirp_ext(double2* bext, const double2* chirp64, int64_t N, int64_t M, hipStream_t stream);
// Rader's convolution kernel b_r = exp(-2 pi i g^r / P), r < P - 1, FP64.
hipError_t launch_gen_rader_b(double2* b, int64_t P, int64_t g, hipStream_t stream);
hipError_t launch_permute_rows(void* dst, const void* src, int precision, int64_t M1, int64_t M2, int64_t rows,
                               hipStream_t stream);
hipError_t launch_finish_bhat(void* bhat, const double2* spec64, int precision, int64_t M, hipStream_t stream);

}  // namespace fftahip
