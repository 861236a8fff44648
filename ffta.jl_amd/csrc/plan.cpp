// plan.cpp -- host planner and C ABI of the transform library.
//
// The reference's planner (FFTA.jl src/callgraph.jl:76-126, CallGraphNode!)
// classifies a length as radix-4 pow2 / pow3 / DFT / Bluestein / composite
// and recurses.  Here the same classification is kept (it is what
// fftahip_plan_describe reports and what the Julia wrapper exposes), and a
// device schedule is emitted beside it:
//   * lengths whose prime factors are all <= 13 run as self-sorting Stockham
//     passes: one in-LDS pass when the whole line fits the workgroup, else a
//     global mixed-radix Stockham over "pass factors" (each pass is a batch
//     of in-LDS line FFTs with the inter-pass twiddle fused into its load);
//   * any other length runs as Bluestein's convolution (src/algos.jl:344-424)
//     over a power-of-two length M >= 2N-1: one kernel when M fits in LDS
//     (chirp-multiply, FFT, spectrum product, FFT, chirp-multiply), else two
//     multi-pass sweeps with the chirp fused into the first load, the
//     spectrum product into the last store of the first sweep, and the final
//     chirp into the last store.  The spectrum FFT(b)/M is computed once per
//     plan in FP64 (FFTA recomputes it every call, src/algos.jl:409).
// All tables (stage twiddles, two-level pass twiddles, chirp, spectrum) are
// generated on the device in FP64 with exact integer angle reduction and
// rounded to the plan's precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fftahip.h"
#include "common.h"
#include "kernels.h"

namespace fftahip {

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static fftahip_status fail(fftahip_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(FFTAHIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ device

struct DevInfo {
  int64_t lds = 65536;
  int64_t cus = 256;
  int64_t warp = 64;
  bool ok = false;
};

static DevInfo& dev_info() {
  static DevInfo info;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return;
    info.lds = (int64_t)prop.sharedMemPerBlock;
    info.cus = prop.multiProcessorCount;
    info.warp = prop.warpSize;
    info.ok = true;
    init_kernels((size_t)info.lds);
  });
  return info;
}

// Dry-run mode (schedule inspection without a device): device pointers are
// replaced by tokens (1<<60 | kind<<52 | offset) and nothing is launched.
static thread_local bool g_dry = false;
static thread_local int64_t g_dry_lds = 0;
enum TokenKind { TK_STAGE = 1, TK_TWLO = 2, TK_TWHI = 3, TK_CHIRP = 4, TK_BHAT = 5, TK_IN = 6, TK_OUT = 7,
                 TK_SCRATCH = 8, TK_BHATROWS = 9, TK_OFFTAB = 10, TK_TW16 = 11, TK_CHIRPD = 12, TK_RTAB = 13 };
static void* token(int kind, int64_t v) {
  return (void*)(uintptr_t)((1ull << 60) | ((uint64_t)kind << 52) | (uint64_t)v);
}

static int64_t env_i64(const char* name, int64_t dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::strtoll(v, nullptr, 10);
}

// ------------------------------------------------------------------ FFTA call graph
// Restatement of src/callgraph.jl:76-126 for reporting (node list as FFTA
// builds it: pre-order, relative child offsets, baked strides).

enum FftaType { F_COMPOSITE = 0, F_DFT = 1, F_POW3 = 2, F_POW2RADIX4 = 3, F_BLUESTEIN = 4 };
static const char* ffta_type_name(int t) {
  switch (t) {
    case F_COMPOSITE: return "COMPOSITE_FFT";
    case F_DFT: return "DFT";
    case F_POW3: return "POW3_FFT";
    case F_POW2RADIX4: return "POW2RADIX4_FFT";
    default: return "BLUESTEIN";
  }
}

struct FftaNode {
  int64_t left, right;
  int type;
  int64_t sz, s_in, s_out;
};

static std::vector<int64_t> prime_factors(int64_t n) {
  std::vector<int64_t> f;
  for (int64_t p = 2; p * p <= n; ++p)
    while (n % p == 0) { f.push_back(p); n /= p; }
  if (n > 1) f.push_back(n);
  return f;
}

static bool is_prime(int64_t n) {
  if (n < 2) return false;
  for (int64_t p = 2; p * p <= n; ++p)
    if (n % p == 0) return false;
  return true;
}

static int64_t ffta_nodes(std::vector<FftaNode>& nodes, int64_t N, int64_t cutoff, int64_t s_in, int64_t s_out) {
  if (N >= 2 && (N & (N - 1)) == 0) {
    nodes.push_back({0, 0, F_POW2RADIX4, N, s_in, s_out});
    return 1;
  }
  {
    int64_t m = N;
    bool pow3 = N >= 3 && N % 3 == 0;
    while (pow3 && m % 3 == 0) m /= 3;
    if (pow3 && m == 1) {
      nodes.push_back({0, 0, F_POW3, N, s_in, s_out});
      return 1;
    }
  }
  if (N == 1 || is_prime(N)) {
    nodes.push_back({0, 0, N < cutoff ? F_DFT : F_BLUESTEIN, N, s_in, s_out});
    return 1;
  }
  std::vector<int64_t> Ns = prime_factors(N);
  int64_t N1 = 1;
  if (Ns[0] == 2 || Ns[0] == 3) {
    for (int64_t p : Ns)
      if (p == Ns[0]) N1 *= p;
  } else {
    const double fs = std::sqrt((double)N);
    int64_t is = (int64_t)fs;
    while (is * is > N) --is;
    while ((is + 1) * (is + 1) <= N) ++is;
    std::vector<int64_t> cp;
    int64_t acc = 1;
    for (int64_t p : Ns) { acc *= p; cp.push_back(acc); }
    int64_t idx = 0;
    for (size_t i = 0; i < cp.size(); ++i)
      if (cp[i] <= is) idx = (int64_t)i;
    N1 = cp[idx];
    if (idx != (int64_t)cp.size() - 1 && ((double)cp[idx + 1] - fs < fs - (double)N1)) N1 = cp[idx + 1];
  }
  const int64_t N2 = N / N1;
  const size_t at = nodes.size();
  nodes.push_back({0, 0, F_DFT, N, s_in, s_out});
  const int64_t l = ffta_nodes(nodes, N1, cutoff, N2, N2 * s_out);
  const int64_t r = ffta_nodes(nodes, N2, cutoff, N1 * s_in, 1);
  nodes[at] = {1, 1 + l, F_COMPOSITE, N, s_in, s_out};
  return 1 + l + r;
}

// ------------------------------------------------------------------ device memory

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  fftahip_status ensure(size_t nbytes) {
    if (nbytes <= bytes) return FFTAHIP_SUCCESS;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (nbytes == 0) return FFTAHIP_SUCCESS;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(FFTAHIP_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(nbytes) + "): " + hipGetErrorString(e));
    }
    bytes = nbytes;
    return FFTAHIP_SUCCESS;
  }
};

// ------------------------------------------------------------------ in-LDS line FFT

constexpr int kEPT = 16;

struct Inner {
  int64_t L = 1;
  std::vector<int> rad;
  int Tl = 1;
  int block = 256;
  // strided tiles of this pass (0: the default rule): LDS budget per tile in
  // bytes and the fewest adjacent lines a tile keeps
  int64_t tile_budget = 0, tile_cmin = 0;
};

// Radices of the in-register codelets (codelets.h): 2..16 in every kernel
// variant, the primes 17, 19, 23 in the FP32 direct-DFT variants; any other
// radix is a prime P >= 17 run by the direct-DFT stage (stage_big in
// device.h).  prec 0 (a length checked for any precision) counts as FP64.
static bool small_radix(int r) {
  switch (r) {
    case 2: case 3: case 4: case 5: case 7: case 8: case 9: case 11: case 13: case 16: return true;
    default: return false;
  }
}
static bool codelet_radix(int r, int prec) { return small_radix(r) || prime_codelet(r, prec); }

// Largest prime the direct-DFT stage takes (larger primes go through a
// convolution whatever BLUESTEIN_CUTOFF says), and the most roots one
// schedule's direct-DFT stages may stage in LDS (their sum of primes; 760
// FP64 entries fit the LDS headroom kept beside the largest line).
constexpr int64_t kMaxDirectPrime = 751;
constexpr int64_t kMaxDirectRoots = 760;
// Direct-DFT stage geometry (device.h): work items of kBigKP output pairs,
// kBigHold per thread.
constexpr int64_t kBigKPHost = 4, kBigHoldHost = 2;
static int64_t lds_budget();
static size_t csize(int prec);

// In-LDS line schedule with the given radices (in stage order): threads per
// line Tl = max over stages of ceil((L / R) / (16 / R)) for codelet radices,
// and for a direct-DFT stage of prime P the work items over kBigHold:
// ceil((L / P) * ceil(((P - 1) / 2) / kBigKP) / kBigHold).
static bool inner_from(int64_t L, const std::vector<int>& rad, Inner& in, int prec) {
  if ((int)rad.size() > kMaxStages) return false;
  int64_t prod = 1, Tl = 1;
  for (int r : rad) {
    if (codelet_radix(r, prec)) {
      const int64_t nb = std::max<int64_t>(1, kEPT / r);
      Tl = std::max(Tl, ((L / r) + nb - 1) / nb);
    } else {
      if (r < 17 || r > kMaxDirectPrime || !is_prime(r)) return false;
      const int64_t groups = ((r - 1) / 2 + kBigKPHost - 1) / kBigKPHost;
      Tl = std::max(Tl, ((L / r) * groups + kBigHoldHost - 1) / kBigHoldHost);
    }
    prod *= r;
  }
  int64_t roots = 0;
  for (int r : rad)
    if (!codelet_radix(r, prec)) roots += r;
  if (prod != L || Tl > 1024 || roots > kMaxDirectRoots) return false;
  // one unpadded line and its staged roots (kBigKP copies per direct-DFT
  // stage) must fit one workgroup's LDS; the pass then takes a padding that
  // still fits (set_inner).  prec 0, the planner's precision-free probes:
  // sized as FP64, so that a probe never admits a schedule the pass's own
  // precision rejects.
  const int64_t esz_fit = (int64_t)csize(prec == 32 ? 32 : 64);
  if (roots > 0 && (L + 1 + kBigKPHost * roots) * esz_fit > lds_budget()) return false;
  in = Inner();
  in.L = L;
  in.rad = rad;
  in.Tl = (int)Tl;
  in.block = Tl <= 256 ? 256 : (Tl <= 512 ? 512 : 1024);
  return true;
}

// Radix schedules chosen by the autotuner, per (precision, line length).
static std::mutex g_tune_mu;
static std::map<std::pair<int, int64_t>, std::vector<int>> g_tuned;

// Strided tile shape chosen by the autotuner for the passes of a transform
// length whose lines are laid across memory (multi-pass Stockham passes and
// strided region dimensions): LDS bytes per tile and the fewest adjacent
// lines a tile keeps.  A larger budget lets a long strided line run in one
// pass with few lines per tile (e.g. 4 lines of 4096: one pass instead of
// two of 64 x 64, one workgroup per CU); the default keeps 40 KB tiles of
// >= 128-byte segments, several workgroups per CU.  Per (precision, length).
struct TileShape {
  int64_t budget = 0;  // 0: the default rule
  int64_t cmin = 0;
};
static std::map<std::pair<int, int64_t>, TileShape> g_tuned_tile;

static TileShape tuned_tile(int prec, int64_t N) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tuned_tile.find({prec, N});
  return it == g_tuned_tile.end() ? TileShape() : it->second;
}

// Primes a length may contain and still run as direct DFT stages: 2 and 3
// always (FFTA classifies powers of 2 and 3 before primes), any other prime
// p when p < BLUESTEIN_CUTOFF (src/callgraph.jl:95-100) -- codelets up to 13,
// the direct-DFT stage from 17 up to kMaxDirectPrime.  Internal lengths (the
// convolution lengths of Bluestein and Rader) use kInternalCutoff.
constexpr int64_t kInternalCutoff = 73;
static bool direct_prime(int64_t p, int64_t cutoff) {
  if (p <= 3) return true;
  return p < cutoff && p <= kMaxDirectPrime;
}

// True when every prime factor of n is a direct prime for `cutoff`.
static bool direct_smooth(int64_t n, int64_t cutoff) {
  for (int64_t p = 2; p * p <= n; ++p) {
    if (n % p) continue;
    if (!direct_prime(p, cutoff)) return false;
    while (n % p == 0) n /= p;
  }
  return n == 1 || direct_prime(n, cutoff);
}

// Radix schedule for an in-LDS line of length L whose primes are direct for
// `cutoff`: the autotuned one for (prec, L) if there is one, else the
// default rule -- primes >= 17 as direct-DFT stages, 13, 11, 7, 5 as single
// stages, powers of two in as few stages as possible with sizes balanced
// (each <= 16), 9s then a 3, largest radix first (so a direct-DFT stage
// comes first and needs no stage twiddles).
static bool make_inner(int64_t L, Inner& in, int prec = 0, int64_t cutoff = kInternalCutoff) {
  if (L < 1 || !direct_smooth(L, cutoff)) return false;
  if (prec) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find({prec, L});
    if (it != g_tuned.end() && inner_from(L, it->second, in, prec)) return true;
  }
  int64_t m = L;
  int e2 = 0, e3 = 0;
  while (m % 2 == 0) { m /= 2; ++e2; }
  while (m % 3 == 0) { m /= 3; ++e3; }
  std::vector<int> rad;
  for (int p : {13, 11, 7, 5})
    while (m % p == 0) { m /= p; rad.push_back(p); }
  for (int64_t p = 17; m > 1 && p * p <= m; p += 2)
    while (m % p == 0) { m /= p; rad.push_back((int)p); }
  if (m > 1) {
    if (m < 17) return false;
    rad.push_back((int)m);
  }
  // powers of two: as few stages as possible, sizes balanced, each <= 16
  if (e2 > 0) {
    int ns = (e2 + 3) / 4;
    int base = e2 / ns, extra = e2 % ns;
    for (int i = 0; i < ns; ++i) rad.push_back(1 << (base + (i < extra ? 1 : 0)));
  }
  while (e3 >= 2) { rad.push_back(9); e3 -= 2; }
  if (e3) rad.push_back(3);
  std::sort(rad.begin(), rad.end(), [](int a, int b) { return a > b; });
  return inner_from(L, rad, in, prec);
}

// Candidate radix schedules of the autotuner for a line of length L: the
// powers of two split into stages of <= 16 (balanced), <= 8 (balanced) or
// greedy 16s; the powers of three as 9s (+3) or as 3s; each multiset in
// descending and ascending order.
static std::vector<std::vector<int>> inner_candidates(int64_t L) {
  std::vector<std::vector<int>> out;
  int64_t m = L;
  int e2 = 0, e3 = 0;
  while (m % 2 == 0) { m /= 2; ++e2; }
  while (m % 3 == 0) { m /= 3; ++e3; }
  std::vector<int> odd;
  for (int p : {13, 11, 7, 5})
    while (m % p == 0) { m /= p; odd.push_back(p); }
  if (m != 1) return out;
  std::vector<std::vector<int>> p2s, p3s;
  auto balanced = [](int e, int cap_log) {
    std::vector<int> r;
    if (e == 0) return r;
    const int ns = (e + cap_log - 1) / cap_log;
    const int base = e / ns, extra = e % ns;
    for (int i = 0; i < ns; ++i) r.push_back(1 << (base + (i < extra ? 1 : 0)));
    return r;
  };
  p2s.push_back(balanced(e2, 4));
  p2s.push_back(balanced(e2, 3));
  {
    std::vector<int> g;
    int e = e2;
    while (e >= 4) { g.push_back(16); e -= 4; }
    if (e) g.push_back(1 << e);
    p2s.push_back(g);
  }
  {
    std::vector<int> r;
    int e = e3;
    while (e >= 2) { r.push_back(9); e -= 2; }
    if (e) r.push_back(3);
    p3s.push_back(r);
  }
  if (e3 >= 2 && e3 <= 6) p3s.push_back(std::vector<int>(e3, 3));
  std::set<std::vector<int>> seen;
  for (auto& a : p2s)
    for (auto& b : p3s) {
      std::vector<int> r = odd;
      r.insert(r.end(), a.begin(), a.end());
      r.insert(r.end(), b.begin(), b.end());
      std::sort(r.begin(), r.end(), [](int x, int y) { return x > y; });
      for (int order = 0; order < 2; ++order) {
        std::vector<int> c = r;
        if (order) std::reverse(c.begin(), c.end());
        Inner tmp;
        if (inner_from(L, c, tmp, 0) && seen.insert(c).second) out.push_back(c);
      }
    }
  return out;
}

static bool smooth13(int64_t n) {
  for (int p : {2, 3, 5, 7, 11, 13})
    while (n % p == 0) n /= p;
  return n == 1;
}

// ------------------------------------------------------------------ length plan

enum Kind { K_SINGLE = 0, K_MULTI = 1, K_BLUE_SINGLE = 2, K_BLUE_MULTI = 3, K_BLUE_3PASS = 4, K_MIXED = 5,
            K_RADER_SINGLE = 6, K_RADER_MIXED = 7 };
static const char* kind_name(int k) {
  switch (k) {
    case K_SINGLE: return "stockham-1pass";
    case K_MULTI: return "stockham-multipass";
    case K_BLUE_SINGLE: return "bluestein-1kernel";
    case K_BLUE_3PASS: return "bluestein-3pass";
    case K_MIXED: return "bluestein-line+stockham";
    case K_RADER_SINGLE: return "rader-1kernel";
    case K_RADER_MIXED: return "rader-line+stockham";
    default: return "bluestein-multipass";
  }
}

struct Decomp {
  int kind = K_SINGLE;
  int64_t M = 0;                // device transform length (N, or Bluestein's M)
  std::vector<int64_t> F;       // pass factors (product M); one entry for single-pass
  std::vector<Inner> inner;     // per pass factor
};

struct Tables;  // per-plan device tables

struct LengthPlan {
  int prec = 32;
  int64_t N = 1;
  int64_t cutoff = 73;
  std::vector<FftaNode> ffta;
  bool blue = false;  // N has a prime factor > 13
  bool mixed = false; // N = S*Q: Q through an in-LDS Bluestein (or Rader) line, S through Stockham passes
  // the convolution of the prime bq runs Rader's algorithm: a cyclic
  // convolution of M = bq - 1 points in one workgroup (instead of Bluestein's
  // M >= 2 bq - 1)
  bool rader = false;
  int64_t rg = 0;     // primitive root of bq (Rader)
  int64_t S = 1;      // 13-smooth part of N
  int64_t bq = 1;     // length that goes through Bluestein (N, or Q when mixed)
  int64_t M = 1;      // padded Bluestein length (>= 2*bq-1), or N
  DevBuf chirp, bhat, bhat_rows, rtab;
  const uint16_t* rtab_ptr = nullptr;  // Rader slot of each element (rtab.p, or a token in dry-run mode)
  const void* chirp_ptr = nullptr;  // chirp.p, or a token in dry-run mode
  const void* bhat_ptr = nullptr;
  const void* bhat_rows_ptr = nullptr;  // spectrum permuted for the fused row pass (M = M1*M2)
  bool blue_ready = false;
  std::map<int, Decomp> decomp;  // key: 0 contiguous lines, 1 strided lines
};

struct Tables {
  int prec = 32;
  std::map<std::vector<int64_t>, std::unique_ptr<DevBuf>> stage;  // (L, radices) -> per-stage twiddles
  std::map<int64_t, std::pair<std::unique_ptr<DevBuf>, int>> two;  // M -> (lo|hi, sh)
  std::map<std::vector<int64_t>, std::unique_ptr<DevBuf>> tw16;     // (M, scale, amax) -> second factors
  std::map<std::vector<int64_t>, std::unique_ptr<DevBuf>> chirp_d;  // (N, step, L) -> chirp steps
};

static size_t csize(int prec) { return prec == 32 ? 8 : 16; }

// Per-stage twiddle table of an in-LDS radix schedule (cached per schedule).
static fftahip_status stage_table(Tables& T, const Inner& in, const void** out, hipStream_t s) {
  std::vector<int64_t> key = {in.L};
  for (int r : in.rad) key.push_back(r);
  if (g_dry) {
    *out = token(TK_STAGE, in.L);
    return FFTAHIP_SUCCESS;
  }
  auto it = T.stage.find(key);
  if (it == T.stage.end()) {
    auto buf = std::make_unique<DevBuf>();
    // L - 1 stage twiddles, then the P roots of each direct-DFT stage
    int64_t n = in.L - 1;
    for (int r : in.rad)
      if (!codelet_radix(r, T.prec)) n += r;
    n = std::max<int64_t>(1, n);
    fftahip_status st = buf->ensure((size_t)n * csize(T.prec));
    if (st) return st;
    RadixList rl;
    std::memset(&rl, 0, sizeof(rl));
    rl.nst = (int32_t)in.rad.size();
    for (size_t i = 0; i < in.rad.size(); ++i) rl.rad[i] = in.rad[i];
    HIP_TRY(launch_gen_stage_twiddles(buf->p, T.prec, rl, n, s));
    it = T.stage.emplace(key, std::move(buf)).first;
  }
  *out = it->second->p;
  return FFTAHIP_SUCCESS;
}

static fftahip_status two_level(Tables& T, int64_t M, const double2** lo, const double2** hi, int* sh_out,
                                hipStream_t s) {
  if (g_dry) {
    int bits = 0;
    while ((1ll << bits) < M) ++bits;
    *sh_out = (bits + 1) / 2;
    *lo = (const double2*)token(TK_TWLO, M);
    *hi = (const double2*)token(TK_TWHI, M);
    return FFTAHIP_SUCCESS;
  }
  auto it = T.two.find(M);
  if (it == T.two.end()) {
    int bits = 0;
    while ((1ll << bits) < M) ++bits;
    const int sh = (bits + 1) / 2;
    const int64_t nlo = 1ll << sh;
    const int64_t nhi = (M + nlo - 1) / nlo;
    auto buf = std::make_unique<DevBuf>();
    fftahip_status st = buf->ensure((size_t)(nlo + nhi) * sizeof(double2));
    if (st) return st;
    double2* plo = (double2*)buf->p;
    HIP_TRY(launch_gen_twiddle_2level(plo, plo + nlo, M, sh, nhi, s));
    it = T.two.emplace(M, std::make_pair(std::move(buf), sh)).first;
  }
  const int sh = it->second.second;
  double2* base = (double2*)it->second.first->p;
  *lo = base;
  *hi = base + (1ll << sh);
  *sh_out = sh;
  return FFTAHIP_SUCCESS;
}

// Second factors of the pass twiddles: tw16[a*16 + j] = w_M^{a*j*scale},
// a < amax, j < 16 (cached per (M, scale, amax)).
static fftahip_status tw16_table(Tables& T, int64_t M, int64_t scale, int64_t amax, const void** out, hipStream_t s) {
  if (g_dry) {
    *out = token(TK_TW16, amax * 16);
    return FFTAHIP_SUCCESS;
  }
  const std::vector<int64_t> key = {M, scale, amax};
  auto it = T.tw16.find(key);
  if (it == T.tw16.end()) {
    auto buf = std::make_unique<DevBuf>();
    fftahip_status st = buf->ensure((size_t)(amax * 16) * csize(T.prec));
    if (st) return st;
    HIP_TRY(launch_gen_tw16(buf->p, T.prec, M, scale, amax, s));
    it = T.tw16.emplace(key, std::move(buf)).first;
  }
  *out = it->second->p;
  return FFTAHIP_SUCCESS;
}

// Largest in-LDS line for a contiguous line set, and the pass-factor cap for
// strided line sets (keeps >= min_lines adjacent lines per workgroup so that
// global accesses stay coalesced).
static int64_t lds_budget() {
  const int64_t cap = env_i64("FFTAHIP_LDS_BUDGET", 0);
  int64_t lds = g_dry ? g_dry_lds : dev_info().lds;
  if (lds < 65536) lds = 65536;
  // headroom for the per-line tables of up to kMaxLines lines (a tile's
  // LDS is its lines plus C * kLineTabBytes, tile_lds_bytes)
  lds -= kLineTabBytes * kMaxLines + 1024;
  return cap > 0 ? std::min(cap, lds) : lds;
}
// Adjacent lines per workgroup for strided passes: 128-byte segments.
static int64_t min_lines(int prec) { return prec == 32 ? 16 : 8; }
// LDS per workgroup for strided (multi-pass) tiles: 40 KB, so that three
// or more workgroups share a CU and hide each other's global-memory latency
// (measured on the batched workload: 3.62 ms per step against 3.77 ms with
// 64 KB tiles; 32 KB and below split lines into more passes).
static int64_t lds_budget();
static int64_t pass_budget() { return std::min<int64_t>(env_i64("FFTAHIP_PASS_LDS", 40960), lds_budget()); }
static int64_t single_max(int prec) {
  // room for the 1-in-16 LDS padding
  int64_t m = lds_budget() / (int64_t)csize(prec) * 16 / 17;
  return std::min<int64_t>(m, 1024 * kEPT);
}
static int64_t pass_max(int prec, const TileShape& ts = TileShape()) {
  if (ts.budget > 0) {
    // the longest line whose padded pitch (<= 17/16 L + 1) fits cmin lines
    const int64_t b = std::min(ts.budget, lds_budget());
    return std::max<int64_t>(16, (b / ((int64_t)csize(prec) * ts.cmin) - 1) * 16 / 17);
  }
  const int64_t dflt = std::min<int64_t>(1024, pass_budget() / ((int64_t)csize(prec) * min_lines(prec)));
  const int64_t v = env_i64("FFTAHIP_PASS_MAX", dflt);
  return std::max<int64_t>(v, 16);
}

// Longest row of the fused Bluestein row pass (one line per workgroup).
static int64_t row_max(int prec) {
  return std::min<int64_t>(single_max(prec), env_i64("FFTAHIP_ROW_MAX", prec == 32 ? 8192 : 4096));
}

// Bluestein length M = M1 * M2 for the three fused passes: column passes of
// length M1 (strided, <= pass_max) and one row pass of length M2 (<= row_max).
// Relative work of a three-pass Bluestein row pass over rows of m2 points
// with schedule `row` (per point: its stages plus about three for the column
// passes and the row pass's own load, spectrum and store phases), rows that
// are not a power of two weighted by FFTAHIP_BLUE_NP2 percent (their
// radix-3/5 stages and partly filled workgroups ran at 1.4-2.2 TB/s against
// 3.0 TB/s for 4096-point rows) and rows that need more than 512 threads by
// FFTAHIP_BLUE_WIDE percent (a 1024-thread workgroup of the full variant
// needs a nearly idle CU and waits behind the step's other launches).
static double row_cost(const Inner& row, int64_t m2) {
  double c = (double)row.rad.size() + 3.0;
  if (m2 & (m2 - 1)) c *= (double)env_i64("FFTAHIP_BLUE_NP2", 150) / 100.0;
  if (row.Tl > 512) c *= (double)env_i64("FFTAHIP_BLUE_WIDE", 150) / 100.0;
  return c;
}

static bool blue3_split(int64_t M, int prec, int64_t* M1, int64_t* M2) {
  const int64_t pmax = pass_max(prec);
  Inner a, b;
  // first choice: columns that run as register line passes, the row of least
  // row_cost (ties to the longest columns)
  const int64_t rmax = env_i64("FFTAHIP_REG", 1) != 0 ? reg_pass_max(prec, 2) : 0;
  double best = 1e300;
  for (int64_t m1 = rmax; m1 >= 8; --m1) {
    if (M % m1) continue;
    const int64_t m2 = M / m1;
    if (m2 > row_max(prec)) break;
    // the rule's schedules (not the tuned table): the split must not change
    // after the plan's spectrum rows are laid out for it
    if (!make_inner(m2, b) || !make_inner(m1, a)) continue;
    const double cost = row_cost(b, m2);
    if (cost < best) {
      best = cost;
      *M1 = m1;
      *M2 = m2;
    }
  }
  if (best < 1e300) return true;
  // else LDS column passes of >= 32 points (enough work per strided tile),
  // then the longest row that fits
  for (int64_t min_m1 : {(int64_t)32, (int64_t)2}) {
    for (int64_t m2 = std::min(M / min_m1, row_max(prec)); m2 >= 2; --m2) {
      if (M % m2) continue;
      const int64_t m1 = M / m2;
      if (m1 < min_m1 || m1 > pmax) continue;
      if (!make_inner(m2, b) || !make_inner(m1, a)) continue;
      *M1 = m1;
      *M2 = m2;
      return true;
    }
  }
  return false;
}

// Split M into pass factors, each <= cap and smooth; balanced greedily.
static bool split_factors(int64_t M, int64_t cap, std::vector<int64_t>& F, int64_t cutoff = kInternalCutoff) {
  std::vector<int64_t> primes = prime_factors(M);
  std::sort(primes.begin(), primes.end(), std::greater<int64_t>());
  // a direct prime above the cap gets a pass of its own (a strided pass then
  // holds fewer adjacent lines per tile)
  if (!primes.empty()) cap = std::max(cap, primes.front());
  for (int p = 2; p <= 8; ++p) {
    std::vector<int64_t> g(p, 1);
    bool ok = true;
    for (int64_t q : primes) {
      int best = -1;
      for (int i = 0; i < p; ++i)
        if (g[i] * q <= cap && (best < 0 || g[i] < g[best])) best = i;
      if (best < 0) { ok = false; break; }
      g[best] *= q;
    }
    if (!ok) continue;
    F.clear();
    for (int64_t x : g)
      if (x > 1) F.push_back(x);
    if (F.empty()) F.push_back(1);
    // every factor must have an in-LDS schedule
    bool sched = true;
    Inner tmp;
    for (int64_t x : F) sched = sched && make_inner(x, tmp, 0, cutoff);
    if (sched) return true;
  }
  return false;
}

static fftahip_status make_decomp_rule(LengthPlan& lp, int strided, Decomp& d);

// Strided passes of a decomposition carry the length's tuned tile shape.
static void apply_tile(Decomp& d, const TileShape& ts) {
  for (Inner& in : d.inner) {
    in.tile_budget = ts.budget;
    in.tile_cmin = ts.cmin;
  }
}

// FP32 convolution lines (Bluestein, Rader, the three-pass row) run their
// two sweeps with the boundary fused into one stage when the schedule starts
// and ends with the same codelet radix: a radix that occurs at least twice
// (the largest such) is moved to both ends, the others keep their order.
static void mid_ends(Inner& in, int prec) {
  if (prec != 32 || env_i64("FFTAHIP_MID_FUSE", 1) == 0 || in.rad.size() < 2) return;
  if (in.rad.front() == in.rad.back() && small_radix(in.rad.front())) return;
  if (!small_radix(in.rad.front())) return;  // a direct-DFT first stage keeps the product in its own phase
  int pick = 0;
  for (int r : in.rad)
    if (small_radix(r) && std::count(in.rad.begin(), in.rad.end(), r) >= 2) pick = std::max(pick, r);
  if (!pick) return;
  std::vector<int> rad{pick};
  bool skipped_first = false, skipped_second = false;
  for (int r : in.rad) {
    if (r == pick && !skipped_first) { skipped_first = true; continue; }
    if (r == pick && !skipped_second) { skipped_second = true; continue; }
    rad.push_back(r);
  }
  rad.push_back(pick);
  Inner tmp;
  if (!inner_from(in.L, rad, tmp, prec)) return;
  tmp.tile_budget = in.tile_budget;
  tmp.tile_cmin = in.tile_cmin;
  in = tmp;
}

static fftahip_status make_decomp(LengthPlan& lp, int strided, Decomp& d) {
  fftahip_status st = make_decomp_rule(lp, strided, d);
  if (st) return st;
  apply_tile(d, tuned_tile(lp.prec, lp.N));
  switch (d.kind) {
    case K_BLUE_SINGLE: case K_RADER_SINGLE: case K_MIXED: case K_RADER_MIXED:
      mid_ends(d.inner[0], lp.prec);
      break;
    case K_BLUE_3PASS:
      mid_ends(d.inner[1], lp.prec);
      break;
    default: break;
  }
  return st;
}

static fftahip_status make_decomp_rule(LengthPlan& lp, int strided, Decomp& d) {
  d = Decomp();
  const int prec = lp.prec;
  const TileShape ts = tuned_tile(prec, lp.N);
  if (lp.mixed) {
    // pass 0: Q-point DFTs as Bluestein lines of M in LDS (or Rader lines of
    // Q - 1); then Stockham over S
    d.kind = lp.rader ? K_RADER_MIXED : K_MIXED;
    d.M = lp.N;
    std::vector<int64_t> FS;
    Inner tmp;
    if (lp.S <= pass_max(prec, ts) && make_inner(lp.S, tmp, prec, lp.cutoff)) FS = {lp.S};
    else if (!split_factors(lp.S, pass_max(prec, ts), FS, lp.cutoff))
      return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no pass decomposition for length " + std::to_string(lp.S));
    d.F = {lp.bq};
    Inner qb;
    if (!make_inner(lp.M, qb, prec)) return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no line schedule for Bluestein length");
    d.inner = {qb};
    for (int64_t f : FS) {
      Inner in;
      if (!make_inner(f, in, prec, lp.cutoff))
        return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no line schedule for pass length " + std::to_string(f));
      d.F.push_back(f);
      d.inner.push_back(in);
    }
    return FFTAHIP_SUCCESS;
  }
  const int64_t target = lp.blue ? lp.M : lp.N;
  // the user length's own primes follow the plan's cutoff; Bluestein's M is internal
  const int64_t cut = lp.blue ? kInternalCutoff : lp.cutoff;
  d.M = target;
  if (lp.rader) {
    // one workgroup per line whatever the layout (a strided line that does
    // not fit pass_max runs with fewer neighbouring lines per tile)
    Inner in;
    if (!make_inner(target, in, prec, cut)) return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no line schedule for Rader");
    d.kind = K_RADER_SINGLE;
    d.F = {target};
    d.inner = {in};
    return FFTAHIP_SUCCESS;
  }
  const int64_t single_cap = strided ? pass_max(prec, ts) : single_max(prec);
  if (target <= single_cap) {
    Inner in;
    if (make_inner(target, in, prec, cut)) {
      d.kind = lp.blue ? K_BLUE_SINGLE : K_SINGLE;
      d.F = {target};
      d.inner = {in};
      return FFTAHIP_SUCCESS;
    }
  }
  int64_t M1, M2;
  if (lp.blue && blue3_split(target, prec, &M1, &M2)) {
    d.kind = K_BLUE_3PASS;
    d.F = {M1, M2};
    Inner a, b;
    if (!make_inner(M1, a, prec) || !make_inner(M2, b, prec))
      return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no line schedule for the Bluestein passes");
    d.inner = {a, b};
    return FFTAHIP_SUCCESS;
  }
  std::vector<int64_t> F;
  if (!split_factors(target, pass_max(prec, ts), F, cut))
    return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no pass decomposition for length " + std::to_string(target));
  d.kind = lp.blue ? K_BLUE_MULTI : K_MULTI;
  d.F = F;
  for (int64_t f : F) {
    Inner in;
    if (!make_inner(f, in, prec, cut))
      return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no line schedule for pass length " + std::to_string(f));
    d.inner.push_back(in);
  }
  return FFTAHIP_SUCCESS;
}

static fftahip_status get_decomp(LengthPlan& lp, int strided, Decomp** out) {
  auto it = lp.decomp.find(strided);
  if (it == lp.decomp.end()) {
    Decomp d;
    fftahip_status st = make_decomp(lp, strided, d);
    if (st) return st;
    it = lp.decomp.emplace(strided, d).first;
  }
  *out = &it->second;
  return FFTAHIP_SUCCESS;
}

// ------------------------------------------------------------------ line sets and passes

// A set of independent lines of one length: up to two affine batch dims, or a
// table of offsets (ragged batches).  Positions within a line are scaled by
// es_in / es_out.
struct LineSet {
  int64_t n[2] = {1, 1};
  int64_t in_str[2] = {0, 0};
  int64_t out_str[2] = {0, 0};
  const int64_t* in_tab = nullptr;   // device; per-line offsets (then n = {1, count})
  const int64_t* out_tab = nullptr;
  int64_t es_in = 1, es_out = 1;
  int64_t count() const { return n[0] * n[1]; }
};

struct Dim {
  int64_t n, a, b, ca, cb;
  bool tw;
  int64_t cm = 0;  // coefficient of the mid-pass spectrum index
};

enum LaunchOp { OP_PASS = 0, OP_COPY = 1 };
struct Launch {
  int op = OP_PASS;  // OP_COPY: copy `bytes` from the input buffer to the output buffer
  bool reg = false;  // register line pass (one line per thread) instead of the LDS line pass
  LinePass lp;
  int block = 256;
  size_t lds = 0;
  int64_t grid = 0;
  size_t bytes = 0;
};

// Line indices, per-slot extents and in-line positions are 32-bit on the
// device (FastDiv needs divisors below 2^31): a line set must stay below this.
constexpr int64_t kMaxLinesPerPass = (int64_t(1) << 31) - 1;

// Fill geometry: dims (<= 3 affine + optional table dim) -> LinePass fields.
// False (nothing usable written) when the line count or an extent reaches
// 2^31.
static bool set_dims(LinePass& p, std::vector<Dim> Din, bool tab) {
  // unit dims carry nothing; the table dim (if any) stays last
  std::vector<Dim> D;
  for (size_t i = 0; i < Din.size(); ++i)
    if (Din[i].n > 1 || (tab && i + 1 == Din.size())) D.push_back(Din[i]);
  if (D.empty()) D.push_back(Dim{1, 0, 0, 0, 0, false, 0});
  // choose slot 0 among the affine dims for coalescing
  const size_t naff = tab ? D.size() - 1 : D.size();
  int pick = -1;
  for (size_t i = 0; i < naff && pick < 0; ++i)
    if (D[i].n > 1 && D[i].a == 1 && p.ae != 1) pick = (int)i;
  for (size_t i = 0; i < naff && pick < 0; ++i)
    if (D[i].n > 1 && D[i].b == 1 && p.be != 1) pick = (int)i;
  if (pick > 0) std::swap(D[0], D[pick]);
  std::vector<Dim> slots(4, Dim{1, 0, 0, 0, 0, false, 0});
  for (size_t i = 0; i < naff; ++i) slots[i] = D[i];
  if (tab) slots[3] = D.back();
  int64_t total = 1;
  for (int i = 0; i < 4; ++i) {
    if (slots[i].n < 1 || slots[i].n > kMaxLinesPerPass) return false;
    total *= slots[i].n;
    if (total > kMaxLinesPerPass) return false;
  }
  p.nlines = total;
  for (int i = 0; i < 3; ++i) {
    p.n[i] = slots[i].n;
    p.divn[i] = make_fastdiv((uint32_t)slots[i].n);
  }
  p.twdim = 0;
  for (int i = 0; i < 4; ++i) {
    p.a[i] = slots[i].a;
    p.b[i] = slots[i].b;
    p.ca[i] = slots[i].ca;
    p.cb[i] = slots[i].cb;
    p.cm[i] = slots[i].cm;
    if (slots[i].tw) p.twdim = i;
  }
  p.col_in = (p.ae != 1 && slots[0].n > 1 && slots[0].a == 1) ? 1 : 0;
  p.col_out = (p.be != 1 && slots[0].n > 1 && slots[0].b == 1) ? 1 : 0;
  return true;
}

static fftahip_status lines_capacity_error() {
  return fail(FFTAHIP_ERR_CAPACITY, "a pass over this array needs 2^31 or more lines (line indices are 32-bit)");
}

// ------------------------------------------------------------------ LDS bank model
// 32 banks of 4 bytes; one 64-lane access takes as many cycles as the most
// loaded bank has distinct words.  Used to pick the LDS padding of a pass.

static int padded(int i, int sh) { return sh ? i + (i >> sh) : i; }

static int access_cycles(const std::vector<int>& elem, int words) {
  int cnt[32] = {0};
  std::vector<int> seen;
  seen.reserve(elem.size() * words);
  for (int e : elem)
    for (int w = 0; w < words; ++w) {
      const int word = e * words + w;
      if (std::find(seen.begin(), seen.end(), word) != seen.end()) continue;
      seen.push_back(word);
      ++cnt[word & 31];
    }
  int m = 0;
  for (int b = 0; b < 32; ++b) m = std::max(m, cnt[b]);
  return m;
}

// Cycles per element of LDS traffic for one line pass with padding sh and
// line pitch ldl (wave 0 pattern of every stage, load and store phases).
static double lds_cost(const Inner& in, int C, int ldl, int sh, int words, bool col_in, bool col_out) {
  const int L = (int)in.L;
  double cost = 0;
  std::vector<int> idx;
  // load / store phases
  for (int side = 0; side < 2; ++side) {
    const bool col = side == 0 ? col_in : col_out;
    idx.clear();
    for (int e = 0; e < 64 && e < C * L; ++e) {
      int c, t;
      if (col) { t = e / C; c = e % C; } else { c = e / L; t = e % L; }
      idx.push_back(c * ldl + padded(t, sh));
    }
    if (!idx.empty()) cost += access_cycles(idx, words) / 64.0;
  }
  int Ns = 1;
  for (int R : in.rad) {
    const int Lr = L / R;
    double sc = 0;
    for (int t = 0; t < R; ++t) {
      std::vector<int> rd, wr;
      for (int tid = 0; tid < 64; ++tid) {
        const int c = tid / in.Tl, j = tid % in.Tl;
        if (c >= C || j >= Lr) continue;
        const int k = j % Ns, base = (j / Ns) * Ns * R + k;
        rd.push_back(c * ldl + padded(j + t * Lr, sh));
        wr.push_back(c * ldl + padded(base + t * Ns, sh));
      }
      if (!rd.empty()) sc += access_cycles(rd, words) + access_cycles(wr, words);
    }
    cost += sc / (64.0 * R) * 1.0;
    Ns *= R;
  }
  return cost;
}

static bool reg_eligible(const LinePass& p, const Inner& in, int prec) {
  if (env_i64("FFTAHIP_REG", 1) == 0) return false;
  if (p.mid_mode != MID_NONE || !p.col_in || !p.col_out) return false;
  if (in.L < 2 || in.L > reg_pass_max(prec, pass_mode(p)) || !smooth13(in.L)) return false;
  return p.n[0] * (int64_t)csize(prec) >= 128;
}

// allow_reg = false keeps the pass in LDS even when its lines qualify for the
// register path: a transform whose lines fit one in-LDS pass then computes
// the same arithmetic (radix schedule, stage twiddles) whether its lines are
// contiguous or strided, so a 1-D plan along any dimension equals the 1-D
// transform of each slice bit for bit (test/onedim/complex_forward.jl:27-33).
static void set_inner(LinePass& p, const Inner& in, int prec, Launch& L, bool allow_reg = true) {
  p.L = (int32_t)in.L;
  p.nst = (int32_t)in.rad.size();
  int64_t Ns = 1;
  for (int s = 0; s < kMaxStages; ++s) {
    if (s < p.nst) {
      const int r = in.rad[s];
      p.rad[s] = r;
      p.nsdiv[s] = make_fastdiv((uint32_t)Ns);
      Ns *= r;
    } else {
      p.rad[s] = 1;
      p.nsdiv[s] = make_fastdiv(1);
    }
  }
  // threads per line: lines of up to 64 threads get a power of two (at most
  // a third of the lanes idle), so that every line lies within one wavefront
  // and the sweeps synchronise per wave (wave_sync) instead of per workgroup
  int64_t Tl = in.Tl;
  bool wave = false;
  if (env_i64("FFTAHIP_WAVE_SYNC", 1) != 0 && Tl <= 64) {
    int64_t t2 = 1;
    while (t2 < Tl) t2 <<= 1;
    if (t2 * 3 <= Tl * 4) {
      Tl = t2;
      wave = true;
    }
  }
  p.Tl = (int32_t)Tl;
  p.wave_sync = wave ? 1 : 0;
  const int64_t esz = (int64_t)csize(prec);
  const bool strided = p.col_in || p.col_out;
  // lines per workgroup: fill 256 threads; strided tiles also want
  // min_lines adjacent lines (128-byte segments) within the pass budget
  int64_t C = std::max<int64_t>(1, 256 / Tl);
  if (strided) C = std::max<int64_t>(C, in.tile_cmin > 0 ? in.tile_cmin : min_lines(prec));
  C = std::min<int64_t>(C, kMaxLines);
  C = std::min<int64_t>(C, std::max<int64_t>(1, 1024 / Tl));
  // fill the workgroup size the tile needs with as many lines as it holds
  // (e.g. 40 threads per line: 25 lines in 1024 threads, not 16 in 640)
  if (env_i64("FFTAHIP_FILL", 1) != 0) {
    const int64_t thr = C * Tl;
    const int64_t blk = thr <= 256 ? 256 : (thr <= 512 ? 512 : 1024);
    C = std::max<int64_t>(C, std::min<int64_t>(kMaxLines, blk / Tl));
  }
  // Rader lines keep x0 and the first sweep's DC term after their positions
  const int64_t side = p.load_mode == LOAD_RADER ? kRaderSide : 0;
  const int64_t padmax = in.L + (in.L - 1) / 16 + 1 + side;  // largest pitch any padding choice needs
  const int64_t pbud = in.tile_budget > 0 ? std::min(in.tile_budget, lds_budget()) : pass_budget();
  const int64_t budget = strided ? std::min(lds_budget(), std::max<int64_t>(pbud, padmax * esz)) : lds_budget();
  // direct-DFT stages stage their roots after the tile's line tables, each
  // stage's P roots repeated kBigKP times
  int64_t roots_bytes = 0;
  for (int r : in.rad)
    if (!codelet_radix(r, prec)) roots_bytes += kBigKPHost * r * esz;
  while (C > 1 && C * padmax * esz + roots_bytes > budget) --C;
  if (C > p.nlines) C = std::max<int64_t>(1, p.nlines);
  // LDS padding: env FFTAHIP_LDS_PAD (-1 = choose by the bank model)
  const int64_t force = env_i64("FFTAHIP_LDS_PAD", -1);
  int best_sh = 0;
  int64_t ldl = in.L;
  double best = 1e30;
  const int words = (int)(esz / 4);
  static std::mutex pad_mu;
  static std::map<std::vector<int64_t>, std::pair<int, int64_t>> pad_cache;
  Inner inw = in;  // the line schedule with the threads per line this pass runs
  inw.Tl = (int)Tl;
  std::vector<int64_t> key = {in.L, Tl, C, words, p.col_in, p.col_out, force, side};
  for (int r : in.rad) key.push_back(r);
  {
    std::lock_guard<std::mutex> g(pad_mu);
    auto it = pad_cache.find(key);
    if (it != pad_cache.end()) {
      best_sh = it->second.first;
      ldl = it->second.second;
      best = 0;
    }
  }
  for (int sh : {0, 4, 5}) {
    if (best == 0) break;
    if (force >= 0 && sh != force) continue;
    for (int extra : {0, 1}) {
      if (C == 1 && extra) continue;
      const int64_t pitch = padded((int)in.L - 1, sh) + 1 + extra + side;
      // a padding the tile (lines and staged roots) no longer fits is out
      if (sh && C * pitch * esz + roots_bytes > lds_budget()) continue;
      const double cst = lds_cost(inw, (int)C, (int)pitch, sh, words, p.col_in, p.col_out);
      if (cst < best - 1e-9) {
        best = cst;
        best_sh = sh;
        ldl = pitch;
      }
    }
  }
  if (force > 0 && best > 1e29) {
    // the forced padding (when the tile still fits it), else none
    best_sh = C * (padded((int)in.L - 1, (int)force) + 1 + side) * esz + roots_bytes <= lds_budget() ? (int)force : 0;
    ldl = padded((int)in.L - 1, best_sh) + 1 + side;
  }
  {
    std::lock_guard<std::mutex> g(pad_mu);
    pad_cache[key] = {best_sh, ldl};
  }
  p.lds_sh = best_sh;
  // row-mode sides go straight between HBM and the first/last radix stage
  // when the threads of one line cover >= 128 contiguous bytes per access
  // and the pass has no table-driven element operations
  const bool plain = p.load_mode == LOAD_PLAIN && p.mid_mode == MID_NONE && p.store_mode == STORE_PLAIN && p.pretw == 0;
  const bool fuse = env_i64("FFTAHIP_FUSE", 1) != 0 && !in.rad.empty() && plain && Tl * esz >= 128;
  // the direct-DFT stage reads and writes LDS only
  const bool big_first = !in.rad.empty() && !codelet_radix(in.rad.front(), prec);
  const bool big_last = !in.rad.empty() && !codelet_radix(in.rad.back(), prec);
  p.fuse_in = (fuse && !p.col_in && !big_first) ? 1 : 0;
  p.fuse_out = (fuse && !p.col_out && !big_last) ? 1 : 0;
  const int64_t threads = C * Tl;
  p.C = (int32_t)C;
  p.ldl = (int32_t)ldl;
  p.divC = make_fastdiv((uint32_t)C);
  p.divL = make_fastdiv((uint32_t)in.L);
  L.block = threads <= 256 ? 256 : (threads <= 512 ? 512 : 1024);
  // stepped row-layout phases: block/C threads per line, coalesced when they
  // cover >= 128 bytes per access or the whole line
  {
    const int64_t Tr = L.block / C;
    const bool wide = env_i64("FFTAHIP_ROWSTEP", 1) != 0 && (Tr * esz >= 128 || Tr >= in.L);
    p.divR = make_fastdiv((uint32_t)Tr);
    p.row_in = (wide && !p.col_in && !p.fuse_in) ? 1 : 0;
    p.row_out = (wide && !p.col_out && !p.fuse_out) ? 1 : 0;
    p.row_mid = wide ? 1 : 0;
    // Rader passes load and store through their stepped paths only
    if (side) {
      p.row_in = p.col_in ? 0 : 1;
      p.row_out = p.col_out ? 0 : 1;
    }
    p.mid_in_stage = (p.mid_mode == MID_BHAT && env_i64("FFTAHIP_MID_STAGE", 1) != 0 && !big_first) ? 1 : 0;
    p.mid_fused = (p.mid_in_stage && prec == 32 && env_i64("FFTAHIP_MID_FUSE", 1) != 0 && !in.rad.empty() &&
                   in.rad.front() == in.rad.back() && small_radix(in.rad.back()))
                      ? 1
                      : 0;
  }
  L.lds = tile_lds_bytes(C, ldl, esz) + (size_t)roots_bytes;
  L.grid = (p.nlines + C - 1) / C;
  p.xcd_remap = (L.grid >= 16 && L.grid % 8 == 0) ? 1 : 0;
  // short lines laid across memory on both sides (>= 128 contiguous bytes
  // per wave access): one line per thread in registers
  if (allow_reg && reg_eligible(p, in, prec)) {
    L.reg = true;
    L.block = kRegBlock;
    L.lds = 0;
    L.grid = (p.nlines + kRegBlock - 1) / kRegBlock;
    p.C = kRegBlock;
    p.fuse_in = p.fuse_out = 0;
    p.xcd_remap = 0;
  }
}

// Pass twiddle layout of a built pass: split the exponent along the index
// that runs across neighbouring lanes (lines when the pass side is laid out
// across lines or the pass runs in registers, else positions in the line),
// and attach the matching second-factor table.
static fftahip_status attach_pass_twiddle(LinePass& p, const Launch& la, Tables& T, int64_t M, hipStream_t s) {
  if (!p.pretw) return FFTAHIP_SUCCESS;
  const bool lanes_over_lines = la.reg || (p.pretw == 1 ? p.col_in != 0 : p.col_out != 0);
  p.tw_split = lanes_over_lines ? 0 : 1;
  int64_t ext;
  if (p.twdim < 3) {
    ext = p.n[p.twdim];
  } else {
    ext = p.nlines / std::max<int64_t>(1, p.n[0] * p.n[1] * p.n[2]);
  }
  const int64_t amax = p.tw_split == 0 ? (int64_t)p.L : ext;
  return tw16_table(T, M, p.tw_scale, amax, &p.tw16, s);
}

// FP64 chirp steps d[t] of a register pass with a chirp on load or store
// (LinePass::chirp_d; cached per (N, step, L)).  Env FFTAHIP_CHIRP_CALC=0
// keeps the table reads.
static fftahip_status attach_chirp_steps(LinePass& p, const Launch& la, Tables& T, hipStream_t s) {
  p.chirp_d = nullptr;
  const bool cin = p.load_mode == LOAD_CHIRP, cout = p.store_mode == STORE_CHIRP;
  if (!la.reg || (!cin && !cout) || env_i64("FFTAHIP_CHIRP_CALC", 1) == 0) return FFTAHIP_SUCCESS;
  if (cin && cout && p.cae != p.cbe) return FFTAHIP_SUCCESS;
  if (cout && p.pretw == 2) return FFTAHIP_SUCCESS;  // the computed store path carries no post-twiddle
  const int64_t step = cin ? p.cae : p.cbe;
  if (step <= 0 || p.chirp_n <= 0) return FFTAHIP_SUCCESS;
  // the kernel reduces a^2 mod 2N and a*step mod N through an FP64 quotient,
  // exact when the line start a and the step are below N (quotients < 2^31)
  {
    const int64_t* c = cin ? p.ca : p.cb;
    int64_t amax = 0;
    int64_t rest = p.nlines;
    for (int d = 0; d < 3; ++d) {
      amax += (std::min(p.n[d], rest) - 1) * std::max<int64_t>(0, c[d]);
      rest = (rest + p.n[d] - 1) / p.n[d];
      if (c[d] < 0) return FFTAHIP_SUCCESS;
    }
    if (c[3] < 0) return FFTAHIP_SUCCESS;
    amax += (rest - 1) * c[3];
    if (amax >= p.chirp_n || step >= p.chirp_n) return FFTAHIP_SUCCESS;
  }
  p.chirp_inv = 1.0 / (double)p.chirp_n;
  if (g_dry) {
    p.chirp_d = (const double2*)token(TK_CHIRPD, p.L);
    return FFTAHIP_SUCCESS;
  }
  const std::vector<int64_t> key = {p.chirp_n, step, p.L};
  auto it = T.chirp_d.find(key);
  if (it == T.chirp_d.end()) {
    auto buf = std::make_unique<DevBuf>();
    fftahip_status st = buf->ensure((size_t)p.L * sizeof(double2));
    if (st) return st;
    HIP_TRY(launch_gen_chirp_steps((double2*)buf->p, step, p.chirp_n, p.L, s));
    it = T.chirp_d.emplace(key, std::move(buf)).first;
  }
  p.chirp_d = (const double2*)it->second->p;
  return FFTAHIP_SUCCESS;
}

static LinePass blank_pass() {
  LinePass p;
  std::memset(&p, 0, sizeof(p));
  p.scale = 1.0;
  p.ae = p.be = 1;
  p.n[0] = p.n[1] = p.n[2] = 1;
  return p;
}

// Build the launches that transform every line of `ls` through `lp`.
// conj: backward sign (conjugate in and out).
// A buffer a pass reads or writes: which one (BufId) and an element offset.
struct BufRef {
  int id;
  int64_t off;
};
static void set_in(LinePass& p, BufRef r) {
  p.in_buf = r.id;
  p.in_off = r.off;
}
static void set_out(LinePass& p, BufRef r) {
  p.out_buf = r.id;
  p.out_off = r.off;
}

// Operations of a Rader line pass over the prime lp.bq: load through the
// slot table, spectrum product between the sweeps, x0 + conj on store; the
// spectrum is indexed by the position alone (one table for every line).
// A Bluestein line of L >= 2 chirp_n - 1 points holds its chirp_n inputs at
// positions t < chirp_n and zeros after them, and its store reads positions
// t < chirp_n only.  When chirp_n <= ((r + 1) / 2) * (L / r) for the first
// radix r, the first stage drops its inputs t >= (r + 1) / 2 (all zero) and
// the load writes no position past that bound; when the same holds for the
// last radix, the last stage drops its outputs t >= (r + 1) / 2.  Needs
// every line to start at chirp index 0 with unit steps on both sides.
static void set_zero_pad(LinePass& p, const Launch& la, int prec) {
  p.zero_pad = 0;
  p.zero_from = 0;
  // a one-stage fused boundary (mid_fused, nst == 1) runs no sweep stage
  if (la.reg || p.nst < 1 || (p.mid_fused && p.nst < 2) || env_i64("FFTAHIP_ZERO_PAD", 1) == 0) return;
  if (p.load_mode != LOAD_CHIRP || p.mid_mode != MID_BHAT || p.store_mode != STORE_CHIRP || p.pretw != 0) return;
  if (p.cae != 1 || p.cbe != 1 || p.mirror_rows) return;
  for (int u = 0; u < 4; ++u)
    if (p.ca[u] != 0 || p.cb[u] != 0) return;
  const int r0 = p.rad[0], rl = p.rad[p.nst - 1];
  const int64_t f0 = (int64_t)((r0 + 1) / 2) * (p.L / r0);
  const int64_t fl = (int64_t)((rl + 1) / 2) * (p.L / rl);
  if (codelet_radix(r0, prec) && zero_pad_radix(r0) && p.chirp_n <= f0) {
    p.zero_pad |= 1;
    p.zero_from = (int32_t)f0;
  }
  if (codelet_radix(rl, prec) && zero_pad_radix(rl) && p.chirp_n <= fl) p.zero_pad |= 2;
}

static void set_rader(LinePass& p, const LengthPlan& lp) {
  p.load_mode = LOAD_RADER;
  p.mid_mode = MID_BHAT;
  p.store_mode = STORE_RADER;
  p.chirp_n = lp.bq;
  p.cae = p.cbe = p.cme = 1;
  for (int u = 0; u < 4; ++u) p.ca[u] = p.cb[u] = p.cm[u] = 0;
  p.bhat = lp.bhat_ptr;
  p.rtab = lp.rtab_ptr;
  p.chirp = nullptr;
}

static fftahip_status build_launches(LengthPlan& lp, Tables& T, const LineSet& ls, BufRef in, BufRef out, BufRef s0,
                                     BufRef s1, bool conj, std::vector<Launch>& L, hipStream_t stream) {
  const int strided = (ls.es_in != 1 || ls.es_out != 1) ? 1 : 0;
  Decomp* d;
  fftahip_status st = get_decomp(lp, strided, &d);
  if (st) return st;
  const int prec = lp.prec;
  const int64_t M = d->M;
  const int64_t B = ls.count();
  const bool tab = ls.in_tab != nullptr;

  if (d->kind == K_SINGLE || d->kind == K_BLUE_SINGLE || d->kind == K_RADER_SINGLE) {
    Launch la;
    LinePass p = blank_pass();
    set_in(p, in);
    set_out(p, out);
    p.ae = ls.es_in;
    p.be = ls.es_out;
    std::vector<Dim> D;
    if (tab) {
      D.push_back(Dim{ls.n[1], 0, 0, 0, 0, false});
      p.in_tab = ls.in_tab;
      p.out_tab = ls.out_tab;
    } else {
      for (int i = 0; i < 2; ++i)
        if (ls.n[i] > 1) D.push_back(Dim{ls.n[i], ls.in_str[i], ls.out_str[i], 0, 0, false});
      if (D.empty()) D.push_back(Dim{1, 0, 0, 0, 0, false});
    }
    if (!set_dims(p, D, tab)) return lines_capacity_error();
    st = stage_table(T, d->inner[0], &p.tw, stream);
    if (st) return st;
    p.conj_in = conj ? 1 : 0;
    p.conj_out = conj ? 1 : 0;
    if (d->kind == K_BLUE_SINGLE) {
      p.load_mode = LOAD_CHIRP;
      p.mid_mode = MID_BHAT;
      p.store_mode = STORE_CHIRP;
      p.chirp_n = lp.bq;
      p.cae = 1;
      p.cbe = 1;
      p.cme = 1;
      p.chirp = lp.chirp_ptr;
      p.bhat = lp.bhat_ptr;
    } else if (d->kind == K_RADER_SINGLE) {
      set_rader(p, lp);
    }
    set_inner(p, d->inner[0], prec, la, false);
    set_zero_pad(p, la, prec);
    la.lp = p;
    L.push_back(la);
    return FFTAHIP_SUCCESS;
  }

  // batch dims: the user side follows the line set, scratch holds transform b at b*M
  auto batch_dims = [&](std::vector<Dim>& D, bool user_in, bool user_out) {
    if (tab) {
      D.push_back(Dim{ls.n[1], user_in ? 0 : M, user_out ? 0 : M, 0, 0, false});
      return;
    }
    int64_t sstride = M;
    for (int b = 0; b < 2; ++b) {
      if (ls.n[b] > 1)
        D.push_back(Dim{ls.n[b], user_in ? ls.in_str[b] : sstride, user_out ? ls.out_str[b] : sstride, 0, 0, false});
      sstride *= ls.n[b];
    }
  };

  if (d->kind == K_BLUE_3PASS) {
    // M = M1*M2, positions n = M2*n1 + n2.
    //  K1: lines n2, FFT over n1 of chirp-multiplied input, post-twiddle w_M^{n2*k1},
    //      store S[k1*M2 + n2];
    //  K2: rows k1 of S, FFT over n2 -> k2, times bhat[k1 + M1*k2], conjugate,
    //      FFT over k2 -> n2, post-twiddle w_M^{k1*n2}, store in place;
    //  K3: lines n2 of S, FFT over k1 -> n1, store conj(chirp[n] * v) at n < N.
    const int64_t M1 = d->F[0], M2 = d->F[1];
    const double2* tlo = nullptr;
    const double2* thi = nullptr;
    int tsh = 0;
    st = two_level(T, M, &tlo, &thi, &tsh, stream);
    if (st) return st;
    for (int k = 0; k < 3; ++k) {
      Launch la;
      LinePass p = blank_pass();
      std::vector<Dim> D;
      p.pretw = (k < 2) ? 2 : 0;
      p.tw_scale = 1;
      p.tw_lo = tlo;
      p.tw_hi = thi;
      p.tw_sh = tsh;
      p.chirp_n = lp.bq;
      p.chirp = lp.chirp_ptr;
      p.bhat = lp.bhat_ptr;
      if (k == 0) {
        set_in(p, in);
        set_out(p, s0);
        p.ae = M2 * ls.es_in;
        p.be = M2;
        D.push_back(Dim{M2, ls.es_in, 1, 1, 0, true});
        batch_dims(D, true, false);
        p.load_mode = LOAD_CHIRP;
        p.cae = M2;
        p.conj_in = conj ? 1 : 0;
      } else if (k == 1) {
        set_in(p, s0);
        set_out(p, s0);
        p.ae = 1;
        p.be = 1;
        D.push_back(Dim{M1, M2, M2, 0, 0, true, M2});
        batch_dims(D, false, false);
        p.mid_mode = MID_BHAT;
        p.cme = 1;
        p.bhat = lp.bhat_rows_ptr;
        // rows above M1/2 read their mirror row backwards (half the table)
        p.mirror_rows = (int32_t)M1;
      } else {
        set_in(p, s0);
        set_out(p, out);
        p.ae = M2;
        p.be = M2 * ls.es_out;
        D.push_back(Dim{M2, 1, ls.es_out, 0, 1, false});
        batch_dims(D, false, true);
        p.store_mode = STORE_CHIRP;
        p.cbe = M2;
        p.conj_out = conj ? 1 : 0;
      }
      if (!set_dims(p, D, tab)) return lines_capacity_error();
      if (k == 0) p.in_tab = tab ? ls.in_tab : nullptr;
      if (k == 2) p.out_tab = tab ? ls.out_tab : nullptr;
      st = stage_table(T, d->inner[k == 1 ? 1 : 0], &p.tw, stream);
      if (st) return st;
      set_inner(p, d->inner[k == 1 ? 1 : 0], prec, la);
      if ((st = attach_pass_twiddle(p, la, T, M, stream))) return st;
      if ((st = attach_chirp_steps(p, la, T, stream))) return st;
      la.lp = p;
      L.push_back(la);
    }
    return FFTAHIP_SUCCESS;
  }

  // multi-pass (Stockham over pass factors), possibly two sweeps (Bluestein)
  const int np = (int)d->F.size();
  const bool blue = d->kind == K_BLUE_MULTI;
  const int total = blue ? 2 * np : np;
  const double2* tlo = nullptr;
  const double2* thi = nullptr;
  int tsh = 0;
  st = two_level(T, M, &tlo, &thi, &tsh, stream);
  if (st) return st;
  // Lines adjacent in memory with strided positions (e.g. the second
  // dimension of a matrix): the scratch keeps the line index fastest as well
  // (element m of line b at b + B*m), so that every pass reads and writes
  // runs of neighbouring lines.  Otherwise scratch holds line b at b*M.
  const bool line_minor = !tab && strided && ls.n[0] > 1 && ls.in_str[0] == 1 && ls.out_str[0] == 1;
  const int64_t es_scr = line_minor ? B : 1;
  for (int i = 0; i < total; ++i) {
    const int s = i % np;
    const int64_t R = d->F[s];
    int64_t Ns = 1;
    for (int u = 0; u < s; ++u) Ns *= d->F[u];
    const int64_t Q = M / (R * Ns);
    const bool first = i == 0, last = i == total - 1;
    Launch la;
    LinePass p = blank_pass();
    set_in(p, first ? in : (((i - 1) % 2 == 0) ? s0 : s1));
    set_out(p, last ? out : ((i % 2 == 0) ? s0 : s1));
    const int64_t ein = first ? ls.es_in : es_scr;
    const int64_t eout = last ? ls.es_out : es_scr;
    p.ae = ein * (M / R);
    p.be = eout * Ns;
    std::vector<Dim> D;
    // k (position within the Ns group) and q (group index)
    D.push_back(Dim{Ns, ein, eout, 1, 1, true});
    D.push_back(Dim{Q, ein * Ns, eout * Ns * R, Ns, Ns * R, false});
    // batch: user side uses the line-set layout, scratch side b*M
    if (tab) {
      Dim bd{ls.n[1], first ? 0 : M, last ? 0 : M, 0, 0, false};
      D.push_back(bd);
      p.in_tab = first ? ls.in_tab : nullptr;
      p.out_tab = last ? ls.out_tab : nullptr;
    } else {
      int64_t sstride = line_minor ? 1 : M;
      for (int b = 0; b < 2; ++b) {
        if (ls.n[b] > 1) {
          D.push_back(Dim{ls.n[b], first ? ls.in_str[b] : sstride, last ? ls.out_str[b] : sstride, 0, 0, false});
        }
        sstride *= ls.n[b];
      }
    }
    if (!set_dims(p, D, tab)) return lines_capacity_error();
    st = stage_table(T, d->inner[s], &p.tw, stream);
    if (st) return st;
    if (Ns > 1) {
      p.pretw = 1;
      p.tw_scale = M / (Ns * R);
      p.tw_lo = tlo;
      p.tw_hi = thi;
      p.tw_sh = tsh;
    }
    p.conj_in = (first && conj) ? 1 : 0;
    p.conj_out = (last && conj) ? 1 : 0;
    if (blue) {
      p.chirp_n = lp.bq;
      p.chirp = lp.chirp_ptr;
      p.bhat = lp.bhat_ptr;
      p.cae = M / R;
      p.cbe = Ns;
      if (first) p.load_mode = LOAD_CHIRP;
      if (i == np - 1) p.store_mode = STORE_BHAT;
      if (last) p.store_mode = STORE_CHIRP;
    }
    if (d->kind == K_RADER_MIXED && s == 0) {
      // the R = Q point DFT of every line by Rader's algorithm: a line of
      // Q - 1 points in LDS
      set_rader(p, lp);
    } else if (d->kind == K_MIXED && s == 0) {
      // the R = Q point DFT of every line as a Bluestein line of lp.M points
      // in LDS: positions t < Q of the line are loaded and stored
      p.chirp_n = lp.bq;
      p.chirp = lp.chirp_ptr;
      p.bhat = lp.bhat_ptr;
      p.load_mode = LOAD_CHIRP;
      p.mid_mode = MID_BHAT;
      p.store_mode = STORE_CHIRP;
      p.cae = 1;
      p.cme = 1;
      p.cbe = 1;
      for (int u = 0; u < 4; ++u) p.ca[u] = p.cb[u] = p.cm[u] = 0;
    }
    set_inner(p, d->inner[s], prec, la);
    set_zero_pad(p, la, prec);
    if ((st = attach_pass_twiddle(p, la, T, M, stream))) return st;
    if ((st = attach_chirp_steps(p, la, T, stream))) return st;
    la.lp = p;
    L.push_back(la);
  }
  (void)B;
  return FFTAHIP_SUCCESS;
}

// Scratch elements per line needed by build_launches.
static fftahip_status scratch_per_line(LengthPlan& lp, int strided, int64_t* elems, int* nbuf) {
  Decomp* d;
  fftahip_status st = get_decomp(lp, strided, &d);
  if (st) return st;
  *elems = d->M;
  if (d->kind == K_SINGLE || d->kind == K_BLUE_SINGLE || d->kind == K_RADER_SINGLE) *nbuf = 0;
  else if (((d->kind == K_MULTI || d->kind == K_MIXED || d->kind == K_RADER_MIXED) && d->F.size() == 2) ||
           d->kind == K_BLUE_3PASS)
    *nbuf = 1;
  else *nbuf = 2;
  return FFTAHIP_SUCCESS;
}

// A launch list with its descriptors resident on the device (uploaded once;
// independent of the arrays it runs on).
struct DevSchedule {
  std::vector<Launch> L;
  DevBuf passes;  // LinePass per launch
  DevBuf zero;    // tile_start of single-group launches
  size_t scratch_need = 0;
};

// Uploads a newly built schedule and waits for everything it depends on:
// the table kernels its build launched on `s` and the descriptor copies.  A
// cached schedule may next run on any stream (the serving loop spreads one
// plan's requests over several non-blocking streams, which do not order
// themselves after the null stream), so nothing of the build may still be
// in flight when this returns.
static fftahip_status upload_schedule(DevSchedule& S, hipStream_t s) {
  std::vector<LinePass> host;
  for (const Launch& la : S.L) host.push_back(la.lp);
  fftahip_status st = S.passes.ensure(std::max<size_t>(1, host.size()) * sizeof(LinePass));
  if (st) return st;
  if ((st = S.zero.ensure(sizeof(uint32_t)))) return st;
  HIP_TRY(hipStreamSynchronize(s));
  if (!host.empty()) HIP_TRY(hipMemcpy(S.passes.p, host.data(), host.size() * sizeof(LinePass), hipMemcpyHostToDevice));
  const uint32_t zero = 0;
  HIP_TRY(hipMemcpy(S.zero.p, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(hipDeviceSynchronize());
  return FFTAHIP_SUCCESS;
}

static fftahip_status run_schedule(const DevSchedule& S, int prec, const BufPtrs& bp, hipStream_t s) {
  for (size_t i = 0; i < S.L.size(); ++i) {
    const Launch& la = S.L[i];
    if (la.op == OP_COPY) {
      if (bp.out != bp.in) HIP_TRY(hipMemcpyAsync(bp.out, bp.in, la.bytes, hipMemcpyDeviceToDevice, s));
      continue;
    }
    if (la.grid <= 0) continue;
    if (la.grid > 0x7fffffffLL) return fail(FFTAHIP_ERR_CAPACITY, "grid too large");
    if (la.reg) {
      HIP_TRY(launch_reg_pass_groups((const LinePass*)S.passes.p + i, (const uint32_t*)S.zero.p, 1, pass_mode(la.lp),
                                     prec, bp, la.grid, s));
      continue;
    }
    HIP_TRY(launch_line_pass_groups((const LinePass*)S.passes.p + i, (const uint32_t*)S.zero.p, 1, pass_variant(la.lp),
                                    prec, la.block, bp, la.lp.xcd_remap, la.lds, la.grid, s));
  }
  return FFTAHIP_SUCCESS;
}

// ------------------------------------------------------------------ Bluestein and Rader spectra

// out = FFT_n(in), forward, in FP64 with a schedule of this library.
static fftahip_status fft64_device(const void* in, void* out, int64_t n, hipStream_t s) {
  LengthPlan sub;
  sub.prec = 64;
  sub.N = n;
  sub.M = n;
  sub.blue = false;
  Tables T64;
  T64.prec = 64;
  int64_t per;
  int nbuf;
  fftahip_status st;
  if ((st = scratch_per_line(sub, 0, &per, &nbuf))) return st;
  DevBuf s0;
  if (nbuf >= 1 && (st = s0.ensure((size_t)per * nbuf * 16))) return st;
  LineSet ls;
  DevSchedule S;
  if ((st = build_launches(sub, T64, ls, BufRef{BUF_IN, 0}, BufRef{BUF_OUT, 0}, BufRef{BUF_SCRATCH, 0},
                           BufRef{BUF_SCRATCH, per}, false, S.L, s)))
    return st;
  if ((st = upload_schedule(S, s))) return st;
  if ((st = run_schedule(S, 64, BufPtrs{in, out, s0.p}, s))) return st;
  // the tables and scratch above are freed on return
  HIP_TRY(hipStreamSynchronize(s));
  return FFTAHIP_SUCCESS;
}

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = (unsigned __int128)r * b % m;
    b = (unsigned __int128)b * b % m;
    e >>= 1;
  }
  return r;
}

// Smallest primitive root of the prime P: g with g^((P-1)/q) != 1 mod P for
// every prime q dividing P - 1.
static int64_t primitive_root(int64_t P) {
  if (P == 2) return 1;
  std::vector<int64_t> qs = prime_factors(P - 1);
  qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
  for (int64_t g = 2; g < P; ++g) {
    bool ok = true;
    for (int64_t q : qs) ok = ok && powmod((uint64_t)g, (uint64_t)((P - 1) / q), (uint64_t)P) != 1;
    if (ok) return g;
  }
  return 0;
}

// Rader's algorithm for the prime P = lp.bq (C. M. Rader, 1968): with g a
// primitive root, a'_q = x_{g^-q} and b_r = w_P^{g^r} (q, r < P - 1),
//   X_0 = x_0 + sum_q a'_q,   X_{g^m} = x_0 + (a' (*) b)_m,
// a cyclic convolution of P - 1 points.  The line pass computes it as
// conj(FFT(conj(FFT(a') * B))) with B = FFT(b) / (P - 1), generated here in
// FP64 and rounded once; rtab[t] = (-log_g t) mod (P - 1) places element t.
static fftahip_status prepare_rader(LengthPlan& lp, hipStream_t s) {
  const int64_t P = lp.bq, L = P - 1;
  lp.rg = primitive_root(P);
  if (lp.rg <= 0 || L > 65535) return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "no Rader table for " + std::to_string(P));
  std::vector<uint16_t> tab((size_t)P, 0);
  uint64_t cur = 1;
  for (int64_t e = 0; e < L; ++e) {
    tab[cur] = (uint16_t)((L - e) % L);
    cur = cur * (uint64_t)lp.rg % (uint64_t)P;
  }
  fftahip_status st;
  if ((st = lp.rtab.ensure((size_t)P * sizeof(uint16_t)))) return st;
  HIP_TRY(hipMemcpy(lp.rtab.p, tab.data(), (size_t)P * sizeof(uint16_t), hipMemcpyHostToDevice));
  if ((st = lp.bhat.ensure((size_t)L * csize(lp.prec)))) return st;
  DevBuf b64, spec;
  if ((st = b64.ensure((size_t)L * sizeof(double2)))) return st;
  if ((st = spec.ensure((size_t)L * sizeof(double2)))) return st;
  HIP_TRY(launch_gen_rader_b((double2*)b64.p, P, lp.rg, s));
  if ((st = fft64_device(b64.p, spec.p, L, s))) return st;
  HIP_TRY(launch_finish_bhat(lp.bhat.p, (const double2*)spec.p, lp.prec, L, s));
  HIP_TRY(hipStreamSynchronize(s));
  lp.rtab_ptr = (const uint16_t*)lp.rtab.p;
  lp.bhat_ptr = lp.bhat.p;
  lp.blue_ready = true;
  return FFTAHIP_SUCCESS;
}

// bhat = FFT_M(b_ext) / M computed with an FP64 plan of the same machinery.
static fftahip_status prepare_bluestein(LengthPlan& lp, Tables& T, hipStream_t s) {
  const int64_t N = lp.bq, M = lp.M;
  fftahip_status st = lp.chirp.ensure((size_t)N * csize(lp.prec));
  if (st) return st;
  st = lp.bhat.ensure((size_t)M * csize(lp.prec));
  if (st) return st;
  DevBuf c64, bext, spec;
  if ((st = c64.ensure((size_t)N * sizeof(double2)))) return st;
  if ((st = bext.ensure((size_t)M * sizeof(double2)))) return st;
  if ((st = spec.ensure((size_t)M * sizeof(double2)))) return st;
  HIP_TRY(launch_gen_chirp(lp.chirp.p, (double2*)c64.p, lp.prec, N, s));
  HIP_TRY(launch_gen_chirp_ext((double2*)bext.p, (const double2*)c64.p, N, M, s));
  if ((st = fft64_device(bext.p, spec.p, M, s))) return st;
  HIP_TRY(launch_finish_bhat(lp.bhat.p, (const double2*)spec.p, lp.prec, M, s));
  int64_t M1, M2;
  if (!lp.mixed && blue3_split(M, lp.prec, &M1, &M2)) {
    // rows 0 .. M1/2 of the row-permuted spectrum; the row pass reads the
    // others mirrored (B[M - k] = B[k])
    const int64_t rows = M1 / 2 + 1;
    if ((st = lp.bhat_rows.ensure((size_t)(rows * M2) * csize(lp.prec)))) return st;
    HIP_TRY(launch_permute_rows(lp.bhat_rows.p, lp.bhat.p, lp.prec, M1, M2, rows, s));
    lp.bhat_rows_ptr = lp.bhat_rows.p;
  }
  HIP_TRY(hipStreamSynchronize(s));
  lp.chirp_ptr = lp.chirp.p;
  lp.bhat_ptr = lp.bhat.p;
  lp.blue_ready = true;
  (void)T;
  return FFTAHIP_SUCCESS;
}

// Bluestein length for an in-LDS line: among 5-smooth M in [2n-1, 1.3(2n-1)]
// that fit one workgroup, the one with the least LDS work, M * (stages + 1.5)
// (a slightly longer M with fewer, larger radices is often cheaper).  0 if
// none fits.
static int64_t best_line_length(int64_t n, int prec) {
  const int64_t lo = 2 * n - 1;
  const int64_t hi = std::min<int64_t>(single_max(prec), lo + (lo * 3) / 10);
  int64_t best = 0;
  double best_cost = 1e300;
  for (int64_t M = lo; M <= hi; ++M) {
    int64_t m = M;
    for (int q : {2, 3, 5})
      while (m % q == 0) m /= q;
    if (m != 1) continue;
    Inner in;
    if (!make_inner(M, in)) continue;
    const double cost = (double)M * ((double)in.rad.size() + 1.5);
    if (cost < best_cost) {
      best_cost = cost;
      best = M;
    }
  }
  return best;
}

// Work of an in-LDS line FFT in stage units per point: 1 per codelet
// stage, P/div per direct-DFT stage of a prime P (its O(P) work per point;
// env FFTAHIP_BIG_DIV, default 8).
static double inner_work(const Inner& in) {
  const double div = (double)std::max<int64_t>(1, env_i64("FFTAHIP_BIG_DIV", 8));
  double w = 0;
  for (int r : in.rad) w += small_radix(r) ? 1.0 : r / div;
  return w;
}

// Rader (a line of P - 1 points) for the convolution prime P, when P - 1
// has an in-LDS schedule, fits one workgroup with its side slots, and its
// sweeps cost no more than the Bluestein line of blue_M points (0: no
// Bluestein line fits, so Bluestein would take three passes through HBM).
// Env FFTAHIP_RADER: 0 turns Rader off, 1 allows direct-DFT stages in
// P - 1, 2 (default) takes only 13-smooth P - 1.
static bool rader_pick(int64_t P, int prec, int64_t blue_M) {
  const int64_t mode = env_i64("FFTAHIP_RADER", 2);
  if (mode == 0) return false;
  if (P < 5 || P - 1 > 65535 || !is_prime(P)) return false;
  const int64_t L = P - 1;
  if (L + kRaderSide > single_max(prec)) return false;
  Inner in;
  if (!make_inner(L, in, prec, mode == 2 ? 17 : kInternalCutoff)) return false;
  Inner b;
  if (blue_M > 0 && make_inner(blue_M, b, prec) && (double)L * inner_work(in) > (double)blue_M * inner_work(b))
    return false;
  return true;
}

static fftahip_status init_length(LengthPlan& lp, int prec, int64_t N, int64_t cutoff) {
  lp.prec = prec;
  lp.N = N;
  lp.cutoff = cutoff;
  lp.ffta.clear();
  ffta_nodes(lp.ffta, N, cutoff, 1, 1);
  // primes p < cutoff (and 2, 3) run as direct DFT stages, as FFTA runs
  // primes below BLUESTEIN_CUTOFF as DFT leaves (src/callgraph.jl:95-100);
  // any other prime factor goes through a convolution
  lp.blue = !direct_smooth(N, cutoff);
  lp.mixed = false;
  lp.rader = false;
  lp.rg = 0;
  lp.S = 1;
  lp.bq = N;
  auto smooth5 = [](int64_t v) {
    for (int q : {2, 3, 5})
      while (v % q == 0) v /= q;
    return v == 1;
  };
  if (lp.blue) {
    // N = S * Q with S the 13-smooth part: when Q's Bluestein line fits in
    // LDS, run Q as a Bluestein pass and S as Stockham passes (FFTA splits
    // such lengths the same way, src/callgraph.jl:102-124, with Bluestein at
    // the prime leaf); otherwise Bluestein on the whole length.
    int64_t S = 1, rest = N, m = N;
    for (int64_t q = 2; q * q <= m; ++q) {
      if (m % q) continue;
      int64_t pk = 1;
      while (m % q == 0) { m /= q; pk *= q; }
      if (direct_prime(q, cutoff)) { S *= pk; rest /= pk; }
    }
    if (m > 1 && direct_prime(m, cutoff)) { S *= m; rest /= m; }
    if (S > 1) {
      const int64_t M = best_line_length(rest, prec);
      if (rader_pick(rest, prec, M)) {
        lp.mixed = true;
        lp.rader = true;
        lp.S = S;
        lp.bq = rest;
        lp.M = rest - 1;
      } else if (M) {
        lp.mixed = true;
        lp.S = S;
        lp.bq = rest;
        lp.M = M;
      }
    }
  }
  if (lp.blue && !lp.mixed && rader_pick(N, prec, best_line_length(N, prec))) {
    lp.rader = true;
    lp.bq = N;
    lp.M = N - 1;
  } else if (lp.blue && !lp.mixed) {
    // smallest 5-smooth length >= 2N-1 that runs in one kernel or three fused
    // passes (FFTA pads to a power of two, src/algos.jl:345; any length >=
    // 2N-1 gives the same convolution)
    int64_t chosen = best_line_length(N, prec);
    int64_t first = 0;
    // three fused passes: among the splittable 5-smooth M up to 15% above
    // the first, the least M * row_cost of the split blue3_split takes
    // (3.56 against 3.61 ms per batched step with the power-of-two weight);
    // env FFTAHIP_BLUE_COST=0 takes the first
    const bool costed = env_i64("FFTAHIP_BLUE_COST", 1) != 0;
    double best_cost = 1e300;
    int64_t limit = chosen ? 0 : 4 * (2 * N - 1);  // one kernel: no three-pass search
    for (int64_t M = 2 * N - 1; M <= limit; ++M) {
      if (!smooth5(M)) continue;
      if (!first) first = M;
      int64_t a, b;
      if (!blue3_split(M, prec, &a, &b)) continue;
      if (!chosen) limit = std::min(limit, M + M * 15 / 100);
      if (!costed) {
        chosen = M;
        break;
      }
      Inner row;
      make_inner(b, row, prec);
      const double cost = (double)M * row_cost(row, b);
      if (cost < best_cost) {
        best_cost = cost;
        chosen = M;
      }
    }
    lp.M = chosen ? chosen : first;
  } else if (!lp.blue) {
    lp.M = N;
  }
  // line indices, in-line positions and pass-twiddle exponents are 32-bit
  if (lp.M >= (int64_t(1) << 31))
    return fail(FFTAHIP_ERR_CAPACITY, "transform length " + std::to_string(N) + " needs more than 2^31 - 1 points");
  return FFTAHIP_SUCCESS;
}

static fftahip_status prepare_length(LengthPlan& lp, Tables& T, hipStream_t s) {
  if (lp.blue && g_dry) {
    if (lp.rader) {
      lp.rg = primitive_root(lp.bq);
      lp.rtab_ptr = (const uint16_t*)token(TK_RTAB, lp.bq);
    } else {
      lp.chirp_ptr = token(TK_CHIRP, lp.bq);
      lp.bhat_rows_ptr = token(TK_BHATROWS, lp.M);
    }
    lp.bhat_ptr = token(TK_BHAT, lp.M);
    lp.blue_ready = true;
    return FFTAHIP_SUCCESS;
  }
  if (lp.blue && !lp.blue_ready) return lp.rader ? prepare_rader(lp, s) : prepare_bluestein(lp, T, s);
  return FFTAHIP_SUCCESS;
}

}  // namespace fftahip

using namespace fftahip;

// ------------------------------------------------------------------ plans

struct fftahip_plan_s {
  int prec = 32;
  int dir = -1;
  int64_t cutoff = 73;
  double scale = 1.0;                             // applied by the pass that writes the result
  std::vector<int64_t> region;                    // sorted, 1-based
  std::vector<std::unique_ptr<LengthPlan>> lens;  // one per region dim
  Tables tables;
  std::map<std::vector<int64_t>, std::unique_ptr<DevSchedule>> scheds;  // per array shape
  // launches of `count` vectors at arbitrary addresses (1-D plans), per
  // count: built over offset tables whose addresses are patched per call
  struct Many {
    std::vector<Launch> L;
    size_t scratch = 0;
  };
  std::map<int, Many> many;
  DevBuf scratch;
  DevBuf hin, hout;  // staging for execute_host
  std::string last_schedule;
};

struct fftahip_batch_s {
  int prec = 32;
  int64_t count = 0;
  std::vector<int64_t> lengths, offsets;
  std::vector<int> dirs;
  int64_t total = 0;  // elements spanned
  std::map<int64_t, std::unique_ptr<LengthPlan>> lens;
  Tables tables;
  struct Group {
    int64_t N;
    int dir;
    std::vector<int64_t> idx;  // transform indices
  };
  std::vector<Group> groups;
  DevBuf offtab;   // per group offsets (device), concatenated
  DevBuf scratch;  // per group scratch, concatenated
  const int64_t* offtab_ptr = nullptr;  // offtab.p, or a token in dry-run mode
  char* scratch_ptr = nullptr;
  std::vector<int64_t> host_offtab;
  std::vector<std::vector<Launch>> per_group;  // built launches (dry-run inspection)
  DevBuf passes;   // device LinePass arrays per launch
  DevBuf tiles;
  struct Phase {
    bool reg;  // register line passes (block kRegBlock, no LDS)
    int block;
    int mode;  // pass_mode for register phases, pass_variant for in-LDS phases
    size_t lds;
    int64_t grid;
    int ngroups;
    size_t pass_off, tile_off;  // element offsets into passes / tiles
    std::vector<int> deps;      // earlier launches holding the previous pass of one of its groups
    int pass = 0;               // which pass of its groups the launch runs
    std::vector<int> members;   // its groups, in tile order (dry-run inspection)
  };
  std::vector<Phase> phases;
  std::vector<LinePass> host_passes;
  std::vector<uint32_t> host_tiles;
  hipGraphExec_t graph = nullptr;
  const void* graph_in = nullptr;
  void* graph_out = nullptr;
  // launches run on nstreams (<= kStreams) streams, ordered only by their
  // dependencies (env FFTAHIP_BATCH_STREAMS, default 4: one per hardware
  // queue HIP gives a process by default)
  static constexpr int kStreams = 8;
  int nstreams = 4;
  hipStream_t streams[kStreams] = {};
  std::vector<hipEvent_t> events;  // one per launch, then fork, then one join per stream
  ~fftahip_batch_s() {
    if (graph) (void)hipGraphExecDestroy(graph);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (hipStream_t q : streams)
      if (q) (void)hipStreamDestroy(q);
  }
};

static fftahip_status plan_shape_schedule(fftahip_plan plan, int ndims, const int64_t* dims, DevSchedule** out,
                                          hipStream_t s);

extern "C" {

const char* fftahip_version(void) { return "fftahip 0.1.0 (gfx950)"; }

const char* fftahip_status_string(fftahip_status s) {
  switch (s) {
    case FFTAHIP_SUCCESS: return "success";
    case FFTAHIP_ERR_DIMENSION_MISMATCH: return "DimensionMismatch";
    case FFTAHIP_ERR_ARGUMENT: return "ArgumentError";
    case FFTAHIP_ERR_DOMAIN: return "DomainError";
    case FFTAHIP_ERR_HIP: return "HIP runtime error";
    case FFTAHIP_ERR_ALLOC: return "allocation failed";
    case FFTAHIP_ERR_NOT_IMPLEMENTED: return "not implemented";
    case FFTAHIP_ERR_CAPACITY: return "capacity exceeded";
    case FFTAHIP_ERR_NOT_READY: return "not ready";
  }
  return "unknown status";
}

const char* fftahip_last_error(void) { return g_err.c_str(); }

fftahip_status fftahip_plan_create(fftahip_plan* plan, int precision, int direction, int ndims, const int64_t* dims,
                                   int nregion, const int64_t* region, int64_t bluestein_cutoff) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "plan pointer is NULL");
  *plan = nullptr;
  if (precision != 32 && precision != 64) return fail(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (direction != -1 && direction != 1) return fail(FFTAHIP_ERR_ARGUMENT, "direction must be -1 or +1");
  if (nregion < 1) return fail(FFTAHIP_ERR_ARGUMENT, "region must name at least one dimension");
  std::vector<int64_t> reg(region, region + nregion);
  std::sort(reg.begin(), reg.end());
  auto p = std::make_unique<fftahip_plan_s>();
  p->prec = precision;
  p->dir = direction;
  p->cutoff = bluestein_cutoff;
  p->region = reg;
  p->tables.prec = precision;
  for (int64_t r : reg) {
    if (r < 1) return fail(FFTAHIP_ERR_ARGUMENT, "arraysize: dimension out of range");
    const int64_t n = r <= ndims ? dims[r - 1] : 1;  // size(x, r) is 1 beyond ndims
    if (n <= 0) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Array length must be strictly positive");
    auto lp = std::make_unique<LengthPlan>();
    init_length(*lp, precision, n, bluestein_cutoff);
    p->lens.push_back(std::move(lp));
  }
  hipStream_t s = nullptr;
  for (auto& lp : p->lens) {
    fftahip_status st = prepare_length(*lp, p->tables, s);
    if (st) return st;
  }
  // the schedule of the creation shape is built and uploaded now (with a
  // device present), so that the first execution, a stream capture and
  // fftahip_execute_noalloc find it ready
  int64_t total = 1;
  for (int i = 0; i < ndims; ++i) total *= dims[i];
  if (!g_dry && total > 0 && dev_info().ok) {
    DevSchedule* S;
    fftahip_status st = plan_shape_schedule(p.get(), ndims, dims, &S, s);
    if (st) return st;
  }
  *plan = p.release();
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_plan_size(fftahip_plan plan, int i, int64_t* size_out) {
  if (!plan || !size_out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  if (i < 1) return fail(FFTAHIP_ERR_DOMAIN, "No non-positive dimensions");
  if (i > (int)plan->lens.size()) {
    *size_out = 1;
    return FFTAHIP_SUCCESS;
  }
  *size_out = plan->lens[i - 1]->N;
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_plan_rank(fftahip_plan plan, int* rank_out) {
  if (!plan || !rank_out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  *rank_out = (int)plan->lens.size();
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_plan_precision(fftahip_plan plan, int* out) {
  if (!plan || !out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  *out = plan->prec;
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_plan_direction(fftahip_plan plan, int* out) {
  if (!plan || !out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  *out = plan->dir;
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_plan_set_scale(fftahip_plan plan, double scale) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  if (!std::isfinite(scale)) return fail(FFTAHIP_ERR_ARGUMENT, "scale must be finite");
  if (scale != plan->scale) {
    plan->scale = scale;
    plan->scheds.clear();  // schedules carry the factor in their last pass
  }
  return FFTAHIP_SUCCESS;
}

static std::string tuple_str(const int64_t* d, int n) {
  std::ostringstream os;
  os << "(";
  for (int i = 0; i < n; ++i) {
    if (i) os << ", ";
    os << d[i];
  }
  if (n == 1) os << ",";
  os << ")";
  return os.str();
}

static std::string axes_str(const int64_t* d, int n) {
  std::ostringstream os;
  os << "(";
  for (int i = 0; i < n; ++i) {
    if (i) os << ", ";
    os << "Base.OneTo(" << d[i] << ")";
  }
  if (n == 1) os << ",";
  os << ")";
  return os.str();
}

fftahip_status fftahip_plan_check(fftahip_plan plan, int nregion, const int64_t* region, int ndims_x,
                                  const int64_t* dims_x, int ndims_y, const int64_t* dims_y) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  const int M = (int)plan->lens.size();
  std::vector<int64_t> psize(M);
  for (int i = 0; i < M; ++i) psize[i] = plan->lens[i]->N;
  auto same = [](int na, const int64_t* a, int nb, const int64_t* b) {
    if (na != nb) return false;
    for (int i = 0; i < na; ++i)
      if (a[i] != b[i]) return false;
    return true;
  };
  const bool axes_equal = same(ndims_x, dims_x, ndims_y, dims_y);
  if (M == 1 && ndims_x == 1 && ndims_y == 1) {
    // src/plan.jl:157-164
    if (!axes_equal)
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "input array has axes " + axes_str(dims_x, 1) +
                                                      ", but output array has axes " + axes_str(dims_y, 1));
    if (psize[0] != dims_x[0])
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "plan has axes " + tuple_str(psize.data(), 1) +
                                                      ", but input array has axes " + tuple_str(dims_x, 1));
    return FFTAHIP_SUCCESS;
  }
  if (M == 1) {
    // src/plan.jl:170-182
    if (!axes_equal)
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "input array has axes " + axes_str(dims_x, ndims_x) +
                                                      ", but output array has axes " + axes_str(dims_y, ndims_y));
    const int64_t R1 = nregion >= 1 ? region[0] : plan->region[0];
    const int64_t xlen = (R1 >= 1 && R1 <= ndims_x) ? dims_x[R1 - 1] : 1;
    if (psize[0] != xlen)
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "plan has size " + std::to_string(psize[0]) +
                                                      ", but input array has size " + std::to_string(xlen) +
                                                      " along region " + std::to_string(R1));
    return FFTAHIP_SUCCESS;
  }
  const int N = ndims_x;
  if (!same(ndims_x, dims_x, ndims_y, dims_y))
    return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "input array has axes " + axes_str(dims_x, ndims_x) +
                                                    ", but output array has axes " + axes_str(dims_y, ndims_y));
  if (M == N) {
    // src/plan.jl:216-222
    if (!same(M, psize.data(), N, dims_x))
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "plan has size " + tuple_str(psize.data(), M) +
                                                      ", but input array has size " + tuple_str(dims_x, N));
    bool valid = nregion == N;
    for (int i = 0; valid && i < nregion; ++i) valid = region[i] == i + 1;
    if (!valid) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Plan region is outside array dimensions.");
    return FFTAHIP_SUCCESS;
  }
  // src/plan.jl:266-272 (also reached for M > N)
  bool valid = nregion == M;
  for (int i = 1; valid && i < nregion; ++i) valid = region[i - 1] < region[i];
  if (!valid) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Region is invalid.");
  if (M > N || region[0] < 1 || region[nregion - 1] > N)
    return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Plan region is outside array dimensions.");
  for (int i = 0; i < M; ++i)
    if (dims_x[region[i] - 1] != psize[i])
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "plan has size " + tuple_str(psize.data(), M) +
                                                      ", but input array has size " + tuple_str(dims_x, N));
  return FFTAHIP_SUCCESS;
}

// Lines along dimension d (0-based) of a column-major array.
static LineSet lines_along(int ndims, const int64_t* dims, int d) {
  int64_t pre = 1, post = 1;
  for (int i = 0; i < d && i < ndims; ++i) pre *= dims[i];
  for (int i = d + 1; i < ndims; ++i) post *= dims[i];
  const int64_t n = d < ndims ? dims[d] : 1;
  LineSet ls;
  ls.es_in = ls.es_out = pre;
  ls.n[0] = pre;
  ls.in_str[0] = ls.out_str[0] = 1;
  ls.n[1] = post;
  ls.in_str[1] = ls.out_str[1] = pre * n;
  return ls;
}

// The launches of mul!(y, p, x) for an array of shape dims: region dims in
// ascending order, the first reading x and every later one working in y
// (src/plan.jl:232-253).  The scratch buffer must hold *scratch_need bytes.
static fftahip_status plan_schedule(fftahip_plan plan, int ndims, const int64_t* dims, std::vector<Launch>& L,
                                    size_t* scratch_need, hipStream_t s) {
  int64_t total = 1;
  for (int i = 0; i < ndims; ++i) total *= dims[i];
  size_t need = 0;
  bool wrote = false;  // y already holds the partial result
  for (size_t k = 0; k < plan->region.size(); ++k) {
    const int d = (int)plan->region[k] - 1;
    if (d >= ndims) {
      // size(x, d) == 1 beyond ndims: a length-1 transform is the identity
      if (!wrote) {
        Launch cp;
        cp.op = OP_COPY;
        cp.lp = blank_pass();
        cp.bytes = (size_t)total * csize(plan->prec);
        L.push_back(cp);
      }
      wrote = true;
      continue;
    }
    LineSet ls = lines_along(ndims, dims, d);
    int64_t per;
    int nbuf;
    fftahip_status st = scratch_per_line(*plan->lens[k], (ls.es_in != 1) ? 1 : 0, &per, &nbuf);
    if (st) return st;
    need = std::max(need, (size_t)(per * ls.count() * nbuf) * csize(plan->prec));
    const BufRef src = wrote ? BufRef{BUF_OUT, 0} : BufRef{BUF_IN, 0};
    wrote = true;
    st = build_launches(*plan->lens[k], plan->tables, ls, src, BufRef{BUF_OUT, 0}, BufRef{BUF_SCRATCH, 0},
                        BufRef{BUF_SCRATCH, per * ls.count()}, plan->dir == 1, L, s);
    if (st) return st;
  }
  // a scaled plan (ifft's 1/N, AbstractFFTs' ScaledPlan) scales in the last
  // pass: it writes every element of the result once, after all earlier ones
  if (plan->scale != 1.0) {
    Launch* last = nullptr;
    for (auto& la : L)
      if (la.op == OP_PASS) last = &la;
    if (!last)
      return fail(FFTAHIP_ERR_NOT_IMPLEMENTED, "a scale factor on a transform without a pass (all region lengths 1)");
    last->lp.scale *= plan->scale;
  }
  *scratch_need = need;
  return FFTAHIP_SUCCESS;
}

// Shape schedule of a plan (built and uploaded on first use of the shape).
static fftahip_status plan_shape_schedule(fftahip_plan plan, int ndims, const int64_t* dims, DevSchedule** out,
                                          hipStream_t s) {
  std::vector<int64_t> key(dims, dims + ndims);
  auto it = plan->scheds.find(key);
  if (it == plan->scheds.end()) {
    auto S = std::make_unique<DevSchedule>();
    fftahip_status st = plan_schedule(plan, ndims, dims, S->L, &S->scratch_need, s);
    if (st) return st;
    if ((st = upload_schedule(*S, s))) return st;
    it = plan->scheds.emplace(key, std::move(S)).first;
  }
  *out = it->second.get();
  return FFTAHIP_SUCCESS;
}

// The argument checks every execution applies before it launches anything:
// the shape must fit the plan as mul!(y, p, x) requires (x and y share dims
// here), with the plan's own region (src/plan.jl:157-164, 170-182,
// 210-222, 260-272).
static fftahip_status check_shape(fftahip_plan plan, int ndims, const int64_t* dims) {
  if (ndims < 0 || (ndims > 0 && !dims)) return fail(FFTAHIP_ERR_ARGUMENT, "dims is NULL");
  for (int i = 0; i < ndims; ++i)
    if (dims[i] < 0) return fail(FFTAHIP_ERR_ARGUMENT, "negative dimension");
  return fftahip_plan_check(plan, (int)plan->region.size(), plan->region.data(), ndims, dims, ndims, dims);
}

fftahip_status fftahip_execute(fftahip_plan plan, int ndims, const int64_t* dims, const void* d_in, void* d_out,
                               void* stream) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  if (fftahip_status cs = check_shape(plan, ndims, dims)) return cs;
  hipStream_t s = (hipStream_t)stream;
  int64_t total = 1;
  for (int i = 0; i < ndims; ++i) total *= dims[i];
  if (total == 0) return FFTAHIP_SUCCESS;
  DevSchedule* Sp;
  fftahip_status st = plan_shape_schedule(plan, ndims, dims, &Sp, s);
  if (st) return st;
  DevSchedule& S = *Sp;
  st = plan->scratch.ensure(S.scratch_need);
  if (st) return st;
  return run_schedule(S, plan->prec, BufPtrs{d_in, d_out, plan->scratch.p}, s);
}

fftahip_status fftahip_execute_host(fftahip_plan plan, int ndims, const int64_t* dims, const void* h_in,
                                    void* h_out) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  if (fftahip_status cs = check_shape(plan, ndims, dims)) return cs;
  int64_t total = 1;
  for (int i = 0; i < ndims; ++i) total *= dims[i];
  const size_t bytes = (size_t)total * csize(plan->prec);
  if (bytes == 0) return FFTAHIP_SUCCESS;
  fftahip_status st;
  if ((st = plan->hin.ensure(bytes))) return st;
  if ((st = plan->hout.ensure(bytes))) return st;
  HIP_TRY(hipMemcpy(plan->hin.p, h_in, bytes, hipMemcpyHostToDevice));
  if ((st = fftahip_execute(plan, ndims, dims, plan->hin.p, plan->hout.p, nullptr))) return st;
  HIP_TRY(hipMemcpy(h_out, plan->hout.p, bytes, hipMemcpyDeviceToHost));
  return FFTAHIP_SUCCESS;
}

// ------------------------------------------------------------------ execution without allocation
// fftahip_execute_noalloc takes its scratch from `scratch`, a fixed area of
// device memory inside the library (every GPU the library runs on holds its
// own copy until the process ends).  The array's name is a host-side handle,
// not a device address: its address on the current device is looked up once
// per device with hipGetSymbolAddress and only that pointer is passed to the
// kernels.  Every pass writes each scratch element before any pass reads it,
// so the area never needs clearing; calls that overlap on one GPU share it
// and race.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::once_flag g_scratch_once[kMaxScratchDevices];
static void* g_scratch_ptr[kMaxScratchDevices];
static hipError_t g_scratch_err[kMaxScratchDevices];

// Address of `scratch` on the current device (resolved once per device).
static fftahip_status noalloc_scratch(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(FFTAHIP_ERR_CAPACITY, "device index " + std::to_string(dev) + " beyond the no-alloc table");
  std::call_once(g_scratch_once[dev], [dev] {
    void* ptr = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? ptr : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
    return fail(FFTAHIP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[dev]));
  *out = g_scratch_ptr[dev];
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_execute_noalloc(fftahip_plan plan, int ndims, const int64_t* dims, const void* d_in,
                                       void* d_out, void* stream) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  if (fftahip_status cs = check_shape(plan, ndims, dims)) return cs;
  int64_t total = 1;
  for (int i = 0; i < ndims; ++i) total *= dims[i];
  if (total == 0) return FFTAHIP_SUCCESS;
  // building a schedule uploads it (an allocation): only shapes whose
  // schedule exists -- the creation shape, or one executed before -- run here
  std::vector<int64_t> key(dims, dims + ndims);
  auto it = plan->scheds.find(key);
  if (it == plan->scheds.end())
    return fail(FFTAHIP_ERR_NOT_READY,
                "no schedule for array shape " + tuple_str(dims, ndims) +
                    " on this plan: create the plan with this shape or run fftahip_execute on it once");
  const DevSchedule& S = *it->second;
  if (S.scratch_need > kNoAllocScratchBytes)
    return fail(FFTAHIP_ERR_CAPACITY, "this transform needs " + std::to_string(S.scratch_need) +
                                          " bytes of scratch; fftahip_execute_noalloc has " +
                                          std::to_string(kNoAllocScratchBytes));
  void* scr = nullptr;
  if (S.scratch_need > 0) {
    fftahip_status st = noalloc_scratch(&scr);
    if (st) return st;
  }
  return run_schedule(S, plan->prec, BufPtrs{d_in, d_out, scr}, (hipStream_t)stream);
}

// ------------------------------------------------------------------ real transforms
// plan_rfft / plan_brfft (src/plan.jl:102-150) and their application
// (src/plan.jl:386-583).  1 region dim d1: even n = 2m runs an m-point
// complex FFT on packed pairs with the split step (rfft after it, brfft
// before it); odd n runs an n-point complex FFT on the complexified input
// (rfft, keeping n/2 + 1 outputs) or on the Hermitian extension (brfft,
// keeping the real part).  2 region dims d1 < d2: rfft along d1, then a
// complex FFT along d2 in place; brfft: a complex backward FFT along d2 into
// scratch, then brfft along d1.

struct fftahip_rplan_s {
  int prec = 32;
  int dir = -1;  // -1 rfft (real in), +1 brfft (complex in, real out)
  int64_t n = 1; // real length along d1
  int d1 = 0, d2 = -1;
  std::vector<int64_t> in_dims, out_dims;
  std::unique_ptr<LengthPlan> lp1, lp2;
  Tables tables;
  DevSchedule S1, S2;
  DevBuf scratch;
  DevBuf hin, hout;
  // scratch regions (complex elements)
  int64_t offA = 0, offB = 0, offC = 0;
  RealGeom pre{}, post{};
  int pre_op = -1, post_op = -1;
  int64_t m = 0;  // complex line length of S1
  const double2* lo = nullptr;
  const double2* hi = nullptr;
  int sh = 0;
  bool s1_from_c = false;  // 2-D brfft: S1's split step reads the d2-transformed copy
};

static RealGeom real_geom(int64_t P, int64_t T, int64_t Q, int64_t in_t, int64_t in_q, int64_t out_t,
                          int64_t out_q) {
  RealGeom g;
  g.P = P;
  g.T = T;
  g.Q = Q;
  g.in_t = in_t;
  g.in_q = in_q;
  g.out_t = out_t;
  g.out_q = out_q;
  g.divP = make_fastdiv((uint32_t)std::max<int64_t>(P, 1));
  g.divT = make_fastdiv((uint32_t)std::max<int64_t>(T, 1));
  return g;
}

static LineSet line_set(int64_t P, int64_t Q, int64_t in_es, int64_t in_q, int64_t out_es, int64_t out_q) {
  LineSet ls;
  ls.n[0] = P;
  ls.in_str[0] = ls.out_str[0] = 1;
  ls.n[1] = Q;
  ls.in_str[1] = in_q;
  ls.out_str[1] = out_q;
  ls.es_in = in_es;
  ls.es_out = out_es;
  return ls;
}

fftahip_status fftahip_rplan_create(fftahip_rplan* plan, int precision, int direction, int ndims,
                                    const int64_t* dims, int64_t len, int nregion, const int64_t* region,
                                    int64_t bluestein_cutoff) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "plan pointer is NULL");
  *plan = nullptr;
  if (precision != 32 && precision != 64) return fail(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (direction != -1 && direction != 1) return fail(FFTAHIP_ERR_ARGUMENT, "direction must be -1 or +1");
  if (nregion != 1 && nregion != 2) return fail(FFTAHIP_ERR_ARGUMENT, "only supports 1D and 2D FFTs");
  std::vector<int64_t> reg(region, region + nregion);
  std::sort(reg.begin(), reg.end());
  for (int64_t r : reg)
    if (r < 1) return fail(FFTAHIP_ERR_ARGUMENT, "arraysize: dimension out of range");
  // dims padded with 1s up to the largest region dim (size(x, r) == 1 beyond ndims)
  int nd = ndims;
  for (int64_t r : reg) nd = std::max<int>(nd, (int)r);
  std::vector<int64_t> D(nd, 1);
  for (int i = 0; i < ndims; ++i) D[i] = dims[i];
  for (int64_t v : D)
    if (v < 0) return fail(FFTAHIP_ERR_ARGUMENT, "negative dimension");
  auto p = std::make_unique<fftahip_rplan_s>();
  p->prec = precision;
  p->dir = direction;
  p->tables.prec = precision;
  p->d1 = (int)reg[0] - 1;
  p->d2 = nregion == 2 ? (int)reg[1] - 1 : -1;
  const int d1 = p->d1, d2 = p->d2;
  const int64_t n = direction == -1 ? D[d1] : len;
  if (n <= 0) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Array length must be strictly positive");
  p->n = n;
  const int64_t h = n / 2 + 1;
  std::vector<int64_t> X = D, Y = D;  // input / output shapes
  if (direction == -1) {
    Y[d1] = h;
  } else {
    if (D[d1] != h) {
      std::ostringstream os;
      if (nregion == 1)
        os << "real 1D plan has size " << n << ". Dimension of input array along region " << d1 + 1
           << " should have size " << h << ", but has size " << D[d1];
      else
        os << "real 2D plan has size (" << n << ", " << D[d2] << "). Transform dimensions of input array are ("
           << D[d1] << ", " << D[d2] << ") but should be (" << h << ", " << D[d2] << ")";
      return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, os.str());
    }
    Y[d1] = n;
  }
  if (d2 >= 0 && D[d2] <= 0) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Array length must be strictly positive");
  p->in_dims = X;
  p->out_dims = Y;
  int64_t P = 1, Q = 1;
  for (int i = 0; i < d1; ++i) P *= D[i];
  for (int i = d1 + 1; i < nd; ++i) Q *= D[i];
  int64_t totalX = 1;
  for (int i = 0; i < nd; ++i) totalX *= X[i];
  if (P * std::max(n, h) * Q >= (int64_t(1) << 31))
    return fail(FFTAHIP_ERR_CAPACITY, "real transform needs more than 2^31 - 1 elements");
  const bool even = n % 2 == 0;
  const int64_t m = even ? n / 2 : n;
  p->m = m;
  p->lp1 = std::make_unique<LengthPlan>();
  fftahip_status st = init_length(*p->lp1, precision, m, bluestein_cutoff);
  if (st) return st;
  hipStream_t s = nullptr;
  if ((st = prepare_length(*p->lp1, p->tables, s))) return st;
  if (d2 >= 0) {
    p->lp2 = std::make_unique<LengthPlan>();
    if ((st = init_length(*p->lp2, precision, D[d2], bluestein_cutoff))) return st;
    if ((st = prepare_length(*p->lp2, p->tables, s))) return st;
  }
  if ((st = two_level(p->tables, n, &p->lo, &p->hi, &p->sh, s))) return st;

  // scratch: [C: 2-D brfft copy of x][A: packed / extended lines][B: FFT out][schedule scratch]
  int64_t off = 0;
  if (direction == 1 && d2 >= 0) {
    p->offC = off;
    off += totalX;
  }
  p->offA = off;
  off += P * m * Q;
  p->offB = off;
  off += P * m * Q;
  const int strided1 = P > 1 ? 1 : 0;
  int64_t per1;
  int nbuf1;
  if ((st = scratch_per_line(*p->lp1, strided1, &per1, &nbuf1))) return st;
  const int64_t s1off = off;
  off += per1 * P * Q * nbuf1;
  const int64_t s2off = off;
  int64_t s2second = off;  // second scratch buffer of the d2 schedule
  if (d2 >= 0) {
    LineSet l2 = lines_along(nd, direction == -1 ? Y.data() : X.data(), d2);
    int64_t per2;
    int nbuf2;
    if ((st = scratch_per_line(*p->lp2, l2.es_in != 1 ? 1 : 0, &per2, &nbuf2))) return st;
    s2second = off + per2 * l2.count();
    off += per2 * l2.count() * nbuf2;
  }
  if ((st = p->scratch.ensure((size_t)std::max<int64_t>(off, 1) * csize(precision)))) return st;
  const BufRef SA{BUF_SCRATCH, p->offA}, SB{BUF_SCRATCH, p->offB}, SC{BUF_SCRATCH, p->offC};
  const BufRef s1a{BUF_SCRATCH, s1off}, s1b{BUF_SCRATCH, s1off + per1 * P * Q};

  if (direction == -1) {
    // ---- rfft along d1 (real X -> complex Y with h along d1)
    if (even) {
      BufRef src = SA;
      LineSet ls;
      if (P == 1) {
        // pairs along a contiguous line are already complex elements
        src = BufRef{BUF_IN, 0};
        ls = line_set(1, Q, 1, m, 1, h);
      } else {
        p->pre_op = REAL_PACK;
        p->pre = real_geom(P, m, Q, P, P * n, P, P * m);
        ls = line_set(P, Q, P, P * m, P, P * h);
      }
      if ((st = build_launches(*p->lp1, p->tables, ls, src, BufRef{BUF_OUT, 0}, s1a, s1b, false, p->S1.L, s)))
        return st;
      p->post_op = REAL_RFFT_SPLIT;
      p->post = real_geom(P, m / 2 + 1, Q, P, P * h, P, P * h);
    } else {
      p->pre_op = REAL_COMPLEXIFY;
      p->pre = real_geom(P, n, Q, P, P * n, P, P * n);
      LineSet ls = line_set(P, Q, P, P * n, P, P * n);
      if ((st = build_launches(*p->lp1, p->tables, ls, SA, SB, s1a, s1b, false, p->S1.L, s))) return st;
      p->post_op = REAL_COPY;
      p->post = real_geom(P, h, Q, P, P * n, P, P * h);
    }
    if (d2 >= 0) {
      LineSet l2 = lines_along(nd, Y.data(), d2);
      if ((st = build_launches(*p->lp2, p->tables, l2, BufRef{BUF_OUT, 0}, BufRef{BUF_OUT, 0},
                               BufRef{BUF_SCRATCH, s2off}, BufRef{BUF_SCRATCH, s2second}, false, p->S2.L, s)))
        return st;
    }
  } else {
    // ---- brfft: [complex backward along d2 into C], then along d1 (complex h -> real n)
    if (d2 >= 0) {
      LineSet l2 = lines_along(nd, X.data(), d2);
      if ((st = build_launches(*p->lp2, p->tables, l2, BufRef{BUF_IN, 0}, SC, BufRef{BUF_SCRATCH, s2off},
                               BufRef{BUF_SCRATCH, s2second}, true, p->S2.L, s)))
        return st;
      p->s1_from_c = true;
    }
    if (even) {
      p->pre_op = REAL_BRFFT_SPLIT;
      p->pre = real_geom(P, m / 2 + 1, Q, P, P * h, P, P * m);
      if (P == 1) {
        LineSet ls = line_set(1, Q, 1, m, 1, m);  // y viewed as complex pairs
        if ((st = build_launches(*p->lp1, p->tables, ls, SA, BufRef{BUF_OUT, 0}, s1a, s1b, true, p->S1.L, s)))
          return st;
      } else {
        LineSet ls = line_set(P, Q, P, P * m, P, P * m);
        if ((st = build_launches(*p->lp1, p->tables, ls, SA, SB, s1a, s1b, true, p->S1.L, s))) return st;
        p->post_op = REAL_UNPACK;
        p->post = real_geom(P, m, Q, P, P * m, P, P * n);
      }
    } else {
      p->pre_op = REAL_HERMITIAN;
      p->pre = real_geom(P, n, Q, P, P * h, P, P * n);
      LineSet ls = line_set(P, Q, P, P * n, P, P * n);
      if ((st = build_launches(*p->lp1, p->tables, ls, SA, SB, s1a, s1b, true, p->S1.L, s))) return st;
      p->post_op = REAL_PART;
      p->post = real_geom(P, n, Q, P, P * n, P, P * n);
    }
  }
  if ((st = upload_schedule(p->S1, s))) return st;
  if ((st = upload_schedule(p->S2, s))) return st;
  *plan = p.release();
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_rplan_output_dims(fftahip_rplan plan, int* ndims_out, int64_t* dims_out) {
  if (!plan || !ndims_out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  *ndims_out = (int)plan->out_dims.size();
  if (dims_out)
    for (size_t i = 0; i < plan->out_dims.size(); ++i) dims_out[i] = plan->out_dims[i];
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_rplan_execute(fftahip_rplan plan, const void* d_in, void* d_out, void* stream) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  for (int64_t v : plan->in_dims)
    if (v == 0) return FFTAHIP_SUCCESS;
  hipStream_t s = (hipStream_t)stream;
  const size_t cs = csize(plan->prec);
  char* scr = (char*)plan->scratch.p;
  const BufPtrs bpx{d_in, d_out, plan->scratch.p};
  fftahip_status st;
  // 2-D brfft: x along d2 first, into C
  if (plan->dir == 1 && plan->d2 >= 0) {
    if ((st = run_schedule(plan->S2, plan->prec, bpx, s))) return st;
  }
  const void* x1 = plan->s1_from_c ? (const void*)(scr + plan->offC * cs) : d_in;
  void* A = scr + plan->offA * cs;
  void* B = scr + plan->offB * cs;
  if (plan->pre_op >= 0) {
    HIP_TRY(launch_real_boundary(plan->pre_op, plan->prec, A, x1, plan->pre, plan->m, plan->lo, plan->hi, plan->sh,
                                 plan->n, s));
  }
  if ((st = run_schedule(plan->S1, plan->prec, BufPtrs{x1, d_out, plan->scratch.p}, s))) return st;
  if (plan->post_op == REAL_RFFT_SPLIT) {
    HIP_TRY(launch_real_boundary(REAL_RFFT_SPLIT, plan->prec, d_out, nullptr, plan->post, plan->m, plan->lo,
                                 plan->hi, plan->sh, plan->n, s));
  } else if (plan->post_op >= 0) {
    HIP_TRY(launch_real_boundary(plan->post_op, plan->prec, d_out, B, plan->post, plan->m, plan->lo, plan->hi,
                                 plan->sh, plan->n, s));
  }
  // 2-D rfft: then along d2, in place in y
  if (plan->dir == -1 && plan->d2 >= 0) {
    if ((st = run_schedule(plan->S2, plan->prec, BufPtrs{d_out, d_out, plan->scratch.p}, s))) return st;
  }
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_rplan_execute_host(fftahip_rplan plan, const void* h_in, void* h_out) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  int64_t ti = 1, to = 1;
  for (int64_t v : plan->in_dims) ti *= v;
  for (int64_t v : plan->out_dims) to *= v;
  const size_t cs = csize(plan->prec), rs = cs / 2;
  const size_t bi = (size_t)ti * (plan->dir == -1 ? rs : cs), bo = (size_t)to * (plan->dir == -1 ? cs : rs);
  if (bi == 0 || bo == 0) return FFTAHIP_SUCCESS;
  fftahip_status st;
  if ((st = plan->hin.ensure(bi))) return st;
  if ((st = plan->hout.ensure(bo))) return st;
  HIP_TRY(hipMemcpy(plan->hin.p, h_in, bi, hipMemcpyHostToDevice));
  if ((st = fftahip_rplan_execute(plan, plan->hin.p, plan->hout.p, nullptr))) return st;
  HIP_TRY(hipMemcpy(h_out, plan->hout.p, bo, hipMemcpyDeviceToHost));
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_rplan_destroy(fftahip_rplan plan) {
  delete plan;
  return FFTAHIP_SUCCESS;
}

static void describe_length(std::ostringstream& os, LengthPlan& lp) {
  os << "{\"n\": " << lp.N << ", \"ffta\": [";
  for (size_t i = 0; i < lp.ffta.size(); ++i) {
    const FftaNode& q = lp.ffta[i];
    if (i) os << ", ";
    os << "{\"type\": \"" << ffta_type_name(q.type) << "\", \"sz\": " << q.sz << ", \"s_in\": " << q.s_in
       << ", \"s_out\": " << q.s_out << ", \"left\": " << q.left << ", \"right\": " << q.right << "}";
  }
  os << "], \"device\": {";
  bool firstk = true;
  for (int strided = 0; strided < 2; ++strided) {
    Decomp* d;
    if (get_decomp(lp, strided, &d) != FFTAHIP_SUCCESS) continue;
    if (!firstk) os << ", ";
    firstk = false;
    os << "\"" << (strided ? "strided" : "contiguous") << "\": {\"kind\": \"" << kind_name(d->kind)
       << "\", \"m\": " << d->M << ", \"passes\": [";
    for (size_t i = 0; i < d->F.size(); ++i) {
      if (i) os << ", ";
      os << "{\"L\": " << d->F[i] << ", \"radices\": [";
      for (size_t r = 0; r < d->inner[i].rad.size(); ++r) os << (r ? ", " : "") << d->inner[i].rad[r];
      os << "], \"threads_per_line\": " << d->inner[i].Tl << ", \"block\": " << d->inner[i].block << "}";
    }
    os << "]}";
  }
  os << "}}";
}

static fftahip_status emit(const std::string& s, char* buf, size_t cap, size_t* needed) {
  if (needed) *needed = s.size() + 1;
  if (buf && cap > 0) {
    const size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_plan_describe(fftahip_plan plan, char* buf, size_t cap, size_t* needed) {
  if (!plan) return fail(FFTAHIP_ERR_ARGUMENT, "NULL plan");
  std::ostringstream os;
  os << "{\"precision\": " << plan->prec << ", \"direction\": " << plan->dir << ", \"lds_budget\": " << lds_budget()
     << ", \"region\": [";
  for (size_t i = 0; i < plan->region.size(); ++i) os << (i ? ", " : "") << plan->region[i];
  os << "], \"dims\": [";
  for (size_t i = 0; i < plan->lens.size(); ++i) {
    if (i) os << ", ";
    describe_length(os, *plan->lens[i]);
  }
  os << "]}";
  return emit(os.str(), buf, cap, needed);
}

fftahip_status fftahip_plan_destroy(fftahip_plan plan) {
  if (plan) {
    (void)hipDeviceSynchronize();
    delete plan;
  }
  return FFTAHIP_SUCCESS;
}

// ------------------------------------------------------------------ batches

static fftahip_status batch_build(fftahip_batch b, hipStream_t s);

fftahip_status fftahip_batch_create(fftahip_batch* batch, int precision, int64_t count, const int64_t* lengths,
                                    const int64_t* offsets, const int* directions, int64_t bluestein_cutoff) {
  if (!batch) return fail(FFTAHIP_ERR_ARGUMENT, "batch pointer is NULL");
  *batch = nullptr;
  if (precision != 32 && precision != 64) return fail(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (count < 0) return fail(FFTAHIP_ERR_ARGUMENT, "count must be >= 0");
  auto b = std::make_unique<fftahip_batch_s>();
  b->prec = precision;
  b->count = count;
  b->tables.prec = precision;
  b->lengths.assign(lengths, lengths + count);
  b->dirs.assign(directions, directions + count);
  b->offsets.resize(count);
  int64_t acc = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (lengths[i] <= 0) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Array length must be strictly positive");
    if (directions[i] != -1 && directions[i] != 1) return fail(FFTAHIP_ERR_ARGUMENT, "direction must be -1 or +1");
    b->offsets[i] = offsets ? offsets[i] : acc;
    acc += lengths[i];
    b->total = std::max(b->total, b->offsets[i] + lengths[i]);
  }
  // group by (length, direction)
  std::map<std::pair<int64_t, int>, size_t> gi;
  for (int64_t i = 0; i < count; ++i) {
    auto key = std::make_pair(b->lengths[i], b->dirs[i]);
    auto it = gi.find(key);
    if (it == gi.end()) {
      it = gi.emplace(key, b->groups.size()).first;
      b->groups.push_back({b->lengths[i], b->dirs[i], {}});
    }
    b->groups[it->second].idx.push_back(i);
  }
  hipStream_t s = nullptr;
  for (auto& g : b->groups) {
    if (!b->lens.count(g.N)) {
      auto lp = std::make_unique<LengthPlan>();
      init_length(*lp, precision, g.N, bluestein_cutoff);
      fftahip_status st = prepare_length(*lp, b->tables, s);
      if (st) return st;
      b->lens.emplace(g.N, std::move(lp));
    }
  }
  // offset tables and scratch
  std::vector<int64_t> offs;
  size_t scratch_elems = 0;
  for (auto& g : b->groups) {
    for (int64_t i : g.idx) offs.push_back(b->offsets[i]);
    int64_t per;
    int nbuf;
    fftahip_status st = scratch_per_line(*b->lens[g.N], 0, &per, &nbuf);
    if (st) return st;
    scratch_elems += (size_t)(per * (int64_t)g.idx.size() * nbuf);
  }
  b->host_offtab = offs;
  if (g_dry) {
    b->offtab_ptr = (const int64_t*)token(TK_OFFTAB, 0);
    b->scratch_ptr = (char*)token(TK_SCRATCH, 0);
    b->scratch.bytes = scratch_elems * csize(precision);
    fftahip_status st = batch_build(b.get(), s);
    if (st) return st;
    *batch = b.release();
    return FFTAHIP_SUCCESS;
  }
  fftahip_status st = b->offtab.ensure(offs.size() * sizeof(int64_t));
  if (st) return st;
  if (!offs.empty()) HIP_TRY(hipMemcpy(b->offtab.p, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  if ((st = b->scratch.ensure(scratch_elems * csize(precision)))) return st;
  b->offtab_ptr = (const int64_t*)b->offtab.p;
  b->scratch_ptr = (char*)b->scratch.p;
  if ((st = batch_build(b.get(), s))) return st;
  HIP_TRY(hipDeviceSynchronize());
  *batch = b.release();
  return FFTAHIP_SUCCESS;
}

// Build the per-phase work lists for the given pointers.
// Build the per-phase work lists (pointer independent; run with BufPtrs).
static fftahip_status batch_build(fftahip_batch b, hipStream_t s) {
  b->phases.clear();
  b->host_passes.clear();
  b->host_tiles.clear();
  std::vector<std::vector<Launch>>& per_group = b->per_group;
  per_group.clear();
  size_t toff = 0;
  int64_t scr = 0;  // element offset of this group's scratch
  for (auto& g : b->groups) {
    LengthPlan& lp = *b->lens[g.N];
    LineSet ls;
    ls.n[0] = 1;
    ls.n[1] = (int64_t)g.idx.size();
    ls.in_tab = b->offtab_ptr + toff;
    ls.out_tab = ls.in_tab;
    toff += g.idx.size();
    int64_t per;
    int nbuf;
    fftahip_status st = scratch_per_line(lp, 0, &per, &nbuf);
    if (st) return st;
    const BufRef s0{BUF_SCRATCH, scr}, s1{BUF_SCRATCH, scr + per * ls.n[1]};
    scr += per * ls.n[1] * nbuf;
    std::vector<Launch> L;
    st = build_launches(lp, b->tables, ls, BufRef{BUF_IN, 0}, BufRef{BUF_OUT, 0}, s0, s1, g.dir == 1, L, s);
    if (st) return st;
    per_group.push_back(L);
  }
  // phase p = the p-th launch of every group; within a phase, one grouped
  // launch per (block size, LDS class) so that small lines are not held to
  // the occupancy of the largest LDS footprint.
  static const size_t kLdsClass[] = {16384, 32768, 65536, 98304, 163840};
  auto lds_class = [](size_t lds) {
    for (size_t c : kLdsClass)
      if (lds <= c) return c;
    return (size_t)163840;
  };
  size_t maxp = 0;
  for (auto& L : per_group) maxp = std::max(maxp, L.size());
  // launch_of[g][ph]: the launch that runs pass ph of group g
  std::vector<std::vector<int>> launch_of(per_group.size());
  for (size_t g = 0; g < per_group.size(); ++g) launch_of[g].assign(per_group[g].size(), -1);
  auto add_deps = [&](fftahip_batch_s::Phase& P, size_t g, size_t ph) {
    for (size_t q = ph; q-- > 0;) {
      const int d = launch_of[g][q];
      if (d < 0) continue;  // an empty pass: depend on the one before it
      if (std::find(P.deps.begin(), P.deps.end(), d) == P.deps.end()) P.deps.push_back(d);
      break;
    }
  };
  // Issue order within a phase: the in-LDS launches first, by kernel variant
  // (full first) and block 1024 -> 256, then the register launches.  The
  // one-workgroup-per-CU launches so take CUs as they free up before the
  // thousands of small register-pass workgroups fill them (measured 3.57 vs
  // 3.62 ms per step eager, 3.49 vs 3.56 ms as a graph, against register
  // launches first and blocks 256 -> 1024: env FFTAHIP_BATCH_ORDER=0).
  const bool old_order = env_i64("FFTAHIP_BATCH_ORDER", 1) == 0;
  // a launch's tile table in host_tiles (its offset returned): one unused
  // word for a single group, else (group, tile within the group) per tile,
  // so that a workgroup finds its pass with one pair of loads
  auto tile_map = [&](const std::vector<uint32_t>& starts, uint32_t ntiles) -> size_t {
    const size_t off = b->host_tiles.size();
    if (starts.size() <= 1) {
      b->host_tiles.push_back(0);
      return off;
    }
    for (size_t g = 0; g < starts.size(); ++g) {
      const uint32_t end = g + 1 < starts.size() ? starts[g + 1] : ntiles;
      for (uint32_t t = starts[g]; t < end; ++t) {
        b->host_tiles.push_back((uint32_t)g);
        b->host_tiles.push_back(t - starts[g]);
      }
    }
    return off;
  };
  for (size_t ph = 0; ph < maxp; ++ph) {
    auto reg_launches = [&](int mode_hi, int mode_lo) {
      // register passes of this phase: one launch per mode, the full
      // (Bluestein) variant first
      for (int mode = mode_hi; mode >= mode_lo; --mode) {
        fftahip_batch_s::Phase P;
        P.reg = true;
        P.block = kRegBlock;
        P.mode = mode;
        P.lds = 0;
        P.ngroups = 0;
        P.pass_off = b->host_passes.size();
        std::vector<uint32_t> starts;
        uint32_t tiles = 0;
        for (size_t g = 0; g < per_group.size(); ++g) {
          auto& L = per_group[g];
          if (ph >= L.size() || !L[ph].reg || L[ph].grid <= 0 || pass_mode(L[ph].lp) != mode) continue;
          b->host_passes.push_back(L[ph].lp);
          starts.push_back(tiles);
          tiles += (uint32_t)L[ph].grid;
          ++P.ngroups;
          if (g_dry) P.members.push_back((int)g);
          add_deps(P, g, ph);
          launch_of[g][ph] = (int)b->phases.size();
        }
        if (P.ngroups == 0) continue;
        P.grid = tiles;
        P.pass = (int)ph;
        P.tile_off = tile_map(starts, tiles);
        b->phases.push_back(P);
      }
    };
    auto lds_launches = [&]() {
      // in-LDS passes: one launch per (kernel variant, block, LDS class),
      // the full variants first
      const std::vector<int> blocks = old_order ? std::vector<int>{256, 512, 1024} : std::vector<int>{1024, 512, 256};
      for (int mode : {kBigVariant + 2, 2, kBigVariant + 1, 1, kBigVariant, 0})
        for (int block : blocks)
          for (size_t cls : kLdsClass) {
            fftahip_batch_s::Phase P;
            P.reg = false;
            P.block = block;
            P.mode = mode;
            P.lds = 0;
            P.ngroups = 0;
            P.pass_off = b->host_passes.size();
            std::vector<uint32_t> starts;
            uint32_t tiles = 0;
            for (size_t g = 0; g < per_group.size(); ++g) {
              auto& L = per_group[g];
              if (ph >= L.size() || L[ph].reg || L[ph].block != block || L[ph].grid <= 0 ||
                  lds_class(L[ph].lds) != cls || pass_variant(L[ph].lp) != mode)
                continue;
              LinePass lp = L[ph].lp;
              lp.xcd_remap = 0;
              b->host_passes.push_back(lp);
              starts.push_back(tiles);
              tiles += (uint32_t)L[ph].grid;
              P.lds = std::max(P.lds, L[ph].lds);
              ++P.ngroups;
              if (g_dry) P.members.push_back((int)g);
              add_deps(P, g, ph);
              launch_of[g][ph] = (int)b->phases.size();
            }
            if (P.ngroups == 0) continue;
            P.grid = tiles;
            P.pass = (int)ph;
            P.tile_off = tile_map(starts, tiles);
            b->phases.push_back(P);
          }
    };
    if (old_order) {
      reg_launches(2, 0);
      lds_launches();
    } else {
      lds_launches();
      reg_launches(2, 0);
    }
  }
  if (g_dry) return FFTAHIP_SUCCESS;
  // the streams and events batch_launch orders the launches with
  if (!b->streams[0])
    b->nstreams = (int)std::min<int64_t>(fftahip_batch_s::kStreams, std::max<int64_t>(1, env_i64("FFTAHIP_BATCH_STREAMS", 4)));
  const int K = b->nstreams;
  if (!b->streams[0])
    for (int k = 0; k < K; ++k) HIP_TRY(hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking));
  if (b->events.size() != b->phases.size() + 1 + K) {
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    b->events.assign(b->phases.size() + 1 + K, nullptr);
    for (auto& e : b->events) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  fftahip_status st = b->passes.ensure(b->host_passes.size() * sizeof(LinePass));
  if (st) return st;
  if ((st = b->tiles.ensure(b->host_tiles.size() * sizeof(uint32_t)))) return st;
  if (!b->host_passes.empty())
    HIP_TRY(hipMemcpy(b->passes.p, b->host_passes.data(), b->host_passes.size() * sizeof(LinePass),
                      hipMemcpyHostToDevice));
  if (!b->host_tiles.empty())
    HIP_TRY(hipMemcpy(b->tiles.p, b->host_tiles.data(), b->host_tiles.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
  return FFTAHIP_SUCCESS;
}

// Eager execution: the launches of a batch form a DAG (launch i waits only
// for the launches that ran the previous pass of its groups), issued on
// kStreams streams that fork from and join back into `s`; independent
// launches (e.g. a bandwidth-bound Stockham pass and an instruction-bound
// Bluestein pass of other lengths) overlap on the device.
static fftahip_status batch_launch(fftahip_batch b, const void* d_in, void* d_out, hipStream_t s) {
  const BufPtrs bp{d_in, d_out, b->scratch_ptr};
  const int K = b->nstreams;
  const size_t np = b->phases.size();
  // streams and events are made by batch_build
  if (!b->streams[0] || b->events.size() != np + 1 + K)
    return fail(FFTAHIP_ERR_HIP, "batch streams/events missing");
  hipEvent_t fork = b->events[np];
  HIP_TRY(hipEventRecord(fork, s));
  for (int k = 0; k < K; ++k) HIP_TRY(hipStreamWaitEvent(b->streams[k], fork, 0));
  std::vector<int> stream_of(np, 0);
  int rr = 0;
  // env FFTAHIP_BATCH_SERIAL=1: every launch on one stream in issue order (A/B:
  // 4.56 ms per step against 3.57 ms on four streams)
  static const bool serial = env_i64("FFTAHIP_BATCH_SERIAL", 0) != 0;
  for (size_t i = 0; i < np; ++i) {
    const auto& P = b->phases[i];
    // a launch with a single predecessor follows it on its stream (no event
    // wait); otherwise the streams are taken in turn
    const int k = serial ? 0 : (P.deps.size() == 1 ? stream_of[P.deps[0]] : (rr++ % K));
    stream_of[i] = k;
    hipStream_t q = b->streams[k];
    for (int d : P.deps)
      if (stream_of[d] != k) HIP_TRY(hipStreamWaitEvent(q, b->events[d], 0));
    if (P.reg) {
      HIP_TRY(launch_reg_pass_groups((const LinePass*)b->passes.p + P.pass_off, (const uint32_t*)b->tiles.p + P.tile_off,
                                     P.ngroups, P.mode, b->prec, bp, P.grid, q));
    } else {
      HIP_TRY(launch_line_pass_groups((const LinePass*)b->passes.p + P.pass_off,
                                      (const uint32_t*)b->tiles.p + P.tile_off, P.ngroups, P.mode, b->prec, P.block,
                                      bp, 0, P.lds, P.grid, q));
    }
    HIP_TRY(hipEventRecord(b->events[i], q));
  }
  for (int k = 0; k < K; ++k) {
    HIP_TRY(hipEventRecord(b->events[np + 1 + k], b->streams[k]));
    HIP_TRY(hipStreamWaitEvent(s, b->events[np + 1 + k], 0));
  }
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_batch_execute(fftahip_batch b, const void* d_in, void* d_out, void* stream) {
  if (!b) return fail(FFTAHIP_ERR_ARGUMENT, "NULL batch");
  if (b->count == 0) return FFTAHIP_SUCCESS;
  return batch_launch(b, d_in, d_out, (hipStream_t)stream);
}

// The launch DAG of a batch as a hipGraph built from kernel nodes: one node
// per launch, an edge from every launch that ran the previous pass of one of
// its groups (the same dependencies the eager path orders with events), so
// independent launches stay free to overlap.  Built without stream capture.
static fftahip_status batch_graph_build(fftahip_batch b, const void* d_in, void* d_out) {
  const BufPtrs bp{d_in, d_out, b->scratch_ptr};
  hipGraph_t g = nullptr;
  HIP_TRY(hipGraphCreate(&g, 0));
  std::vector<hipGraphNode_t> nodes(b->phases.size(), nullptr);
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < b->phases.size() && e == hipSuccess; ++i) {
    const auto& P = b->phases[i];
    const KernelRef k = P.reg ? reg_kernel(b->prec, P.mode) : lds_kernel(b->prec, P.block, P.mode);
    if (!k.fn) {
      e = hipErrorInvalidValue;
      break;
    }
    LaunchArgs a((const LinePass*)b->passes.p + P.pass_off, (const uint32_t*)b->tiles.p + P.tile_off, P.ngroups, bp,
                 P.grid);
    hipKernelNodeParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.func = const_cast<void*>(k.fn);
    kp.gridDim = dim3(padded_grid(P.grid));
    kp.blockDim = dim3(k.block);
    kp.sharedMemBytes = (unsigned)P.lds;
    kp.kernelParams = a.args;  // argument values are copied into the node
    std::vector<hipGraphNode_t> deps;
    for (int d : P.deps) deps.push_back(nodes[d]);
    e = hipGraphAddKernelNode(&nodes[i], g, deps.empty() ? nullptr : deps.data(), deps.size(), &kp);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph = nullptr;
    return fail(FFTAHIP_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
  }
  b->graph_in = d_in;
  b->graph_out = d_out;
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_batch_execute_graph(fftahip_batch b, const void* d_in, void* d_out, void* stream) {
  if (!b) return fail(FFTAHIP_ERR_ARGUMENT, "NULL batch");
  hipStream_t s = (hipStream_t)stream;
  if (b->count == 0) return FFTAHIP_SUCCESS;
  if (b->graph && (b->graph_in != d_in || b->graph_out != d_out)) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (!b->graph) {
    fftahip_status st = batch_graph_build(b, d_in, d_out);
    if (st) return st;
  }
  HIP_TRY(hipGraphLaunch(b->graph, s));
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_batch_describe(fftahip_batch b, char* buf, size_t cap, size_t* needed) {
  if (!b) return fail(FFTAHIP_ERR_ARGUMENT, "NULL batch");
  std::ostringstream os;
  os << "{\"precision\": " << b->prec << ", \"count\": " << b->count << ", \"groups\": " << b->groups.size()
     << ", \"lengths\": " << b->lens.size() << ", \"phases\": [";
  for (size_t i = 0; i < b->phases.size(); ++i) {
    auto& P = b->phases[i];
    os << (i ? ", " : "") << "{\"reg\": " << (P.reg ? 1 : 0) << ", \"mode\": " << P.mode << ", \"block\": " << P.block << ", \"lds\": " << P.lds << ", \"grid\": " << P.grid
       << ", \"groups\": " << P.ngroups << "}";
  }
  os << "], \"scratch_bytes\": " << b->scratch.bytes << "}";
  return emit(os.str(), buf, cap, needed);
}

fftahip_status fftahip_batch_destroy(fftahip_batch b) {
  if (b) {
    (void)hipDeviceSynchronize();
    delete b;
  }
  return FFTAHIP_SUCCESS;
}

// ------------------------------------------------------------------ helpers

fftahip_status fftahip_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(FFTAHIP_ERR_ARGUMENT, "NULL pointer");
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail(FFTAHIP_ERR_ALLOC, hipGetErrorString(e));
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_free(void* ptr) {
  HIP_TRY(hipFree(ptr));
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_memcpy_htod(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return FFTAHIP_SUCCESS;
}
fftahip_status fftahip_device_info(int64_t* lds_per_block, int64_t* cu_count, int64_t* warp_size) {
  DevInfo& d = dev_info();
  if (!d.ok) return fail(FFTAHIP_ERR_HIP, "no HIP device");
  if (lds_per_block) *lds_per_block = d.lds;
  if (cu_count) *cu_count = d.cus;
  if (warp_size) *warp_size = d.warp;
  return FFTAHIP_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------------ schedule inspection
// Dry-run dumps used by the CPU test-suite to check pass geometry (index
// maps, twiddle exponents, buffer bounds) without a device.

static void json_ptr(std::ostringstream& os, const char* name, const void* p) {
  const uint64_t v = (uint64_t)(uintptr_t)p;
  if ((v >> 60) == 1) {
    os << "\"" << name << "\": [" << ((v >> 52) & 0xff) << ", " << (v & ((1ull << 52) - 1)) << "]";
  } else {
    os << "\"" << name << "\": " << (p ? "[-1, 0]" : "null");
  }
}

static void json_buf(std::ostringstream& os, const char* name, int id, int64_t off, int prec) {
  const int kind = id == BUF_IN ? TK_IN : (id == BUF_OUT ? TK_OUT : TK_SCRATCH);
  os << "\"" << name << "\": [" << kind << ", " << off * (int64_t)csize(prec) << "]";
}

static void json_launch(std::ostringstream& os, const Launch& la, int prec) {
  const LinePass& p = la.lp;
  os << "{\"op\": " << la.op << ", \"reg\": " << (la.reg ? 1 : 0) << ", \"block\": " << la.block << ", \"lds\": " << la.lds << ", \"grid\": " << la.grid
     << ", \"bytes\": " << la.bytes << ", ";
  json_buf(os, "in", la.op == OP_COPY ? BUF_IN : p.in_buf, la.op == OP_COPY ? 0 : p.in_off, prec);
  os << ", ";
  json_buf(os, "out", la.op == OP_COPY ? BUF_OUT : p.out_buf, la.op == OP_COPY ? 0 : p.out_off, prec);
  os << ", \"L\": " << p.L << ", \"rad\": [";
  for (int i = 0; i < p.nst; ++i) os << (i ? ", " : "") << p.rad[i];
  os << "], \"C\": " << p.C << ", \"Tl\": " << p.Tl << ", \"ldl\": " << p.ldl << ", \"lds_sh\": " << p.lds_sh << ", \"col_in\": " << p.col_in
     << ", \"col_out\": " << p.col_out << ", \"fuse_in\": " << p.fuse_in << ", \"fuse_out\": " << p.fuse_out << ", \"row_in\": " << p.row_in
     << ", \"row_mid\": " << p.row_mid << ", \"row_out\": " << p.row_out << ", \"nlines\": " << p.nlines << ", \"n\": [" << p.n[0] << ", " << p.n[1]
     << ", " << p.n[2] << "], \"a\": [" << p.a[0] << ", " << p.a[1] << ", " << p.a[2] << ", " << p.a[3]
     << "], \"b\": [" << p.b[0] << ", " << p.b[1] << ", " << p.b[2] << ", " << p.b[3] << "], \"ae\": " << p.ae
     << ", \"be\": " << p.be << ", ";
  json_ptr(os, "in_tab", p.in_tab);
  os << ", ";
  json_ptr(os, "out_tab", p.out_tab);
  os << ", ";
  json_ptr(os, "tw", p.tw);
  os << ", ";
  json_ptr(os, "tw16", p.tw16);
  os << ", \"tw_split\": " << p.tw_split << ", \"pretw\": " << p.pretw << ", \"twdim\": " << p.twdim << ", \"tw_sh\": " << p.tw_sh
     << ", \"tw_scale\": " << p.tw_scale << ", ";
  json_ptr(os, "tw_lo", p.tw_lo);
  os << ", ";
  json_ptr(os, "tw_hi", p.tw_hi);
  os << ", \"conj_in\": " << p.conj_in << ", \"conj_out\": " << p.conj_out << ", \"scale\": " << p.scale
     << ", \"load_mode\": " << p.load_mode << ", \"mid_mode\": " << p.mid_mode << ", \"mirror_rows\": " << p.mirror_rows << ", \"mid_in_stage\": " << p.mid_in_stage << ", \"mid_fused\": " << p.mid_fused << ", \"wave_sync\": " << p.wave_sync << ", \"zero_pad\": " << p.zero_pad << ", \"zero_from\": " << p.zero_from << ", \"store_mode\": " << p.store_mode
     << ", \"xcd_remap\": " << p.xcd_remap << ", \"chirp_n\": " << p.chirp_n << ", \"ca\": [" << p.ca[0] << ", "
     << p.ca[1] << ", " << p.ca[2] << ", " << p.ca[3] << "], \"cb\": [" << p.cb[0] << ", " << p.cb[1] << ", "
     << p.cb[2] << ", " << p.cb[3] << "], \"cm\": [" << p.cm[0] << ", " << p.cm[1] << ", " << p.cm[2] << ", "
     << p.cm[3] << "], \"cae\": " << p.cae << ", \"cbe\": " << p.cbe << ", \"cme\": " << p.cme << ", ";
  json_ptr(os, "chirp", p.chirp);
  os << ", ";
  json_ptr(os, "chirp_d", p.chirp_d);
  os << ", ";
  json_ptr(os, "rtab", p.rtab);
  os << ", ";
  json_ptr(os, "bhat", p.bhat);
  os << "}";
}

struct DryGuard {
  DryGuard(int64_t lds) {
    g_dry = true;
    g_dry_lds = lds;
  }
  ~DryGuard() {
    g_dry = false;
    g_dry_lds = 0;
  }
};

extern "C" {

// Schedule of mul!(y, plan_fft(x, region), x) for an array of shape dims on a
// device with `lds_per_block` bytes of LDS, as JSON: {"scratch_bytes", "launches"}.
fftahip_status fftahip_debug_schedule(int precision, int direction, int ndims, const int64_t* dims, int nregion,
                                      const int64_t* region, int64_t bluestein_cutoff, int64_t lds_per_block,
                                      char* buf, size_t cap, size_t* needed) {
  DryGuard g(lds_per_block);
  fftahip_plan plan = nullptr;
  fftahip_status st =
      fftahip_plan_create(&plan, precision, direction, ndims, dims, nregion, region, bluestein_cutoff);
  if (st) return st;
  std::unique_ptr<fftahip_plan_s> owner(plan);
  std::vector<Launch> L;
  size_t need = 0;
  st = plan_schedule(plan, ndims, dims, L, &need, nullptr);
  if (st) return st;
  std::ostringstream os;
  os << "{\"scratch_bytes\": " << need << ", \"launches\": [";
  for (size_t i = 0; i < L.size(); ++i) {
    if (i) os << ", ";
    json_launch(os, L[i], precision);
  }
  os << "]}";
  return emit(os.str(), buf, cap, needed);
}

// Grouped schedule of a ragged batch: {"scratch_bytes", "offtab", "groups": [[launch...]...]}.
fftahip_status fftahip_debug_batch_schedule(int precision, int64_t count, const int64_t* lengths,
                                            const int* directions, int64_t lds_per_block, char* buf, size_t cap,
                                            size_t* needed) {
  DryGuard g(lds_per_block);
  fftahip_batch b = nullptr;
  fftahip_status st = fftahip_batch_create(&b, precision, count, lengths, nullptr, directions, 73);
  if (st) return st;
  std::unique_ptr<fftahip_batch_s> owner(b);
  std::ostringstream os;
  os << "{\"scratch_bytes\": " << b->scratch.bytes << ", \"offtab\": [";
  for (size_t i = 0; i < b->host_offtab.size(); ++i) os << (i ? ", " : "") << b->host_offtab[i];
  os << "], \"groups\": [";
  for (size_t gi = 0; gi < b->per_group.size(); ++gi) {
    os << (gi ? ", " : "") << "[";
    for (size_t i = 0; i < b->per_group[gi].size(); ++i) {
      if (i) os << ", ";
      json_launch(os, b->per_group[gi][i], precision);
    }
    os << "]";
  }
  os << "], \"phases\": " << b->phases.size() << ", \"launches\": [";
  for (size_t i = 0; i < b->phases.size(); ++i) {
    const auto& P = b->phases[i];
    os << (i ? ", " : "") << "{\"reg\": " << (P.reg ? 1 : 0) << ", \"block\": " << P.block << ", \"mode\": " << P.mode
       << ", \"lds\": " << P.lds << ", \"grid\": " << P.grid << ", \"groups\": " << P.ngroups
       << ", \"pass\": " << P.pass << ", \"deps\": [";
    for (size_t d = 0; d < P.deps.size(); ++d) os << (d ? ", " : "") << P.deps[d];
    os << "], \"members\": [";
    for (size_t m = 0; m < P.members.size(); ++m) os << (m ? ", " : "") << P.members[m];
    os << "]}";
  }
  os << "]}";
  return emit(os.str(), buf, cap, needed);
}

// ------------------------------------------------------------------ autotuner

fftahip_status fftahip_autotune_lines(int precision, int64_t count, const int64_t* lengths, int64_t* changed_out) {
  if (precision != 32 && precision != 64) return fail(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (count < 0 || (count > 0 && !lengths)) return fail(FFTAHIP_ERR_ARGUMENT, "NULL lengths");
  const size_t esz = csize(precision);
  // 16 MB per buffer: the passes run from the last-level cache, so the timing
  // ranks the candidates by their instruction and LDS cost (the part the
  // radix schedule changes) rather than by HBM bandwidth
  const int64_t elems = precision == 32 ? (int64_t(1) << 21) : (int64_t(1) << 20);
  DevBuf a, b;
  fftahip_status st;
  if ((st = a.ensure((size_t)elems * esz))) return st;
  if ((st = b.ensure((size_t)elems * esz))) return st;
  HIP_TRY(hipMemset(a.p, 0, (size_t)elems * esz));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int64_t changed = 0;
  std::set<int64_t> done;
  st = FFTAHIP_SUCCESS;
  for (int64_t i = 0; i < count && st == FFTAHIP_SUCCESS; ++i) {
    const int64_t L = lengths[i];
    if (L < 2 || L > single_max(precision) || !smooth13(L) || !done.insert(L).second) continue;
    std::vector<std::vector<int>> cands = inner_candidates(L);
    if (cands.size() < 2) continue;
    Inner dflt;
    make_inner(L, dflt);
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tuned.erase({precision, L});
    }
    double best_ms = 1e30;
    std::vector<int> best;
    for (const auto& c : cands) {
      {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tuned[{precision, L}] = c;
      }
      const int64_t dims[2] = {L, std::max<int64_t>(1, elems / L)};
      const int64_t region[1] = {1};
      fftahip_plan pl = nullptr;
      if ((st = fftahip_plan_create(&pl, precision, -1, 2, dims, 1, region, 73))) break;
      float ms = 0;
      st = fftahip_execute(pl, 2, dims, a.p, b.p, nullptr);
      if (st == FFTAHIP_SUCCESS) {
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
        int reps = 0;
        for (; e == hipSuccess && reps < 20; ++reps)
          if ((st = fftahip_execute(pl, 2, dims, a.p, b.p, nullptr))) break;
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess && st == FFTAHIP_SUCCESS) st = fail(FFTAHIP_ERR_HIP, hipGetErrorString(e));
        ms /= std::max(reps, 1);
      }
      fftahip_plan_destroy(pl);
      if (st) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    if (st || best.empty() || best == dflt.rad) {
      g_tuned.erase({precision, L});
    } else {
      g_tuned[{precision, L}] = best;
      ++changed;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (changed_out) *changed_out = changed;
  return st;
}

// Strided tile shapes the shape autotuner tries (LDS bytes per tile, fewest
// adjacent lines): the default rule first, then larger tiles with fewer
// lines, which let longer strided lines run in one pass.
static const TileShape kTileCands[] = {{0, 0}, {71680, 8}, {71680, 4}, {143360, 8}, {143360, 4}, {143360, 2}};

fftahip_status fftahip_autotune_shape(int precision, int ndims, const int64_t* dims, int nregion,
                                      const int64_t* region, int64_t* changed_out) {
  if (precision != 32 && precision != 64) return fail(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (ndims < 1 || !dims || nregion < 1 || !region) return fail(FFTAHIP_ERR_ARGUMENT, "NULL or empty shape");
  int64_t total = 1;
  for (int i = 0; i < ndims; ++i) {
    if (dims[i] <= 0) return fail(FFTAHIP_ERR_DIMENSION_MISMATCH, "Array length must be strictly positive");
    total *= dims[i];
  }
  // the transform lengths of the region (one tile shape for all of them)
  std::set<int64_t> lens;
  for (int k = 0; k < nregion; ++k) {
    if (region[k] < 1 || region[k] > ndims) return fail(FFTAHIP_ERR_ARGUMENT, "Plan region is outside array dimensions.");
    lens.insert(dims[region[k] - 1]);
  }
  const size_t bytes = (size_t)total * csize(precision);
  DevBuf a, b;
  fftahip_status st;
  if ((st = a.ensure(bytes))) return st;
  if ((st = b.ensure(bytes))) return st;
  HIP_TRY(hipMemset(a.p, 0, bytes));
  HIP_TRY(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  auto set_all = [&](const TileShape& ts) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (int64_t N : lens) {
      if (ts.budget > 0) g_tuned_tile[{precision, N}] = ts;
      else g_tuned_tile.erase({precision, N});
    }
  };
  double best_ms = 1e30;
  TileShape best;
  std::set<std::string> seen;  // candidates giving the same launches are timed once
  for (const TileShape& c : kTileCands) {
    set_all(c);
    fftahip_plan pl = nullptr;
    if ((st = fftahip_plan_create(&pl, precision, -1, ndims, dims, nregion, region, 73))) break;
    std::string key;
    {
      std::vector<Launch> L;
      size_t need = 0;
      if ((st = plan_schedule(pl, ndims, dims, L, &need, nullptr))) {
        fftahip_plan_destroy(pl);
        break;
      }
      std::ostringstream os;
      for (const Launch& la : L) os << la.grid << ',' << la.block << ',' << la.lds << ',' << la.lp.L << ';';
      key = os.str();
    }
    if (!seen.insert(key).second) {
      fftahip_plan_destroy(pl);
      continue;
    }
    float ms = 0;
    st = fftahip_execute(pl, ndims, dims, a.p, b.p, nullptr);
    if (st == FFTAHIP_SUCCESS) {
      hipError_t e = hipDeviceSynchronize();
      int reps = 0;
      if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
      for (; e == hipSuccess && reps < 10; ++reps)
        if ((st = fftahip_execute(pl, ndims, dims, a.p, b.p, nullptr))) break;
      if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess && st == FFTAHIP_SUCCESS) st = fail(FFTAHIP_ERR_HIP, hipGetErrorString(e));
      ms /= std::max(reps, 1);
    }
    fftahip_plan_destroy(pl);
    if (st) break;
    // a larger tile must win by 3% (run-to-run noise) to replace the rule
    if (ms < (c.budget > 0 ? 0.97 : 1.0) * best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_all(st ? TileShape() : best);
  if (changed_out) *changed_out = (!st && best.budget > 0) ? (int64_t)lens.size() : 0;
  return st;
}

fftahip_status fftahip_autotune_save(const char* path) {
  if (!path) return fail(FFTAHIP_ERR_ARGUMENT, "NULL path");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(FFTAHIP_ERR_ARGUMENT, std::string("cannot write ") + path);
  std::fprintf(f, "# fftahip tuning table\n");
  std::fprintf(f, "# radix schedule:     precision line-length radices (stage order)\n");
  std::fprintf(f, "# strided tile shape: S precision transform-length lds-bytes-per-tile fewest-lines\n");
  std::lock_guard<std::mutex> g(g_tune_mu);
  for (const auto& kv : g_tuned) {
    std::fprintf(f, "%d %lld", kv.first.first, (long long)kv.first.second);
    for (int r : kv.second) std::fprintf(f, " %d", r);
    std::fprintf(f, "\n");
  }
  for (const auto& kv : g_tuned_tile)
    std::fprintf(f, "S %d %lld %lld %lld\n", kv.first.first, (long long)kv.first.second, (long long)kv.second.budget,
                 (long long)kv.second.cmin);
  std::fclose(f);
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_autotune_load(const char* path, int64_t* entries_out) {
  if (!path) return fail(FFTAHIP_ERR_ARGUMENT, "NULL path");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(FFTAHIP_ERR_ARGUMENT, std::string("cannot read ") + path);
  char line[512];
  int64_t n = 0;
  std::lock_guard<std::mutex> g(g_tune_mu);
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    if (line[0] == 'S') {
      std::istringstream is(line + 1);
      int prec;
      long long N, budget, cmin;
      if (!(is >> prec >> N >> budget >> cmin)) continue;
      if ((prec != 32 && prec != 64) || N < 1 || budget < 4096 || budget > (1 << 20) || cmin < 1 || cmin > kMaxLines)
        continue;  // skip malformed entries
      g_tuned_tile[{prec, (int64_t)N}] = TileShape{budget, cmin};
      ++n;
      continue;
    }
    std::istringstream is(line);
    int prec;
    long long L;
    if (!(is >> prec >> L)) continue;
    std::vector<int> rad;
    int r;
    while (is >> r) rad.push_back(r);
    Inner tmp;
    if ((prec != 32 && prec != 64) || !inner_from(L, rad, tmp, prec)) continue;  // skip malformed entries
    g_tuned[{prec, (int64_t)L}] = rad;
    ++n;
  }
  std::fclose(f);
  if (entries_out) *entries_out = n;
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_autotune_clear(void) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tuned.clear();
  g_tuned_tile.clear();
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_autotune_get_tile(int precision, int64_t length, int64_t* lds_bytes, int64_t* min_lines_out) {
  if (!lds_bytes || !min_lines_out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  const TileShape ts = tuned_tile(precision, length);
  *lds_bytes = ts.budget;
  *min_lines_out = ts.cmin;
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_autotune_get(int precision, int64_t length, int* nrad, int* radices) {
  if (!nrad) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  Inner in;
  if (!make_inner(length, in, precision)) return fail(FFTAHIP_ERR_ARGUMENT, "no in-LDS schedule for this length");
  *nrad = (int)in.rad.size();
  if (radices)
    for (size_t i = 0; i < in.rad.size(); ++i) radices[i] = in.rad[i];
  return FFTAHIP_SUCCESS;
}

// Device read-back of the chirp values the plans use (the same device
// function as the chirp tables): entries n0 .. n0 + count - 1 of length N in
// the given precision.
fftahip_status fftahip_debug_chirp(int precision, int64_t N, int64_t n0, int64_t count, void* host_out) {
  if (precision != 32 && precision != 64) return fail(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (N < 1 || N >= (int64_t(1) << 31) || n0 < 0 || count < 0 || n0 + count > N)
    return fail(FFTAHIP_ERR_ARGUMENT, "need 0 <= n0 <= n0 + count <= N < 2^31");
  if (count == 0) return FFTAHIP_SUCCESS;
  if (!host_out) return fail(FFTAHIP_ERR_ARGUMENT, "NULL output");
  DevBuf d;
  fftahip_status st;
  if ((st = d.ensure((size_t)count * csize(precision)))) return st;
  HIP_TRY(launch_gen_chirp_range(d.p, precision, n0, count, N, nullptr));
  HIP_TRY(hipMemcpy(host_out, d.p, (size_t)count * csize(precision), hipMemcpyDeviceToHost));
  return FFTAHIP_SUCCESS;
}

// Device read-back of a register pass's FP64 chirp steps d[t] = exp(i pi
// ((t s mod 2N)^2 mod 2N) / N), t < L.
fftahip_status fftahip_debug_chirp_steps(int64_t N, int64_t s, int L, double* host_out) {
  if (N < 1 || N >= (int64_t(1) << 31) || s < 0 || s >= 2 * N || L < 1 || L > 4096 || !host_out)
    return fail(FFTAHIP_ERR_ARGUMENT, "need 1 <= N < 2^31, 0 <= s < 2N, 1 <= L <= 4096, output");
  DevBuf d;
  fftahip_status st;
  if ((st = d.ensure((size_t)L * sizeof(double2)))) return st;
  HIP_TRY(launch_gen_chirp_steps((double2*)d.p, s, N, L, nullptr));
  HIP_TRY(hipMemcpy(host_out, d.p, (size_t)L * sizeof(double2), hipMemcpyDeviceToHost));
  return FFTAHIP_SUCCESS;
}

// Host evaluation of the register passes' 64-bit remainder (mod_q31).
fftahip_status fftahip_debug_mod_q31(const uint64_t* m, uint64_t d, uint64_t* r, int64_t count) {
  if (d == 0 || d >= (uint64_t(1) << 32)) return fail(FFTAHIP_ERR_ARGUMENT, "need 0 < d < 2^32");
  if (count > 0 && (!m || !r)) return fail(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  const double inv = 1.0 / (double)d;
  for (int64_t i = 0; i < count; ++i) r[i] = mod_q31(m[i], d, inv);
  return FFTAHIP_SUCCESS;
}

// Host evaluation of the kernels' division by a runtime constant.
fftahip_status fftahip_debug_fastdiv(uint32_t d, const uint32_t* n, uint32_t* q, int64_t count) {
  if (d == 0) return fail(FFTAHIP_ERR_ARGUMENT, "division by zero");
  const FastDiv f = make_fastdiv(d);
  for (int64_t i = 0; i < count; ++i) q[i] = fdiv(n[i], f);
  return FFTAHIP_SUCCESS;
}

}  // extern "C"

// Library-internal entry points for the serving loop (declared in
// plan_internal.h; C++ linkage, outside the extern "C" blocks).
namespace fftahip {
fftahip_status plan_scratch_bytes(fftahip_plan plan, int ndims, const int64_t* dims, size_t* bytes) {
  if (fftahip_status cs = check_shape(plan, ndims, dims)) return cs;
  DevSchedule* S;
  fftahip_status st = plan_shape_schedule(plan, ndims, dims, &S, nullptr);
  if (st) return st;
  *bytes = S->scratch_need;
  return FFTAHIP_SUCCESS;
}
fftahip_status plan_execute_scratch(fftahip_plan plan, int ndims, const int64_t* dims, const void* d_in, void* d_out,
                                    void* scratch, hipStream_t s) {
  if (fftahip_status cs = check_shape(plan, ndims, dims)) return cs;
  DevSchedule* S;
  fftahip_status st = plan_shape_schedule(plan, ndims, dims, &S, s);
  if (st) return st;
  return run_schedule(*S, plan->prec, BufPtrs{d_in, d_out, scratch}, s);
}

// Stand-ins for the offset tables of a many-vector schedule (never
// dereferenced: replaced by each call's table addresses before upload).
static const int64_t* const kManyTabIn = reinterpret_cast<const int64_t*>(uintptr_t(0x1000));
static const int64_t* const kManyTabOut = reinterpret_cast<const int64_t*>(uintptr_t(0x2000));

static fftahip_status many_schedule(fftahip_plan plan, int count, fftahip_plan_s::Many** out, hipStream_t s) {
  auto it = plan->many.find(count);
  if (it == plan->many.end()) {
    if (plan->lens.size() != 1) return fail(FFTAHIP_ERR_ARGUMENT, "many-vector execution needs a 1-D plan");
    LengthPlan& lp = *plan->lens[0];
    LineSet ls;
    ls.n[0] = 1;
    ls.n[1] = count;
    ls.in_tab = kManyTabIn;
    ls.out_tab = kManyTabOut;
    int64_t per;
    int nbuf;
    fftahip_status st = scratch_per_line(lp, 0, &per, &nbuf);
    if (st) return st;
    const size_t tables_before = plan->tables.stage.size() + plan->tables.two.size() + plan->tables.tw16.size() +
                                 plan->tables.chirp_d.size();
    fftahip_plan_s::Many m;
    m.scratch = (size_t)(per * count * nbuf) * csize(plan->prec);
    if ((st = build_launches(lp, plan->tables, ls, BufRef{BUF_IN, 0}, BufRef{BUF_OUT, 0}, BufRef{BUF_SCRATCH, 0},
                             BufRef{BUF_SCRATCH, per * count}, plan->dir == 1, m.L, s)))
      return st;
    if (plan->scale != 1.0) {
      Launch* last = nullptr;
      for (auto& la : m.L)
        if (la.op == OP_PASS) last = &la;
      if (last) last->lp.scale *= plan->scale;
    }
    it = plan->many.emplace(count, std::move(m)).first;
    // tables the build generated on `s` must exist before another stream
    // uses them (usually the plan's own tables already cover the schedule)
    if (plan->tables.stage.size() + plan->tables.two.size() + plan->tables.tw16.size() +
            plan->tables.chirp_d.size() != tables_before)
      HIP_TRY(hipStreamSynchronize(s));
  }
  *out = &it->second;
  return FFTAHIP_SUCCESS;
}

// Device memory of one many-vector execution: the pass descriptors, the two
// offset tables and a zero tile start, 256-byte aligned.
static size_t many_desc_bytes(size_t nlaunch, int count) {
  const size_t b = nlaunch * sizeof(LinePass) + 2 * (size_t)count * sizeof(int64_t) + sizeof(uint32_t);
  return (b + 255) & ~(size_t)255;
}

fftahip_status plan_many_bytes(fftahip_plan plan, int count, size_t* scratch, size_t* desc, hipStream_t s) {
  fftahip_plan_s::Many* m;
  fftahip_status st = many_schedule(plan, count, &m, s);
  if (st) return st;
  *scratch = m->scratch;
  *desc = many_desc_bytes(m->L.size(), count);
  return FFTAHIP_SUCCESS;
}

fftahip_status plan_execute_many(fftahip_plan plan, int count, const void* const* ins, void* const* outs,
                                 void* scratch, void* desc, hipStream_t s) {
  fftahip_plan_s::Many* m;
  fftahip_status st = many_schedule(plan, count, &m, s);
  if (st) return st;
  const size_t esz = csize(plan->prec);
  // element offsets from the lowest address of each side
  uintptr_t bin = UINTPTR_MAX, bout = UINTPTR_MAX;
  for (int i = 0; i < count; ++i) {
    bin = std::min(bin, (uintptr_t)ins[i]);
    bout = std::min(bout, (uintptr_t)outs[i]);
  }
  const size_t nl = m->L.size();
  std::vector<unsigned char> host(many_desc_bytes(nl, count), 0);
  LinePass* hp = reinterpret_cast<LinePass*>(host.data());
  int64_t* tin = reinterpret_cast<int64_t*>(host.data() + nl * sizeof(LinePass));
  int64_t* tout = tin + count;
  for (int i = 0; i < count; ++i) {
    const uintptr_t di = (uintptr_t)ins[i] - bin, dout = (uintptr_t)outs[i] - bout;
    if (di % esz || dout % esz) return fail(FFTAHIP_ERR_ARGUMENT, "array addresses not element aligned");
    tin[i] = (int64_t)(di / esz);
    tout[i] = (int64_t)(dout / esz);
  }
  unsigned char* dd = (unsigned char*)desc;
  const int64_t* d_tin = reinterpret_cast<const int64_t*>(dd + nl * sizeof(LinePass));
  const int64_t* d_tout = d_tin + count;
  const uint32_t* d_zero = reinterpret_cast<const uint32_t*>(d_tout + count);
  for (size_t i = 0; i < nl; ++i) {
    hp[i] = m->L[i].lp;
    if (hp[i].in_tab == kManyTabIn) hp[i].in_tab = d_tin;
    if (hp[i].out_tab == kManyTabOut) hp[i].out_tab = d_tout;
  }
  // pageable source: the copy is staged before the call returns
  HIP_TRY(hipMemcpyAsync(desc, host.data(), host.size(), hipMemcpyHostToDevice, s));
  const BufPtrs bp{(const void*)bin, (void*)bout, scratch};
  for (size_t i = 0; i < nl; ++i) {
    const Launch& la = m->L[i];
    if (la.op != OP_PASS || la.grid <= 0) continue;
    const LinePass* dpass = reinterpret_cast<const LinePass*>(dd) + i;
    if (la.reg)
      HIP_TRY(launch_reg_pass_groups(dpass, d_zero, 1, pass_mode(la.lp), plan->prec, bp, la.grid, s));
    else
      HIP_TRY(launch_line_pass_groups(dpass, d_zero, 1, pass_variant(la.lp), plan->prec, la.block, bp, 0, la.lds,
                                      la.grid, s));
  }
  return FFTAHIP_SUCCESS;
}

fftahip_status set_error(fftahip_status s, const std::string& msg) { return fail(s, msg); }
}  // namespace fftahip

