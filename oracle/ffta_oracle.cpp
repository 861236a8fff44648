// ffta_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of FFTA.jl's complex fft/bfft transform path, used as the
// parity oracle for the HIP library and as the timed CPU baseline.  It is
// never linked into, loaded by, or called from the product library; only the
// tests, the smoke check and the benchmark's cpu_baseline leg use it.
//
// What it restates (paths relative to the FFTA.jl source tree):
//   planner            src/callgraph.jl:76-126   (CallGraphNode!)
//   Singleton twiddles src/singleton_twiddle.jl:25-39
//   fft! dispatch      src/algos.jl:5-34
//   composite CT       src/algos.jl:51-118
//   O(N^2) DFT leaf    src/algos.jl:135-160
//   radix-4 pow2       src/algos.jl:207-280
//   radix-3 pow3       src/algos.jl:299-341
//   Bluestein          src/algos.jl:344-424
//   N-D / 2-D order    src/plan.jl:210-257, 287-348 (copy, then dims in order)
//
// Arithmetic is done in the plan's real type R (float for ComplexF32, double
// for ComplexF64) exactly where FFTA does, including the Singleton recurrence
// in R.  sincospi is evaluated one precision up and rounded to R, which
// agrees with Julia's sincospi to within an ulp.  Complex products use the
// plain 4-multiply formula (Julia's Complex `*`), built with
// -ffp-contract=off.  Rounding-level bits therefore match FFTA to O(eps), not
// bit for bit (FFTA's @fastmath / @muladd contraction choices are not
// reproducible without Julia).
//
// Also here (truth for pinning the restatement):
//   naive O(N^2) DFT in long double with exact (u*v mod n) angle reduction,
//   following test/runtests.jl:12-42 but one precision up.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <thread>
#include <atomic>
#include <algorithm>

namespace {

enum NodeType : int64_t {  // src/callgraph.jl:3
  COMPOSITE_FFT = 0, DFT = 1, POW3_FFT = 2, POW2RADIX4_FFT = 3, BLUESTEIN = 4
};

struct Node {  // src/callgraph.jl:18-25
  int64_t left, right, type, sz, s_in, s_out;
};

template <class R> struct Cx { R re, im; };

template <class R> inline Cx<R> add(Cx<R> a, Cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R> inline Cx<R> sub(Cx<R> a, Cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R> inline Cx<R> neg(Cx<R> a) { return {-a.re, -a.im}; }
template <class R> inline Cx<R> mul(Cx<R> a, Cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R> inline Cx<R> conj(Cx<R> a) { return {a.re, -a.im}; }

// sincospi one precision up, rounded to R.
inline void sincospi_r(double x, double* s, double* c) {
  const long double pi = 3.141592653589793238462643383279502884L;
  long double lx = (long double)x;
  *s = (double)sinl(pi * lx);
  *c = (double)cosl(pi * lx);
}
inline void sincospi_r(float x, float* s, float* c) {
  const double pi = 3.141592653589793238462643383279502884;
  double dx = (double)x;
  *s = (float)sin(pi * dx);
  *c = (float)cos(pi * dx);
}

// src/singleton_twiddle.jl:25-30: (s,c) = sincospi(-hfreq); z = (2s^2, 2sc).
template <class R> inline Cx<R> singleton_params(R hfreq) {
  R s, c;
  sincospi_r((R)(-hfreq), &s, &c);
  R alpha = (R)2 * s * s;
  R beta = (R)2 * s * c;
  return {alpha, beta};
}
// src/singleton_twiddle.jl:36-39: w <- w - z*w.
template <class R> inline Cx<R> singleton_step(Cx<R> w, Cx<R> z) { return sub(w, mul(z, w)); }

// ---------------------------------------------------------------- planner
bool isprime64(int64_t n) {
  if (n < 2) return false;
  if (n % 2 == 0) return n == 2;
  if (n % 3 == 0) return n == 3;
  for (int64_t f = 5; f * f <= n; f += 6)
    if (n % f == 0 || n % (f + 2) == 0) return false;
  return true;
}
bool is_even_pow2(int64_t n) { return n >= 2 && (n & (n - 1)) == 0; }
bool is_pow3(int64_t n) {
  if (n < 3 || n % 3 != 0) return false;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}
int64_t isqrt64(int64_t n) {
  int64_t r = (int64_t)std::sqrt((double)n);
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

// src/callgraph.jl:76-126.  Returns the number of nodes pushed; error -1 for N<=0.
int64_t build_node(std::vector<Node>& nodes, int64_t N, int64_t cutoff, int64_t s_in, int64_t s_out) {
  if (N <= 0) return -1;
  if (is_even_pow2(N)) {
    nodes.push_back({0, 0, POW2RADIX4_FFT, N, s_in, s_out});
    return 1;
  } else if (is_pow3(N)) {
    nodes.push_back({0, 0, POW3_FFT, N, s_in, s_out});
    return 1;
  } else if (N == 1 || isprime64(N)) {
    nodes.push_back({0, 0, N < cutoff ? DFT : BLUESTEIN, N, s_in, s_out});
    return 1;
  }
  // Primes.factor: ascending primes with multiplicity.
  std::vector<int64_t> Ns;
  {
    int64_t m = N;
    for (int64_t p = 2; p * p <= m; ++p)
      while (m % p == 0) { Ns.push_back(p); m /= p; }
    if (m > 1) Ns.push_back(m);
  }
  int64_t N1;
  int64_t p0 = Ns[0];
  if (p0 == 2 || p0 == 3) {
    N1 = 1;
    for (int64_t p : Ns) if (p == p0) N1 *= p;
  } else {
    double N_fsqrt = std::sqrt((double)N);
    int64_t N_isqrt = isqrt64(N);
    std::vector<int64_t> cp(Ns.size());
    int64_t acc = 1;
    for (size_t i = 0; i < Ns.size(); ++i) { acc *= Ns[i]; cp[i] = acc; }
    // searchsortedlast(cp, N_isqrt): last index with cp[i] <= N_isqrt (0-based here).
    int64_t idx = -1;
    for (size_t i = 0; i < cp.size(); ++i) if (cp[i] <= N_isqrt) idx = (int64_t)i;
    N1 = cp[idx];
    if (idx != (int64_t)cp.size() - 1 &&
        ((double)cp[idx + 1] - N_fsqrt < (N_fsqrt - (double)N1)))
      N1 = cp[idx + 1];
  }
  int64_t N2 = N / N1;
  size_t sz = nodes.size();
  nodes.push_back({0, 0, DFT, N, s_in, s_out});
  int64_t left_len = build_node(nodes, N1, cutoff, N2, N2 * s_out);
  int64_t right_len = build_node(nodes, N2, cutoff, N1 * s_in, 1);
  nodes[sz] = {1, 1 + left_len, COMPOSITE_FFT, N, s_in, s_out};
  return 1 + left_len + right_len;
}

// ---------------------------------------------------------------- kernels
template <class R>
struct Graph {
  std::vector<Node> nodes;
  std::vector<std::vector<Cx<R>>> workspace;  // per node (composite: N, else empty)
};

template <class R>
Graph<R> make_graph(int64_t N, int64_t cutoff) {
  Graph<R> g;
  build_node(g.nodes, N, cutoff, 1, 1);
  g.workspace.resize(g.nodes.size());
  for (size_t i = 0; i < g.nodes.size(); ++i)
    if (g.nodes[i].type == COMPOSITE_FFT) g.workspace[i].resize(g.nodes[i].sz);
  return g;
}

// Indices below are 0-based (FFTA's are 1-based; start offsets shift by one).

// src/algos.jl:135-160
template <class R>
void fft_dft(Cx<R>* out, const Cx<R>* in, int64_t N, int64_t so, int64_t sto, int64_t si, int64_t sti, int dir) {
  Cx<R> tmp = in[si];
  for (int64_t j = 1; j < N; ++j) tmp = add(tmp, in[si + j * sti]);
  out[so] = tmp;
  for (int64_t j = 1; j < N; ++j) {
    tmp = in[si];
    Cx<R> zj = singleton_params<R>((R)(dir * (R)j) / (R)N);
    Cx<R> wk = {(R)1, (R)0};
    for (int64_t k = 1; k < N; ++k) {
      wk = singleton_step(wk, zj);
      tmp = add(tmp, mul(wk, in[si + k * sti]));
    }
    out[so + j * sto] = tmp;
  }
}

// src/algos.jl:207-280
template <class R>
void fft_pow2_radix4(Cx<R>* out, const Cx<R>* in, int64_t N, int64_t so, int64_t sto, int64_t si, int64_t sti, int dir) {
  if (N == 2) {
    out[so] = add(in[si], in[si + sti]);
    out[so + sto] = sub(in[si], in[si + sti]);
    return;
  }
  const Cx<R> minusi = {(R)0, (R)(-dir)};
  if (N == 4) {
    Cx<R> xee = in[si], xoe = in[si + sti], xeo = in[si + 2 * sti], xoo = in[si + 3 * sti];
    Cx<R> xee_p_xeo = add(xee, xeo), xee_m_xeo = sub(xee, xeo);
    Cx<R> xoe_p_xoo = add(xoe, xoo);
    Cx<R> xoe_m_xoo = mul(neg(sub(xoe, xoo)), minusi);
    out[so] = add(xee_p_xeo, xoe_p_xoo);
    out[so + sto] = add(xee_m_xeo, xoe_m_xoo);
    out[so + 2 * sto] = sub(xee_p_xeo, xoe_p_xoo);
    out[so + 3 * sto] = sub(xee_m_xeo, xoe_m_xoo);
    return;
  }
  int64_t m = N / 4;
  fft_pow2_radix4<R>(out, in, m, so, sto, si, sti * 4, dir);
  fft_pow2_radix4<R>(out, in, m, so + m * sto, sto, si + sti, sti * 4, dir);
  fft_pow2_radix4<R>(out, in, m, so + 2 * m * sto, sto, si + 2 * sti, sti * 4, dir);
  fft_pow2_radix4<R>(out, in, m, so + 3 * m * sto, sto, si + 3 * sti, sti * 4, dir);
  Cx<R> z1 = singleton_params<R>((R)(dir * (R)1) / (R)N);
  Cx<R> z2 = singleton_params<R>((R)(dir * (R)2) / (R)N);
  Cx<R> z3 = singleton_params<R>((R)(dir * (R)3) / (R)N);
  Cx<R> wkoe = {1, 0}, wkeo = {1, 0}, wkoo = {1, 0};
  for (int64_t k = 0; k < m; ++k) {
    int64_t kee = so + k * sto, koe = so + (k + m) * sto, keo = so + (k + 2 * m) * sto, koo = so + (k + 3 * m) * sto;
    Cx<R> y_kee = out[kee], y_koe = out[koe], y_keo = out[keo], y_koo = out[koo];
    Cx<R> t_koe = mul(y_koe, wkoe), t_keo = mul(y_keo, wkeo), t_koo = mul(y_koo, wkoo);
    Cx<R> a = add(y_kee, t_keo), b = sub(y_kee, t_keo);
    Cx<R> c = add(t_koe, t_koo);
    Cx<R> d = mul(neg(sub(t_koe, t_koo)), minusi);
    out[kee] = add(a, c);
    out[koe] = add(b, d);
    out[keo] = sub(a, c);
    out[koo] = sub(b, d);
    wkoe = singleton_step(wkoe, z1);
    wkeo = singleton_step(wkeo, z2);
    wkoo = singleton_step(wkoo, z3);
  }
}

// src/algos.jl:299-341
template <class R>
void fft_pow3(Cx<R>* out, const Cx<R>* in, int64_t N, int64_t so, int64_t sto, int64_t si, int64_t sti, Cx<R> minus120, int dir) {
  Cx<R> plus120 = conj(minus120);
  if (N == 3) {
    Cx<R> a = in[si], b = in[si + sti], c = in[si + 2 * sti];
    out[so] = add(add(a, b), c);
    out[so + sto] = add(add(a, mul(b, plus120)), mul(c, minus120));
    out[so + 2 * sto] = add(add(a, mul(b, minus120)), mul(c, plus120));
    return;
  }
  int64_t Np = N / 3;
  fft_pow3<R>(out, in, Np, so, sto, si, sti * 3, minus120, dir);
  fft_pow3<R>(out, in, Np, so + Np * sto, sto, si + sti, sti * 3, minus120, dir);
  fft_pow3<R>(out, in, Np, so + 2 * Np * sto, sto, si + 2 * sti, sti * 3, minus120, dir);
  Cx<R> z1 = singleton_params<R>((R)(dir * (R)1) / (R)N);
  Cx<R> z2 = singleton_params<R>((R)(dir * (R)2) / (R)N);
  Cx<R> wk1 = {1, 0}, wk2 = {1, 0};
  for (int64_t k = 0; k < Np; ++k) {
    int64_t k0 = so + sto * k, k1 = so + sto * (k + Np), k2 = so + sto * (k + 2 * Np);
    Cx<R> y0 = out[k0], y1 = out[k1], y2 = out[k2];
    Cx<R> t1 = mul(y1, wk1), t2 = mul(y2, wk2);
    out[k0] = add(add(y0, t1), t2);
    out[k1] = add(add(y0, mul(t1, plus120)), mul(t2, minus120));
    out[k2] = add(add(y0, mul(t1, minus120)), mul(t2, plus120));
    wk1 = singleton_step(wk1, z1);
    wk2 = singleton_step(wk2, z2);
  }
}

// src/algos.jl:344-367.  The chirp is generated in Float64 (Int/Int -> Float64,
// cispi in Float64) and then converted to R.
template <class R>
struct BlueScratch {
  std::vector<Cx<R>> tmp, a, b;
  int64_t L;
};
template <class R>
BlueScratch<R> prealloc_blue(int64_t N, int dir) {
  BlueScratch<R> s;
  int64_t L = 1;
  while (L < 2 * N - 1) L <<= 1;
  s.L = L;
  s.tmp.assign(L, {0, 0});
  s.a.assign(L, {0, 0});
  s.b.assign(L, {0, 0});
  int64_t sgn = -dir;
  int64_t p = 0;
  for (int64_t i = 1; i <= N; ++i) {
    double sn, cs;
    sincospi_r((double)(sgn * p) / (double)N, &sn, &cs);
    s.b[i - 1] = {(R)cs, (R)sn};
    p += 2 * i - 1;
    if (p > N) p -= 2 * N;
  }
  for (int64_t j = 0; j < N - 1 && L - 1 - j >= 0; ++j) s.b[L - 1 - j] = s.b[1 + j];
  return s;
}

// src/algos.jl:389-424
template <class R>
void fft_bluestein(Cx<R>* out, const Cx<R>* in, int dir, int64_t N, int64_t so, int64_t sto, int64_t si, int64_t sti,
                   BlueScratch<R>& s) {
  int64_t L = s.L;
  Cx<R>* tmp = s.tmp.data();
  Cx<R>* a = s.a.data();
  Cx<R>* b = s.b.data();
  for (int64_t i = N; i < L; ++i) { a[i] = {0, 0}; tmp[i] = {0, 0}; }
  for (int64_t i = 0; i < N; ++i) a[i] = mul(in[si + i * sti], conj(b[i]));
  fft_pow2_radix4<R>(tmp, a, L, 0, 1, 0, 1, +1);
  fft_pow2_radix4<R>(a, b, L, 0, 1, 0, 1, +1);
  for (int64_t i = 0; i < L; ++i) tmp[i] = mul(tmp[i], a[i]);
  fft_pow2_radix4<R>(a, tmp, L, 0, 1, 0, 1, -1);
  for (int64_t i = 0; i < N; ++i) {
    Cx<R> v = mul(conj(b[i]), a[i]);
    tmp[i] = {v.re / (R)L, v.im / (R)L};
  }
  for (int64_t i = 0; i < N; ++i) out[so + i * sto] = tmp[i];
}

template <class R> void fft_node(Cx<R>* out, const Cx<R>* in, int64_t so, int64_t si, int dir, Graph<R>& g, int64_t idx);

// src/algos.jl:51-118
template <class R>
void fft_composite(Cx<R>* out, const Cx<R>* in, int64_t so, int64_t si, int dir, Graph<R>& g, int64_t idx) {
  const Node root = g.nodes[idx];
  int64_t li = idx + root.left, ri = idx + root.right;
  const Node left = g.nodes[li], right = g.nodes[ri];
  int64_t N = root.sz, N1 = left.sz, N2 = right.sz;
  int64_t s_in = root.s_in, s_out = root.s_out;
  Cx<R>* tmp = g.workspace[idx].data();
  bool rblue = right.type == BLUESTEIN, lblue = left.type == BLUESTEIN;
  BlueScratch<R> rs, ls;
  if (rblue) rs = prealloc_blue<R>(N2, dir);
  for (int64_t j1 = 0; j1 < N1; ++j1) {
    int64_t r_si = si + j1 * s_in;
    int64_t r_so = N2 * j1;
    if (rblue) fft_bluestein<R>(tmp, in, dir, N2, r_so, right.s_out, r_si, right.s_in, rs);
    else fft_node<R>(tmp, in, r_so, r_si, dir, g, ri);
    if (j1 > 0) {
      Cx<R> zj1 = singleton_params<R>((R)(dir * (R)j1) / (R)N);
      Cx<R> wk2 = {1, 0};
      for (int64_t k2 = 1; k2 < N2; ++k2) {
        wk2 = singleton_step(wk2, zj1);
        tmp[r_so + k2] = mul(tmp[r_so + k2], wk2);
      }
    }
  }
  if (lblue) ls = prealloc_blue<R>(N1, dir);
  for (int64_t k2 = 0; k2 < N2; ++k2) {
    int64_t l_so = so + k2 * s_out;
    int64_t l_si = k2;
    if (lblue) fft_bluestein<R>(out, tmp, dir, N1, l_so, left.s_out, l_si, left.s_in, ls);
    else fft_node<R>(out, tmp, l_so, l_si, dir, g, li);
  }
}

// src/algos.jl:5-34
template <class R>
void fft_node(Cx<R>* out, const Cx<R>* in, int64_t so, int64_t si, int dir, Graph<R>& g, int64_t idx) {
  const Node n = g.nodes[idx];
  switch (n.type) {
    case COMPOSITE_FFT: fft_composite<R>(out, in, so, si, dir, g, idx); break;
    case DFT: fft_dft<R>(out, in, n.sz, so, n.s_out, si, n.s_in, dir); break;
    case POW2RADIX4_FFT: fft_pow2_radix4<R>(out, in, n.sz, so, n.s_out, si, n.s_in, dir); break;
    case POW3_FFT: {
      // cispi(T(2)/3) in R; forward keeps it, backward conjugates.
      R s, c;
      sincospi_r((R)((R)2 / (R)3), &s, &c);
      Cx<R> m120 = {c, s};
      if (dir != -1) m120 = conj(m120);
      fft_pow3<R>(out, in, n.sz, so, n.s_out, si, n.s_in, m120, dir);
      break;
    }
    case BLUESTEIN: {
      BlueScratch<R> s = prealloc_blue<R>(n.sz, dir);
      fft_bluestein<R>(out, in, dir, n.sz, so, n.s_out, si, n.s_in, s);
      break;
    }
  }
}

// One 1-D transform of contiguous data (src/plan.jl:157-167).
template <class R>
void fft1d(const Cx<R>* in, Cx<R>* out, int64_t n, int dir, int64_t cutoff) {
  Graph<R> g = make_graph<R>(n, cutoff);
  fft_node<R>(out, in, 0, 0, dir, g, 0);
}

// N-D: copy, then for each region dim in ascending order gather/fft/scatter
// every line (src/plan.jl:210-257, 287-348).  dims are column-major sizes.
template <class R>
void fftn(const Cx<R>* in, Cx<R>* out, int ndims, const int64_t* dims, int nreg, const int64_t* region, int dir,
          int64_t cutoff) {
  int64_t total = 1;
  for (int d = 0; d < ndims; ++d) total *= dims[d];
  std::memmove(out, in, sizeof(Cx<R>) * total);
  std::vector<int64_t> reg(region, region + nreg);
  std::sort(reg.begin(), reg.end());
  for (int64_t r1 : reg) {
    int dim = (int)r1 - 1;
    int64_t n = dims[dim];
    int64_t pre = 1, post = 1;
    for (int d = 0; d < dim; ++d) pre *= dims[d];
    for (int d = dim + 1; d < ndims; ++d) post *= dims[d];
    Graph<R> g = make_graph<R>(n, cutoff);
    std::vector<Cx<R>> ib(n), ob(n);
    for (int64_t ip = 0; ip < post; ++ip)
      for (int64_t ie = 0; ie < pre; ++ie) {
        int64_t base = ie + ip * pre * n;
        for (int64_t j = 0; j < n; ++j) ib[j] = out[base + j * pre];
        fft_node<R>(ob.data(), ib.data(), 0, 0, dir, g, 0);
        for (int64_t j = 0; j < n; ++j) out[base + j * pre] = ob[j];
      }
  }
}

template <class R>
void batch_run(int64_t count, const int64_t* lengths, const int64_t* offsets, const int* dirs, const Cx<R>* in,
               Cx<R>* out, int nthreads, int64_t cutoff) {
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    for (;;) {
      int64_t i = next.fetch_add(1);
      if (i >= count) break;
      fft1d<R>(in + offsets[i], out + offsets[i], lengths[i], dirs[i], cutoff);
    }
  };
  if (nthreads <= 1) { worker(); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
}

}  // namespace

extern "C" {

// Planner tree: writes up to `cap` nodes as 6 int64 each
// (left, right, type, sz, s_in, s_out); returns node count or -1 for N<=0.
int64_t oracle_plan_tree(int64_t n, int64_t cutoff, int64_t* out, int64_t cap) {
  if (n <= 0) return -1;
  std::vector<Node> nodes;
  build_node(nodes, n, cutoff, 1, 1);
  for (size_t i = 0; i < nodes.size() && (int64_t)i < cap; ++i) {
    const Node& q = nodes[i];
    int64_t* o = out + 6 * i;
    o[0] = q.left; o[1] = q.right; o[2] = q.type; o[3] = q.sz; o[4] = q.s_in; o[5] = q.s_out;
  }
  return (int64_t)nodes.size();
}

// Bluestein chirp index sequence p_n (src/algos.jl:354-358): p_0..p_{n-1}.
void oracle_chirp_index(int64_t n, int64_t* p_out) {
  int64_t p = 0;
  for (int64_t i = 1; i <= n; ++i) {
    p_out[i - 1] = p;
    p += 2 * i - 1;
    if (p > n) p -= 2 * n;
  }
}

// Bluestein chirp b_n in Float64 for direction dir (src/algos.jl:353-359).
void oracle_chirp(int64_t n, int dir, double* out) {
  int64_t sgn = -dir, p = 0;
  for (int64_t i = 1; i <= n; ++i) {
    double s, c;
    sincospi_r((double)(sgn * p) / (double)n, &s, &c);
    out[2 * (i - 1)] = c;
    out[2 * (i - 1) + 1] = s;
    p += 2 * i - 1;
    if (p > n) p -= 2 * n;
  }
}

// 1-D transforms of interleaved complex data; return 0 or -1 (n<=0).
int oracle_fft_c64(const double* in, double* out, int64_t n, int dir, int64_t cutoff) {
  if (n <= 0) return -1;
  fft1d<double>((const Cx<double>*)in, (Cx<double>*)out, n, dir, cutoff);
  return 0;
}
int oracle_fft_c32(const float* in, float* out, int64_t n, int dir, int64_t cutoff) {
  if (n <= 0) return -1;
  fft1d<float>((const Cx<float>*)in, (Cx<float>*)out, n, dir, cutoff);
  return 0;
}

// N-D transform over a (1-based) region of a column-major array.
int oracle_fftn_c64(const double* in, double* out, int ndims, const int64_t* dims, int nreg, const int64_t* region,
                    int dir, int64_t cutoff) {
  for (int d = 0; d < ndims; ++d) if (dims[d] <= 0) return -1;
  fftn<double>((const Cx<double>*)in, (Cx<double>*)out, ndims, dims, nreg, region, dir, cutoff);
  return 0;
}
int oracle_fftn_c32(const float* in, float* out, int ndims, const int64_t* dims, int nreg, const int64_t* region,
                    int dir, int64_t cutoff) {
  for (int d = 0; d < ndims; ++d) if (dims[d] <= 0) return -1;
  fftn<float>((const Cx<float>*)in, (Cx<float>*)out, ndims, dims, nreg, region, dir, cutoff);
  return 0;
}

// Many independent 1-D transforms, one transform per thread (CPU baseline).
void oracle_fft_batch_c32(int64_t count, const int64_t* lengths, const int64_t* offsets, const int* dirs,
                          const float* in, float* out, int nthreads, int64_t cutoff) {
  batch_run<float>(count, lengths, offsets, dirs, (const Cx<float>*)in, (Cx<float>*)out, nthreads, cutoff);
}
void oracle_fft_batch_c64(int64_t count, const int64_t* lengths, const int64_t* offsets, const int* dirs,
                          const double* in, double* out, int nthreads, int64_t cutoff) {
  batch_run<double>(count, lengths, offsets, dirs, (const Cx<double>*)in, (Cx<double>*)out, nthreads, cutoff);
}

// Truth: naive DFT in long double with the angle reduced exactly,
// X[u] = sum_v x[v] * exp(dir * 2*pi*i * ((u*v) mod n) / n).
void oracle_naive_dft_ld(const double* in, double* out, int64_t n, int dir) {
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<long double> cs(n), sn(n);
  for (int64_t m = 0; m < n; ++m) {
    long double a = 2.0L * pi * (long double)m / (long double)n;
    cs[m] = cosl(a);
    sn[m] = (long double)dir * sinl(a);
  }
  for (int64_t u = 0; u < n; ++u) {
    long double re = 0, im = 0;
    for (int64_t v = 0; v < n; ++v) {
      int64_t m = (int64_t)(((__int128)u * v) % n);
      long double xr = in[2 * v], xi = in[2 * v + 1];
      re += xr * cs[m] - xi * sn[m];
      im += xr * sn[m] + xi * cs[m];
    }
    out[2 * u] = (double)re;
    out[2 * u + 1] = (double)im;
  }
}

}  // extern "C"
