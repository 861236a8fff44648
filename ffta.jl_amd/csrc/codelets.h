// codelets.h -- forward in-register DFTs of radix 2, 3, 4, 5, 7, 8, 9, 11,
// 13, 16, 17, 19 and 23 (X[k] = sum_n x[n] exp(-2 pi i n k / R)), and of every 13-smooth
// composite length as a two-factor split of those.  The backward
// transform is obtained by conjugating on load and on store, so only the
// forward sign exists on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <utility>
#include "trig_tables.h"

namespace fftahip {

template <class R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};

template <class R> __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> __device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class R> __device__ __forceinline__ cx<R> cconj(cx<R> a) { return {a.x, -a.y}; }
// multiply by -i
template <class R> __device__ __forceinline__ cx<R> cmul_mi(cx<R> a) { return {a.y, -a.x}; }
template <class R> __device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.x * s, a.y * s}; }
// multiply by exp(-2 pi i m / N) with compile-time m, N via Trig<N>
template <class R, int N, int M>
__device__ __forceinline__ cx<R> ctw(cx<R> a) {
  constexpr int m = M % N;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == N) {
    return cmul_mi(a);
  } else if constexpr (2 * m == N) {
    return {-a.x, -a.y};
  } else if constexpr (4 * m == 3 * N) {
    return {-a.y, a.x};
  } else {
    const R c = (R)Trig<N>::c[m];
    const R s = -(R)Trig<N>::s[m];
    return {a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

template <class R>
__device__ __forceinline__ void dft2(cx<R>* v) {
  cx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <class R>
__device__ __forceinline__ void dft3(cx<R>* v) {
  const R s3 = (R)Trig<3>::s[1];
  cx<R> a = v[0], b = v[1], c = v[2];
  cx<R> t1 = cadd(b, c);
  cx<R> t2 = {a.x - (R)0.5 * t1.x, a.y - (R)0.5 * t1.y};
  cx<R> d = csub(b, c);
  cx<R> t3 = cmul_mi(cscale(d, s3));
  v[0] = cadd(a, t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <class R>
__device__ __forceinline__ void dft4(cx<R>* v) {
  cx<R> a = v[0], b = v[1], c = v[2], d = v[3];
  cx<R> s0 = cadd(a, c), d0 = csub(a, c);
  cx<R> s1 = cadd(b, d), d1 = cmul_mi(csub(b, d));
  v[0] = cadd(s0, s1);
  v[1] = cadd(d0, d1);
  v[2] = csub(s0, s1);
  v[3] = csub(d0, d1);
}

// Odd prime R: pair x_j with x_{R-j}.
template <class R, int P>
__device__ __forceinline__ void dft_prime(cx<R>* v) {
  constexpr int H = (P - 1) / 2;
  cx<R> s[H], d[H];
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    s[j - 1] = cadd(v[j], v[P - j]);
    d[j - 1] = csub(v[j], v[P - j]);
  }
  cx<R> x0 = v[0];
  cx<R> sum = x0;
#pragma unroll
  for (int j = 0; j < H; ++j) sum = cadd(sum, s[j]);
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    cx<R> A = x0;
    cx<R> B = {(R)0, (R)0};
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const int m = (j * k) % P;
      const R c = (R)Trig<P>::c[m];
      const R sn = (R)Trig<P>::s[m];
      A.x += c * s[j - 1].x;
      A.y += c * s[j - 1].y;
      B.x += sn * d[j - 1].x;
      B.y += sn * d[j - 1].y;
    }
    // X_k = A - i B, X_{P-k} = A + i B
    v[k] = {A.x + B.y, A.y - B.x};
    v[P - k] = {A.x - B.y, A.y + B.x};
  }
  v[0] = sum;
}

// Composite P*Q from P-point and Q-point pieces (input index q + Q*p,
// output index k1 + P*k2).
template <class R, int P> __device__ __forceinline__ void dft_any(cx<R>* v);

template <class R, int N, int K1, int... qs>
__device__ __forceinline__ void twiddle_row(cx<R>* t, std::integer_sequence<int, qs...>) {
  ((t[qs] = ctw<R, N, qs * K1>(t[qs])), ...);
}

template <class R, int P, int Q, int K1>
__device__ __forceinline__ void split_row(cx<R>* v, const cx<R>* y) {
  cx<R> t[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) t[q] = y[q + Q * K1];
  twiddle_row<R, P * Q, K1>(t, std::make_integer_sequence<int, Q>{});
  dft_any<R, Q>(t);
#pragma unroll
  for (int k2 = 0; k2 < Q; ++k2) v[K1 + P * k2] = t[k2];
}

template <class R, int P, int Q, int... k1s>
__device__ __forceinline__ void split_rows(cx<R>* v, const cx<R>* y, std::integer_sequence<int, k1s...>) {
  (split_row<R, P, Q, k1s>(v, y), ...);
}

template <class R, int P, int Q>
__device__ __forceinline__ void dft_split(cx<R>* v) {
  constexpr int N = P * Q;
  cx<R> y[N];
  // Q DFTs of length P over x[q + Q*p]
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    cx<R> t[P];
#pragma unroll
    for (int p = 0; p < P; ++p) t[p] = v[q + Q * p];
    dft_any<R, P>(t);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) y[q + Q * k1] = t[k1];
  }
  // twiddle w_N^{q*k1}, then P DFTs of length Q over q
  split_rows<R, P, Q>(v, y, std::make_integer_sequence<int, P>{});
}

// First factor of a composite length without a codelet of its own: the first
// codelet radix of (16, 9, 8, 13, 11, 7, 5, 4, 3, 2) that divides N.
constexpr int split_radix(int N) {
  const int order[10] = {16, 9, 8, 13, 11, 7, 5, 4, 3, 2};
  for (int r : order)
    if (N % r == 0 && N > r) return r;
  return N;
}

template <class R, int P>
__device__ __forceinline__ void dft_any(cx<R>* v) {
  if constexpr (P == 1) {
  } else if constexpr (P == 2) {
    dft2<R>(v);
  } else if constexpr (P == 3) {
    dft3<R>(v);
  } else if constexpr (P == 4) {
    dft4<R>(v);
  } else if constexpr (P == 8) {
    dft_split<R, 4, 2>(v);
  } else if constexpr (P == 9) {
    dft_split<R, 3, 3>(v);
  } else if constexpr (P == 16) {
    dft_split<R, 4, 4>(v);
  } else if constexpr (P == 5 || P == 7 || P == 11 || P == 13 || P == 17 || P == 19 || P == 23) {
    dft_prime<R, P>(v);
  } else {
    // 13-smooth composite (register line passes): two-factor split
    static_assert(split_radix(P) < P, "no codelet split for this length");
    dft_split<R, split_radix(P), P / split_radix(P)>(v);
  }
}

}  // namespace fftahip
