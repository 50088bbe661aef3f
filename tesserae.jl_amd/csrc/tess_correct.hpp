// Corrected bases on the device: WLS(kernel, poly) and
// KernelCorrection(kernel, poly) (src/Basis/wls.jl, src/Basis/kernelcorrection.jl,
// src/Basis/polynomials.jl).  A corrected weight is the kernel weight times an
// affine (Linear) or multilinear (MultiLinear) polynomial of the node offset,
// fitted per particle so that the weights reproduce linear fields:
//   w_i = N_i q_0(i),  ∇w_i = N_i q_{1..D}(i),  q(i) = M⁻¹ P(x_i - x_p),
//   M = Σ_i N_i P(x_i - x_p) P(x_i - x_p)ᵀ  (filter[i] = 0 drops node i).
// Three forms cover every branch the reference takes:
//   CORR_PLAIN    the kernel itself (KernelCorrection on a whole support);
//   CORR_AXES     a product of per-axis 1-D fits: a MultiLinear fit of a
//                 quadratic or cubic B-spline without a filter factorises by
//                 axis (wls.jl:79-99, kernelcorrection.jl:53-63), and the
//                 closed-form moments of a whole B-spline support (wls.jl:30-49)
//                 are the special case where every axis is whole.  The axis
//                 jets are rewritten in place, so the tensor-product code of the
//                 plain kernel forms w and ∇w unchanged;
//   CORR_GENERAL  the full (D+1)- or 2^D-term fit (wls.jl:51-77, 131-158).
// Offsets enter in units of h (δ̂ = (x_i - x_p)/h), which leaves w unchanged
// and scales the fitted derivative by h⁻¹ exactly as the plain kernel's
// derivative is scaled; the fit is then O(1)-conditioned in every precision.
#pragma once

#include "tess_device.hpp"

namespace tess {

enum { CORR_PLAIN = 0, CORR_AXES = 1, CORR_GENERAL = 2 };

__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

__host__ __device__ constexpr bool is_corrected(int basis) {
  return (basis & (TESS_WLS | TESS_KERNEL_CORRECTION)) != 0;
}

// The form of one particle's weights, following the reference's dispatch:
//   KernelCorrection: the kernel on a whole support that the filter passes
//     (kernelcorrection.jl:27-35, deferred :41-49), else WLS(kernel, poly);
//   WLS(kernel, Linear) of a quadratic or cubic B-spline: closed-form moments
//     on a whole support that the filter passes (stored, wls.jl:30-40) or a
//     whole support without a filter (deferred, wls.jl:131-138);
//   WLS(kernel, MultiLinear) of a quadratic or cubic B-spline without a
//     filter: the per-axis fits (wls.jl:79-99; the deferred transfer's general
//     fit, wls.jl:139-148, equals it: the moment matrix of a product kernel over
//     a box of nodes is the Kronecker product of the axis moment matrices);
//   anything else: the general fit.
__device__ __forceinline__ int corr_mode(int corr, int kernel, bool full, bool has_filter, bool filter_all,
                                         bool stored) {
  const bool ml = (corr & TESS_POLY_MULTILINEAR) != 0;
  const bool qc = kernel == TESS_BSPLINE_QUADRATIC || kernel == TESS_BSPLINE_CUBIC;
  if (corr & TESS_KERNEL_CORRECTION) {
    if (full && filter_all) return CORR_PLAIN;
    return (ml && qc && !has_filter) ? CORR_AXES : CORR_GENERAL;
  }
  if (!ml && qc && full && (stored ? filter_all : !has_filter)) return CORR_AXES;
  if (ml && qc && !has_filter) return CORR_AXES;
  return CORR_GENERAL;
}

// One axis of CORR_AXES: on a whole axis support the closed form
// M̂⁻¹ = diag(1, 12/(n+1)) (the normalised second moment of a degree-n
// B-spline is (n+1)/12, wls.jl:41-49): w = N, w' = N c δ̂; on a truncated axis the
// 2×2 fit of (1, δ̂).  val holds N on entry; val and der (normalised
// derivative) hold the corrected axis values on exit.  Entries past cnt are
// zeroed.
template <typename T, int SW>
__device__ __forceinline__ void wls_axis(int cnt, int degree, const T* dh, T* val, T* der) {
  if (cnt == SW) {
    const T c = T(12) / T(degree + 1);
#pragma unroll
    for (int a = 0; a < SW; ++a) der[a] = val[a] * (c * dh[a]);
    return;
  }
  T m0 = T(0), m1 = T(0), m2 = T(0);
#pragma unroll
  for (int a = 0; a < SW; ++a) {
    const T n = a < cnt ? val[a] : T(0);
    m0 += n;
    m1 += n * dh[a];
    m2 += (n * dh[a]) * dh[a];
  }
  const T rdet = T(1) / (m0 * m2 - m1 * m1);
  const T i00 = m2 * rdet, i01 = -m1 * rdet, i11 = m0 * rdet;
#pragma unroll
  for (int a = 0; a < SW; ++a) {
    const T n = a < cnt ? val[a] : T(0);
    val[a] = n * (i00 + i01 * dh[a]);
    der[a] = n * (i01 + i11 * dh[a]);
  }
}

// P(δ̂) of Polynomial(Linear()) (NP = D+1: 1, x1..xD) or
// Polynomial(MultiLinear()) (NP = 2^D: 1, x1, x2, [x3,] x1x2, [x2x3, x3x1,
// x1x2x3]), in the order of polynomials.jl:32-46.
template <typename T, int D, int NP>
__device__ __forceinline__ void poly_hat(const T* d, T* P) {
  P[0] = T(1);
#pragma unroll
  for (int k = 0; k < D; ++k) P[1 + k] = d[k];
  if constexpr (NP == 4 && D == 2) {
    P[3] = d[0] * d[1];
  } else if constexpr (NP == 8) {
    P[4] = d[0] * d[1];
    P[5] = d[1] * d[2];
    P[6] = d[2] * d[0];
    P[7] = d[0] * d[1] * d[2];
  }
}

// Rows 0..D of M̂⁻¹ (the only ones w and ∇w use: P₀ = e₀ and ∇P₀ = e_{1..D} at
// the particle, polynomials.jl:35-46) for the general fit over the window.
// M̂ = Σ w P̂P̂ᵀ with w = Π_d N_d × filter is symmetric positive definite when the
// fit is well posed; it is factorised by Cholesky and solved for e₀..e_D (a
// singular M̂ gives non-finite weights, as inv(M) does in the reference).
// val/dh: axis values N and offsets δ̂ per window position; lo, cnt: the
// window; filter: NULL or one byte per mesh node.
template <typename T, int D, int SW, int NP>
__device__ void general_fit(const T (&val)[D][SW], const T (&dh)[D][SW], const int* lo, const int* cnt,
                            const int32_t* dims, const uint8_t* __restrict__ filter, T (&C)[D + 1][NP]) {
  T M[NP][NP];
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int j = 0; j < NP; ++j) M[i][j] = T(0);
  constexpr int NOFF = ipow(SW, D);
#pragma unroll
  for (int o = 0; o < NOFF; ++o) {
    int aa[D], r = o;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      aa[d] = r % SW;
      r /= SW;
      ok = ok && aa[d] < cnt[d];
    }
    if (!ok) continue;
    T w = val[0][aa[0]], dd[D];
#pragma unroll
    for (int d = 1; d < D; ++d) w = w * val[d][aa[d]];
    if (filter) {
      int64_t lin = lo[D - 1] + aa[D - 1];
#pragma unroll
      for (int d = D - 2; d >= 0; --d) lin = lin * dims[d] + (lo[d] + aa[d]);
      w = filter[lin] ? w : T(0);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) dd[d] = dh[d][aa[d]];
    T P[NP];
    poly_hat<T, D, NP>(dd, P);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const T wp = w * P[i];
#pragma unroll
      for (int j = 0; j <= i; ++j) M[i][j] += wp * P[j];
    }
  }
  // Cholesky M̂ = L Lᵀ in place (lower triangle), with the reciprocal diagonal
  T rd[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    T s = M[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= M[j][k] * M[j][k];
    const T Ljj = sqrt(s);
    rd[j] = T(1) / Ljj;
#pragma unroll
    for (int i = j + 1; i < NP; ++i) {
      T t = M[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= M[i][k] * M[j][k];
      M[i][j] = t * rd[j];
    }
  }
#pragma unroll
  for (int r = 0; r <= D; ++r) {
    T z[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      T t = (i == r) ? T(1) : T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) t -= M[i][k] * z[k];
      z[i] = t * rd[i];
    }
#pragma unroll
    for (int i = NP - 1; i >= 0; --i) {
      T t = z[i];
#pragma unroll
      for (int k = i + 1; k < NP; ++k) t -= M[k][i] * C[r][k];
      C[r][i] = t * rd[i];
    }
  }
}

// w and ∇w of one node under the general fit: q̂ = C P̂(δ̂), w = N q̂₀,
// ∇w_c = N q̂_{1+c} h⁻¹ (N = Π_d N_d × filter).
template <typename T, int D, int SW, int NP>
__device__ __forceinline__ void general_node(const T (&val)[D][SW], const T (&dh)[D][SW], const int* aa, bool pass,
                                             const T (&C)[D + 1][NP], T hinv, T& w, T* g) {
  T n = val[0][aa[0]], dd[D];
#pragma unroll
  for (int d = 1; d < D; ++d) n = n * val[d][aa[d]];
  n = pass ? n : T(0);
#pragma unroll
  for (int d = 0; d < D; ++d) dd[d] = dh[d][aa[d]];
  T P[NP];
  poly_hat<T, D, NP>(dd, P);
  T q[D + 1];
#pragma unroll
  for (int r = 0; r <= D; ++r) {
    T s = C[r][0] * P[0];
#pragma unroll
    for (int j = 1; j < NP; ++j) s += C[r][j] * P[j];
    q[r] = s;
  }
  w = n * q[0];
#pragma unroll
  for (int c = 0; c < D; ++c) g[c] = (n * q[1 + c]) * hinv;
}

// filter[node] of window position aa (true without a filter).
template <int D>
__device__ __forceinline__ bool filter_pass(const uint8_t* __restrict__ filter, const int32_t* dims, const int* lo,
                                            const int* aa) {
  if (!filter) return true;
  int64_t lin = lo[D - 1] + aa[D - 1];
#pragma unroll
  for (int d = D - 2; d >= 0; --d) lin = lin * dims[d] + (lo[d] + aa[d]);
  return filter[lin] != 0;
}

// alltrue(filter, window) (basis_weight.jl:593-607).
template <int D, int SW>
__device__ __forceinline__ bool filter_alltrue(const uint8_t* __restrict__ filter, const int32_t* dims, const int* lo,
                                               const int* cnt) {
  if (!filter) return true;
  bool all = true;
  constexpr int NOFF = ipow(SW, D);
  for (int o = 0; o < NOFF; ++o) {
    int aa[D], r = o;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      aa[d] = r % SW;
      r /= SW;
      ok = ok && aa[d] < cnt[d];
    }
    if (ok) all = all && filter_pass<D>(filter, dims, lo, aa);
  }
  return all;
}

// Per-particle state of a corrected basis in a transfer or an update!: the
// form, the fitted rows for CORR_GENERAL (Linear or MultiLinear), and the
// normalised offsets.  prepare() takes the nodewise axis values in val/der and
// the offsets ξ_a = (x_p - x_a)/h used to evaluate them; for CORR_AXES it
// rewrites val/der in place.
template <typename T, int D, int SW>
struct Corrected {
  static constexpr int NL = D + 1;
  static constexpr int NM = 1 << D;
  int mode;
  bool ml;
  T dh[D][SW];
  T CL[NL][NL];
  T CM[D + 1][NM];

  __device__ __forceinline__ void prepare(int corr, int kernel, bool stored, const uint8_t* __restrict__ filter,
                                          const int32_t* dims, const int* lo, const int* cnt, const T (&xi)[D][SW],
                                          T (&val)[D][SW], T (&der)[D][SW]) {
    bool full = true;
#pragma unroll
    for (int d = 0; d < D; ++d) full = full && cnt[d] == SW;
    const bool fa = filter_alltrue<D, SW>(filter, dims, lo, cnt);
    mode = corr_mode(corr, kernel, full, filter != nullptr, fa, stored);
    ml = (corr & TESS_POLY_MULTILINEAR) != 0;
    if (mode == CORR_PLAIN) return;
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int a = 0; a < SW; ++a) dh[d][a] = -xi[d][a];
    if (mode == CORR_AXES) {
#pragma unroll
      for (int d = 0; d < D; ++d) wls_axis<T, SW>(cnt[d], degree_of(kernel), dh[d], val[d], der[d]);
      return;
    }
    if (ml) {
      T C[D + 1][NM];
      general_fit<T, D, SW, NM>(val, dh, lo, cnt, dims, filter, C);
#pragma unroll
      for (int r = 0; r <= D; ++r)
#pragma unroll
        for (int j = 0; j < NM; ++j) CM[r][j] = C[r][j];
    } else {
      general_fit<T, D, SW, NL>(val, dh, lo, cnt, dims, filter, CL);
    }
  }

  // w and ∇w (world units) of window position aa under CORR_GENERAL
  __device__ __forceinline__ void node(const T (&val)[D][SW], const int* aa, bool pass, T hinv, T& w, T* g) const {
    if (ml) {
      T C[D + 1][NM];
#pragma unroll
      for (int r = 0; r <= D; ++r)
#pragma unroll
        for (int j = 0; j < NM; ++j) C[r][j] = CM[r][j];
      general_node<T, D, SW, NM>(val, dh, aa, pass, C, hinv, w, g);
    } else {
      general_node<T, D, SW, NL>(val, dh, aa, pass, CL, hinv, w, g);
    }
  }
};

}  // namespace tess
