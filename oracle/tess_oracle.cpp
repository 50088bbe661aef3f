// tess_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of Tesserae.jl v0.7.5's explicit MPM step (the README
// program, README.md:44-87) used as the parity checker for libtesserae_hip and
// as the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library; the product path never
// does.  Every function names the reference file:line it follows.
//
// Parity status of each part (what pins it, beyond this restatement):
//   mesh geometry, windows, findblock ........ pinned by test/mesh.jl:27-65 KATs
//   B-spline / uGIMP values and gradients .... pinned by closed forms, partition of
//                                              unity, linear reproduction, jet KATs
//                                              (test/basis.jl:226-346), brute force
//   CPU block partition ...................... pinned by brute-force grouping and
//                                              colour-disjointness (test/partition.jl:53-158)
//   P2G / grid update / G2P .................. pinned by conservation invariants, affine
//                                              reproduction and a brute-force dense loop
//   Neo-Hookean stress ....................... pinned by the strain-energy derivative
//                                              (J σ = ∂ψ/∂F Fᵀ) with an LU determinant;
//                                              Tensorial.jl's det rounding order unpinned
//
// Arithmetic order follows the reference where it is visible in its source:
// Horner evaluation with muladd (Julia evalpoly), left-associated tensor
// products, hoisted particle factors (p2g.jl:272-307), per-particle temporaries
// in G2P (g2p.jl:123-144).  Tensorial.jl's internal summation order for
// Mat*Vec and det is not in the reference tree; left-to-right sums are used.
//
// Build: g++ -O2 -std=c++17 -fPIC -shared -pthread (see oracle/Makefile).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../include/tesserae_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <typename T>
struct Fld {
  T* p;
  int64_t es, cs;
  Fld() : p(nullptr), es(0), cs(0) {}
  explicit Fld(const tess_field& f) : p(static_cast<T*>(f.ptr)), es(f.elem_stride), cs(f.comp_stride) {}
  T& operator()(int64_t e, int c = 0) const { return p[e * es + c * cs]; }
  bool ok() const { return p != nullptr; }
};

// CartesianMesh as its device copy holds it (cartesian.jl:38-42, adapt.jl:53-57):
// xmin, h, h⁻¹ cast to T; node coordinate T(xmin + i*h).
template <typename T, int D>
struct Mesh {
  int64_t dims[D];
  T xmin[D];
  T h, h_inv;
  double xmin_d[D], h_d;
  int L;
  explicit Mesh(const tess_mesh& m) {
    for (int d = 0; d < D; ++d) {
      dims[d] = m.dims[d];
      xmin[d] = T(m.xmin[d]);
      xmin_d[d] = m.xmin[d];
    }
    h = T(m.h);
    h_inv = T(m.h_inv);
    h_d = m.h;
    L = m.block_size_log2;
  }
  T node(int d, int64_t i) const { return T(xmin_d[d] + double(i) * h_d); }
  int64_t nnodes() const {
    int64_t n = 1;
    for (int d = 0; d < D; ++d) n *= dims[d];
    return n;
  }
  int64_t lin(const int64_t* I) const {
    int64_t l = 0;
    for (int d = D - 1; d >= 0; --d) l = l * dims[d] + I[d];
    return l;
  }
  int64_t nblocks(int d) const { return ((dims[d] - 1) >> L) + 1; }  // cartesian.jl:241-243
};

// ---- basis (bspline.jl, gimp.jl, basis_weight.jl) -------------------------

constexpr int support_width(int basis) {
  basis &= 0xff;  // the kernel of a corrected basis
  return basis == TESS_BSPLINE_LINEAR ? 2 : basis == TESS_BSPLINE_QUADRATIC ? 3 : basis == TESS_BSPLINE_CUBIC ? 4 : 3;
}
constexpr int degree_of(int basis) { return (basis & 0xff) == TESS_UGIMP ? 2 : (basis & 0xff); }

// Julia evalpoly(x, (a0, a1, ...)) = muladd Horner from the top coefficient.
template <typename T>
T horner(T x, const T* a, int n) {
  T ex = a[n - 1];
  for (int i = n - 2; i >= 0; --i) ex = std::fma(x, ex, a[i]);
  return ex;
}

// Coefficient tables of _bspline_local_coeffs (bspline.jl:45-56) and their first
// derivative (:58-66), ascending powers; trailing zeros trimmed as in
// _evalpoly_expr (:68-75).  Rationals enter as T(num)/T(den).
template <typename T>
void bspline_axis_jet(int deg, T xi_local, T* val, T* der) {
  auto c = [](double num, double den) { return den == 1 ? T(num) : T(num) / T(den); };
  if (deg == 1) {
    T a0[2] = {c(1, 1), c(-1, 1)}, a1[2] = {c(0, 1), c(1, 1)};
    T d0[1] = {c(-1, 1)}, d1[1] = {c(1, 1)};
    val[0] = horner(xi_local, a0, 2); val[1] = horner(xi_local, a1, 2);
    der[0] = horner(xi_local, d0, 1); der[1] = horner(xi_local, d1, 1);
  } else if (deg == 2) {
    T a0[3] = {c(1, 2), c(-1, 1), c(1, 2)}, a1[3] = {c(1, 2), c(1, 1), c(-1, 1)}, a2[3] = {c(0, 1), c(0, 1), c(1, 2)};
    T d0[2] = {c(-1, 1), c(1, 1)}, d1[2] = {c(1, 1), c(-2, 1)}, d2[2] = {c(0, 1), c(1, 1)};
    val[0] = horner(xi_local, a0, 3); val[1] = horner(xi_local, a1, 3); val[2] = horner(xi_local, a2, 3);
    der[0] = horner(xi_local, d0, 2); der[1] = horner(xi_local, d1, 2); der[2] = horner(xi_local, d2, 2);
  } else {
    T a0[4] = {c(1, 6), c(-1, 2), c(1, 2), c(-1, 6)}, a1[4] = {c(2, 3), c(0, 1), c(-1, 1), c(1, 2)};
    T a2[4] = {c(1, 6), c(1, 2), c(1, 2), c(-1, 2)}, a3[4] = {c(0, 1), c(0, 1), c(0, 1), c(1, 6)};
    T d0[3] = {c(-1, 2), c(1, 1), c(-1, 2)}, d1[3] = {c(0, 1), c(-2, 1), c(3, 2)};
    T d2[3] = {c(1, 2), c(1, 1), c(-3, 2)}, d3[3] = {c(0, 1), c(0, 1), c(1, 2)};
    val[0] = horner(xi_local, a0, 4); val[1] = horner(xi_local, a1, 4);
    val[2] = horner(xi_local, a2, 4); val[3] = horner(xi_local, a3, 4);
    der[0] = horner(xi_local, d0, 3); der[1] = horner(xi_local, d1, 3);
    der[2] = horner(xi_local, d2, 3); der[3] = horner(xi_local, d3, 3);
  }
}

// Piecewise value and its ForwardDiff derivative (bspline.jl:176-189,
// gimp.jl:28-33, differentiated as basis_weight.jl:35-37 does): abs(Dual) is
// signbit(x) ? -x : x, x^2 / x^3 are literal powers.
template <typename T>
void piecewise_jet(int basis, T xi, T l, T* val, T* der) {
  T s = std::signbit(xi) ? T(-1) : T(1);
  T a = std::fabs(xi);
  if (basis == TESS_BSPLINE_LINEAR) {
    if (a < T(1)) { *val = T(1) - a; *der = -s; }
    else { *val = T(0); *der = T(0); }
  } else if (basis == TESS_BSPLINE_QUADRATIC) {
    if (a < T(0.5)) {
      *val = (T(3) - T(4) * (a * a)) / T(4);
      *der = -(T(4) * (T(2) * a * s)) / T(4);
    } else if (a < T(1.5)) {
      T u = T(3) - T(2) * a;
      *val = (u * u) / T(8);
      *der = (T(2) * u * (T(-2) * s)) / T(8);
    } else { *val = T(0); *der = T(0); }
  } else if (basis == TESS_BSPLINE_CUBIC) {
    if (a < T(1)) {
      T a2 = a * a, a3 = a * a * a;
      *val = (T(3) * a3 - T(6) * a2 + T(4)) / T(6);
      *der = (T(3) * (T(3) * a2 * s) - T(6) * (T(2) * a * s)) / T(6);
    } else if (a < T(2)) {
      T u = T(2) - a;
      *val = (u * u * u) / T(6);
      *der = (T(3) * (u * u) * (-s)) / T(6);
    } else { *val = T(0); *der = T(0); }
  } else {  // uGIMP, gimp.jl:28-33 with l normalised by h
    if (a < l / T(2)) {
      *val = T(1) - (T(4) * (a * a) + l * l) / (T(4) * l);
      *der = -(T(4) * (T(2) * a * s)) / (T(4) * l);
    } else if (a < T(1) - l / T(2)) {
      *val = T(1) - a; *der = -s;
    } else if (a < T(1) + l / T(2)) {
      T u = T(1) + l / T(2) - a;
      *val = (u * u) / (T(2) * l);
      *der = (T(2) * u * (-s)) / (T(2) * l);
    } else { *val = T(0); *der = T(0); }
  }
}

// One particle's BasisWeight: window (0-based inclusive, clamped) and the values
// at the local column-major positions of the clipped box inside sw^D storage.
template <typename T, int D>
struct PW {
  static constexpr int MAXSW = 4;
  int64_t lo[D], hi[D];    // clamped window, 0-based inclusive; empty if hi<lo
  int64_t start[D];        // unclamped start (0-based)
  bool empty;
  bool full;
  int sw;
  T w[MAXSW * MAXSW * MAXSW];
  T gw[MAXSW * MAXSW * MAXSW][D];
};

template <typename T, int D>
bool isinside_xi(const T* xi, const Mesh<T, D>& mesh) {  // cartesian.jl:157-163, 227-234
  for (int d = 0; d < D; ++d) {
    T f = std::floor(xi[d]);
    if (!(T(0) <= f && f <= T(mesh.dims[d] - 2))) return false;
  }
  return true;
}

// supportnodes (bspline.jl:27-40; gimp.jl:20-24 -> cartesian.jl:190-199).
template <typename T, int D>
void support_window(int basis, const T* x, T lp, const Mesh<T, D>& mesh, PW<T, D>& pw) {
  T xi[D];
  for (int d = 0; d < D; ++d) xi[d] = (x[d] - mesh.xmin[d]) * mesh.h_inv;  // normalize, cartesian.jl:146-150
  pw.sw = support_width(basis);
  pw.empty = !isinside_xi<T, D>(xi, mesh);
  if (pw.empty) {
    for (int d = 0; d < D; ++d) { pw.lo[d] = 0; pw.hi[d] = -1; pw.start[d] = 0; }
    pw.full = false;
    return;
  }
  bool full = true;
  for (int d = 0; d < D; ++d) {
    int64_t s, e;
    if (basis == TESS_UGIMP) {
      T l = lp / mesh.h;  // _normalized_particle_length, gimp.jl:25
      T r = T(1) + l / T(2);
      s = int64_t(std::floor(xi[d] - r)) + 1;
      e = int64_t(std::floor(xi[d] + r));
    } else {
      int n = degree_of(basis);
      T offset = T(n - 1) / T(2);
      s = int64_t(std::floor(xi[d] - offset));
      e = s + n;
    }
    pw.start[d] = s;
    pw.lo[d] = std::max<int64_t>(s, 0);
    pw.hi[d] = std::min<int64_t>(e, mesh.dims[d] - 1);
    if (pw.hi[d] - pw.lo[d] + 1 != pw.sw) full = false;
  }
  pw.full = full;
}

// update_basis_weight! (basis_weight.jl:636-642): fast full-support B-spline
// fill (bspline.jl:146-170) or the nodewise jet (basis_weight.jl:624-632).
// `deferred` selects the transfer-time evaluation of deferred weights, which is
// always nodewise (basis_weight.jl:442-446 -> bspline.jl:204-205).
template <typename T, int D>
void eval_weights(int basis, const T* x, T lp, const Mesh<T, D>& mesh, bool deferred, PW<T, D>& pw) {
  support_window<T, D>(basis, x, lp, mesh, pw);
  if (pw.empty) return;
  const int sw = pw.sw;
  int64_t ext[D];
  for (int d = 0; d < D; ++d) ext[d] = pw.hi[d] - pw.lo[d] + 1;
  T v1[D][PW<T, D>::MAXSW], d1[D][PW<T, D>::MAXSW];
  if (pw.full && basis != TESS_UGIMP && !deferred) {
    for (int d = 0; d < D; ++d) {
      T xi = (x[d] - mesh.xmin[d]) * mesh.h_inv;
      int n = degree_of(basis);
      T y = (n % 2 == 1) ? xi : xi - T(0.5);
      T loc = y - std::floor(y);  // fract
      bspline_axis_jet<T>(n, loc, v1[d], d1[d]);
    }
  } else {
    T l = basis == TESS_UGIMP ? lp / mesh.h : T(0);
    for (int d = 0; d < D; ++d)
      for (int64_t a = 0; a < ext[d]; ++a) {
        T xi = (x[d] - mesh.node(d, pw.lo[d] + a)) * mesh.h_inv;  // separable_nodal_basis_jet :49-60
        piecewise_jet<T>(basis, xi, l, &v1[d][a], &d1[d][a]);
      }
  }
  // local column-major positions of the clipped box inside the sw^D storage
  int64_t cnt = 1;
  for (int d = 0; d < D; ++d) cnt *= ext[d];
  for (int64_t k = 0; k < cnt; ++k) {
    int64_t a[D], r = k;
    for (int d = 0; d < D; ++d) { a[d] = r % ext[d]; r /= ext[d]; }
    int64_t slot = 0;
    for (int d = D - 1; d >= 0; --d) slot = slot * sw + a[d];
    T w = v1[0][a[0]];
    for (int d = 1; d < D; ++d) w = w * v1[d][a[d]];
    pw.w[slot] = w;
    for (int c = 0; c < D; ++c) {
      T g = (c == 0) ? d1[0][a[0]] : v1[0][a[0]];
      for (int d = 1; d < D; ++d) g = g * ((d == c) ? d1[d][a[d]] : v1[d][a[d]]);
      pw.gw[slot][c] = g * mesh.h_inv;
    }
  }
}

// Iterate the window column-major (eachindex(window)), calling f(slot, node_linear).
template <typename T, int D, typename Fn>
void for_window(const PW<T, D>& pw, const Mesh<T, D>& mesh, Fn&& f) {
  if (pw.empty) return;
  int64_t ext[D];
  int64_t cnt = 1;
  for (int d = 0; d < D; ++d) { ext[d] = pw.hi[d] - pw.lo[d] + 1; cnt *= ext[d]; }
  for (int64_t k = 0; k < cnt; ++k) {
    int64_t a[D], I[D], r = k;
    for (int d = 0; d < D; ++d) { a[d] = r % ext[d]; r /= ext[d]; I[d] = pw.lo[d] + a[d]; }
    int64_t slot = 0;
    for (int d = D - 1; d >= 0; --d) slot = slot * pw.sw + a[d];
    f(int(slot), mesh.lin(I));
  }
}

// ---- corrected bases (wls.jl, kernelcorrection.jl, polynomials.jl) --------

constexpr int kernel_of(int basis) { return basis & 0xff; }

// value(poly, x) (polynomials.jl:32-46): Linear (1, x1..xD); MultiLinear in 2D
// (1, x1, x2, x1x2), in 3D (1, x1, x2, x3, x1x2, x2x3, x3x1, x1x2x3).
template <typename T, int D>
int poly_value(bool multilinear, const T* x, T* P) {
  P[0] = T(1);
  for (int d = 0; d < D; ++d) P[1 + d] = x[d];
  if (!multilinear) return D + 1;
  if (D == 2) {
    P[3] = x[0] * x[1];
    return 4;
  }
  P[4] = x[0] * x[1];
  P[5] = x[1] * x[2];
  P[6] = x[2] * x[0];
  P[7] = x[0] * x[1] * x[2];
  return 8;
}

// inv(M) of an n×n matrix (row-major): Gauss-Jordan elimination with partial
// pivoting in T.  The reference calls Tensorial.jl's inv (not in its tree); a
// singular M gives non-finite entries in both.
template <typename T>
void invert(int n, const T* M, T* Minv) {
  T a[8][16];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) a[i][j] = j < n ? M[i * n + j] : (j - n == i ? T(1) : T(0));
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i][k]) > std::fabs(a[piv][k])) piv = i;
    if (piv != k)
      for (int j = 0; j < 2 * n; ++j) std::swap(a[k][j], a[piv][j]);
    const T p = a[k][k];
    for (int j = 0; j < 2 * n; ++j) a[k][j] = a[k][j] / p;
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      const T f = a[i][k];
      for (int j = 0; j < 2 * n; ++j) a[i][j] = a[i][j] - f * a[k][j];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Minv[i * n + j] = a[i][n + j];
}

// filter[i] over the window (alltrue, basis_weight.jl:593-607); NULL = Trues.
template <typename T, int D>
bool filter_alltrue(const uint8_t* filter, const PW<T, D>& pw, const Mesh<T, D>& mesh) {
  if (!filter) return true;
  bool all = true;
  for_window<T, D>(pw, mesh, [&](int, int64_t i) { all = all && filter[i] != 0; });
  return all;
}

// Nodewise kernel value (nodal_basis_jet(Order(0), kernel, pt, mesh, i),
// basis_weight.jl:49-60): the left-associated product of the axis values.
template <typename T, int D>
T kernel_value(int kernel, const T* x, T lp, const Mesh<T, D>& mesh, const int64_t* I) {
  const T l = kernel == TESS_UGIMP ? lp / mesh.h : T(0);
  T w = T(1);
  for (int d = 0; d < D; ++d) {
    T val, der;
    piecewise_jet<T>(kernel, (x[d] - mesh.node(d, I[d])) * mesh.h_inv, l, &val, &der);
    w = d == 0 ? val : w * val;
  }
  return w;
}

// apply_wls_correction! (wls.jl:66-77): wq = w·(M⁻¹P(xᵢ-xₚ)); w ← wq⊡P₀ = wq[0],
// ∇w ← wq⊡∇P₀ = wq[1..D] (the jets of Linear / MultiLinear at 0, polynomials.jl:35-46).
template <typename T, int D>
void apply_correction(bool multilinear, const T* x, const Mesh<T, D>& mesh, const T* Minv, int np, PW<T, D>& pw) {
  for_window<T, D>(pw, mesh, [&](int s, int64_t i) {
    int64_t I[D], r = i;
    for (int d = 0; d < D; ++d) { I[d] = r % mesh.dims[d]; r /= mesh.dims[d]; }
    T dx[D], P[8], q[8];
    for (int d = 0; d < D; ++d) dx[d] = mesh.node(d, I[d]) - x[d];
    poly_value<T, D>(multilinear, dx, P);
    for (int a = 0; a < np; ++a) {
      T acc = Minv[a * np] * P[0];
      for (int b = 1; b < np; ++b) acc = acc + Minv[a * np + b] * P[b];
      q[a] = acc;
    }
    const T w = pw.w[s];
    pw.w[s] = w * q[0];
    for (int c = 0; c < D; ++c) pw.gw[s][c] = w * q[1 + c];
  });
}

// update_wls_values! (wls.jl:51-64) and the deferred state's fit (:131-148):
// M = Σ w P⊗P over the window with w = kernel value × filter, then the correction.
template <typename T, int D>
void wls_general(int kernel, bool multilinear, const T* x, T lp, const Mesh<T, D>& mesh, const uint8_t* filter,
                 PW<T, D>& pw) {
  const int np = multilinear ? (1 << D) : D + 1;
  T M[64] = {};
  for_window<T, D>(pw, mesh, [&](int s, int64_t i) {
    int64_t I[D], r = i;
    for (int d = 0; d < D; ++d) { I[d] = r % mesh.dims[d]; r /= mesh.dims[d]; }
    T w = kernel_value<T, D>(kernel, x, lp, mesh, I);
    if (filter) w = w * T(filter[i] ? 1 : 0);
    pw.w[s] = w;
    T dx[D], P[8];
    for (int d = 0; d < D; ++d) dx[d] = mesh.node(d, I[d]) - x[d];
    poly_value<T, D>(multilinear, dx, P);
    for (int a = 0; a < np; ++a)
      for (int b = 0; b < np; ++b) M[a * np + b] += w * P[a] * P[b];
  });
  T Minv[64];
  invert<T>(np, M, Minv);
  apply_correction<T, D>(multilinear, x, mesh, Minv, np, pw);
}

// Closed-form moments of a whole B-spline support (wls.jl:34-49):
// M⁻¹ = diag(1, 12/(n+1)/h², ...), applied to the nodewise kernel values.
template <typename T, int D>
void wls_closed(int kernel, const T* x, const Mesh<T, D>& mesh, PW<T, D>& pw) {
  const int np = D + 1;
  T Minv[16] = {};
  Minv[0] = T(1);
  const T c = T(12) / T(degree_of(kernel) + 1) / (mesh.h * mesh.h);
  for (int d = 0; d < D; ++d) Minv[(1 + d) * np + 1 + d] = c;
  for_window<T, D>(pw, mesh, [&](int s, int64_t i) {
    int64_t I[D], r = i;
    for (int d = 0; d < D; ++d) { I[d] = r % mesh.dims[d]; r /= mesh.dims[d]; }
    pw.w[s] = kernel_value<T, D>(kernel, x, T(0), mesh, I);
  });
  apply_correction<T, D>(false, x, mesh, Minv, np, pw);
}

// MultiLinear fit of a quadratic or cubic B-spline without a filter
// (wls.jl:79-99, deferred :166-189): dim independent 1-D WLS(kernel, Linear)
// fits (closed form on a whole axis support, else the 2×2 fit), then the
// tensor product of the axis values and derivatives.
template <typename T, int D>
void wls_separable(int kernel, const T* x, const Mesh<T, D>& mesh, PW<T, D>& pw) {
  const int sw = pw.sw;
  T v1[D][PW<T, D>::MAXSW], d1[D][PW<T, D>::MAXSW];
  int64_t ext[D];
  for (int d = 0; d < D; ++d) {
    ext[d] = pw.hi[d] - pw.lo[d] + 1;
    T N[PW<T, D>::MAXSW], dx[PW<T, D>::MAXSW];
    for (int64_t a = 0; a < ext[d]; ++a) {
      const T xi = mesh.node(d, pw.lo[d] + a);
      T der;
      piecewise_jet<T>(kernel, (x[d] - xi) * mesh.h_inv, T(0), &N[a], &der);
      dx[a] = xi - x[d];
    }
    T Minv[4];
    if (ext[d] == sw) {
      Minv[0] = T(1);
      Minv[1] = Minv[2] = T(0);
      Minv[3] = T(12) / T(degree_of(kernel) + 1) / (mesh.h * mesh.h);
    } else {
      T M[4] = {};
      for (int64_t a = 0; a < ext[d]; ++a) {
        const T P[2] = {T(1), dx[a]};
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) M[i * 2 + j] += N[a] * P[i] * P[j];
      }
      invert<T>(2, M, Minv);
    }
    for (int64_t a = 0; a < ext[d]; ++a) {
      const T q0 = Minv[0] * T(1) + Minv[1] * dx[a];
      const T q1 = Minv[2] * T(1) + Minv[3] * dx[a];
      v1[d][a] = N[a] * q0;
      d1[d][a] = N[a] * q1;
    }
  }
  int64_t cnt = 1;
  for (int d = 0; d < D; ++d) cnt *= ext[d];
  for (int64_t k = 0; k < cnt; ++k) {
    int64_t a[D], r = k;
    for (int d = 0; d < D; ++d) { a[d] = r % ext[d]; r /= ext[d]; }
    int64_t slot = 0;
    for (int d = D - 1; d >= 0; --d) slot = slot * sw + a[d];
    T w = v1[0][a[0]];
    for (int d = 1; d < D; ++d) w = w * v1[d][a[d]];
    pw.w[slot] = w;
    for (int c = 0; c < D; ++c) {
      T g = (c == 0) ? d1[0][a[0]] : v1[0][a[0]];
      for (int d = 1; d < D; ++d) g = g * ((d == c) ? d1[d][a[d]] : v1[d][a[d]]);
      pw.gw[slot][c] = g;
    }
  }
}

// One particle's corrected weights.  Which branch applies follows the
// reference's dispatch:
//   KernelCorrection: the plain kernel when the window is whole and the filter
//     passes all of it (kernelcorrection.jl:27-35, deferred :41-49); otherwise
//     WLS(kernel, poly) with the filter, which for MultiLinear over a quadratic
//     or cubic B-spline without a filter is the separable fit (:53-63,
//     wls.jl:79-99);
//   WLS(kernel, Linear) over a quadratic or cubic B-spline: closed-form moments
//     on a whole window that the filter passes (stored, wls.jl:30-40) or a
//     whole window without a filter (deferred, :131-138);
//   WLS(kernel, MultiLinear) over a quadratic or cubic B-spline, stored, no
//     filter: the separable fit (wls.jl:79-99);
//   otherwise the general fit (wls.jl:51-64, deferred :139-148).
template <typename T, int D>
void eval_corrected(int basis, const T* x, T lp, const Mesh<T, D>& mesh, bool deferred, const uint8_t* filter,
                    PW<T, D>& pw) {
  const int kernel = kernel_of(basis);
  const bool ml = (basis & TESS_POLY_MULTILINEAR) != 0;
  const bool qc = kernel == TESS_BSPLINE_QUADRATIC || kernel == TESS_BSPLINE_CUBIC;
  support_window<T, D>(kernel, x, lp, mesh, pw);
  if (pw.empty) return;
  const bool ft = filter_alltrue<T, D>(filter, pw, mesh);
  if (basis & TESS_KERNEL_CORRECTION) {
    if (pw.full && ft) {
      eval_weights<T, D>(kernel, x, lp, mesh, deferred, pw);
    } else if (ml && qc && !filter) {
      wls_separable<T, D>(kernel, x, mesh, pw);
    } else {
      wls_general<T, D>(kernel, ml, x, lp, mesh, filter, pw);
    }
    return;
  }
  if (!ml && qc && pw.full && (deferred ? !filter : ft)) {
    wls_closed<T, D>(kernel, x, mesh, pw);
  } else if (ml && qc && !filter && !deferred) {
    wls_separable<T, D>(kernel, x, mesh, pw);
  } else {
    wls_general<T, D>(kernel, ml, x, lp, mesh, filter, pw);
  }
}

// Plain kernels ignore the filter (their deferred state is `nothing`,
// basis_weight.jl:442-446); corrected bases read it.
template <typename T, int D>
void eval_any(int basis, const T* x, T lp, const Mesh<T, D>& mesh, bool deferred, const uint8_t* filter,
              PW<T, D>& pw) {
  if (basis & (TESS_WLS | TESS_KERNEL_CORRECTION))
    eval_corrected<T, D>(basis, x, lp, mesh, deferred, filter, pw);
  else
    eval_weights<T, D>(basis, x, lp, mesh, deferred, pw);
}

template <typename T, int D>
void load_vec(const Fld<T>& f, int64_t e, T* out) {
  for (int c = 0; c < D; ++c) out[c] = f(e, c);
}

// Weights for particle p: from stored arrays (weight_prop_view) or evaluated.
template <typename T, int D>
void particle_weights(int basis, const Mesh<T, D>& mesh, const tess_particles& P, int64_t p,
                      const tess_weights* W, PW<T, D>& pw) {
  Fld<T> X(P.x), Lf(P.l);
  T x[D];
  load_vec<T, D>(X, p, x);
  T lp = Lf.ok() ? Lf(p) : T(0);
  if (!W || !W->w.ptr) {  // deferred; a weights struct without storage carries the recorded filter
    eval_any<T, D>(basis, x, lp, mesh, true, W ? W->filter : nullptr, pw);
    return;
  }
  // stored: window and values as update! wrote them (runtime.jl:15-18, 39-44)
  pw.sw = support_width(basis);
  const int64_t* win = W->window + p * 2 * D;
  pw.empty = false;
  for (int d = 0; d < D; ++d) {
    pw.lo[d] = win[2 * d] - 1;
    pw.hi[d] = win[2 * d + 1] - 1;
    if (pw.hi[d] < pw.lo[d]) pw.empty = true;
  }
  int64_t nsl = 1;
  for (int d = 0; d < D; ++d) nsl *= pw.sw;
  Fld<T> Wf(W->w), Gf(W->grad_w);
  for (int64_t s = 0; s < nsl; ++s) {
    pw.w[s] = Wf(s + nsl * p);
    for (int c = 0; c < D; ++c) pw.gw[s][c] = Gf(s + nsl * p, c);
  }
}

// ---- the README program ----------------------------------------------------

// Zero the `=` targets (p2g.jl:159, zero.jl:78-79).
template <typename T, int D>
void zero_grid(const Mesh<T, D>& mesh, const tess_grid& G) {
  Fld<T> m(G.m), mv(G.mv), f(G.f);
  int64_t nn = mesh.nnodes();
  for (int64_t i = 0; i < nn; ++i) {
    m(i) = T(0);
    for (int c = 0; c < D; ++c) { mv(i, c) = T(0); f(i, c) = T(0); }
  }
}

// The @∑ half for one particle (p2g.jl:145-178 with hoisting :272-307):
// m += w*mₚ ; mv += w*(mₚ*v) ; f += ((-V)*σ)·∇w.
template <typename T, int D>
void p2g_particle(int basis, const Mesh<T, D>& mesh, const tess_grid& G, const tess_particles& P,
                  const tess_weights* W, const tess_params& prm, int64_t p) {
  Fld<T> m(G.m), mv(G.mv), f(G.f);
  Fld<T> V(P.V), v(P.v), sg(P.sigma);
  PW<T, D> pw;
  particle_weights<T, D>(basis, mesh, P, p, W, pw);
  if (pw.empty) return;
  const T mp = T(prm.mp);
  T mpv[D], A[D * D];
  for (int c = 0; c < D; ++c) mpv[c] = mp * v(p, c);
  const T negV = -V(p);
  for (int k = 0; k < D * D; ++k) A[k] = negV * sg(p, k);
  for_window<T, D>(pw, mesh, [&](int s, int64_t i) {
    const T w = pw.w[s];
    m(i) += w * mp;
    for (int c = 0; c < D; ++c) mv(i, c) += w * mpv[c];
    for (int a = 0; a < D; ++a) {
      T acc = A[a] * pw.gw[s][0];
      for (int b = 1; b < D; ++b) acc = acc + A[a + D * b] * pw.gw[s][b];
      f(i, a) += acc;
    }
  });
}

// Grid-node half (p2g.jl:377-395 -> foreach.jl:140-180), guarded division as in
// test/transfer.jl:102-104 with m⁻¹ in Float64, gravity and slip walls.
template <typename T, int D>
void grid_update_node(const Mesh<T, D>& mesh, const tess_grid& G, const tess_params& prm, int64_t i) {
  Fld<T> m(G.m), mv(G.mv), f(G.f), vn(G.vn), v(G.v);
  const double md = double(m(i));
  const double minv = md > 0 ? 1.0 / md : 0.0;
  const T dt = T(prm.dt);
  int64_t I[D], r = i;
  for (int d = 0; d < D; ++d) { I[d] = r % mesh.dims[d]; r /= mesh.dims[d]; }
  for (int c = 0; c < D; ++c) {
    T vnc = T(double(mv(i, c)) * minv);
    T acc = T(double(f(i, c)) * minv) + (md > 0 ? T(prm.gravity[c]) : T(0));
    T vc = vnc + acc * dt;
    bool wall = ((prm.slip_walls >> (2 * c)) & 1) && I[c] == 0;
    wall = wall || (((prm.slip_walls >> (2 * c + 1)) & 1) && I[c] == mesh.dims[c] - 1);
    if (wall) vc = T(0);
    vn(i, c) = vnc;
    v(i, c) = vc;
  }
}

template <typename T, int D>
double det_d(const T* F) {
  if (D == 2) return double(F[0]) * double(F[3]) - double(F[2]) * double(F[1]);
  // F[a + 3b] = F_{ab}; cofactor expansion along the first row
  double f11 = F[0], f21 = F[1], f31 = F[2], f12 = F[3], f22 = F[4], f32 = F[5], f13 = F[6], f23 = F[7], f33 = F[8];
  return f11 * (f22 * f33 - f23 * f32) - f12 * (f21 * f33 - f23 * f31) + f13 * (f21 * f32 - f22 * f31);
}

// G2P for one particle (g2p.jl:110-157) with the README particle equations.
template <typename T, int D>
void g2p_particle(int basis, const Mesh<T, D>& mesh, const tess_grid& G, const tess_particles& P,
                  const tess_weights* W, const tess_params& prm, int64_t p) {
  Fld<T> gvn(G.vn), gv(G.v);
  Fld<T> X(P.x), V(P.V), v(P.v), gradv(P.grad_v), Ff(P.F), sg(P.sigma), LE(P.logE);
  PW<T, D> pw;
  particle_weights<T, D>(basis, mesh, P, p, W, pw);
  const T dt = T(prm.dt);
  T dv[D] = {}, L[D * D] = {}, dx[D] = {};
  for_window<T, D>(pw, mesh, [&](int s, int64_t i) {
    const T w = pw.w[s];
    T vi[D];
    for (int c = 0; c < D; ++c) vi[c] = gv(i, c);
    for (int c = 0; c < D; ++c) dv[c] += w * (vi[c] - gvn(i, c));
    for (int b = 0; b < D; ++b)
      for (int a = 0; a < D; ++a) L[a + D * b] += vi[a] * pw.gw[s][b];
    for (int c = 0; c < D; ++c) dx[c] += (w * vi[c]) * dt;
  });
  for (int c = 0; c < D; ++c) v(p, c) += dv[c];
  for (int k = 0; k < D * D; ++k) gradv(p, k) = L[k];
  for (int c = 0; c < D; ++c) X(p, c) += dx[c];
  // F = (I + ∇v Δt) F
  T Fo[D * D], Mx[D * D], Fn[D * D];
  for (int k = 0; k < D * D; ++k) Fo[k] = Ff(p, k);
  for (int b = 0; b < D; ++b)
    for (int a = 0; a < D; ++a) Mx[a + D * b] = (a == b ? T(1) : T(0)) + L[a + D * b] * dt;
  for (int b = 0; b < D; ++b)
    for (int a = 0; a < D; ++a) {
      T acc = Mx[a] * Fo[D * b];
      for (int k = 1; k < D; ++k) acc = acc + Mx[a + D * k] * Fo[k + D * b];
      Fn[a + D * b] = acc;
    }
  for (int k = 0; k < D * D; ++k) Ff(p, k) = Fn[k];
  // J, ln J, σ in Float64
  const double J = det_d<T, D>(Fn);
  double lam = prm.lambda, mu = prm.mu;
  if (LE.ok()) {
    double s = std::exp(double(LE(p)));
    lam *= s;
    mu *= s;
  }
  const double lnJ = std::log(J);
  for (int b = 0; b < D; ++b)
    for (int a = 0; a < D; ++a) {
      double bb = 0;
      for (int k = 0; k < D; ++k) bb += double(Fn[a + D * k]) * double(Fn[b + D * k]);
      double s = mu * (bb - (a == b ? 1.0 : 0.0)) + (a == b ? lam * lnJ : 0.0);
      sg(p, a + D * b) = T(s / J);
    }
  V(p) = T(prm.V0 * J);
}

template <typename T, int D>
int p2g_seq(int basis, const tess_mesh& tm, const tess_grid& G, const tess_particles& P, const tess_weights* W,
            const tess_params& prm) {
  Mesh<T, D> mesh(tm);
  zero_grid<T, D>(mesh, G);
  for (int64_t p = 0; p < P.n; ++p) p2g_particle<T, D>(basis, mesh, G, P, W, prm, p);  // p2g.jl:322-324
  const int64_t nn = mesh.nnodes();
  for (int64_t i = 0; i < nn; ++i) grid_update_node<T, D>(mesh, G, prm, i);
  return TESS_OK;
}

template <typename T, int D>
int g2p_seq(int basis, const tess_mesh& tm, const tess_grid& G, const tess_particles& P, const tess_weights* W,
            const tess_params& prm) {
  Mesh<T, D> mesh(tm);
  for (int64_t p = 0; p < P.n; ++p) g2p_particle<T, D>(basis, mesh, G, P, W, prm, p);
  return TESS_OK;
}

template <typename T, int D>
int update_weights_impl(int basis, const tess_mesh& tm, const tess_particles& P, const tess_weights& W,
                        int64_t p0, int64_t p1, bool deferred = false) {
  Mesh<T, D> mesh(tm);
  Fld<T> X(P.x), Lf(P.l), Wf(W.w), Gf(W.grad_w);
  const int sw = support_width(basis);
  int64_t nsl = 1;
  for (int d = 0; d < D; ++d) nsl *= sw;
  for (int64_t p = p0; p < p1; ++p) {
    T x[D];
    load_vec<T, D>(X, p, x);
    PW<T, D> pw;
    eval_any<T, D>(basis, x, Lf.ok() ? Lf(p) : T(0), mesh, deferred, W.filter, pw);
    int64_t* win = W.window + p * 2 * D;
    for (int d = 0; d < D; ++d) {
      if (pw.empty) { win[2 * d] = 1; win[2 * d + 1] = 0; }
      else { win[2 * d] = pw.lo[d] + 1; win[2 * d + 1] = pw.hi[d] + 1; }
    }
    if (pw.empty) continue;
    for_window<T, D>(pw, mesh, [&](int s, int64_t) {
      Wf(s + nsl * p) = pw.w[s];
      for (int c = 0; c < D; ++c) Gf(s + nsl * p, c) = pw.gw[s][c];
    });
  }
  return TESS_OK;
}

// ---- CPU block partition (cpu_block_strategy.jl:62-219) --------------------

template <typename T, int D>
int64_t findblock_lin(const Mesh<T, D>& mesh, const T* x, int64_t* nb) {  // cartesian.jl:277-285
  int64_t b[D];
  for (int d = 0; d < D; ++d) {
    T f = std::floor((x[d] - mesh.xmin[d]) * mesh.h_inv);
    int64_t c0 = int64_t(f);
    if (!(0 <= c0 && c0 <= mesh.dims[d] - 2)) return -1;
    b[d] = c0 >> mesh.L;
  }
  int64_t l = 0;
  for (int d = D - 1; d >= 0; --d) l = l * nb[d] + b[d];
  return l;
}

template <int D>
int block_color(int64_t lin, const int64_t* nb) {  // :197-203, 0-based colour
  int c = 0;
  for (int d = 0; d < D; ++d) {
    int64_t bd = lin % nb[d];
    lin /= nb[d];
    c += int(bd & 1) << d;
  }
  return c;
}

struct PartitionOut {
  std::vector<int32_t> pidx;           // 0-based particle ids in block order
  std::vector<int64_t> start, stop;    // per block [start, stop) into pidx, start==stop if empty
  std::vector<std::vector<int64_t>> groups;  // active blocks per colour, linear order
  int64_t nassigned = 0;
};

template <typename T, int D>
void partition_impl(const Mesh<T, D>& mesh, const Fld<T>& X, int64_t n, PartitionOut& out) {
  int64_t nb[D], nbt = 1;
  for (int d = 0; d < D; ++d) { nb[d] = mesh.nblocks(d); nbt *= nb[d]; }
  std::vector<int64_t> blk(n), count(nbt, 0);
  for (int64_t p = 0; p < n; ++p) {
    T x[D];
    load_vec<T, D>(X, p, x);
    blk[p] = findblock_lin<T, D>(mesh, x, nb);
    if (blk[p] >= 0) count[blk[p]]++;
  }
  out.groups.assign(1 << D, {});
  for (int64_t b = 0; b < nbt; ++b)
    if (count[b]) out.groups[block_color<D>(b, nb)].push_back(b);
  out.start.assign(nbt, 0);
  out.stop.assign(nbt, 0);
  int64_t assigned = 0;
  for (auto& g : out.groups)
    for (int64_t b : g) { out.start[b] = assigned; assigned += count[b]; out.stop[b] = assigned; }
  out.nassigned = assigned;
  out.pidx.assign(assigned, 0);
  std::vector<int64_t> cursor(out.start);
  for (int64_t p = 0; p < n; ++p)
    if (blk[p] >= 0) out.pidx[cursor[blk[p]]++] = int32_t(p);
}

// ---- threaded CPU baseline (region.jl:207-256, g2p.jl:104-108) -------------

template <typename Fn>
void parallel_for(int nthreads, int64_t n, Fn&& fn) {
  if (nthreads <= 1 || n < 2) { fn(int64_t(0), n); return; }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    int64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &fn] { fn(a, b); });
  }
  for (auto& x : th) x.join();
}

template <typename T, int D>
int step_threaded_impl(int basis, const tess_mesh& tm, const tess_grid& G, const tess_particles& P,
                       const tess_params& prm, int nthreads, int nsteps, tess_weights& W) {
  Mesh<T, D> mesh(tm);
  Fld<T> X(P.x);
  if (support_width(basis) > (1 << mesh.L))
    return fail(TESS_ERR_STATE, "Block size for Partition is too small for the basis");
  const int64_t n = P.n, nn = mesh.nnodes();
  for (int s = 0; s < nsteps; ++s) {
    // update!(wgts, pcls, mesh): @threaded over particles, stored weights
    parallel_for(nthreads, n, [&](int64_t a, int64_t b) { update_weights_impl<T, D>(basis, tm, P, W, a, b); });
    // update!(partition, x)
    PartitionOut part;
    partition_impl<T, D>(mesh, X, n, part);
    // @P2G with partition: parallel zero, one phase per colour, grid half
    parallel_for(nthreads, nn, [&](int64_t a, int64_t b) {
      Fld<T> m(G.m), mv(G.mv), f(G.f);
      for (int64_t i = a; i < b; ++i) {
        m(i) = T(0);
        for (int c = 0; c < D; ++c) { mv(i, c) = T(0); f(i, c) = T(0); }
      }
    });
    for (auto& group : part.groups) {
      parallel_for(nthreads, int64_t(group.size()), [&](int64_t a, int64_t b) {
        for (int64_t k = a; k < b; ++k) {
          int64_t blk = group[k];
          for (int64_t j = part.start[blk]; j < part.stop[blk]; ++j)
            p2g_particle<T, D>(basis, mesh, G, P, &W, prm, part.pidx[j]);
        }
      });
    }
    parallel_for(nthreads, nn, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i) grid_update_node<T, D>(mesh, G, prm, i);
    });
    // @G2P: threaded over particles
    parallel_for(nthreads, n, [&](int64_t a, int64_t b) {
      for (int64_t p = a; p < b; ++p) g2p_particle<T, D>(basis, mesh, G, P, &W, prm, p);
    });
  }
  return TESS_OK;
}

// ---- adjoint of one README step (no reference counterpart) ------------------
//
// The engine's differentiation extension: the reverse of one step with
// deferred weights and per-particle E = E0*exp(logE).  The state of a step is
// (x, v, F) with V = V0 J and σ = (μ(FFᵀ-I) + λ lnJ I)/J functions of F, so
// P = V σ = V0 (μ(FFᵀ-I) + λ lnJ I).  Given the adjoint (x̄', v̄', F̄') of the
// state after the step, this returns the adjoint before it and adds ∂/∂logE.
// Written sequentially over a dense grid with textbook B-spline jets; its
// finite-difference check lives in tests/test_oracle_adjoint.py.

// Centred B-spline of the basis at node-relative coordinate r = ξ - i, with
// its first and second derivatives in r.
inline void bspline_jet2(int basis, double r, double& N, double& dN, double& ddN) {
  const double a = std::fabs(r), sg = std::signbit(r) ? -1.0 : 1.0;
  N = dN = ddN = 0.0;
  if (basis == TESS_BSPLINE_LINEAR) {
    if (a < 1.0) { N = 1.0 - a; dN = -sg; }
  } else if (basis == TESS_BSPLINE_QUADRATIC) {
    if (a < 0.5) { N = 0.75 - r * r; dN = -2.0 * r; ddN = -2.0; }
    else if (a < 1.5) { const double u = 1.5 - a; N = 0.5 * u * u; dN = -sg * u; ddN = 1.0; }
  } else {  // cubic
    if (a < 1.0) { N = 0.5 * a * a * a - r * r + 2.0 / 3.0; dN = 1.5 * sg * r * r - 2.0 * r; ddN = 3.0 * a - 2.0; }
    else if (a < 2.0) { const double u = 2.0 - a; N = u * u * u / 6.0; dN = -sg * 0.5 * u * u; ddN = u; }
  }
}

// w, ∇w, ∇∇w (in x) of the nodes of particle p's window, column-major.
template <typename T, int D>
struct Jets2 {
  int64_t lo[D], ext[D], count;
  bool empty;
  double N[D][4], dN[D][4], ddN[D][4];
  template <typename Fn>
  void each(const Mesh<T, D>& mesh, Fn&& f) const {
    for (int64_t k = 0; k < count; ++k) {
      int64_t a[D], I[D], r = k;
      for (int d = 0; d < D; ++d) { a[d] = r % ext[d]; r /= ext[d]; I[d] = lo[d] + a[d]; }
      double w = 1.0, g[D], H[D][D];
      for (int d = 0; d < D; ++d) w *= N[d][a[d]];
      for (int c = 0; c < D; ++c) {
        g[c] = 1.0;
        for (int d = 0; d < D; ++d) g[c] *= (d == c) ? dN[d][a[d]] : N[d][a[d]];
        g[c] *= 1.0 / mesh.h_d;
        for (int e = 0; e < D; ++e) {
          double h = 1.0;
          for (int d = 0; d < D; ++d) {
            if (c == e) h *= (d == c) ? ddN[d][a[d]] : N[d][a[d]];
            else h *= (d == c || d == e) ? dN[d][a[d]] : N[d][a[d]];
          }
          H[c][e] = h / (mesh.h_d * mesh.h_d);
        }
      }
      f(mesh.lin(I), w, g, H);
    }
  }
};

template <typename T, int D>
void particle_jets2(int basis, const Mesh<T, D>& mesh, const T* x, Jets2<T, D>& J) {
  PW<T, D> pw;
  support_window<T, D>(basis, x, T(0), mesh, pw);
  J.empty = pw.empty;
  J.count = 0;
  if (pw.empty) return;
  J.count = 1;
  for (int d = 0; d < D; ++d) {
    J.lo[d] = pw.lo[d];
    J.ext[d] = pw.hi[d] - pw.lo[d] + 1;
    J.count *= J.ext[d];
    for (int64_t a = 0; a < J.ext[d]; ++a) {
      const double r = (double(x[d]) - (mesh.xmin_d[d] + double(pw.lo[d] + a) * mesh.h_d)) / mesh.h_d;
      bspline_jet2(basis, r, J.N[d][a], J.dN[d][a], J.ddN[d][a]);
    }
  }
}

template <typename T, int D>
void inverse_d(const double* F, double* Fi) {  // column-major
  if (D == 2) {
    const double det = F[0] * F[3] - F[2] * F[1];
    Fi[0] = F[3] / det; Fi[1] = -F[1] / det; Fi[2] = -F[2] / det; Fi[3] = F[0] / det;
    return;
  }
  auto f = [&](int r, int c) { return F[r + 3 * c]; };
  double cof[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cof[r + 3 * c] = f(r1, c1) * f(r2, c2) - f(r1, c2) * f(r2, c1);
    }
  const double det = f(0, 0) * cof[0] + f(0, 1) * cof[3] + f(0, 2) * cof[6];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Fi[r + 3 * c] = cof[c + 3 * r] / det;  // adjugate = cofactorᵀ
}

template <typename T, int D>
int step_adjoint_impl(int basis, const tess_mesh& tm, const tess_grid& G, const tess_particles& P,
                      const tess_params& prm, double* ax, double* av, double* aF, double* ath) {
  if (basis == TESS_UGIMP || (basis & ~0xff)) return fail(TESS_ERR_UNSUPPORTED, "adjoint: B-spline bases only");
  Mesh<T, D> mesh(tm);
  p2g_seq<T, D>(basis, tm, G, P, nullptr, prm);  // the forward grid of this step
  Fld<T> X(P.x), v(P.v), Ff(P.F), V(P.V), sg(P.sigma), LE(P.logE);
  Fld<T> gm(G.m), gmv(G.mv), gf(G.f), gvn(G.vn), gv(G.v);
  const int64_t nn = mesh.nnodes(), n = P.n;
  const double dt = double(T(prm.dt)), mp = double(T(prm.mp));
  std::vector<double> dvb(nn * D, 0.0), vgb(nn * D, 0.0);
  // particle adjoints scattered to the nodes: Δv̄_i = Σ w v̄', v̄_i = Σ (L̄ ∇w + Δt w x̄')
  for (int64_t p = 0; p < n; ++p) {
    T x[D];
    load_vec<T, D>(X, p, x);
    Jets2<T, D> J;
    particle_jets2<T, D>(basis, mesh, x, J);
    if (J.empty) continue;
    double Lb[D * D];
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc += aF[p * D * D + r + D * k] * double(Ff(p, c + D * k));
        Lb[r + D * c] = dt * acc;  // L̄ = Δt F̄' Fᵀ
      }
    J.each(mesh, [&](int64_t i, double w, const double* g, const double (*)[D]) {
      for (int r = 0; r < D; ++r) {
        dvb[i * D + r] += w * av[p * D + r];
        double acc = dt * w * ax[p * D + r];
        for (int c = 0; c < D; ++c) acc += Lb[r + D * c] * g[c];
        vgb[i * D + r] += acc;
      }
    });
  }
  // grid update reversed: f̄, q̄ (of mv), m̄ per node, walls included
  std::vector<double> fb(nn * D, 0.0), qb(nn * D, 0.0), mb(nn, 0.0);
  for (int64_t i = 0; i < nn; ++i) {
    const double m = double(gm(i));
    if (!(m > 0)) continue;
    int64_t I[D], r = i;
    for (int d = 0; d < D; ++d) { I[d] = r % mesh.dims[d]; r /= mesh.dims[d]; }
    double acc_m = 0.0;
    for (int c = 0; c < D; ++c) {
      bool wall = ((prm.slip_walls >> (2 * c)) & 1) && I[c] == 0;
      wall = wall || (((prm.slip_walls >> (2 * c + 1)) & 1) && I[c] == mesh.dims[c] - 1);
      const double a = wall ? 0.0 : vgb[i * D + c] + dvb[i * D + c];
      const double vnb = a - dvb[i * D + c];
      fb[i * D + c] = a * dt / m;
      qb[i * D + c] = vnb / m;
      acc_m += -a * dt * double(gf(i, c)) / (m * m) - vnb * double(gmv(i, c)) / (m * m);
    }
    mb[i] = acc_m;
  }
  // nodes gathered back to the particles
  for (int64_t p = 0; p < n; ++p) {
    T x[D];
    load_vec<T, D>(X, p, x);
    Jets2<T, D> J;
    particle_jets2<T, D>(basis, mesh, x, J);
    if (J.empty) continue;
    double F[D * D], Fb1[D * D], Lb[D * D], Pm[D * D], vp[D];
    for (int k = 0; k < D * D; ++k) {
      F[k] = double(Ff(p, k));
      Fb1[k] = aF[p * D * D + k];
      Pm[k] = double(V(p)) * double(sg(p, k));
    }
    for (int c = 0; c < D; ++c) vp[c] = double(v(p, c));
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc += Fb1[r + D * k] * F[c + D * k];
        Lb[r + D * c] = dt * acc;
      }
    double gx[D] = {}, gvq[D] = {}, Pb[D * D] = {}, L[D * D] = {};
    J.each(mesh, [&](int64_t i, double w, const double* g, const double (*H)[D]) {
      double vg[D], dv[D], s = 0.0;
      for (int c = 0; c < D; ++c) {
        vg[c] = double(gv(i, c));
        dv[c] = vg[c] - double(gvn(i, c));
      }
      for (int c = 0; c < D; ++c)
        s += av[p * D + c] * dv[c] + dt * ax[p * D + c] * vg[c] + mp * qb[i * D + c] * vp[c];
      s += mp * mb[i];
      double u[D];  // L̄ᵀ vg - Pᵀ f̄
      for (int c = 0; c < D; ++c) {
        double acc = 0.0;
        for (int r = 0; r < D; ++r) acc += Lb[r + D * c] * vg[r] - Pm[r + D * c] * fb[i * D + r];
        u[c] = acc;
      }
      for (int c = 0; c < D; ++c) {
        double acc = g[c] * s;
        for (int e = 0; e < D; ++e) acc += H[c][e] * u[e];
        gx[c] += acc;
        gvq[c] += w * qb[i * D + c];
      }
      for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
          Pb[r + D * c] -= fb[i * D + r] * g[c];
          L[r + D * c] += vg[r] * g[c];
        }
    });
    double lam = prm.lambda, mu = prm.mu;
    if (LE.ok()) {
      const double e = std::exp(double(LE(p)));
      lam *= e;
      mu *= e;
    }
    double Fi[D * D];
    inverse_d<T, D>(F, Fi);
    double trPb = 0.0, PbP = 0.0;
    for (int c = 0; c < D; ++c) trPb += Pb[c + D * c];
    for (int k = 0; k < D * D; ++k) PbP += Pb[k] * Pm[k];
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) {
        double acc = Fb1[r + D * c];  // (I + Δt L)ᵀ F̄'
        for (int k = 0; k < D; ++k) acc += dt * L[k + D * r] * Fb1[k + D * c];
        double sym = 0.0;  // μ (P̄ + P̄ᵀ) F
        for (int k = 0; k < D; ++k) sym += (Pb[r + D * k] + Pb[k + D * r]) * F[k + D * c];
        acc += prm.V0 * (mu * sym + lam * trPb * Fi[c + D * r]);
        aF[p * D * D + r + D * c] = acc;
      }
    for (int c = 0; c < D; ++c) {
      ax[p * D + c] += gx[c];
      av[p * D + c] += mp * gvq[c];
    }
    if (ath) ath[p] += PbP;
  }
  return TESS_OK;
}

template <typename T, int D>
struct RunAdjoint {
  static int run(const tess_mesh& m, int basis, const tess_grid* G, const tess_particles* P, const tess_params* prm,
                 double* ax, double* av, double* aF, double* ath) {
    return step_adjoint_impl<T, D>(basis, m, *G, *P, *prm, ax, av, aF, ath);
  }
};

template <template <typename, int> class Fn, typename... Args>
int dispatch(const tess_mesh* m, Args&&... args) {
  if (!m) return fail(TESS_ERR_ARGUMENT, "mesh is NULL");
  if (m->dim == 2 && m->dtype == TESS_FLOAT64) return Fn<double, 2>::run(*m, args...);
  if (m->dim == 3 && m->dtype == TESS_FLOAT64) return Fn<double, 3>::run(*m, args...);
  if (m->dim == 2 && m->dtype == TESS_FLOAT32) return Fn<float, 2>::run(*m, args...);
  if (m->dim == 3 && m->dtype == TESS_FLOAT32) return Fn<float, 3>::run(*m, args...);
  return fail(TESS_ERR_UNSUPPORTED, "unsupported dim/dtype");
}

// A kernel, optionally wrapped in exactly one of WLS / KernelCorrection, the
// polynomial flag only with a wrapper.
bool basis_ok(int basis) {
  const int k = kernel_of(basis), flags = basis & ~0xff;
  const bool kernel = k == TESS_BSPLINE_LINEAR || k == TESS_BSPLINE_QUADRATIC || k == TESS_BSPLINE_CUBIC ||
                      k == TESS_UGIMP;
  const int wrap = flags & (TESS_WLS | TESS_KERNEL_CORRECTION);
  const bool ok_flags = (flags & ~(TESS_WLS | TESS_KERNEL_CORRECTION | TESS_POLY_MULTILINEAR)) == 0 &&
                        (wrap == 0 ? flags == 0 : wrap != (TESS_WLS | TESS_KERNEL_CORRECTION));
  return kernel && ok_flags;
}

template <typename T, int D>
struct RunP2G {
  static int run(const tess_mesh& m, int basis, const tess_grid* G, const tess_particles* P, const tess_weights* W,
                 const tess_params* prm) {
    return p2g_seq<T, D>(basis, m, *G, *P, W, *prm);
  }
};
template <typename T, int D>
struct RunG2P {
  static int run(const tess_mesh& m, int basis, const tess_grid* G, const tess_particles* P, const tess_weights* W,
                 const tess_params* prm) {
    return g2p_seq<T, D>(basis, m, *G, *P, W, *prm);
  }
};
// @G2P2G of the README program (src/Transfer/g2p2g.jl:100-159): one set of
// weights, evaluated at the positions before the G2P half, serves both halves
// (the P2G half does not rebind them, g2p2g.jl:107-114).  The G2P half updates
// every particle; the P2G half zeroes m, mv, f and scatters the updated v, V,
// σ with those weights in particle order; the grid-node half follows.  That is
// the reference test's manual loop (test/transfer.jl:267-314): G2P, then P2G
// with the same weights.  Deferred (W == NULL) evaluates the weights into
// scratch first.
template <typename T, int D>
struct RunG2P2G {
  static int run(const tess_mesh& m, int basis, const tess_grid* G, const tess_particles* P, const tess_weights* W,
                 const tess_params* prm) {
    std::vector<T> w, gw;
    std::vector<int64_t> win;
    tess_weights own{};
    if (!W || !W->w.ptr) {
      own.filter = W ? W->filter : nullptr;
      int64_t nsl = 1;
      for (int d = 0; d < D; ++d) nsl *= support_width(basis);
      w.assign(size_t(nsl * P->n), T(0));
      gw.assign(size_t(nsl * P->n * D), T(0));
      win.assign(size_t(2 * D * P->n), 0);
      own.w = tess_field{w.data(), 1, 0};
      own.grad_w = tess_field{gw.data(), D, 1};
      own.window = win.data();
      const int st = update_weights_impl<T, D>(basis, m, *P, own, 0, P->n, true);
      if (st != TESS_OK) return st;
      W = &own;
    }
    const int st = g2p_seq<T, D>(basis, m, *G, *P, W, *prm);
    if (st != TESS_OK) return st;
    return p2g_seq<T, D>(basis, m, *G, *P, W, *prm);
  }
};
template <typename T, int D>
struct RunStep {
  static int run(const tess_mesh& m, int basis, const tess_grid* G, const tess_particles* P, const tess_params* prm,
                 int nsteps) {
    for (int s = 0; s < nsteps; ++s) {
      p2g_seq<T, D>(basis, m, *G, *P, nullptr, *prm);
      g2p_seq<T, D>(basis, m, *G, *P, nullptr, *prm);
    }
    return TESS_OK;
  }
};
template <typename T, int D>
struct RunWeights {
  static int run(const tess_mesh& m, int basis, const tess_particles* P, const tess_weights* W) {
    return update_weights_impl<T, D>(basis, m, *P, *W, 0, P->n);
  }
};
template <typename T, int D>
struct RunPartition {
  static int run(const tess_mesh& m, tess_field x, int64_t n, int32_t* pidx, int32_t* starts, int32_t* stops,
                 int64_t* nassigned) {
    Mesh<T, D> mesh(m);
    PartitionOut out;
    partition_impl<T, D>(mesh, Fld<T>(x), n, out);
    for (int64_t j = 0; j < out.nassigned; ++j) pidx[j] = out.pidx[j] + 1;  // 1-based ids
    for (size_t b = 0; b < out.start.size(); ++b) {
      bool nonempty = out.stop[b] > out.start[b];
      starts[b] = nonempty ? int32_t(out.start[b] + 1) : 0;
      stops[b] = nonempty ? int32_t(out.stop[b]) : 0;
    }
    *nassigned = out.nassigned;
    return TESS_OK;
  }
};
// update_sparsity!(spinds, xₚ, mesh) on the CPU (src/Grids/sparsity.jl:23-40):
// per-particle block ids (:120-139), occupied blocks, the ±1-block
// neighbourhood scattered from each occupied block (:141-155), then the
// sequential numbering of active blocks (:53-60).  *changed compares the new
// active set with the numbering passed in (:262-267).
template <typename T, int D>
struct RunSparsity {
  static int run(const tess_mesh& m, tess_field x, int64_t n, int32_t* numbering, int64_t* storage_len,
                 int32_t* changed) {
    Mesh<T, D> mesh(m);
    int64_t nb[D], nbt = 1;
    for (int d = 0; d < D; ++d) { nb[d] = mesh.nblocks(d); nbt *= nb[d]; }
    Fld<T> X(x);
    std::vector<uint8_t> occupied(nbt, 0), active(nbt, 0);
    for (int64_t p = 0; p < n; ++p) {
      T xp[D];
      load_vec<T, D>(X, p, xp);
      const int64_t b = findblock_lin<T, D>(mesh, xp, nb);
      if (b >= 0) occupied[b] = 1;
    }
    for (int64_t b = 0; b < nbt; ++b) {
      if (!occupied[b]) continue;
      int64_t I[D], r = b;
      for (int d = 0; d < D; ++d) { I[d] = r % nb[d]; r /= nb[d]; }
      int64_t lo[D], hi[D], ext[D], tot = 1;
      for (int d = 0; d < D; ++d) {
        lo[d] = std::max<int64_t>(I[d] - 1, 0);
        hi[d] = std::min<int64_t>(I[d] + 1, nb[d] - 1);
        ext[d] = hi[d] - lo[d] + 1;
        tot *= ext[d];
      }
      for (int64_t k = 0; k < tot; ++k) {
        int64_t q = k, lin = 0, mul = 1;
        for (int d = 0; d < D; ++d) { lin += (lo[d] + q % ext[d]) * mul; q /= ext[d]; mul *= nb[d]; }
        active[lin] = 1;
      }
    }
    int32_t count = 0;
    bool diff = false;
    for (int64_t b = 0; b < nbt; ++b) {
      diff = diff || ((numbering[b] != 0) != (active[b] != 0));
      numbering[b] = active[b] ? ++count : 0;
    }
    *storage_len = int64_t(count) << (mesh.L * D);
    if (changed) *changed = diff ? 1 : 0;
    return TESS_OK;
  }
};

template <typename T, int D>
struct RunThreaded {
  static int run(const tess_mesh& m, int basis, const tess_grid* G, const tess_particles* P, const tess_params* prm,
                 int nthreads, int nsteps, tess_weights* W) {
    return step_threaded_impl<T, D>(basis, m, *G, *P, *prm, nthreads, nsteps, *W);
  }
};
template <typename T, int D>
struct RunJet {
  static int run(const tess_mesh& m, int basis, const double* x, double l, int deferred, const uint8_t* filter,
                 int64_t* window, double* w, double* gw) {
    Mesh<T, D> mesh(m);
    T xt[D];
    for (int d = 0; d < D; ++d) xt[d] = T(x[d]);
    PW<T, D> pw;
    eval_any<T, D>(basis, xt, T(l), mesh, deferred != 0, filter, pw);
    for (int d = 0; d < D; ++d) {
      window[2 * d] = pw.empty ? 1 : pw.lo[d] + 1;
      window[2 * d + 1] = pw.empty ? 0 : pw.hi[d] + 1;
    }
    int64_t nsl = 1;
    for (int d = 0; d < D; ++d) nsl *= pw.sw;
    for (int64_t s = 0; s < nsl; ++s) { w[s] = 0; for (int c = 0; c < D; ++c) gw[s * D + c] = 0; }
    if (!pw.empty)
      for_window<T, D>(pw, mesh, [&](int s, int64_t) {
        w[s] = double(pw.w[s]);
        for (int c = 0; c < D; ++c) gw[s * D + c] = double(pw.gw[s][c]);
      });
    return TESS_OK;
  }
};

}  // namespace

extern "C" {

const char* oref_last_error(void) { return g_err.c_str(); }

// One particle's BasisWeight (update!(bw, pt, mesh), basis_weight.jl:634-654):
// window as dim (start, stop) 1-based pairs, w and ∇w over the sw^dim storage
// (entries outside the window are zero).  deferred=1 gives the transfer-time
// (nodewise) evaluation.
int oref_basis_weight(const tess_mesh* m, int basis, const double* x, double l, int deferred, int64_t* window,
                      double* w, double* gw) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunJet>(m, basis, x, l, deferred, static_cast<const uint8_t*>(nullptr), window, w, gw);
}

// The same with a node filter (update!(bw, pt, mesh, filter),
// basis_weight.jl:683-695): one byte per mesh node, column-major; corrected
// bases only.
int oref_basis_weight_filtered(const tess_mesh* m, int basis, const double* x, double l, int deferred,
                               const uint8_t* filter, int64_t* window, double* w, double* gw) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  if (filter && !(basis & (TESS_WLS | TESS_KERNEL_CORRECTION)))
    return fail(TESS_ERR_ARGUMENT, "the basis does not support filtered updates");
  return dispatch<RunJet>(m, basis, x, l, deferred, filter, window, w, gw);
}

// 1-D jet(Order(1), kernel, ξ[, l]) (basis_weight.jl:35-37).
int oref_jet1d(int basis, double xi, double l, double* val, double* der) {
  if (!basis_ok(basis) || (basis & ~0xff)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  piecewise_jet<double>(basis, xi, l, val, der);
  return TESS_OK;
}

// findblock(x, mesh) (cartesian.jl:277-285): 1-based Cartesian block index, or
// all zeros when x is outside.
int oref_findblock(const tess_mesh* m, const double* x, int64_t* blk) {
  if (!m) return fail(TESS_ERR_ARGUMENT, "mesh");
  const int D = m->dim;
  for (int d = 0; d < D; ++d) {
    double f = std::floor((x[d] - m->xmin[d]) * m->h_inv);
    int64_t c0 = int64_t(f);
    if (!(0 <= c0 && c0 <= m->dims[d] - 2)) {
      for (int e = 0; e < D; ++e) blk[e] = 0;
      return TESS_OK;
    }
    blk[d] = (c0 >> m->block_size_log2) + 1;
  }
  return TESS_OK;
}

// supportnodes(x, r, mesh) (cartesian.jl:190-199), 1-based ranges.
int oref_supportnodes_r(const tess_mesh* m, const double* x, double r, int64_t* window) {
  const int D = m->dim;
  bool inside = true;
  double xi[3];
  for (int d = 0; d < D; ++d) {
    xi[d] = (x[d] - m->xmin[d]) * m->h_inv;
    double f = std::floor(xi[d]);
    if (!(0 <= f && f <= double(m->dims[d] - 2))) inside = false;
  }
  for (int d = 0; d < D; ++d) {
    if (!inside) { window[2 * d] = 1; window[2 * d + 1] = 0; continue; }
    int64_t s = int64_t(std::floor(xi[d] - r)) + 2, e = int64_t(std::floor(xi[d] + r)) + 1;
    window[2 * d] = std::max<int64_t>(s, 1);
    window[2 * d + 1] = std::min<int64_t>(e, m->dims[d]);
  }
  return TESS_OK;
}

int oref_update_weights(int basis, const tess_mesh* m, const tess_particles* P, const tess_weights* W) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  if (W && W->filter && !(basis & (TESS_WLS | TESS_KERNEL_CORRECTION)))  // require_filtered_update_support
    return fail(TESS_ERR_ARGUMENT, "the basis does not support filtered updates");
  return dispatch<RunWeights>(m, basis, P, W);
}

int oref_partition_update(const tess_mesh* m, tess_field x, int64_t n, int32_t* pidx, int32_t* starts,
                          int32_t* stops, int64_t* nassigned) {
  return dispatch<RunPartition>(m, x, n, pidx, starts, stops, nassigned);
}

int oref_sparsity_update(const tess_mesh* m, tess_field x, int64_t n, int32_t* numbering, int64_t* storage_len,
                         int32_t* changed) {
  return dispatch<RunSparsity>(m, x, n, numbering, storage_len, changed);
}

// Reverse of one README step (deferred weights): the particle state is the one at
// the start of the step (unchanged); ax/av/aF (Float64, Julia layout) hold the
// adjoint after the step on entry and before it on return; ath += ∂/∂logE.
// The grid receives the step's forward P2G.
int oref_step_adjoint(int basis, const tess_mesh* m, const tess_grid* G, const tess_particles* P,
                      const tess_params* prm, double* ax, double* av, double* aF, double* ath) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunAdjoint>(m, basis, G, P, prm, ax, av, aF, ath);
}

int oref_p2g(int basis, const tess_mesh* m, const tess_grid* G, const tess_particles* P, const tess_weights* W,
             const tess_params* prm) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunP2G>(m, basis, G, P, W, prm);
}

int oref_g2p(int basis, const tess_mesh* m, const tess_grid* G, const tess_particles* P, const tess_weights* W,
             const tess_params* prm) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunG2P>(m, basis, G, P, W, prm);
}

int oref_g2p2g(int basis, const tess_mesh* m, const tess_grid* G, const tess_particles* P, const tess_weights* W,
               const tess_params* prm) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunG2P2G>(m, basis, G, P, W, prm);
}

// nsteps of the README step with deferred weights, sequential order.
int oref_step(int basis, const tess_mesh* m, const tess_grid* G, const tess_particles* P, const tess_params* prm,
              int nsteps) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunStep>(m, basis, G, P, prm, nsteps);
}

// The reference CPU algorithm with threads: stored weights, CPU block
// partition, colour-phased P2G, threaded G2P.  W is caller-provided storage.
int oref_step_threaded(int basis, const tess_mesh* m, const tess_grid* G, const tess_particles* P,
                       const tess_params* prm, int nthreads, int nsteps, tess_weights* W) {
  if (!basis_ok(basis)) return fail(TESS_ERR_UNSUPPORTED, "basis");
  return dispatch<RunThreaded>(m, basis, G, P, prm, nthreads, nsteps, W);
}

}  // extern "C"
