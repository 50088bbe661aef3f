// Device-side building blocks shared by the gfx950 kernels of libtesserae_hip:
// strided field access, the Cartesian mesh as its device copy holds it, and the
// 1-D kernel jets of the supported bases.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tesserae_hip.h"

namespace tess {

// Component c of element e of a caller-owned field (StructArray component).
template <typename T>
struct DField {
  T* p;
  int64_t es, cs;
  __host__ __device__ T& operator()(int64_t e, int c = 0) const { return p[e * es + c * cs]; }
};

template <typename T>
inline DField<T> dfield(const tess_field& f) {
  return DField<T>{static_cast<T*>(f.ptr), f.elem_stride, f.comp_stride};
}

// CartesianMesh{dim,T} on the device (src/Mesh/cartesian.jl:38-42 after
// src/adapt.jl:53-57): T(xmin), T(h), T(h_inv); node i sits at T(xmin + i*h).
template <typename T, int D>
struct DMesh {
  int32_t dims[D];
  T xmin[D];
  T h, h_inv;
  double xmin_d[D];
  double h_d, hinv_d;  // h and 1/h in Float64 (1/h rounded once, on the host)
  __device__ T node(int d, int64_t i) const { return T(xmin_d[d] + double(i) * h_d); }
  __device__ int64_t lin(const int* I) const {
    int64_t l = I[D - 1];
    for (int d = D - 2; d >= 0; --d) l = l * dims[d] + I[d];
    return l;
  }
};

template <typename T, int D>
inline DMesh<T, D> dmesh(const tess_mesh& m) {
  DMesh<T, D> r;
  for (int d = 0; d < D; ++d) {
    r.dims[d] = int32_t(m.dims[d]);
    r.xmin[d] = T(m.xmin[d]);
    r.xmin_d[d] = m.xmin[d];
  }
  r.h = T(m.h);
  r.h_inv = T(m.h_inv);
  r.h_d = m.h;
  r.hinv_d = 1.0 / m.h;
  return r;
}

__host__ __device__ constexpr int support_width(int basis) {
  return basis == TESS_BSPLINE_LINEAR ? 2 : basis == TESS_BSPLINE_QUADRATIC ? 3 : basis == TESS_BSPLINE_CUBIC ? 4 : 3;
}
__host__ __device__ constexpr int degree_of(int basis) { return basis == TESS_UGIMP ? 2 : basis; }

template <typename T>
__device__ __forceinline__ T sq(T a) { return a * a; }

// Piecewise kernel value and its derivative in the normalised coordinate ξ
// (src/Basis/bspline.jl:176-189, src/Basis/gimp.jl:28-33), differentiated the way
// the reference's dual numbers do (abs(x) = signbit(x) ? -x : x).  l is the
// uGIMP particle length divided by h.
template <typename T, int BASIS>
__device__ __forceinline__ void piecewise_jet(T xi, T l, T& val, T& der) {
  const T s = signbit(xi) ? T(-1) : T(1);
  const T a = fabs(xi);
  if constexpr (BASIS == TESS_BSPLINE_LINEAR) {
    const bool in = a < T(1);
    val = in ? T(1) - a : T(0);
    der = in ? -s : T(0);
  } else if constexpr (BASIS == TESS_BSPLINE_QUADRATIC) {
    if (a < T(0.5)) {
      val = (T(3) - T(4) * (a * a)) / T(4);
      der = -(T(4) * (T(2) * a * s)) / T(4);
    } else if (a < T(1.5)) {
      const T u = T(3) - T(2) * a;
      val = (u * u) / T(8);
      der = (T(2) * u * (T(-2) * s)) / T(8);
    } else {
      val = T(0);
      der = T(0);
    }
  } else if constexpr (BASIS == TESS_BSPLINE_CUBIC) {
    if (a < T(1)) {
      const T a2 = a * a, a3 = a * a * a;
      val = (T(3) * a3 - T(6) * a2 + T(4)) / T(6);
      der = (T(3) * (T(3) * a2 * s) - T(6) * (T(2) * a * s)) / T(6);
    } else if (a < T(2)) {
      const T u = T(2) - a;
      val = (u * u * u) / T(6);
      der = (T(3) * (u * u) * (-s)) / T(6);
    } else {
      val = T(0);
      der = T(0);
    }
  } else {
    if (a < l / T(2)) {
      val = T(1) - (T(4) * (a * a) + l * l) / (T(4) * l);
      der = -(T(4) * (T(2) * a * s)) / (T(4) * l);
    } else if (a < T(1) - l / T(2)) {
      val = T(1) - a;
      der = -s;
    } else if (a < T(1) + l / T(2)) {
      const T u = T(1) + l / T(2) - a;
      val = (u * u) / (T(2) * l);
      der = (T(2) * u * (-s)) / (T(2) * l);
    } else {
      val = T(0);
      der = T(0);
    }
  }
}

// Julia evalpoly: muladd Horner from the highest coefficient.
template <typename T>
__device__ __forceinline__ T horner2(T x, T a0, T a1) { return fma(x, a1, a0); }
template <typename T>
__device__ __forceinline__ T horner3(T x, T a0, T a1, T a2) { return fma(x, fma(x, a2, a1), a0); }
template <typename T>
__device__ __forceinline__ T horner4(T x, T a0, T a1, T a2, T a3) { return fma(x, fma(x, fma(x, a3, a2), a1), a0); }

// Full-support fast path (src/Basis/bspline.jl:45-100): values and derivatives
// of the local polynomials at ξ_local = fract(ξ) (odd degree) or fract(ξ-½).
template <typename T, int BASIS>
__device__ __forceinline__ void bspline_local_jet(T loc, T* val, T* der) {
  if constexpr (BASIS == TESS_BSPLINE_LINEAR) {
    val[0] = horner2(loc, T(1), T(-1));
    val[1] = horner2(loc, T(0), T(1));
    der[0] = T(-1);
    der[1] = T(1);
  } else if constexpr (BASIS == TESS_BSPLINE_QUADRATIC) {
    const T half = T(1) / T(2);
    val[0] = horner3(loc, half, T(-1), half);
    val[1] = horner3(loc, half, T(1), T(-1));
    val[2] = horner3(loc, T(0), T(0), half);
    der[0] = horner2(loc, T(-1), T(1));
    der[1] = horner2(loc, T(1), T(-2));
    der[2] = horner2(loc, T(0), T(1));
  } else {
    const T s6 = T(1) / T(6), h2 = T(1) / T(2), t3 = T(2) / T(3), th = T(3) / T(2);
    val[0] = horner4(loc, s6, T(-1) / T(2), h2, T(-1) / T(6));
    val[1] = horner4(loc, t3, T(0), T(-1), h2);
    val[2] = horner4(loc, s6, h2, h2, T(-1) / T(2));
    val[3] = horner4(loc, T(0), T(0), T(0), s6);
    der[0] = horner3(loc, T(-1) / T(2), T(1), T(-1) / T(2));
    der[1] = horner3(loc, T(0), T(-2), th);
    der[2] = horner3(loc, h2, T(1), T(-3) / T(2));
    der[3] = horner3(loc, T(0), T(0), h2);
  }
}

// Support window of one particle along one axis, 0-based inclusive and clamped
// to the mesh (src/Basis/bspline.jl:27-40; uGIMP through
// src/Mesh/cartesian.jl:190-199 with r = 1 + l/2, src/Basis/gimp.jl:20-24).
template <typename T, int BASIS>
__device__ __forceinline__ void axis_window(T xi, T l, int dim_nodes, int& lo, int& hi) {
  int s, e;
  if constexpr (BASIS == TESS_UGIMP) {
    const T r = T(1) + l / T(2);
    s = int(floor(xi - r)) + 1;
    e = int(floor(xi + r));
  } else {
    constexpr int n = degree_of(BASIS);
    const T offset = T(n - 1) / T(2);
    s = int(floor(xi - offset));
    e = s + n;
  }
  lo = s > 0 ? s : 0;
  hi = e < dim_nodes - 1 ? e : dim_nodes - 1;
}

// isinside (src/Mesh/cartesian.jl:157-163, 227-234): 0 <= floor(ξ_d) <= dims_d-2.
template <typename T, int D>
__device__ __forceinline__ bool inside_xi(const T* xi, const DMesh<T, D>& m) {
  bool in = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T f = floor(xi[d]);
    in = in && (T(0) <= f) && (f <= T(m.dims[d] - 2));
  }
  return in;
}

}  // namespace tess
