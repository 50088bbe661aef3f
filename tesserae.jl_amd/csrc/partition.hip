// Stored basis weights in the reference layout, the CPU block partition
// reproduced on the device, and particle reordering.
//   update!(weights, particles, mesh)  src/Basis/basis_weight.jl:714-752
//   update!(partition, x)              src/Parallel/cpu_block_strategy.jl:62-219
//   reorder_particles!                 src/Parallel/cpu_block_strategy.jl:248-355
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "tess_correct.hpp"
#include "tess_device.hpp"
#include "tess_internal.hpp"

namespace tess {
namespace {

template <int N, int E>
struct IPow {
  static constexpr int v = N * IPow<N, E - 1>::v;
};
template <int N>
struct IPow<N, 0> {
  static constexpr int v = 1;
};

template <typename T, int D>
struct WArgs {
  DMesh<T, D> mesh;
  DField<T> x, l, w, gw;
  int64_t* window;
  int64_t n;
  uint32_t* err;
  int32_t corr;            // corrected basis bits (CORR kernels)
  const uint8_t* filter;   // node filter of update!(..., filter), NULL = Trues
};

// Per particle: support window, then either the full-support polynomial fill
// (bspline.jl:146-170) or the nodewise jet (basis_weight.jl:624-632).  CORR:
// a corrected basis, whose plain branch is the kernel's and whose other
// branches correct the nodewise jets (wls.jl:19-128, kernelcorrection.jl:26-35).
template <typename T, int D, int BASIS, bool CORR = false>
__global__ __launch_bounds__(256) void k_update_weights(WArgs<T, D> a) {
  constexpr int SW = support_width(BASIS);
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  T x[D], xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    x[d] = a.x(p, d);
    xi[d] = (x[d] - a.mesh.xmin[d]) * a.mesh.h_inv;
  }
  int64_t* win = a.window + p * 2 * D;
  if (!inside_xi<T, D>(xi, a.mesh)) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      win[2 * d] = 1;
      win[2 * d + 1] = 0;
    }
    return;
  }
  T l = T(0);
  if constexpr (BASIS == TESS_UGIMP) {
    l = a.l(p) / a.mesh.h;
    if (!(T(0) <= l && l <= T(1))) {
      atomicOr(a.err, 1u);
      return;
    }
  }
  int lo[D], hi[D], cnt[D];
  bool whole = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    axis_window<T, BASIS>(xi[d], l, a.mesh.dims[d], lo[d], hi[d]);
    cnt[d] = hi[d] - lo[d] + 1;
    whole = whole && cnt[d] == SW;
  }
  bool plain = true;
  if constexpr (CORR) {
    const bool fa = filter_alltrue<D, SW>(a.filter, a.mesh.dims, lo, cnt);
    plain = corr_mode(a.corr, BASIS, whole, a.filter != nullptr, fa, true) == CORR_PLAIN;
  }
  const bool full = BASIS != TESS_UGIMP && whole && plain;
  T val[D][SW], der[D][SW];
  if (full) {
    if constexpr (BASIS != TESS_UGIMP) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const T y = (degree_of(BASIS) % 2 == 1) ? xi[d] : xi[d] - T(0.5);
        bspline_local_jet<T, BASIS>(y - floor(y), val[d], der[d]);
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int k = 0; k < SW; ++k) {
        if (k <= hi[d] - lo[d]) {
          const T xin = (x[d] - a.mesh.node(d, lo[d] + k)) * a.mesh.h_inv;
          piecewise_jet<T, BASIS>(xin, l, val[d][k], der[d][k]);
        } else {
          val[d][k] = T(0);
          der[d][k] = T(0);
        }
      }
  }
  const T hinv = a.mesh.h_inv;
  Corrected<T, D, SW> cs;
  cs.mode = CORR_PLAIN;
  if constexpr (CORR) {
    if (!plain) {
      T xin[D][SW];
#pragma unroll
      for (int d = 0; d < D; ++d)
#pragma unroll
        for (int k = 0; k < SW; ++k) xin[d][k] = (x[d] - a.mesh.node(d, lo[d] + k)) * hinv;
      cs.prepare(a.corr, BASIS, true, a.filter, a.mesh.dims, lo, cnt, xin, val, der);
    }
  }
#pragma unroll
  for (int o = 0; o < IPow<SW, D>::v; ++o) {
    int aa[D], r = o;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      aa[d] = r % SW;
      r /= SW;
      ok = ok && aa[d] <= hi[d] - lo[d];
    }
    if (!ok) continue;
    const int64_t e = int64_t(o) + int64_t(IPow<SW, D>::v) * p;
    if (CORR && cs.mode == CORR_GENERAL) {
      T w, g[D];
      cs.node(val, aa, filter_pass<D>(a.filter, a.mesh.dims, lo, aa), hinv, w, g);
      a.w(e) = w;
#pragma unroll
      for (int c = 0; c < D; ++c) a.gw(e, c) = g[c];
      continue;
    }
    T w = val[0][aa[0]];
#pragma unroll
    for (int d = 1; d < D; ++d) w = w * val[d][aa[d]];
    a.w(e) = w;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      T g = (c == 0) ? der[0][aa[0]] : val[0][aa[0]];
#pragma unroll
      for (int d = 1; d < D; ++d) g = g * ((d == c) ? der[d][aa[d]] : val[d][aa[d]]);
      a.gw(e, c) = g * hinv;
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    win[2 * d] = lo[d] + 1;
    win[2 * d + 1] = hi[d] + 1;
  }
}

// ---- partition ----------------------------------------------------------------

// key = colour * nblocks + linear block id (cpu_block_strategy.jl:197-219: groups
// by colour, blocks in linear order inside a group), or `outside` past the end.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_partition_keys(DMesh<T, D> mesh, DField<T> X, int64_t n, int L,
                                                         int3 nb, uint32_t nbt, uint32_t* __restrict__ keys) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int nbd[3] = {nb.x, nb.y, nb.z};
  uint32_t lin = 0, color = 0;
  bool inside = true;
  int b[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T f = floor((X(p, d) - mesh.xmin[d]) * mesh.h_inv);  // _findblock, cartesian.jl:277-285
    inside = inside && T(0) <= f && f <= T(mesh.dims[d] - 2);
    b[d] = inside ? (int(f) >> L) : 0;
  }
  if (!inside) {
    keys[p] = (uint32_t(1) << D) * nbt;
    return;
  }
#pragma unroll
  for (int d = D - 1; d >= 0; --d) lin = lin * uint32_t(nbd[d]) + uint32_t(b[d]);
#pragma unroll
  for (int d = 0; d < D; ++d) color += uint32_t(b[d] & 1) << d;  // block_color, :197-203
  keys[p] = color * nbt + lin;
}

__global__ void k_partition_runs(const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ vals, int64_t n,
                                 uint32_t nbt, uint32_t outside, int32_t* __restrict__ pidx,
                                 int32_t* __restrict__ starts, int32_t* __restrict__ stops,
                                 uint32_t* __restrict__ count) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = skeys[i];
  const bool first = i == 0 || skeys[i - 1] != k;
  const bool last = i == n - 1 || skeys[i + 1] != k;
  if (k < outside) {
    const uint32_t blk = k % nbt;
    pidx[i] = int32_t(vals[i]) + 1;
    if (first) starts[blk] = int32_t(i) + 1;
    if (last) stops[blk] = int32_t(i) + 1;
    if (i == n - 1) *count = uint32_t(n);
  } else if (first) {
    *count = uint32_t(i);
  }
}

// ---- reorder ------------------------------------------------------------------

// per 256-thread block: number of i with pidx[i] == pidx[i-1] + 1
__global__ __launch_bounds__(256) void k_contiguity(const int32_t* __restrict__ pidx, int64_t na,
                                                    uint32_t* __restrict__ per_block) {
  __shared__ uint32_t s[4];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x + 1;
  const bool hit = i < na && pidx[i] == pidx[i - 1] + 1;
  const uint64_t m = __ballot(hit);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = uint32_t(__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) per_block[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ void k_mark_seen(const int32_t* __restrict__ pidx, int64_t na, int64_t n, uint32_t* __restrict__ seen,
                            uint32_t* __restrict__ err) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const int64_t p = int64_t(pidx[i]) - 1;
  if (p < 0 || p >= n) {
    atomicOr(err, 1u);
    return;
  }
  if (atomicAdd(&seen[p], 1u) != 0) atomicOr(err, 2u);
}

__global__ void k_build_perm(const int32_t* __restrict__ pidx, int64_t na, int64_t n,
                             const uint32_t* __restrict__ seen, const uint32_t* __restrict__ unseen_pos,
                             uint32_t* __restrict__ perm) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < na) perm[i] = uint32_t(pidx[i] - 1);
  if (i < n && !seen[i]) perm[na + unseen_pos[i]] = uint32_t(i);
}

__global__ void k_not(const uint32_t* __restrict__ seen, uint32_t* __restrict__ out, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = seen[i] ? 0u : 1u;
}

template <typename U>
__global__ void k_gather_field(const U* __restrict__ src, int64_t es, int64_t cs, int ncomp,
                               const uint32_t* __restrict__ perm, int64_t n, U* __restrict__ tmp) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t p = perm[k];
  for (int c = 0; c < ncomp; ++c) tmp[c * n + k] = src[p * es + c * cs];
}

template <typename U>
__global__ void k_scatter_back(U* __restrict__ dst, int64_t es, int64_t cs, int ncomp, const U* __restrict__ tmp,
                               int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < ncomp; ++c) dst[k * es + c * cs] = tmp[c * n + k];
}

__global__ void k_iota1(int32_t* __restrict__ pidx, int64_t na) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < na) pidx[i] = int32_t(i) + 1;
}

inline unsigned nblk(int64_t n, int t) { return unsigned(std::max<int64_t>(1, ceil_div(n, t))); }

template <typename T, int D>
void update_weights_typed(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_particles& P,
                          const tess_weights& W) {
  WArgs<T, D> a;
  a.mesh = dmesh<T, D>(mesh);
  a.x = dfield<T>(P.x);
  a.l = dfield<T>(P.l);
  a.w = dfield<T>(W.w);
  a.gw = dfield<T>(W.grad_w);
  a.window = W.window;
  a.n = P.n;
  a.corr = basis & ~0xff;
  a.filter = W.filter;
  uint32_t* err = ctx->counters.as<uint32_t>(4);
  a.err = err;
  TESS_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), ctx->stream));
  if (P.n == 0) return;
  const dim3 g(nblk(P.n, 256)), b(256);
  const int kernel = basis & 0xff;
  const bool corr = is_corrected(basis);
#define TESS_UPD(K)                                                                           \
  if (corr)                                                                                   \
    hipLaunchKernelGGL((k_update_weights<T, D, K, true>), g, b, 0, ctx->stream, a);           \
  else                                                                                        \
    hipLaunchKernelGGL((k_update_weights<T, D, K>), g, b, 0, ctx->stream, a);
  switch (kernel) {
    case TESS_BSPLINE_LINEAR:
      TESS_UPD(TESS_BSPLINE_LINEAR)
      break;
    case TESS_BSPLINE_QUADRATIC:
      TESS_UPD(TESS_BSPLINE_QUADRATIC)
      break;
    case TESS_BSPLINE_CUBIC:
      TESS_UPD(TESS_BSPLINE_CUBIC)
      break;
    default:
      TESS_UPD(TESS_UGIMP)
      break;
  }
#undef TESS_UPD
  TESS_HIP(hipGetLastError());
  if (kernel == TESS_UGIMP) {
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
    TESS_HIP(hipMemcpyAsync(hc, err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    if (hc[0]) raise(TESS_ERR_ARGUMENT, "uGIMP requires 0 ≤ pt.l ≤ spacing(mesh)");
  }
}

template <typename T, int D>
void partition_typed(tess_ctx* ctx, const tess_mesh& mesh, tess_field x, int64_t n, int32_t* pidx, int32_t* starts,
                     int32_t* stops, int64_t* nassigned) {
  const int L = mesh.block_size_log2;
  int nbd[3] = {1, 1, 1};
  uint64_t nbt = 1;
  for (int d = 0; d < D; ++d) {
    nbd[d] = int(((mesh.dims[d] - 1) >> L) + 1);  // nblocks, cartesian.jl:241-243
    nbt *= uint64_t(nbd[d]);
  }
  const uint64_t outside = (uint64_t(1) << D) * nbt;
  if (outside >= uint64_t(UINT32_MAX) || n >= int64_t(INT32_MAX))
    raise(TESS_ERR_ARGUMENT, "Partition block count exceeds packed block id capacity.");
  TESS_HIP(hipMemsetAsync(starts, 0, nbt * sizeof(int32_t), ctx->stream));
  TESS_HIP(hipMemsetAsync(stops, 0, nbt * sizeof(int32_t), ctx->stream));
  uint32_t* cnt = ctx->counters.as<uint32_t>(4);
  TESS_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), ctx->stream));
  if (n > 0) {
    uint32_t* keys = ctx->keys.as<uint32_t>(size_t(n));
    uint32_t* skeys = ctx->keys_sorted.as<uint32_t>(size_t(n));
    uint32_t* vals = ctx->vals.as<uint32_t>(size_t(n));
    hipLaunchKernelGGL((k_partition_keys<T, D>), dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, dmesh<T, D>(mesh),
                       dfield<T>(x), n, L, make_int3(nbd[0], nbd[1], nbd[2]), uint32_t(nbt), keys);
    TESS_HIP(hipGetLastError());
    int bits = 0;
    while ((uint64_t(1) << bits) <= outside) ++bits;
    radix_sort_pairs(ctx, keys, nullptr, skeys, vals, n, bits);
    hipLaunchKernelGGL(k_partition_runs, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, skeys, vals, n,
                       uint32_t(nbt), uint32_t(outside), pidx, starts, stops, cnt);
    TESS_HIP(hipGetLastError());
    // the internal tile binning used the same buffers
    ctx->bin.valid = false;
    ctx->bin.state_valid = false;
  }
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
  TESS_HIP(hipMemcpyAsync(hc, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
  *nassigned = n > 0 ? int64_t(hc[0]) : 0;
}

// ---- partition without allocation --------------------------------------------
//
// tess_partition_update_noalloc runs the partition above over a fixed work area
// instead of the context's growing buffers: keys, sorted keys, values, the sort's
// ping-pong keys and values (n words each), the digit histograms, the scan's
// block sums and one count.  The area is a file-scope device array sized for
// kNoAllocMaxParticles; every GPU that loads the library holds its own copy
// until the process ends.  Calls on one GPU share it: calls that overlap on a
// device race.

constexpr int64_t kNoAllocMaxParticles = TESS_PARTITION_NOALLOC_MAX;
static_assert(kNoAllocMaxParticles % 4096 == 0, "the histogram area counts whole 4096-key sort chunks");
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kNoAllocWords = size_t(kNoAllocMaxParticles) * sizeof(uint32_t);
// 4096-key sort chunks × 256 digits; the scan of those words has 16 blocks of 4096
constexpr size_t kNoAllocHistBytes = size_t(kNoAllocMaxParticles / 4096) * 256 * sizeof(uint32_t);
constexpr size_t kNoAllocScanBytes = 4096;
static_assert(kNoAllocMaxParticles / 4096 * 256 <= 4096 * 4096, "the histogram scan takes two levels");
constexpr size_t kNoAllocScratchBytes = 5 * align_up(kNoAllocWords) + align_up(kNoAllocHistBytes) + kNoAllocScanBytes +
                                        kAlign;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_store_nassigned(const uint32_t* __restrict__ count, int64_t n, int64_t* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = n > 0 ? int64_t(*count) : 0;
}

template <typename T, int D>
void partition_noalloc_typed(const tess_mesh& mesh, tess_field x, int64_t n, int32_t* pidx, int32_t* starts,
                             int32_t* stops, int64_t* nassigned, hipStream_t stream, uint8_t* area) {
  const int L = mesh.block_size_log2;
  int nbd[3] = {1, 1, 1};
  uint64_t nbt = 1;
  for (int d = 0; d < D; ++d) {
    nbd[d] = int(((mesh.dims[d] - 1) >> L) + 1);  // nblocks, cartesian.jl:241-243
    nbt *= uint64_t(nbd[d]);
  }
  const uint64_t outside = (uint64_t(1) << D) * nbt;
  if (outside >= uint64_t(UINT32_MAX) || n >= int64_t(INT32_MAX))
    raise(TESS_ERR_ARGUMENT, "Partition block count exceeds packed block id capacity.");
  // carve the area: five n-word arrays, histograms, scan block sums, the count
  uint8_t* q = area;
  auto take = [&](size_t bytes) {
    uint8_t* r = q;
    q += align_up(bytes);
    return r;
  };
  uint32_t* keys = reinterpret_cast<uint32_t*>(take(kNoAllocWords));
  uint32_t* skeys = reinterpret_cast<uint32_t*>(take(kNoAllocWords));
  uint32_t* vals = reinterpret_cast<uint32_t*>(take(kNoAllocWords));
  SortWork w;
  w.stream = stream;
  w.keys_alt = reinterpret_cast<uint32_t*>(take(kNoAllocWords));
  w.vals_alt = reinterpret_cast<uint32_t*>(take(kNoAllocWords));
  w.hist = reinterpret_cast<uint32_t*>(take(kNoAllocHistBytes));
  w.scan_tmp = take(kNoAllocScanBytes);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t)));
  // fill kernels rather than memset nodes, so that a captured graph replays them
  hipLaunchKernelGGL(k_fill_i32, dim3(nblk(int64_t(nbt), 256)), dim3(256), 0, stream, starts, int64_t(nbt), 0);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_fill_i32, dim3(nblk(int64_t(nbt), 256)), dim3(256), 0, stream, stops, int64_t(nbt), 0);
  TESS_HIP(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL((k_partition_keys<T, D>), dim3(nblk(n, 256)), dim3(256), 0, stream, dmesh<T, D>(mesh),
                       dfield<T>(x), n, L, make_int3(nbd[0], nbd[1], nbd[2]), uint32_t(nbt), keys);
    TESS_HIP(hipGetLastError());
    int bits = 0;
    while ((uint64_t(1) << bits) <= outside) ++bits;
    radix_sort_pairs_ws(w, keys, nullptr, skeys, vals, n, bits);
    // exactly one thread writes the count when n > 0
    hipLaunchKernelGGL(k_partition_runs, dim3(nblk(n, 256)), dim3(256), 0, stream, skeys, vals, n, uint32_t(nbt),
                       uint32_t(outside), pidx, starts, stops, cnt);
    TESS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_store_nassigned, dim3(1), dim3(64), 0, stream, cnt, n, nassigned);
  TESS_HIP(hipGetLastError());
}

template <typename U>
void permute_field(tess_ctx* ctx, const tess_field& f, int ncomp, const uint32_t* perm, int64_t n) {
  if (!f.ptr || n == 0) return;
  U* tmp = ctx->reorder_tmp.as<U>(size_t(n) * ncomp);
  hipLaunchKernelGGL((k_gather_field<U>), dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                     static_cast<const U*>(f.ptr), f.elem_stride, f.comp_stride, ncomp, perm, n, tmp);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_scatter_back<U>), dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, static_cast<U*>(f.ptr),
                     f.elem_stride, f.comp_stride, ncomp, tmp, n);
  TESS_HIP(hipGetLastError());
}

}  // namespace

void launch_update_weights(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_particles& P,
                           const tess_weights& W) {
  if (mesh.dim == 2 && mesh.dtype == TESS_FLOAT32) return update_weights_typed<float, 2>(ctx, basis, mesh, P, W);
  if (mesh.dim == 3 && mesh.dtype == TESS_FLOAT32) return update_weights_typed<float, 3>(ctx, basis, mesh, P, W);
  if (mesh.dim == 2 && mesh.dtype == TESS_FLOAT64) return update_weights_typed<double, 2>(ctx, basis, mesh, P, W);
  if (mesh.dim == 3 && mesh.dtype == TESS_FLOAT64) return update_weights_typed<double, 3>(ctx, basis, mesh, P, W);
  raise(TESS_ERR_UNSUPPORTED, "unsupported (dim, dtype)");
}

void run_partition_update(tess_ctx* ctx, const tess_mesh& mesh, tess_field x, int64_t n, int32_t* pidx,
                          int32_t* starts, int32_t* stops, int64_t* nassigned) {
  if (mesh.dim == 2 && mesh.dtype == TESS_FLOAT32)
    return partition_typed<float, 2>(ctx, mesh, x, n, pidx, starts, stops, nassigned);
  if (mesh.dim == 3 && mesh.dtype == TESS_FLOAT32)
    return partition_typed<float, 3>(ctx, mesh, x, n, pidx, starts, stops, nassigned);
  if (mesh.dim == 2 && mesh.dtype == TESS_FLOAT64)
    return partition_typed<double, 2>(ctx, mesh, x, n, pidx, starts, stops, nassigned);
  if (mesh.dim == 3 && mesh.dtype == TESS_FLOAT64)
    return partition_typed<double, 3>(ctx, mesh, x, n, pidx, starts, stops, nassigned);
  raise(TESS_ERR_UNSUPPORTED, "unsupported (dim, dtype)");
}

int64_t partition_noalloc_max_particles() { return kNoAllocMaxParticles; }

void noalloc_scratch_resolve() {
  int dev = 0;
  TESS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) raise(TESS_ERR_DEVICE, "device ordinal beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* p = nullptr;
  // each device holds its own copy of the array at its own address; host code
  // uses this address, never the array's name
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    raise(TESS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
}

void run_partition_update_noalloc(const tess_mesh& mesh, tess_field x, int64_t n, int32_t* pidx, int32_t* starts,
                                  int32_t* stops, int64_t* nassigned, hipStream_t stream) {
  if (n > kNoAllocMaxParticles)
    raise(TESS_ERR_ARGUMENT, "tess_partition_update_noalloc handles at most " + std::to_string(kNoAllocMaxParticles) +
                                 " particles, got " + std::to_string(n));
  noalloc_scratch_resolve();
  int dev = 0;
  TESS_HIP(hipGetDevice(&dev));
  uint8_t* area = static_cast<uint8_t*>(g_scratch_ptr[dev]);
  if (mesh.dim == 2 && mesh.dtype == TESS_FLOAT32)
    return partition_noalloc_typed<float, 2>(mesh, x, n, pidx, starts, stops, nassigned, stream, area);
  if (mesh.dim == 3 && mesh.dtype == TESS_FLOAT32)
    return partition_noalloc_typed<float, 3>(mesh, x, n, pidx, starts, stops, nassigned, stream, area);
  if (mesh.dim == 2 && mesh.dtype == TESS_FLOAT64)
    return partition_noalloc_typed<double, 2>(mesh, x, n, pidx, starts, stops, nassigned, stream, area);
  if (mesh.dim == 3 && mesh.dtype == TESS_FLOAT64)
    return partition_noalloc_typed<double, 3>(mesh, x, n, pidx, starts, stops, nassigned, stream, area);
  raise(TESS_ERR_UNSUPPORTED, "unsupported (dim, dtype)");
}

void permute_particles(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, const uint32_t* perm) {
  const int64_t n = P.n;
  const int D = mesh.dim;
  if (mesh.dtype == TESS_FLOAT32) {
    permute_field<uint32_t>(ctx, P.x, D, perm, n);
    permute_field<uint32_t>(ctx, P.v, D, perm, n);
    permute_field<uint32_t>(ctx, P.V, 1, perm, n);
    permute_field<uint32_t>(ctx, P.grad_v, D * D, perm, n);
    permute_field<uint32_t>(ctx, P.F, D * D, perm, n);
    permute_field<uint32_t>(ctx, P.sigma, D * D, perm, n);
    permute_field<uint32_t>(ctx, P.l, 1, perm, n);
    permute_field<uint32_t>(ctx, P.logE, 1, perm, n);
  } else {
    permute_field<uint64_t>(ctx, P.x, D, perm, n);
    permute_field<uint64_t>(ctx, P.v, D, perm, n);
    permute_field<uint64_t>(ctx, P.V, 1, perm, n);
    permute_field<uint64_t>(ctx, P.grad_v, D * D, perm, n);
    permute_field<uint64_t>(ctx, P.F, D * D, perm, n);
    permute_field<uint64_t>(ctx, P.sigma, D * D, perm, n);
    permute_field<uint64_t>(ctx, P.l, 1, perm, n);
    permute_field<uint64_t>(ctx, P.logE, 1, perm, n);
  }
}

void run_reorder(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, int32_t* pidx, int64_t na,
                 double threshold, int32_t* reordered) {
  *reordered = 0;
  const int64_t n = P.n;
  if (!(threshold >= 0.0 && threshold <= 1.0)) raise(TESS_ERR_ARGUMENT, "threshold must be in [0, 1].");
  if (threshold == 0.0) return;
  if (na > n)
    raise(TESS_ERR_STATE, "reorder_particles!: The block assignment contains more particle IDs than exist");
  uint32_t* cnt = ctx->counters.as<uint32_t>(4);
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
  bool go = threshold == 1.0;
  if (!go) {
    double score = 1.0;
    if (na > 1) {
      const unsigned nb = nblk(na - 1, 256);
      uint32_t* per_block = ctx->per_tile.as<uint32_t>(nb);
      hipLaunchKernelGGL(k_contiguity, dim3(nb), dim3(256), 0, ctx->stream, pidx, na, per_block);
      TESS_HIP(hipGetLastError());
      sum_u32(ctx, per_block, nb, cnt);
      TESS_HIP(hipMemcpyAsync(hc, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
      TESS_HIP(hipStreamSynchronize(ctx->stream));
      score = double(hc[0]) / double(na - 1);
    }
    go = score < threshold;
  }
  if (!go || n == 0) return;
  uint32_t* perm = ctx->keys_alt.as<uint32_t>(size_t(n));
  uint32_t* seen = ctx->vals_alt.as<uint32_t>(size_t(n));
  uint32_t* pos = ctx->keys.as<uint32_t>(size_t(n));
  TESS_HIP(hipMemsetAsync(seen, 0, size_t(n) * sizeof(uint32_t), ctx->stream));
  TESS_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), ctx->stream));
  if (na > 0) {
    hipLaunchKernelGGL(k_mark_seen, dim3(nblk(na, 256)), dim3(256), 0, ctx->stream, pidx, na, n, seen, cnt);
    TESS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_not, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, seen, pos, n);
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(ctx, pos, n, nullptr);
  TESS_HIP(hipMemcpyAsync(hc, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
  if (hc[0] & 1u) raise(TESS_ERR_STATE, "reorder_particles!: particle ID is out of range.");
  if (hc[0] & 2u) raise(TESS_ERR_STATE, "reorder_particles!: particle is duplicated in the block assignment.");
  hipLaunchKernelGGL(k_build_perm, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, pidx, na, n, seen, pos, perm);
  TESS_HIP(hipGetLastError());
  permute_particles(ctx, mesh, P, perm);
  if (na > 0) {
    hipLaunchKernelGGL(k_iota1, dim3(nblk(na, 256)), dim3(256), 0, ctx->stream, pidx, na);
    TESS_HIP(hipGetLastError());
  }
  ctx->bin.valid = false;
    ctx->bin.state_valid = false;
  *reordered = 1;
}

}  // namespace tess
