// SpIndices block activation and numbering from particle positions
// (update_sparsity!(spinds, xₚ, mesh), src/Grids/sparsity.jl:23-40,53-60,
// 120-238).  Blocks of 2^L nodes per axis; a block holding a particle is
// occupied, every block within ±1 block of an occupied one is active, and the
// active blocks are numbered 1, 2, ... in column-major block order.
#include <hip/hip_runtime.h>

#include "tess_device.hpp"
#include "tess_internal.hpp"

namespace tess {
namespace {

// counter slots of this file inside ctx->counters (uint32)
enum { SP_COUNT = 32, SP_CHANGED = 33 };

template <int D>
struct BlockGrid {
  int32_t nb[D];
  int64_t total;
};

// occupied[b] = 1 for the block of every particle inside the mesh
// (findblock, src/Mesh/cartesian.jl:277-285: cell0 = floor((x-xmin)h⁻¹),
// inside iff 0 <= cell0 <= dims-2, block = cell0 >> L).
template <typename T, int D>
__global__ __launch_bounds__(256) void k_sp_occupied(DField<T> x, int64_t n, DMesh<T, D> mesh, int L,
                                                     BlockGrid<D> bg, uint32_t* __restrict__ occ) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int64_t b = 0;
  bool inside = true;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    const T f = floor((x(p, d) - mesh.xmin[d]) * mesh.h_inv);
    inside = inside && T(0) <= f && f <= T(mesh.dims[d] - 2);
    const int c = inside ? int(f) : 0;
    b = b * bg.nb[d] + (c >> L);
  }
  if (inside) occ[b] = 1u;  // every writer stores the same value
}

// act[b] = 1 if any block of the 3^dim neighbourhood is occupied
// (_activate_block_neighborhood!, sparsity.jl:141-145); flags a change of the
// active set against the numbering passed in (sparsity.jl:262-267).
template <int D>
__global__ __launch_bounds__(256) void k_sp_activity(const uint32_t* __restrict__ occ, BlockGrid<D> bg,
                                                     const int32_t* __restrict__ old_numbering,
                                                     uint32_t* __restrict__ act, uint32_t* __restrict__ scan,
                                                     uint32_t* __restrict__ changed) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool a = false, diff = false;
  if (b < bg.total) {
    int I[D];
    int64_t r = b;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      I[d] = int(r % bg.nb[d]);
      r /= bg.nb[d];
    }
    constexpr int NN = (D == 3) ? 27 : 9;
    for (int k = 0; k < NN && !a; ++k) {
      int q = k;
      int64_t nb = 0;
      bool ok = true;
#pragma unroll
      for (int d = D - 1; d >= 0; --d) {
        int div = 1;
        for (int e = 0; e < d; ++e) div *= 3;
        const int J = I[d] + (q / div) % 3 - 1;
        ok = ok && J >= 0 && J < bg.nb[d];
        nb = nb * bg.nb[d] + J;
      }
      a = ok && occ[nb] != 0u;
    }
    act[b] = a ? 1u : 0u;
    scan[b] = a ? 1u : 0u;
    diff = (old_numbering[b] != 0) != a;
  }
  // one store per wave that saw a change
  if (__ballot(diff) != 0 && (threadIdx.x & 63) == 0) *changed = 1u;
}

__global__ __launch_bounds__(256) void k_sp_number(const uint32_t* __restrict__ act,
                                                   const uint32_t* __restrict__ excl, int64_t total,
                                                   int32_t* __restrict__ numbering) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b < total) numbering[b] = act[b] ? int32_t(excl[b] + 1u) : 0;
}

template <int D>
BlockGrid<D> block_grid(const tess_mesh& mesh) {
  BlockGrid<D> bg;
  bg.total = 1;
  for (int d = 0; d < D; ++d) {
    bg.nb[d] = int32_t(((mesh.dims[d] - 1) >> mesh.block_size_log2) + 1);
    bg.total *= bg.nb[d];
  }
  return bg;
}

template <typename T, int D>
void sparsity_impl(tess_ctx* ctx, const tess_mesh& mesh, tess_field x, int64_t n, tess_sparsity& sp,
                   int32_t* changed) {
  const BlockGrid<D> bg = block_grid<D>(mesh);
  const int L = mesh.block_size_log2;
  const int64_t nbt = bg.total;
  if (nbt >= (int64_t(1) << 31)) raise(TESS_ERR_UNSUPPORTED, "block table larger than 2^31 blocks");
  uint32_t* occ = ctx->sp_occ.as<uint32_t>(size_t(nbt));
  uint32_t* act = ctx->sp_act.as<uint32_t>(size_t(nbt));
  uint32_t* scan = ctx->sp_scan.as<uint32_t>(size_t(nbt));
  uint32_t* counters = ctx->counters.as<uint32_t>(64);
  TESS_HIP(hipMemsetAsync(occ, 0, size_t(nbt) * sizeof(uint32_t), ctx->stream));
  TESS_HIP(hipMemsetAsync(counters + SP_COUNT, 0, 2 * sizeof(uint32_t), ctx->stream));
  if (n > 0) {
    const DMesh<T, D> dm = dmesh<T, D>(mesh);
    hipLaunchKernelGGL((k_sp_occupied<T, D>), dim3(unsigned(ceil_div(n, 256))), dim3(256), 0, ctx->stream,
                       dfield<T>(x), n, dm, L, bg, occ);
    TESS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((k_sp_activity<D>), dim3(unsigned(ceil_div(nbt, 256))), dim3(256), 0, ctx->stream, occ, bg,
                     static_cast<const int32_t*>(sp.blocknumbering), act, scan, counters + SP_CHANGED);
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(ctx, scan, nbt, counters + SP_COUNT);
  hipLaunchKernelGGL(k_sp_number, dim3(unsigned(ceil_div(nbt, 256))), dim3(256), 0, ctx->stream, act, scan, nbt,
                     sp.blocknumbering);
  TESS_HIP(hipGetLastError());
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
  TESS_HIP(hipMemcpyAsync(hc + SP_COUNT, counters + SP_COUNT, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
  sp.storage_len = int64_t(hc[SP_COUNT]) << (L * D);
  if (changed) *changed = hc[SP_CHANGED] ? 1 : 0;
}

}  // namespace

void run_sparsity_update(tess_ctx* ctx, const tess_mesh& mesh, tess_field x, int64_t n, tess_sparsity& sp,
                         int32_t* changed) {
  if (mesh.dtype == TESS_FLOAT32) {
    if (mesh.dim == 2) sparsity_impl<float, 2>(ctx, mesh, x, n, sp, changed);
    else sparsity_impl<float, 3>(ctx, mesh, x, n, sp, changed);
  } else {
    if (mesh.dim == 2) sparsity_impl<double, 2>(ctx, mesh, x, n, sp, changed);
    else sparsity_impl<double, 3>(ctx, mesh, x, n, sp, changed);
  }
}

}  // namespace tess
