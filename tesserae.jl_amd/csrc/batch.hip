// Batches of independent scenes (the engine's batching extension; no
// counterpart in the reference).  Scenes of one dimension, dtype and basis but
// their own mesh, particle count, constants and Δt are stepped together: the
// tiles and particles of all scenes share one index space (scene s owns tiles
// [toff_s, toff_s+1) and particles [poff_s, poff_s+1)), so each kernel of a
// step — keys, sort, tile runs, P2G, gather, G2P — is one launch for the whole
// batch.  Tile kernels look their scene up from the tile index
// (scene_of/bind_scene in transfer_impl.hpp).  An autotuner times the launch
// variants (waves per tile in P2G and G2P) on the batch itself.
#include "transfer_impl.hpp"

#include <algorithm>
#include <vector>

namespace tess {
namespace {

template <typename T, int D>
SceneRec<T, D> scene_record(const tess_scene& sc, int64_t toff, int64_t poff) {
  SceneRec<T, D> r;
  std::memset(static_cast<void*>(&r), 0, sizeof(r));
  const Args<T, D> a = make_args<T, D>(sc.mesh, &sc.grid, sc.particles, nullptr, &sc.params);
  r.mesh = a.mesh;
  r.x = a.x;
  r.v = a.v;
  r.V = a.V;
  r.gradv = a.gradv;
  r.F = a.F;
  r.sigma = a.sigma;
  r.l = a.l;
  r.logE = a.logE;
  r.gm = a.gm;
  r.gmv = a.gmv;
  r.gf = a.gf;
  r.gvn = a.gvn;
  r.gv = a.gv;
  r.mp = a.mp;
  r.dt = a.dt;
  r.lambda = a.lambda;
  r.mu = a.mu;
  r.V0 = a.V0;
  for (int d = 0; d < D; ++d) {
    r.gravity[d] = a.gravity[d];
    r.ntiles[d] = a.ntiles[d];
  }
  r.slip = a.slip;
  r.toff = toff;
  r.poff = poff;
  r.n = sc.particles.n;
  r.nnodes = 1;
  for (int d = 0; d < D; ++d) r.nnodes *= sc.mesh.dims[d];
  return r;
}

template <typename T, int D>
void batch_create_impl(tess_ctx* ctx, tess_batch& b) {
  std::vector<SceneRec<T, D>> recs;
  for (int s = 0; s < b.nscenes; ++s) recs.push_back(scene_record<T, D>(b.scenes[size_t(s)], b.toff[size_t(s)], b.poff[size_t(s)]));
  void* p = b.recs.get(recs.size() * sizeof(SceneRec<T, D>));
  TESS_HIP(hipMemcpyAsync(p, recs.data(), recs.size() * sizeof(SceneRec<T, D>), hipMemcpyHostToDevice, ctx->stream));
  TESS_HIP(hipMemcpyAsync(b.d_toff.get(b.toff.size() * 8), b.toff.data(), b.toff.size() * 8, hipMemcpyHostToDevice,
                          ctx->stream));
  TESS_HIP(hipMemcpyAsync(b.d_poff.get(b.poff.size() * 8), b.poff.data(), b.poff.size() * 8, hipMemcpyHostToDevice,
                          ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
}

template <typename T, int D>
Args<T, D> batch_args(tess_ctx* ctx, tess_batch& b) {
  Args<T, D> a;
  std::memset(static_cast<void*>(&a), 0, sizeof(a));
  a.scenes = static_cast<const SceneRec<T, D>*>(b.recs.ptr);
  a.scene_toff = static_cast<const int64_t*>(b.d_toff.ptr);
  a.scene_poff = static_cast<const int64_t*>(b.d_poff.ptr);
  a.nscenes = b.nscenes;
  a.n = b.ntot;
  fill_binning_ptrs<T, D>(ctx, a);
  return a;
}

// Full sort of the batch's keys every step (the batch-wide tile space), then the
// tile runs, active tiles and touched node blocks, as bin_particles does.
template <typename T, int D, int BASIS>
void batch_bin_run(tess_ctx* ctx, tess_batch& B, bool keys_from_g2p);

template <typename T, int D, int BASIS>
void batch_bin(tess_ctx* ctx, tess_batch& B, bool keys_from_g2p) {
  batch_bin_run<T, D, BASIS>(ctx, B, keys_from_g2p);
  if (TESS_DEBUG && B.ntot > 0) {
    const Binning& b = ctx->bin;
    debug_check_binning(ctx->stream, static_cast<const uint32_t*>(ctx->keys_sorted.ptr),
                        static_cast<const uint32_t*>(ctx->vals.ptr), static_cast<const uint32_t*>(ctx->tile_start.ptr),
                        static_cast<const uint32_t*>(ctx->tile_end.ptr),
                        static_cast<const uint32_t*>(ctx->active_tiles.ptr), B.ntot, B.tiles_total, b.n_active,
                        b.n_inside, "batch_bin");
  }
}

template <typename T, int D, int BASIS>
void batch_bin_run(tess_ctx* ctx, tess_batch& B, bool keys_from_g2p) {
  const int64_t n = B.ntot;
  const uint64_t ntiles = uint64_t(B.tiles_total);
  Binning& b = ctx->bin;
  const bool incremental = keys_from_g2p && B.keys_ready && B.binned && b.n == n && b.ntiles == int64_t(ntiles) &&
                           n > 0;
  B.keys_ready = false;
  b.valid = false;
  b.state_valid = false;  // the single-scene incremental state does not describe a batch
  b.keys_ready = false;
  uint32_t* counters = ctx->counters.as<uint32_t>(CNT_SLOTS);
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
  const int kbits = key_bits_for(ntiles * 64);
  int64_t movers = -1;
  if (incremental) {
    // the last G2P counted the particles whose key changed; with none the
    // previous order and tile lists stand
    TESS_HIP(hipMemcpyAsync(hc, counters, CNT_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    movers_readback(ctx);
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    movers = movers_total(ctx);
    if (movers == 0) {
      ctx->last_active_nodes = int64_t(hc[CNT_NODES]);
      B.kept_binnings += 1;
      b.valid = true;
      return;
    }
  }
  TESS_HIP(hipMemsetAsync(counters, 0, CNT_BIN_SLOTS * sizeof(uint32_t), ctx->stream));
  uint32_t* keys = ctx->keys.as<uint32_t>(size_t(std::max<int64_t>(n, 1)));
  uint32_t* skeys = ctx->keys_sorted.as<uint32_t>(size_t(std::max<int64_t>(n, 1)));
  uint32_t* perm = ctx->vals.as<uint32_t>(size_t(std::max<int64_t>(n, 1)));
  uint32_t* tstart = ctx->tile_start.as<uint32_t>(ntiles);
  uint32_t* tend = ctx->tile_end.as<uint32_t>(ntiles);
  uint32_t* tflag = ctx->tile_flag.as<uint32_t>(ntiles);
  uint32_t* tslot = ctx->tile_slot.as<uint32_t>(ntiles);
  uint32_t* active = ctx->active_tiles.as<uint32_t>(ntiles);
  uint32_t* nbflag = ctx->nb_flag.as<uint32_t>(ntiles);
  uint32_t* nbslot = ctx->nb_slot.as<uint32_t>(ntiles);
  uint32_t* nblist = ctx->nb_list.as<uint32_t>(ntiles);
  Args<T, D> a = batch_args<T, D>(ctx, B);
  const int tok = prof_begin(ctx, "binning", 0.0);
  if (incremental) {
    // movers merged into the previous sorted order
    uint32_t* nk = ctx->inc_nk.as<uint32_t>(size_t(n));
    uint32_t* spos = ctx->inc_spos.as<uint32_t>(size_t(n) + 1);
    TESS_HIP(hipMemsetAsync(spos + n, 0, sizeof(uint32_t), ctx->stream));
    exclusive_scan_u32(ctx, spos, n + 1, counters + CNT_STAY);
    TESS_HIP(hipMemcpyAsync(hc + CNT_STAY, counters + CNT_STAY, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t M = n - int64_t(hc[CNT_STAY]);
    if (M > 0) merge_movers(ctx, nk, spos, n, M, kbits);
    skeys = static_cast<uint32_t*>(ctx->keys_sorted.ptr);
    perm = static_cast<uint32_t*>(ctx->vals.ptr);
  } else {
    hipLaunchKernelGGL((k_bin_keys<T, D, BASIS, false, true>), dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a,
                       keys, uint32_t(ntiles), counters);
    TESS_HIP(hipGetLastError());
    radix_sort_pairs(ctx, keys, nullptr, skeys, perm, n, kbits);
  }
  TESS_HIP(hipMemsetAsync(tstart, 0, ntiles * sizeof(uint32_t), ctx->stream));
  TESS_HIP(hipMemsetAsync(tend, 0, ntiles * sizeof(uint32_t), ctx->stream));
  hipLaunchKernelGGL(k_tile_runs, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, skeys, n, uint32_t(ntiles),
                     tstart, tend, counters);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_tile_flags, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, tstart, tend,
                     uint32_t(ntiles), tflag, tslot);
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(ctx, tslot, int64_t(ntiles), counters + CNT_ACTIVE);
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, tflag, tslot,
                     uint32_t(ntiles), active);
  TESS_HIP(hipGetLastError());
  TESS_HIP(hipMemsetAsync(nbflag, 0, ntiles * sizeof(uint32_t), ctx->stream));
  hipLaunchKernelGGL((k_nb_mark_batch<T, D>), dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, a, tflag,
                     uint32_t(ntiles), nbflag);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, nbflag, nbslot,
                     int64_t(ntiles));
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(ctx, nbslot, int64_t(ntiles), counters + CNT_NB);
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, nbflag, nbslot,
                     uint32_t(ntiles), nblist);
  TESS_HIP(hipGetLastError());
  prof_end(ctx, tok);
  TESS_HIP(hipMemcpyAsync(hc, counters, CNT_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
  if (hc[CNT_ERR] & ERR_GIMP_LENGTH) raise(TESS_ERR_ARGUMENT, "uGIMP requires 0 ≤ pt.l ≤ spacing(mesh)");
  b.n_inside = n > 0 ? int64_t(hc[CNT_INSIDE]) : 0;
  ctx->last_active_nodes = int64_t(hc[CNT_NODES]);  // written by the previous step's gather
  b.n_active = int64_t(hc[CNT_ACTIVE]);
  b.n_nodeblocks = int64_t(hc[CNT_NB]);
  b.ntiles = int64_t(ntiles);
  b.n = n;
  b.valid = true;
  B.binned = true;
}

// ---- layered reorder of a batch ------------------------------------------------
//
// Each scene's particles into the layered order of its tiles (the order the
// tile kernels read contiguously), its particles outside the mesh after them
// in their previous order.  Every particle stays in its scene's arrays: the
// layered slot L of a scene-s particle becomes global index L + (particles
// outside the mesh in scenes before s), since the scene's inside particles
// start at sorted position poff_s minus those.

__global__ void k_mark_outside(const uint32_t* __restrict__ perm, int64_t from, int64_t n, uint32_t* __restrict__ flag) {
  const int64_t j = from + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) flag[perm[j]] = 1u;
}

// new global index of the particle at each sorted position (pout) and the
// gather order (new index -> old index)
template <typename T, int D>
__global__ __launch_bounds__(256) void k_batch_new_index(Args<T, D> a, const uint32_t* __restrict__ layered,
                                                         const uint32_t* __restrict__ R, uint32_t* __restrict__ order,
                                                         uint32_t* __restrict__ pout) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < a.n) {
    const uint32_t g = a.perm[j];
    const int s = scene_of(a.scene_poff, a.nscenes, int64_t(g));
    const int64_t p0 = a.scene_poff[s], p1 = a.scene_poff[s + 1];
    const uint32_t dst = j < a.n_inside ? uint32_t(int64_t(layered[j]) + int64_t(R[p0]))
                                        : uint32_t(p1 - int64_t(R[p1] - R[p0]) + int64_t(R[g] - R[p0]));
    order[dst] = g;
    pout[j] = dst;
  }
}

template <typename T, int D>
__device__ __forceinline__ DField<T> scene_field(const SceneRec<T, D>& r, int fid) {
  switch (fid) {
    case 0: return r.x;
    case 1: return r.v;
    case 2: return r.V;
    case 3: return r.gradv;
    case 4: return r.F;
    case 5: return r.sigma;
    case 6: return r.l;
    default: return r.logE;
  }
}

// components of the particle fields x, v, V, ∇v, F, σ, l, log E
template <int D>
constexpr int kFieldComps[8] = {D, D, 1, D * D, D * D, D * D, 1, 1};
template <int D>
constexpr int kStateWidth = 2 * D + 2 + 3 * D * D + 1;

// With `order`: every field of particle order[i] into row i of the packed
// scratch (field-major: field f's rows start at its component offset times n).
// Without: row i back into particle i.  One scene lookup per particle.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_batch_state(Args<T, D> a, T* __restrict__ scratch,
                                                     const uint32_t* __restrict__ order) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int s = scene_of(a.scene_poff, a.nscenes, i);
  const SceneRec<T, D>& r = a.scenes[s];
  const int64_t li = (order ? int64_t(order[i]) : i) - r.poff;  // the permutation stays within the scene
  int64_t off = 0;
#pragma unroll
  for (int fid = 0; fid < 8; ++fid) {
    const int nc = kFieldComps<D>[fid];
    const DField<T> f = scene_field<T, D>(r, fid);
    if (f.p) {
      T* row = scratch + off * a.n + i * nc;
      if (order) {
        for (int c = 0; c < nc; ++c) row[c] = f(li, c);
      } else {
        for (int c = 0; c < nc; ++c) f(li, c) = row[c];
      }
    }
    off += nc;
  }
}

// Reorders the scenes' particles when the batch's contiguity score is below
// `threshold`, or at every call when threshold >= 1 (reorder_particles!,
// src/Parallel/cpu_block_strategy.jl:222-278).  The score is the fraction of
// neighbours in the layered order whose storage indices are consecutive,
// counted by the previous step's P2G, or on a call's first step by
// k_layered_order; the binning's permutation follows the particles.
template <typename T, int D>
void batch_reorder(tess_ctx* ctx, tess_batch& B, double threshold) {
  const Binning& b = ctx->bin;
  const int64_t n = B.ntot;
  if (n == 0 || b.n_active == 0) return;
  Args<T, D> a = batch_args<T, D>(ctx, B);
  uint32_t* layered = ctx->keys_alt.as<uint32_t>(size_t(n));
  uint32_t* order = ctx->keys.as<uint32_t>(size_t(n));  // the raw keys are dead after the sort
  uint32_t* pout = ctx->inc_mr.as<uint32_t>(size_t(n));
  uint32_t* R = ctx->inc_mks.as<uint32_t>(size_t(n) + 1);
  uint32_t* broken = ctx->reorder_cnt.as<uint32_t>(size_t(b.n_active));
  uint32_t* counters = ctx->counters.as<uint32_t>(CNT_SLOTS);
  const uint32_t* sk = static_cast<const uint32_t*>(ctx->keys_sorted.ptr);
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
  const bool by_p2g_count = threshold < 1.0 && B.counted;
  if (by_p2g_count) {
    // the previous step's P2G counted the broken pairs; the binning's readback
    // brought the count to the host
    if (!(layered_contiguity(int64_t(hc[CNT_MISPLACED]), B.counted_pairs) < threshold)) return;
  }
  const bool count_here = threshold < 1.0 && !B.counted;
  hipLaunchKernelGGL((k_layered_order<T, D>), dim3(unsigned(b.n_active)), dim3(64), 0, ctx->stream, a, sk, pout,
                     layered, count_here ? broken : nullptr);
  TESS_HIP(hipGetLastError());
  if (count_here) {
    sum_u32(ctx, broken, b.n_active, counters + CNT_MISPLACED);
    TESS_HIP(hipMemcpyAsync(hc + CNT_MISPLACED, counters + CNT_MISPLACED, sizeof(uint32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    if (!(layered_contiguity(int64_t(hc[CNT_MISPLACED]), b.n_inside - b.n_active) < threshold)) return;
  }
  TESS_HIP(hipMemsetAsync(R, 0, (size_t(n) + 1) * sizeof(uint32_t), ctx->stream));
  if (n > b.n_inside) {
    hipLaunchKernelGGL(k_mark_outside, dim3(blocks_for(n - b.n_inside, 256)), dim3(256), 0, ctx->stream, a.perm,
                       b.n_inside, n, R);
    TESS_HIP(hipGetLastError());
  }
  exclusive_scan_u32(ctx, R, n + 1, nullptr);
  hipLaunchKernelGGL((k_batch_new_index<T, D>), dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a, layered, R,
                     order, pout);
  TESS_HIP(hipGetLastError());
  T* scratch = ctx->reorder_tmp.as<T>(size_t(n) * kStateWidth<D>);
  hipLaunchKernelGGL((k_batch_state<T, D>), dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a, scratch,
                     static_cast<const uint32_t*>(order));
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_batch_state<T, D>), dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a, scratch,
                     static_cast<const uint32_t*>(nullptr));
  TESS_HIP(hipGetLastError());
  TESS_HIP(hipMemcpyAsync(ctx->vals.ptr, pout, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
  B.reorders += 1;
  if (by_p2g_count) B.reorders_by_count += 1;
}

template <typename T, int D, int BASIS, int NP>
void launch_p2g(tess_ctx* ctx, const Args<T, D>& a, int64_t n_active, const uint32_t* sk, uint32_t* broken) {
  hipLaunchKernelGGL((k_p2g_fast<T, D, BASIS, NP, true, true>), dim3(xcd_grid(n_active)), dim3(64 * NP), 0,
                     ctx->stream, a, sk, broken);
  TESS_HIP(hipGetLastError());
}

template <typename T, int D, int BASIS, int NW>
void launch_g2p(tess_ctx* ctx, const Args<T, D>& a, int64_t n_active, const uint32_t* sk, uint32_t* nk, uint32_t* stay,
                uint32_t* movers) {
  const uint32_t nt = uint32_t(ctx->bin.ntiles);
  hipLaunchKernelGGL((k_g2p_fast<T, D, BASIS, true, NW, true>), dim3(xcd_grid(n_active)), dim3(64 * NW), 0,
                     ctx->stream, a, sk, nk, stay, nt, movers);
  TESS_HIP(hipGetLastError());
}

template <typename T, int D, int BASIS>
void batch_step_impl(tess_ctx* ctx, tess_batch& B, int nsteps) {
  constexpr int SW = support_width(BASIS);
  B.counted = false;  // the context's counters may have served other calls since
  if (B.last_graph) {
    // graph-mode steps wrote node blocks the eager stale list does not name
    B.zeroed = false;
    B.last_graph = false;
  }
  for (int s = 0; s < nsteps; ++s) {
    // from the second step of the call on, the previous G2P wrote the keys
    batch_bin<T, D, BASIS>(ctx, B, s > 0);
    if (B.reorder_threshold > 0.0) {
      const int tok = prof_begin(ctx, "reorder", 0.0);
      batch_reorder<T, D>(ctx, B, B.reorder_threshold);
      prof_end(ctx, tok);
    }
    const Binning& b = ctx->bin;
    Args<T, D> a = batch_args<T, D>(ctx, B);
    // zero what this step does not overwrite: every grid the first time, then
    // the node blocks written last step and not now
    if (!B.zeroed) {
      hipLaunchKernelGGL((k_zero_grid_batch<T, D>), dim3(blocks_for(B.max_nodes, 256), unsigned(B.nscenes)), dim3(256),
                         0, ctx->stream, a);
      TESS_HIP(hipGetLastError());
      B.zeroed = true;
    } else if (B.prev_nb > 0) {
      hipLaunchKernelGGL((k_zero_stale<T, D, true>), dim3(unsigned(B.prev_nb)), dim3(64), 0, ctx->stream, a,
                         static_cast<const uint32_t*>(B.prev_list.ptr), B.prev_nb);
      TESS_HIP(hipGetLastError());
    }
    const uint32_t* sk = static_cast<const uint32_t*>(ctx->keys_sorted.ptr);
    uint32_t* counters = static_cast<uint32_t*>(ctx->counters.ptr);
    if (b.n_active > 0) {
      a.partials = ctx->partials.as<T>(size_t(b.n_active) * Part<D, SW>::SIZE);
      const int tok = prof_begin(ctx, "k_p2g_tile", double(b.n_inside) * sizeof(T) * (2 * D + 1 + D * D));
      // with a threshold in (0, 1) P2G counts the broken layered pairs for the
      // next step's reorder test
      const bool count = B.reorder_threshold > 0.0 && B.reorder_threshold < 1.0;
      uint32_t* broken = count ? ctx->reorder_cnt.as<uint32_t>(size_t(b.n_active)) : nullptr;
      if (D == 3 || (B.variant & 1)) launch_p2g<T, D, BASIS, SW>(ctx, a, b.n_active, sk, broken);
      else launch_p2g<T, D, BASIS, 1>(ctx, a, b.n_active, sk, broken);
      prof_end(ctx, tok);
      if (count) {
        sum_u32(ctx, broken, b.n_active, counters + CNT_MISPLACED);
        B.counted_pairs = b.n_inside - b.n_active;
      }
      B.counted = count;
    } else {
      B.counted = false;
    }
    if (b.n_nodeblocks > 0) {
      uint32_t* per_nb = ctx->per_nodeblock.as<uint32_t>(size_t(b.n_nodeblocks));
      const int tok = prof_begin(ctx, "k_gather_grid", 0.0);
      hipLaunchKernelGGL((k_gather_grid<T, D, SW, true, true>), dim3(xcd_grid(b.n_nodeblocks)), dim3(64), 0, ctx->stream,
                         a, per_nb);
      TESS_HIP(hipGetLastError());
      prof_end(ctx, tok);
      sum_u32(ctx, per_nb, b.n_nodeblocks, counters + CNT_NODES);
      uint32_t* prev = B.prev_list.as<uint32_t>(size_t(b.n_nodeblocks));
      TESS_HIP(hipMemcpyAsync(prev, ctx->nb_list.ptr, size_t(b.n_nodeblocks) * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, ctx->stream));
    }
    B.prev_nb = b.n_nodeblocks;
    // next-step keys at the sorted positions while the call goes on
    uint32_t* nk = nullptr;
    uint32_t* stay = nullptr;
    uint32_t* mv = nullptr;
    if (s + 1 < nsteps && B.ntot > 0) {
      nk = ctx->inc_nk.as<uint32_t>(size_t(B.ntot));
      stay = ctx->inc_spos.as<uint32_t>(size_t(B.ntot) + 1);
      mv = movers_counters(ctx);
      movers_zero(ctx);
    }
    if (b.n_active > 0) {
      const int tok = prof_begin(ctx, "k_g2p_tile",
                                 double(b.n_inside) * sizeof(T) * (2 * D + D * D + 2 * D + 3 * D * D + 1));
      if (B.variant & 4) launch_g2p<T, D, BASIS, 1>(ctx, a, b.n_active, sk, nk, stay, mv);
      else if (B.variant & 2) launch_g2p<T, D, BASIS, 2>(ctx, a, b.n_active, sk, nk, stay, mv);
      else launch_g2p<T, D, BASIS, 4>(ctx, a, b.n_active, sk, nk, stay, mv);
      prof_end(ctx, tok);
    }
    if (B.ntot > b.n_inside) {
      hipLaunchKernelGGL((k_g2p_outside<T, D, true>), dim3(blocks_for(B.ntot - b.n_inside, 256)), dim3(256), 0,
                         ctx->stream, a, sk, nk, stay);
      TESS_HIP(hipGetLastError());
    }
    B.keys_ready = nk != nullptr;
  }
  // the single-scene zeroing bookkeeping no longer describes any grid
  ctx->last_grid_m = nullptr;
  ctx->bin.valid = false;
}

// ---- graph mode ----------------------------------------------------------------
//
// With bit 3 of the variant a batch steps on a private context (its own stream
// and scratch, so no other call can move the buffers a captured step points
// at).  Its step reads nothing back to the host: the binning sorts every
// particle again (every launch size fixed by the batch), the counts of inside
// particles, active tiles and touched node blocks stay on the device, and the
// tile and node-block kernels are persistent launches (a grid sized to the
// device by its occupancy) that walk those lists up to the device-side counts.
// Zeroing is done by kernels.  The step is captured once into a hipGraph and
// replayed; the reorder test, which needs the host, runs between chunks of
// graph_chunk replays.

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) p[i] = v;
}

void fill_u32(hipStream_t st, uint32_t* p, int64_t n, uint32_t v) {
  if (n <= 0) return;
  const unsigned g = unsigned(std::min<int64_t>(blocks_for(n, 256), 4096));
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, st, p, n, v);
  TESS_HIP(hipGetLastError());
}

// The nodes of the node blocks listed by the previous step and not flagged by
// this one, back to zero (the count read on the device).
template <typename T, int D>
__global__ __launch_bounds__(64) void k_zero_stale_persist(Args<T, D> a, const uint32_t* __restrict__ prev_list,
                                                           const uint32_t* __restrict__ prev_count) {
  using G = Geo<D>;
  const int64_t np = int64_t(*prev_count);
  for (int64_t j = blockIdx.x; j < np; j += gridDim.x) {
    const uint32_t B = prev_list[j];
    if (a.nb_flag[B]) continue;
    uint32_t poff;
    const uint32_t Bl = bind_tile<true>(a, B, poff);
    int Bi[D], I[D];
    unlin<D>(Bl, a.ntiles, Bi);
    int r = int(threadIdx.x);
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      I[d] = Bi[d] * G::BW + r % G::BW;
      r /= G::BW;
      in = in && I[d] < a.mesh.dims[d];
    }
    if (in) {
      const int64_t i = grid_slot(a, I);
      if (i >= 0) zero_node(a, i);
    }
  }
}

// This step's node-block list and count become the next step's stale list.
__global__ __launch_bounds__(256) void k_save_list(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                                                   uint32_t* __restrict__ prev_list, uint32_t* __restrict__ prev_count) {
  const int64_t n = int64_t(*count);
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    prev_list[i] = list[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) *prev_count = uint32_t(n);
}

// Particles outside their mesh (sorted positions from the device count of
// inside particles to n): the update with zero grid increments.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_g2p_outside_persist(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                             uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                             const uint32_t* __restrict__ n_inside) {
  const int64_t from = int64_t(*n_inside);
  for (int64_t j = from + int64_t(blockIdx.x) * 256 + threadIdx.x; j < a.n; j += int64_t(gridDim.x) * 256) {
    nk[j] = skeys[j];  // the position does not change: the key stays
    stay[j] = 1u;
    uint32_t pid = a.perm[j];
    pid -= uint32_t(bind_scene(a, scene_of(a.scene_poff, a.nscenes, pid)).poff);
    T dv[D], L[D * D], dx[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dv[c] = T(0);
      dx[c] = T(0);
    }
#pragma unroll
    for (int k = 0; k < D * D; ++k) L[k] = T(0);
    particle_update<T, D>(a, pid, dv, L, dx);
  }
}

int device_cus(int dev) {
  int c = 0;
  TESS_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(c, 1);
}

// Workgroups of a persistent launch: twice what the occupancy query reports
// the CUs hold at once, at most one per work item, a multiple of 8.
unsigned persist_grid(const void* kern, int threads, int64_t max_items, int cus) {
  int per = 0;
  TESS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0));
  int64_t g = 2 * int64_t(std::max(per, 1)) * cus;
  g = std::min<int64_t>(g, int64_t(xcd_grid(std::max<int64_t>(max_items, 1))));
  return unsigned(((g + 7) / 8) * 8);
}

template <typename T, int D, int BASIS, int NP>
const void* p2g_persist_fn() {
  return reinterpret_cast<const void*>(&k_p2g_fast_persist<T, D, BASIS, NP, true, true>);
}
template <typename T, int D, int BASIS, int NW>
const void* g2p_persist_fn() {
  return reinterpret_cast<const void*>(&k_g2p_fast_persist<T, D, BASIS, true, NW, true>);
}

template <typename T, int D, int BASIS>
void batch_graph_grids(tess_ctx* g, tess_batch& B) {
  constexpr int SW = support_width(BASIS);
  const int cus = device_cus(g->device);
  const int64_t nt = B.tiles_total;
  const bool split = D == 3 || (B.variant & 1);
  B.grid_p2g = split ? persist_grid(p2g_persist_fn<T, D, BASIS, SW>(), 64 * SW, nt, cus)
                     : persist_grid(p2g_persist_fn<T, D, BASIS, 1>(), 64, nt, cus);
  if (B.variant & 4) B.grid_g2p = persist_grid(g2p_persist_fn<T, D, BASIS, 1>(), 64, nt, cus);
  else if (B.variant & 2) B.grid_g2p = persist_grid(g2p_persist_fn<T, D, BASIS, 2>(), 128, nt, cus);
  else B.grid_g2p = persist_grid(g2p_persist_fn<T, D, BASIS, 4>(), 256, nt, cus);
  B.grid_gather = persist_grid(reinterpret_cast<const void*>(&k_gather_grid_persist<T, D, SW, true, true>), 64, nt,
                               cus);
  B.grid_zero = persist_grid(reinterpret_cast<const void*>(&k_zero_stale_persist<T, D>), 64, nt, cus);
  B.grid_out = persist_grid(reinterpret_cast<const void*>(&k_g2p_outside_persist<T, D>), 256,
                            ceil_div(std::max<int64_t>(B.ntot, 1), int64_t(256)), cus);
}

__global__ void k_mover_count(const uint32_t* __restrict__ stay_total, int64_t n, uint32_t* __restrict__ movers) {
  if (threadIdx.x == 0) *movers = uint32_t(n - int64_t(*stay_total));
}

// Graph mode's binning double buffer: (sorted keys, permutation) in set 0 =
// (keys_sorted, vals) or set 1 = (skeys2, perm2).
struct BinSets {
  uint32_t* keys[2];
  uint32_t* perm[2];
};
inline BinSets bin_sets(tess_ctx* g, int64_t n) {
  const size_t m = size_t(std::max<int64_t>(n, 1));
  return BinSets{{g->keys_sorted.as<uint32_t>(m), g->skeys2.as<uint32_t>(m)},
                 {g->vals.as<uint32_t>(m), g->perm2.as<uint32_t>(m)}};
}
constexpr int CNT_MOVERS = 12;  // counter slot of the mover count (graph mode)

// One step of the batch on context g, every launch size fixed and every count
// on the device.  full: the binning sorts every particle's key into set 0
// (the step that starts each call); else it merges the particles whose key
// the previous step's G2P changed (written with the stayer flags at the
// sorted positions of set `in`) into set 1-in, the form that is captured.
// Either way the G2P writes the next keys and stayer flags.
template <typename T, int D, int BASIS>
void batch_graph_step(tess_ctx* g, tess_batch& B, bool full, int in) {
  constexpr int SW = support_width(BASIS);
  const int64_t n = B.ntot;
  const uint64_t ntiles = uint64_t(B.tiles_total);
  const size_t m = size_t(std::max<int64_t>(n, 1));
  hipStream_t st = g->stream;
  uint32_t* counters = g->counters.as<uint32_t>(CNT_SLOTS);
  uint32_t* keys = g->keys.as<uint32_t>(m);
  const BinSets sets = bin_sets(g, n);
  uint32_t* nk = g->inc_nk.as<uint32_t>(m);
  uint32_t* spos = g->inc_spos.as<uint32_t>(m + 1);
  uint32_t* mkey = g->inc_mkey.as<uint32_t>(m);
  uint32_t* mj = g->inc_mj.as<uint32_t>(m);
  uint32_t* mks = g->inc_mks.as<uint32_t>(m + 1);  // the reorder uses it with n + 1 words
  uint32_t* mr = g->inc_mr.as<uint32_t>(m);
  uint32_t* mjs = g->inc_mjs.as<uint32_t>(m);
  const int64_t nsb = blocks_for(std::max<int64_t>(n, 1), STAY_BLOCK);
  int64_t* ranges = g->inc_rng.as<int64_t>(size_t(2 * nsb));
  uint32_t* movers = movers_counters(g);  // G2P's mover tally (unused here)
  uint32_t* tstart = g->tile_start.as<uint32_t>(ntiles);
  uint32_t* tend = g->tile_end.as<uint32_t>(ntiles);
  uint32_t* tflag = g->tile_flag.as<uint32_t>(ntiles);
  uint32_t* tslot = g->tile_slot.as<uint32_t>(ntiles);
  uint32_t* active = g->active_tiles.as<uint32_t>(ntiles);
  uint32_t* nbflag = g->nb_flag.as<uint32_t>(ntiles);
  uint32_t* nbslot = g->nb_slot.as<uint32_t>(ntiles);
  uint32_t* nblist = g->nb_list.as<uint32_t>(ntiles);
  uint32_t* per_nb = g->per_nodeblock.as<uint32_t>(ntiles);
  uint32_t* prev_list = B.g_prev_list.as<uint32_t>(ntiles);
  uint32_t* prev_count = B.g_prev_count.as<uint32_t>(64);
  T* partials = g->partials.as<T>(size_t(ntiles) * Part<D, SW>::SIZE);
  // the scans' work area at its largest use (spos, the tile lists, the sort's
  // histograms, the reorder's ranks), so no later call moves it
  g->scan_tmp.get(std::max({scan_work_bytes(n + 1), scan_work_bytes(int64_t(ntiles)),
                            scan_work_bytes(int64_t(sort_hist_words(std::max<int64_t>(n, 1))))}));
  const int out = full ? 0 : 1 - in;
  uint32_t* skeys = sets.keys[out];
  uint32_t* perm = sets.perm[out];
  Args<T, D> a = batch_args<T, D>(g, B);
  a.perm = perm;
  a.partials = partials;
  const int kbits = key_bits_for(ntiles * 64);
  fill_u32(st, counters, CNT_BIN_SLOTS, 0u);
  if (n > 0) {
    if (full) {
      hipLaunchKernelGGL((k_bin_keys<T, D, BASIS, false, true>), dim3(blocks_for(n, 256)), dim3(256), 0, st, a, keys,
                         uint32_t(ntiles), counters);
      TESS_HIP(hipGetLastError());
      radix_sort_pairs(g, keys, nullptr, skeys, perm, n, kbits);
    } else {
      // stayers keep their order; the movers, sorted by (key, previous
      // position), are merged in (the mover count M stays on the device)
      const uint32_t* M = counters + CNT_MOVERS;
      fill_u32(st, spos + n, 1, 0u);
      exclusive_scan_u32(g, spos, n + 1, counters + CNT_STAY);
      hipLaunchKernelGGL(k_mover_count, dim3(1), dim3(64), 0, st, static_cast<const uint32_t*>(counters + CNT_STAY), n,
                         counters + CNT_MOVERS);
      TESS_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_compact_movers, dim3(blocks_for(n, 256)), dim3(256), 0, st, static_cast<const uint32_t*>(nk),
                         static_cast<const uint32_t*>(spos), n, mkey, mj);
      TESS_HIP(hipGetLastError());
      radix_sort_pairs(g, mkey, nullptr, mks, mr, n, kbits, M);
      hipLaunchKernelGGL(k_place_movers, dim3(blocks_for(n, 256)), dim3(256), 0, st, static_cast<const uint32_t*>(mks),
                         static_cast<const uint32_t*>(mr), static_cast<const uint32_t*>(mj),
                         static_cast<const uint32_t*>(sets.keys[in]), static_cast<const uint32_t*>(sets.perm[in]),
                         static_cast<const uint32_t*>(spos), n, n, mjs, skeys, perm, M);
      TESS_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_stay_ranges, dim3(blocks_for(nsb, 256)), dim3(256), 0, st,
                         static_cast<const uint32_t*>(sets.keys[in]), n, static_cast<const uint32_t*>(mks),
                         static_cast<const uint32_t*>(mjs), n, nsb, ranges, M);
      TESS_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_place_stayers, dim3(unsigned(nsb)), dim3(STAY_BLOCK), 0, st,
                         static_cast<const uint32_t*>(sets.keys[in]), static_cast<const uint32_t*>(sets.perm[in]),
                         static_cast<const uint32_t*>(spos), n, static_cast<const uint32_t*>(mks),
                         static_cast<const uint32_t*>(mjs), static_cast<const int64_t*>(ranges), skeys, perm);
      TESS_HIP(hipGetLastError());
    }
  }
  fill_u32(st, tstart, int64_t(ntiles), 0u);
  fill_u32(st, tend, int64_t(ntiles), 0u);
  if (n > 0) {
    hipLaunchKernelGGL(k_tile_runs, dim3(blocks_for(n, 256)), dim3(256), 0, st, skeys, n, uint32_t(ntiles), tstart,
                       tend, counters);
    TESS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_tile_flags, dim3(blocks_for(ntiles, 256)), dim3(256), 0, st, tstart, tend, uint32_t(ntiles),
                     tflag, tslot);
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(g, tslot, int64_t(ntiles), counters + CNT_ACTIVE);
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(ntiles, 256)), dim3(256), 0, st, tflag, tslot, uint32_t(ntiles),
                     active);
  TESS_HIP(hipGetLastError());
  fill_u32(st, nbflag, int64_t(ntiles), 0u);
  hipLaunchKernelGGL((k_nb_mark_batch<T, D>), dim3(blocks_for(ntiles, 256)), dim3(256), 0, st, a, tflag,
                     uint32_t(ntiles), nbflag);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(ntiles, 256)), dim3(256), 0, st, nbflag, nbslot, int64_t(ntiles));
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(g, nbslot, int64_t(ntiles), counters + CNT_NB);
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(ntiles, 256)), dim3(256), 0, st, nbflag, nbslot, uint32_t(ntiles),
                     nblist);
  TESS_HIP(hipGetLastError());
  // grids: the previous step's stale node blocks, P2G, the grid-node half
  hipLaunchKernelGGL((k_zero_stale_persist<T, D>), dim3(B.grid_zero), dim3(64), 0, st, a,
                     static_cast<const uint32_t*>(prev_list), static_cast<const uint32_t*>(prev_count));
  TESS_HIP(hipGetLastError());
  const uint32_t* sk = skeys;
  const uint32_t* n_active = counters + CNT_ACTIVE;
  const uint32_t* n_nb = counters + CNT_NB;
  // bit 4: persistent grids; else one workgroup per possible tile or node block
  const bool persist = (B.variant & 16) != 0;
  const dim3 all_tiles(xcd_grid(int64_t(ntiles)));
  uint32_t* none = nullptr;
  if (D == 3 || (B.variant & 1)) {
    if (persist)
      hipLaunchKernelGGL((k_p2g_fast_persist<T, D, BASIS, SW, true, true>), dim3(B.grid_p2g), dim3(64 * SW), 0, st, a,
                         sk, none, n_active);
    else
      hipLaunchKernelGGL((k_p2g_fast_grid<T, D, BASIS, SW, true, true>), all_tiles, dim3(64 * SW), 0, st, a, sk, none,
                         n_active);
  } else {
    if (persist)
      hipLaunchKernelGGL((k_p2g_fast_persist<T, D, BASIS, 1, true, true>), dim3(B.grid_p2g), dim3(64), 0, st, a, sk,
                         none, n_active);
    else
      hipLaunchKernelGGL((k_p2g_fast_grid<T, D, BASIS, 1, true, true>), all_tiles, dim3(64), 0, st, a, sk, none,
                         n_active);
  }
  TESS_HIP(hipGetLastError());
  if (persist)
    hipLaunchKernelGGL((k_gather_grid_persist<T, D, SW, true, true>), dim3(B.grid_gather), dim3(64), 0, st, a, per_nb,
                       n_nb);
  else
    hipLaunchKernelGGL((k_gather_grid_grid<T, D, SW, true, true>), all_tiles, dim3(64), 0, st, a, per_nb, n_nb);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_save_list, dim3(unsigned(std::min<int64_t>(blocks_for(int64_t(ntiles), 256), 1024))), dim3(256),
                     0, st, static_cast<const uint32_t*>(nblist), n_nb, prev_list, prev_count);
  TESS_HIP(hipGetLastError());
  // G2P over the active tiles (next keys and stayer flags at the sorted
  // positions), then the particles outside every mesh
  const uint32_t nt32 = uint32_t(ntiles);
  const int nw = (B.variant & 4) ? 1 : (B.variant & 2) ? 2 : 4;
  auto g2p = [&](auto nwc) {
    constexpr int NW = decltype(nwc)::value;
    if (persist)
      hipLaunchKernelGGL((k_g2p_fast_persist<T, D, BASIS, true, NW, true>), dim3(B.grid_g2p), dim3(64 * NW), 0, st, a,
                         sk, nk, spos, nt32, movers, n_active);
    else
      hipLaunchKernelGGL((k_g2p_fast_grid<T, D, BASIS, true, NW, true>), all_tiles, dim3(64 * NW), 0, st, a, sk, nk,
                         spos, nt32, movers, n_active);
  };
  if (nw == 1) g2p(std::integral_constant<int, 1>{});
  else if (nw == 2) g2p(std::integral_constant<int, 2>{});
  else g2p(std::integral_constant<int, 4>{});
  TESS_HIP(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL((k_g2p_outside_persist<T, D>), dim3(B.grid_out), dim3(256), 0, st, a, sk, nk, spos,
                       static_cast<const uint32_t*>(counters + CNT_INSIDE));
    TESS_HIP(hipGetLastError());
  }
}

void graph_release(tess_batch& B) {
  for (int p = 0; p < 2; ++p) {
    if (B.graph_exec[p]) (void)hipGraphExecDestroy(B.graph_exec[p]);
    if (B.graph[p]) (void)hipGraphDestroy(B.graph[p]);
    B.graph_exec[p] = nullptr;
    B.graph[p] = nullptr;
  }
  B.captured_variant = -1;
  B.graph_deps.clear();
}

// Every buffer a captured step points at, with its address at the capture.
std::vector<std::pair<const Buffer*, void*>> graph_buffers(tess_ctx* g, tess_batch& B) {
  const Buffer* bufs[] = {&g->counters,  &g->keys,       &g->keys_sorted,  &g->vals,         &g->skeys2,
                          &g->perm2,     &g->keys_alt,   &g->vals_alt,     &g->hist,         &g->scan_tmp,
                          &g->inc_nk,    &g->inc_spos,   &g->inc_mkey,     &g->inc_mj,       &g->inc_mks,
                          &g->inc_mr,    &g->inc_mjs,    &g->inc_rng,      &g->movers,       &g->tile_start,
                          &g->tile_end,  &g->tile_flag,  &g->tile_slot,    &g->active_tiles, &g->nb_flag,
                          &g->nb_slot,   &g->nb_list,    &g->per_nodeblock, &g->partials,    &B.g_prev_list,
                          &B.g_prev_count, &B.recs,      &B.d_toff,        &B.d_poff};
  std::vector<std::pair<const Buffer*, void*>> v;
  for (const Buffer* b : bufs) v.emplace_back(b, b->ptr);
  return v;
}

bool graph_deps_hold(const tess_batch& B) {
  for (const auto& d : B.graph_deps)
    if (d.first->ptr != d.second) return false;
  return !B.graph_deps.empty();
}

// The reorder test between replays, on the binning the last step left in set
// `cur` (its counts read back), exactly as an eager step would run it.
template <typename T, int D>
void graph_reorder_check(tess_ctx* g, tess_batch& B, int cur) {
  uint32_t* hc = reinterpret_cast<uint32_t*>(g->counters_host);
  TESS_HIP(hipMemcpyAsync(hc, g->counters.as<uint32_t>(CNT_SLOTS), CNT_SLOTS * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, g->stream));
  TESS_HIP(hipStreamSynchronize(g->stream));
  Binning& b = g->bin;
  b.n = B.ntot;
  b.ntiles = B.tiles_total;
  b.n_inside = B.ntot > 0 ? int64_t(hc[CNT_INSIDE]) : 0;
  b.n_active = int64_t(hc[CNT_ACTIVE]);
  b.n_nodeblocks = int64_t(hc[CNT_NB]);
  b.valid = true;
  if (cur == 1) {
    std::swap(g->keys_sorted, g->skeys2);
    std::swap(g->vals, g->perm2);
  }
  B.counted = false;
  try {
    batch_reorder<T, D>(g, B, B.reorder_threshold);
  } catch (...) {
    if (cur == 1) {
      std::swap(g->keys_sorted, g->skeys2);
      std::swap(g->vals, g->perm2);
    }
    throw;
  }
  if (cur == 1) {
    std::swap(g->keys_sorted, g->skeys2);
    std::swap(g->vals, g->perm2);
  }
  b.valid = false;
}

template <typename T, int D, int BASIS>
void batch_graph_run(tess_ctx* ctx, tess_batch& B, int nsteps) {
  if (nsteps <= 0) return;
  for (auto& e : B.graph_ev)
    if (!e) {
      hipEvent_t ev = nullptr;
      const hipError_t err = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (err == hipSuccess) e = ev;  // kept before any error is raised
      TESS_HIP(err);
    }
  if (!B.gctx) {
    hipStream_t gs = nullptr;
    const hipError_t err = hipStreamCreateWithFlags(&gs, hipStreamNonBlocking);
    try {
      TESS_HIP(err);
      B.gctx = ctx_new(ctx->device, gs);
    } catch (...) {
      if (gs) (void)hipStreamDestroy(gs);
      throw;
    }
  }
  tess_ctx* g = B.gctx;
  hipStream_t gs = g->stream;
  // the caller's work on its stream comes first
  TESS_HIP(hipEventRecord(B.graph_ev[0], ctx->stream));
  TESS_HIP(hipStreamWaitEvent(gs, B.graph_ev[0], 0));
  uint32_t* prev_count = B.g_prev_count.as<uint32_t>(64);
  B.g_prev_list.as<uint32_t>(size_t(std::max<int64_t>(B.tiles_total, 1)));
  if (!B.last_graph) {
    // eager steps wrote node blocks this mode's stale list does not name
    Args<T, D> a = batch_args<T, D>(g, B);
    hipLaunchKernelGGL((k_zero_grid_batch<T, D>), dim3(blocks_for(B.max_nodes, 256), unsigned(B.nscenes)), dim3(256),
                       0, gs, a);
    TESS_HIP(hipGetLastError());
    fill_u32(gs, prev_count, 1, 0u);
    B.last_graph = true;
  }
  if (B.captured_variant != B.variant) graph_release(B);
  batch_graph_grids<T, D, BASIS>(g, B);
  // each call starts with an ordinary step whose binning sorts every particle
  // (the positions may have changed since the last call); it leaves set 0
  batch_graph_step<T, D, BASIS>(g, B, true, 0);
  int done = 1, cur = 0, since_test = 0;
  while (done < nsteps) {
    if (!graph_deps_hold(B)) {
      // first use, another variant, or a buffer moved: capture both parities
      // (capturing launches nothing)
      graph_release(B);
      TESS_HIP(hipStreamSynchronize(gs));
      for (int p = 0; p < 2; ++p) {
        TESS_HIP_EXACT(hipStreamBeginCapture(gs, hipStreamCaptureModeRelaxed));
        try {
          batch_graph_step<T, D, BASIS>(g, B, false, p);
        } catch (...) {
          hipGraph_t dead = nullptr;
          (void)hipStreamEndCapture(gs, &dead);
          if (dead) (void)hipGraphDestroy(dead);
          graph_release(B);
          throw;
        }
        TESS_HIP_EXACT(hipStreamEndCapture(gs, &B.graph[p]));
        TESS_HIP(hipGraphInstantiate(&B.graph_exec[p], B.graph[p], nullptr, nullptr, 0));
      }
      B.graph_deps = graph_buffers(g, B);
      B.captured_variant = B.variant;
      B.captures += 1;
    }
    if (B.reorder_threshold > 0.0 && B.ntot > 0 && since_test == 0) {
      graph_reorder_check<T, D>(g, B, cur);
      if (!graph_deps_hold(B)) {
        // the test moved a buffer the graphs use: start over from the positions
        batch_graph_step<T, D, BASIS>(g, B, true, 0);
        ++done;
        cur = 0;
        continue;
      }
    }
    TESS_HIP(hipGraphLaunch(B.graph_exec[cur], gs));
    cur = 1 - cur;
    ++done;
    B.graph_replays += 1;
    since_test = (since_test + 1) % std::max(B.graph_chunk, 1);
  }
  TESS_HIP(hipEventRecord(B.graph_ev[1], gs));
  TESS_HIP(hipStreamWaitEvent(ctx->stream, B.graph_ev[1], 0));
  // the error flags of the call's binnings (one readback per call)
  uint32_t* hc = reinterpret_cast<uint32_t*>(g->counters_host);
  TESS_HIP(hipMemcpyAsync(hc, g->counters.as<uint32_t>(CNT_SLOTS), CNT_SLOTS * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, gs));
  TESS_HIP(hipStreamSynchronize(gs));
  if (hc[CNT_ERR] & ERR_GIMP_LENGTH) raise(TESS_ERR_ARGUMENT, "uGIMP requires 0 ≤ pt.l ≤ spacing(mesh)");
  if (TESS_DEBUG && B.ntot > 0) {
    // the last step's binning, in set `cur` (the graphs' merges checked through it)
    const BinSets sets = bin_sets(g, B.ntot);
    debug_check_binning(gs, sets.keys[cur], sets.perm[cur], static_cast<const uint32_t*>(g->tile_start.ptr),
                        static_cast<const uint32_t*>(g->tile_end.ptr),
                        static_cast<const uint32_t*>(g->active_tiles.ptr), B.ntot, B.tiles_total,
                        int64_t(hc[CNT_ACTIVE]), int64_t(hc[CNT_INSIDE]), "batch graph step");
  }
  // the eager path's binning and zeroing bookkeeping no longer hold
  B.keys_ready = false;
  B.binned = false;
  B.counted = false;
  g->bin.valid = false;
  ctx->bin.valid = false;
}

template <typename T, int D, int B_, bool S>
struct BatchStepOp {
  static void run(tess_ctx* ctx, const tess_mesh&, tess_batch& b, int nsteps) {
    if constexpr (S) raise(TESS_ERR_UNSUPPORTED, "batches evaluate the basis inside the transfers");
    else if (b.variant & 8) {
      try {
        batch_graph_run<T, D, B_>(ctx, b, nsteps);
      } catch (...) {
        // a failed call leaves nothing of its own running on the batch's
        // stream: the caller may reuse the scene arrays at once
        if (b.gctx) (void)hipStreamSynchronize(b.gctx->stream);
        throw;
      }
    }
    else batch_step_impl<T, D, B_>(ctx, b, nsteps);
  }
};
template <typename T, int D, int B_, bool S>
struct BatchCreateOp {
  static void run(tess_ctx* ctx, const tess_mesh&, tess_batch& b) {
    if constexpr (!S) batch_create_impl<T, D>(ctx, b);
  }
};

bool julia_fields(const tess_scene& sc, int D) {
  auto pok = [](const tess_field& f, int nc) { return !f.ptr || (f.comp_stride == 1 && f.elem_stride == nc); };
  const tess_particles& P = sc.particles;
  const tess_grid& G = sc.grid;
  return pok(P.x, D) && pok(P.v, D) && pok(P.V, 1) && pok(P.grad_v, D * D) && pok(P.F, D * D) &&
         pok(P.sigma, D * D) && pok(P.l, 1) && pok(P.logE, 1) && pok(G.m, 1) && pok(G.mv, D) && pok(G.f, D) &&
         pok(G.vn, D) && pok(G.v, D);
}

}  // namespace

void run_batch_create(tess_ctx* ctx, int basis, int nscenes, const tess_scene* scenes, tess_batch** out) {
  auto* b = new tess_batch();
  b->seen_epoch = error_epoch();
  try {
    b->basis = basis;
    b->nscenes = nscenes;
    b->dim = scenes[0].mesh.dim;
    b->dtype = scenes[0].mesh.dtype;
    b->scenes.assign(scenes, scenes + nscenes);
    b->toff.assign(size_t(nscenes) + 1, 0);
    b->poff.assign(size_t(nscenes) + 1, 0);
    const int LT = b->dim == 3 ? Geo<3>::LT : Geo<2>::LT;
    for (int s = 0; s < nscenes; ++s) {
      const tess_scene& sc = scenes[s];
      if (sc.mesh.dim != b->dim || sc.mesh.dtype != b->dtype)
        raise(TESS_ERR_ARGUMENT, "all scenes of a batch share the mesh dimension and dtype");
      if (!julia_fields(sc, b->dim))
        raise(TESS_ERR_UNSUPPORTED, "batched scenes use the Julia-native field layout");
      int64_t nt = 1, nn = 1;
      for (int d = 0; d < b->dim; ++d) {
        nt *= ((sc.mesh.dims[d] - 1) >> LT) + 1;
        nn *= sc.mesh.dims[d];
      }
      b->toff[size_t(s) + 1] = b->toff[size_t(s)] + nt;
      b->poff[size_t(s) + 1] = b->poff[size_t(s)] + sc.particles.n;
      b->max_nodes = std::max(b->max_nodes, nn);
    }
    b->ntot = b->poff.back();
    b->tiles_total = b->toff.back();
    if (uint64_t(b->tiles_total) * 64 >= uint64_t(UINT32_MAX) || b->ntot >= int64_t(UINT32_MAX))
      raise(TESS_ERR_ARGUMENT, "the batch exceeds the 32-bit binning range");
    dispatch<BatchCreateOp>(basis, scenes[0].mesh, false, ctx, scenes[0].mesh, *b);
  } catch (...) {
    b->recs.release();
    b->d_toff.release();
    b->d_poff.release();
    delete b;
    throw;
  }
  *out = b;
}

void run_batch_step(tess_ctx* ctx, tess_batch* b, int nsteps) {
  const uint64_t ep = error_epoch();
  if (b->seen_epoch != ep) {
    // a call failed since this batch's last step: its binning, zeroing lists
    // and graph-mode device state may describe work cut short, so the next
    // step zeroes every grid and bins from scratch in either mode
    if (b->gctx) {
      (void)hipStreamSynchronize(b->gctx->stream);
      b->gctx->bin.valid = false;
      b->gctx->bin.state_valid = false;
    }
    b->zeroed = false;
    b->binned = false;
    b->keys_ready = false;
    b->counted = false;
    b->last_graph = false;
    b->seen_epoch = ep;
  }
  dispatch<BatchStepOp>(b->basis, b->scenes[0].mesh, false, ctx, b->scenes[0].mesh, *b, nsteps);
}

void run_batch_set_variant(tess_batch* b, int variant) { b->variant = variant; }

void run_batch_set_reorder(tess_batch* b, double threshold) { b->reorder_threshold = threshold; }
int run_batch_variant(const tess_batch* b) { return b->variant; }

void run_batch_stats(const tess_batch* b, int64_t* out, int n) {
  const int64_t vals[6] = {b->reorders, b->reorders_by_count, b->kept_binnings, b->variant, b->graph_replays,
                           b->captures};
  for (int k = 0; k < n && k < 6; ++k) out[k] = vals[k];
}

void run_batch_destroy(tess_batch* b) {
  if (b->gctx) {
    hipStream_t gs = b->gctx->stream;
    graph_release(*b);
    ctx_free(b->gctx);  // synchronises its stream
    if (gs) (void)hipStreamDestroy(gs);
    b->gctx = nullptr;
  }
  for (hipEvent_t e : b->graph_ev)
    if (e) (void)hipEventDestroy(e);
  b->g_prev_list.release();
  b->g_prev_count.release();
  b->recs.release();
  b->d_toff.release();
  b->d_poff.release();
  b->prev_list.release();
  delete b;
}

}  // namespace tess
