// Kernels and launchers of the README step on gfx950 (shared by the forward
// entry points in transfer.hip and the adjoint in adjoint.hip): tile binning,
// deterministic P2G, fused grid update, G2P.  Reference: README.md:66-87; @P2G lowering
// src/Transfer/p2g.jl:145-178 (hoisting :272-307), grid-node half :377-395;
// @G2P lowering src/Transfer/g2p.jl:110-157; deferred weights
// src/Transfer/runtime.jl:21-24,55-130.
//
// Data flow of one P2G (no global float atomics; every LDS add instruction
// touches each node at most once, so the sums follow a fixed order):
//   k_bin_keys      particle -> tile whose footprint holds its support window
//   radix sort      stable (tile, particle id) order -> perm
//   k_tile_runs / k_tile_flags / scan / k_active_list   tile ranges + slots
//   k_nb_mark / scan / k_nb_list    node blocks touched by some footprint
//   k_p2g_tile      one wave per tile: lanes own the 64 window-start classes of
//                   the tile, accumulate their particles' contributions in
//                   registers and add them into an LDS tile of the
//                   (BW+3)^dim footprint (distinct targets per instruction),
//                   then store the tile to a per-tile partial buffer
//   k_gather_grid   one workgroup per node block: each node sums the <= 2^dim
//                   tile partials covering it in a fixed order, then applies
//                   the grid-node half (Float64 m⁻¹, vⁿ, v, gravity, slip walls)
// G2P: k_g2p_tile stages the footprint's vⁿ, v in LDS and updates the tile's
// particles; k_g2p_outside handles particles outside the mesh (empty window).
// @G2P2G: the G2P kernel, then the P2G kernels at the kept positions; or, on
// request, k_g2p2g_fast: the G2P half and the P2G half's tile partials in one
// sweep over the particles, then k_gather_grid as for P2G.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "tess_correct.hpp"
#include "tess_device.hpp"
#include "tess_internal.hpp"

namespace tess {
namespace {

// Tile geometry: 64 window-start classes per tile in both dimensions.
template <int D>
struct Geo;
template <>
struct Geo<3> {
  static constexpr int LT = 2, BW = 4, TS = 7, NFOOT = 343, NCLASS = 64, GNODES = 64;
  // padded LDS footprint, one array per component: node (x, y, z) at
  // x + PY*y + PZ*z.  For the 64 class lanes (4x4x4) at any stencil offset
  // the 64 addresses fall in distinct banks of a 64-bank LDS (two per bank
  // with 32 banks), which the unpadded pitches 7/49 do not give.
  static constexpr int PY = 12, PZ = 80, PSZ = TS * PZ;
};
template <>
struct Geo<2> {
  static constexpr int LT = 3, BW = 8, TS = 11, NFOOT = 121, NCLASS = 64, GNODES = 64;
  static constexpr int PY = 24, PZ = 0, PSZ = TS * PY;  // 8x8 class lanes, as in 3D
};

// padded LDS index of footprint node (x, y[, z])
template <int D>
__device__ __forceinline__ constexpr int pad_node(int x, int y, int z) {
  return x + Geo<D>::PY * y + Geo<D>::PZ * z;
}

template <int N, int E>
struct Pow;

// A tile's P2G partial in HBM: the nodes its classes reach, E = BW + SW - 1
// per axis from the footprint origin, x fastest, NV values per node
// (m, mv[D], f[D]).
template <int D, int SW>
struct Part {
  static constexpr int E = Geo<D>::BW + SW - 1;
  static constexpr int NV = 1 + 2 * D;
  static constexpr int NODES = Pow<E, D>::v;
  static constexpr int SIZE = NODES * NV;
};

template <int N, int E>
struct Pow {
  static constexpr int v = N * Pow<N, E - 1>::v;
};

template <int N>
struct Pow<N, 0> {
  static constexpr int v = 1;
};

// device counters: binning totals (zeroed per binning), error flags, and the
// misplaced-particle count of the last P2G (zeroed per P2G)
enum {
  CNT_INSIDE = 0, CNT_ACTIVE = 1, CNT_NB = 2, CNT_ERR = 3, CNT_BIN_SLOTS = 4,
  CNT_MISPLACED = 8, CNT_NODES = 9, CNT_STAY = 10, CNT_P2GERR = 11, CNT_SLOTS = 16
};
// G2P counts the particles whose key changed into MOVER_LANES counters, one per
// 128-byte line, chosen by workgroup (no single contended address)
constexpr int MOVER_LANES = 8, MOVER_STRIDE = 32;

inline uint32_t* movers_counters(tess_ctx* ctx) {
  return ctx->movers.as<uint32_t>(size_t(MOVER_LANES * MOVER_STRIDE));
}
// zero them (before a G2P that counts)
inline void movers_zero(tess_ctx* ctx) {
  TESS_HIP(hipMemsetAsync(movers_counters(ctx), 0, MOVER_LANES * MOVER_STRIDE * sizeof(uint32_t), ctx->stream));
}
// enqueue their readback into the pinned slots after the counters
inline void movers_readback(tess_ctx* ctx) {
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host) + 64;
  TESS_HIP(hipMemcpyAsync(hc, movers_counters(ctx), MOVER_LANES * MOVER_STRIDE * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
}
// the sum, after a synchronisation that followed movers_readback
inline int64_t movers_total(tess_ctx* ctx) {
  const uint32_t* hc = reinterpret_cast<const uint32_t*>(ctx->counters_host) + 64;
  int64_t t = 0;
  for (int i = 0; i < MOVER_LANES; ++i) t += hc[i * MOVER_STRIDE];
  return t;
}
enum { ERR_GIMP_LENGTH = 1, ERR_WINDOW = 2, ERR_INACTIVE = 4 };

template <typename T, int D>
struct SceneRec;

template <typename T, int D>
struct Args {
  DMesh<T, D> mesh;
  DField<T> x, v, V, gradv, F, sigma, l, logE;
  DField<T> gm, gmv, gf, gvn, gv;
  const int64_t* window;
  DField<T> w, gw;
  // corrected basis of deferred weights (0: the plain kernel) and its node filter
  int32_t corr;
  const uint8_t* filter;
  const uint32_t* perm;
  const uint32_t* tile_start;
  const uint32_t* tile_end;
  const uint32_t* active;
  const uint32_t* tile_flag;
  const uint32_t* tile_slot;
  const uint32_t* nb_list;
  const uint32_t* nb_flag;
  T* partials;
  // when set, G2P also writes each particle's position before the update here
  // (dense, D per particle): the positions the P2G half of a @G2P2G pass reads
  T* x_keep;
  // SpGrid: block numbering (nullptr for a dense grid), block_size_log2, blocks per axis
  const int32_t* numbering;
  int32_t spL;
  int32_t nblk[D];
  uint32_t* p2g_err;
  // batches of scenes: per-scene records and the first global tile / particle
  // of each scene (nscenes+1 entries); nullptr for a single scene
  const SceneRec<T, D>* scenes;
  const int64_t* scene_toff;
  const int64_t* scene_poff;
  int32_t nscenes;
  int64_t n, n_inside, n_active, n_nb;
  int32_t ntiles[D];
  T mp, dt;
  double lambda, mu, V0;
  T gravity[D];
  int32_t slip;
};

// One scene of a batch: what differs between scenes in Args, plus where the
// scene's tiles and particles start in the batch-wide index spaces.
template <typename T, int D>
struct SceneRec {
  DMesh<T, D> mesh;
  DField<T> x, v, V, gradv, F, sigma, l, logE;
  DField<T> gm, gmv, gf, gvn, gv;
  T mp, dt;
  double lambda, mu, V0;
  T gravity[D];
  int32_t slip;
  int32_t ntiles[D];
  int64_t toff, poff, n, nnodes;
};

// Largest s with off[s] <= g (off ascending, off[0] = 0).
__device__ __forceinline__ int scene_of(const int64_t* __restrict__ off, int ns, int64_t g) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Replace the scene-dependent part of a by scene s of the batch.
template <typename T, int D>
__device__ __forceinline__ const SceneRec<T, D>& bind_scene(Args<T, D>& a, int s) {
  const SceneRec<T, D>& r = a.scenes[s];
  a.mesh = r.mesh;
  a.x = r.x;
  a.v = r.v;
  a.V = r.V;
  a.gradv = r.gradv;
  a.F = r.F;
  a.sigma = r.sigma;
  a.l = r.l;
  a.logE = r.logE;
  a.gm = r.gm;
  a.gmv = r.gmv;
  a.gf = r.gf;
  a.gvn = r.gvn;
  a.gv = r.gv;
  a.mp = r.mp;
  a.dt = r.dt;
  a.lambda = r.lambda;
  a.mu = r.mu;
  a.V0 = r.V0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    a.gravity[d] = r.gravity[d];
    a.ntiles[d] = r.ntiles[d];
  }
  a.slip = r.slip;
  return r;
}

// A tile kernel's tile: in a batch, binds the scene holding global tile gtile
// and returns the scene-local tile, with the scene's first particle in poff.
template <bool BATCH, typename T, int D>
__device__ __forceinline__ uint32_t bind_tile(Args<T, D>& a, uint32_t gtile, uint32_t& poff) {
  poff = 0;
  if constexpr (!BATCH) {
    return gtile;
  } else {
    const SceneRec<T, D>& r = bind_scene(a, scene_of(a.scene_toff, a.nscenes, gtile));
    poff = uint32_t(r.poff);
    return gtile - uint32_t(r.toff);
  }
}

// Workgroup i -> logical index; consecutive logical indices stay on one XCD
// (workgroups are dealt to the 8 XCDs round-robin).
__device__ __forceinline__ int64_t xcd_remap(int64_t i, int64_t N) {
  const int64_t per = (N + 7) >> 3;
  return (i & 7) * per + (i >> 3);
}
inline unsigned xcd_grid(int64_t N) { return unsigned(((N + 7) >> 3) << 3); }

// Workgroup i -> work item of the tile kernels.  One scene: xcd_remap (each
// XCD on a contiguous run of tiles, which share node blocks in its L2).  A
// batch (TESS_BATCH_SPREAD): consecutive items, dealt round-robin to the XCDs,
// since one scene's tiles carry loads that differ from the next scene's by
// orders of magnitude and contiguous runs would leave XCDs with unequal work.
#ifndef TESS_BATCH_SPREAD
#define TESS_BATCH_SPREAD 1
#endif
template <bool BATCH>
__device__ __forceinline__ int64_t item_of(int64_t i, int64_t N) {
  if constexpr (BATCH && TESS_BATCH_SPREAD) return i;
  return xcd_remap(i, N);
}

template <int D>
__device__ __forceinline__ void unlin(int64_t l, const int32_t* dims, int* I) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    I[d] = int(l % dims[d]);
    l /= dims[d];
  }
}

// Element of grid node I in the caller's grid arrays: the column-major node
// index of a dense grid, or on an SpGrid the SpArray storage index
// (k-1)*2^(L*dim) + local of its block's number k (src/Grids/spindices.jl:89,
// 96-102, 133-140), -1 when the block is inactive.
template <typename T, int D>
__device__ __forceinline__ int64_t grid_slot(const Args<T, D>& a, const int* I) {
  if (a.numbering == nullptr) return a.mesh.lin(I);
  const int L = a.spL;
  int64_t b = 0, loc = 0;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    b = b * a.nblk[d] + (I[d] >> L);
    loc = (loc << L) | (I[d] & ((1 << L) - 1));
  }
  const int32_t k = a.numbering[b];
  return k > 0 ? (int64_t(k - 1) << (L * D)) + loc : int64_t(-1);
}

// SpGrid: every node of a class window [lo, lo+sw-1] (clipped to the mesh)
// must lie in an active block, the reference's "inactive SpGrid support node"
// check (src/Transfer/runtime.jl:155-161).  The class window contains the
// window of every particle of the class.
template <typename T, int D, int SW>
__device__ __forceinline__ void check_sparse_support(const Args<T, D>& a, const int* origin, const int* ccoord,
                                                     int ccount) {
  if (a.numbering == nullptr || ccount == 0) return;
  const int L = a.spL;
  int blo[D], ext[D], tot = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int lo = max(origin[d] + ccoord[d], 0);  // a class holding particles has lo >= 0
    const int hi = min(lo + SW - 1, int(a.mesh.dims[d]) - 1);
    blo[d] = lo >> L;
    ext[d] = (hi >> L) - blo[d] + 1;
    tot *= ext[d];
  }
  bool bad = false;
  for (int k = 0; k < tot; ++k) {
    int r = k;
    int64_t b = 0, mul = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      b += int64_t(blo[d] + r % ext[d]) * mul;
      r /= ext[d];
      mul *= a.nblk[d];
    }
    bad = bad || a.numbering[b] == 0;
  }
  if (bad) atomicOr(a.p2g_err, uint32_t(ERR_INACTIVE));
}

// Particle window (0-based, clamped) from its position (deferred weights) or
// from stored weights; false if the window is empty.
template <typename T, int D, int BASIS, bool STORED>
__device__ __forceinline__ bool particle_window(const Args<T, D>& a, int64_t p, T* x, T& l, int* lo, int* hi,
                                                uint32_t* err) {
#pragma unroll
  for (int d = 0; d < D; ++d) x[d] = a.x(p, d);
  l = T(0);
  if constexpr (BASIS == TESS_UGIMP) {
    l = a.l(p) / a.mesh.h;  // _normalized_particle_length, gimp.jl:25
  }
  if constexpr (STORED) {
    bool ok = true, bad = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lo[d] = int(a.window[p * 2 * D + 2 * d]) - 1;
      hi[d] = int(a.window[p * 2 * D + 2 * d + 1]) - 1;
      ok = ok && hi[d] >= lo[d];
      bad = bad || hi[d] - lo[d] >= support_width(BASIS) || lo[d] < 0 || hi[d] > a.mesh.dims[d] - 1;
    }
    if (ok && bad) {
      if (err) atomicOr(err, uint32_t(ERR_WINDOW));
      return false;
    }
    return ok;
  } else {
    T xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = (x[d] - a.mesh.xmin[d]) * a.mesh.h_inv;
    if (!inside_xi<T, D>(xi, a.mesh)) return false;
    if constexpr (BASIS == TESS_UGIMP) {
      if (!(T(0) <= l && l <= T(1)) && err) atomicOr(err, uint32_t(ERR_GIMP_LENGTH));
    }
#pragma unroll
    for (int d = 0; d < D; ++d) axis_window<T, BASIS>(xi[d], l, a.mesh.dims[d], lo[d], hi[d]);
    return true;
  }
}

// key = tile * 64 + class: the tile whose footprint [BW*t-1, BW*t+BW+1] holds
// the particle's window, and the window start relative to the footprint
// origin; ntiles*64 for particles with an empty window.
template <typename T, int D, int BASIS, bool STORED>
__device__ __forceinline__ uint32_t particle_key(const Args<T, D>& a, int64_t p, uint32_t ntiles, uint32_t* err) {
  T x[D], l;
  int lo[D], hi[D];
  if (!particle_window<T, D, BASIS, STORED>(a, p, x, l, lo, hi, err)) return ntiles * 64u;
  uint32_t t = 0, c = 0;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    t = t * uint32_t(a.ntiles[d]) + uint32_t((lo[d] + 1) >> Geo<D>::LT);
    c = c * uint32_t(Geo<D>::BW) + uint32_t((lo[d] + 1) & (Geo<D>::BW - 1));
  }
  return t * 64u + c;
}

// The same key from a position and particle length held in registers
// (deferred weights); equals particle_key of the stored position.
template <typename T, int D, int BASIS>
__device__ __forceinline__ uint32_t key_of_position(const Args<T, D>& a, const T* x, T l, uint32_t ntiles) {
  T xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = (x[d] - a.mesh.xmin[d]) * a.mesh.h_inv;
  if (!inside_xi<T, D>(xi, a.mesh)) return ntiles * 64u;
  uint32_t t = 0, c = 0;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    int lo, hi;
    axis_window<T, BASIS>(xi[d], l, a.mesh.dims[d], lo, hi);
    t = t * uint32_t(a.ntiles[d]) + uint32_t((lo + 1) >> Geo<D>::LT);
    c = c * uint32_t(Geo<D>::BW) + uint32_t((lo + 1) & (Geo<D>::BW - 1));
  }
  return t * 64u + c;
}

template <typename T, int D, int BASIS, bool STORED, bool BATCH = false>
__global__ __launch_bounds__(256) void k_bin_keys(Args<T, D> a, uint32_t* __restrict__ keys, uint32_t ntiles,
                                                  uint32_t* __restrict__ counters) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  if constexpr (BATCH) {  // batch: scene-local key, shifted to the scene's first global tile
    const SceneRec<T, D>& r = bind_scene(a, scene_of(a.scene_poff, a.nscenes, p));
    uint32_t nl = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) nl *= uint32_t(r.ntiles[d]);
    const uint32_t k = particle_key<T, D, BASIS, STORED>(a, p - r.poff, nl, counters + CNT_ERR);
    keys[p] = k == nl * 64u ? ntiles * 64u : k + uint32_t(r.toff) * 64u;
    return;
  }
  keys[p] = particle_key<T, D, BASIS, STORED>(a, p, ntiles, counters + CNT_ERR);
}

// ---- incremental re-binning ---------------------------------------------------
// New keys are taken in the previous sorted order j; the new order is the
// stable sort of that order by the new keys.  Particles whose key is unchanged
// ("stayers") keep their relative order; the few "movers" are sorted among
// themselves and merged.  The full radix sort of (new key, previous order) is
// the fallback when many particles move; both give the same permutation.

template <typename T, int D, int BASIS, bool STORED>
__global__ __launch_bounds__(256) void k_keys_sorted(Args<T, D> a, const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ skeys, uint32_t ntiles,
                                                     uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                     uint32_t* __restrict__ counters) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j > a.n) return;
  if (j == a.n) {
    stay[j] = 0;  // scan slot holding the stayer total
    return;
  }
  const uint32_t k = particle_key<T, D, BASIS, STORED>(a, perm[j], ntiles, counters + CNT_ERR);
  nk[j] = k;
  stay[j] = k == skeys[j] ? 1u : 0u;
}

__global__ void k_compact_movers(const uint32_t* __restrict__ nk, const uint32_t* __restrict__ spos, int64_t n,
                                 uint32_t* __restrict__ mkey, uint32_t* __restrict__ mj) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n || spos[j + 1] != spos[j]) return;
  const int64_t r = j - spos[j];
  mkey[r] = nk[j];
  mj[r] = uint32_t(j);
}

// number of i in [lo, hi) with (k[i], jj[i]) < (K, J) for lexicographically sorted pairs
__device__ __forceinline__ int64_t lower_bound_pair(const uint32_t* k, const uint32_t* jj, int64_t lo, int64_t hi,
                                                    uint32_t K, uint32_t J) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const bool less = k[mid] < K || (k[mid] == K && jj[mid] < J);
    if (less) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int64_t bound_key(const uint32_t* k, int64_t n, uint32_t K, bool upper) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const bool go = upper ? k[mid] <= K : k[mid] < K;
    if (go) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Sorted mover i (key K, previous position j) lands after every stayer with a
// smaller key or an equal key and smaller previous position, and after the
// movers before it.
// (M_dev, when set: the mover count read on the device, at most M.)
__global__ void k_place_movers(const uint32_t* __restrict__ mks, const uint32_t* __restrict__ mr,
                               const uint32_t* __restrict__ mj, const uint32_t* __restrict__ skeys_old,
                               const uint32_t* __restrict__ perm_old, const uint32_t* __restrict__ spos, int64_t n,
                               int64_t M, uint32_t* __restrict__ mjs, uint32_t* __restrict__ skeys_new,
                               uint32_t* __restrict__ perm_new, const uint32_t* __restrict__ M_dev) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (M_dev) M = int64_t(*M_dev);
  if (i >= M) return;
  const uint32_t K = mks[i];
  const int64_t j = mj[mr[i]];
  mjs[i] = uint32_t(j);
  const int64_t lb = bound_key(skeys_old, n, K, false), ub = bound_key(skeys_old, n, K, true);
  const int64_t c = j < lb ? lb : (j > ub ? ub : j);
  const int64_t pos = i + int64_t(spos[c]);
  skeys_new[pos] = K;
  perm_new[pos] = perm_old[j];
}

// Stayer j moves back by the movers removed before it and forward by the
// movers inserted before it (sorted movers with a pair (key, j') < (key, j)).
constexpr int STAY_BLOCK = 256, STAY_CAP = 2048;
// The movers [ranges[2b], ranges[2b+1]) that fall among the stayers of
// k_place_stayers' workgroup b, one thread per workgroup (the binary searches
// run in parallel instead of at the start of every workgroup).
__global__ void k_stay_ranges(const uint32_t* __restrict__ skeys_old, int64_t n, const uint32_t* __restrict__ mks,
                              const uint32_t* __restrict__ mjs, int64_t M, int64_t nblocks,
                              int64_t* __restrict__ ranges, const uint32_t* __restrict__ M_dev) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  if (M_dev) M = int64_t(*M_dev);
  const int64_t j0 = b * STAY_BLOCK;
  const int64_t j1 = j0 + STAY_BLOCK < n ? j0 + STAY_BLOCK : n;
  const int64_t lo = lower_bound_pair(mks, mjs, 0, M, skeys_old[j0], uint32_t(j0));
  ranges[2 * b] = lo;
  ranges[2 * b + 1] = lower_bound_pair(mks, mjs, lo, M, skeys_old[j1 - 1], uint32_t(j1));
}

__global__ __launch_bounds__(STAY_BLOCK) void k_place_stayers(
    const uint32_t* __restrict__ skeys_old, const uint32_t* __restrict__ perm_old, const uint32_t* __restrict__ spos,
    int64_t n, const uint32_t* __restrict__ mks, const uint32_t* __restrict__ mjs, const int64_t* __restrict__ ranges,
    uint32_t* __restrict__ skeys_new, uint32_t* __restrict__ perm_new) {
  __shared__ uint32_t s_k[STAY_CAP], s_j[STAY_CAP];
  const int64_t j0 = int64_t(blockIdx.x) * STAY_BLOCK;
  const int64_t j1 = j0 + STAY_BLOCK < n ? j0 + STAY_BLOCK : n;
  const int64_t a0 = ranges[2 * int64_t(blockIdx.x)], cnt = ranges[2 * int64_t(blockIdx.x) + 1] - a0;
  const bool in_lds = cnt <= STAY_CAP;
  if (in_lds)
    for (int64_t t = threadIdx.x; t < cnt; t += STAY_BLOCK) {
      s_k[t] = mks[a0 + t];
      s_j[t] = mjs[a0 + t];
    }
  __syncthreads();
  const int64_t j = j0 + threadIdx.x;
  if (j >= j1 || spos[j + 1] == spos[j]) return;  // past the end, or a mover
  const uint32_t K = skeys_old[j];
  const int64_t ins = in_lds ? a0 + lower_bound_pair(s_k, s_j, 0, cnt, K, uint32_t(j))
                             : lower_bound_pair(mks, mjs, a0, a0 + cnt, K, uint32_t(j));
  const int64_t pos = int64_t(spos[j]) + ins;
  skeys_new[pos] = K;
  perm_new[pos] = perm_old[j];
}

__global__ void k_tile_runs(const uint32_t* __restrict__ skeys, int64_t n, uint32_t ntiles,
                            uint32_t* __restrict__ tile_start, uint32_t* __restrict__ tile_end,
                            uint32_t* __restrict__ counters) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = skeys[i] >> 6;  // tile
  const bool first = i == 0 || (skeys[i - 1] >> 6) != k;
  const bool last = i == n - 1 || (skeys[i + 1] >> 6) != k;
  if (k < ntiles) {
    if (first) tile_start[k] = uint32_t(i);
    if (last) tile_end[k] = uint32_t(i + 1);
    if (i == n - 1) counters[CNT_INSIDE] = uint32_t(n);
  } else if (first) {
    counters[CNT_INSIDE] = uint32_t(i);
  }
}

__global__ void k_tile_flags(const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_end,
                             uint32_t ntiles, uint32_t* __restrict__ flag, uint32_t* __restrict__ slot) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t f = tile_end[t] > tile_start[t] ? 1u : 0u;
  flag[t] = f;
  slot[t] = f;
}

__global__ void k_compact(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ slot, uint32_t count,
                          uint32_t* __restrict__ list) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count) return;
  if (flag[t]) list[slot[t]] = uint32_t(t);
}

// Node block B holds nodes BW*B .. BW*B+BW-1 per axis; tile t's footprint
// [BW*t-1, BW*t+BW+1] touches node blocks t-1, t, t+1.
template <int D>
__global__ void k_nb_mark(const uint32_t* __restrict__ tile_flag, uint32_t ntiles, int3 nt,
                          uint32_t* __restrict__ nb_flag) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ntiles || !tile_flag[t]) return;
  const int dims[3] = {nt.x, nt.y, nt.z};
  int I[3] = {0, 0, 0};
  int64_t r = t;
  for (int d = 0; d < D; ++d) {
    I[d] = int(r % dims[d]);
    r /= dims[d];
  }
  for (int dz = (D == 3 ? -1 : 0); dz <= (D == 3 ? 1 : 0); ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int J[3] = {I[0] + dx, I[1] + dy, I[2] + dz};
        bool ok = true;
        for (int d = 0; d < D; ++d) ok = ok && J[d] >= 0 && J[d] < dims[d];
        if (!ok) continue;
        int64_t l = J[D - 1];
        for (int d = D - 2; d >= 0; --d) l = l * dims[d] + J[d];
        nb_flag[l] = 1u;
      }
}

// The same for a batch: neighbours are taken in the tile's own scene.
template <typename T, int D>
__global__ void k_nb_mark_batch(Args<T, D> a, const uint32_t* __restrict__ tile_flag, uint32_t ntiles,
                                uint32_t* __restrict__ nb_flag) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ntiles || !tile_flag[t]) return;
  uint32_t poff;
  const uint32_t lt = bind_tile<true>(a, uint32_t(t), poff);
  const uint32_t toff = uint32_t(t) - lt;
  int I[3] = {0, 0, 0};
  int64_t r = lt;
  for (int d = 0; d < D; ++d) {
    I[d] = int(r % a.ntiles[d]);
    r /= a.ntiles[d];
  }
  for (int dz = (D == 3 ? -1 : 0); dz <= (D == 3 ? 1 : 0); ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int J[3] = {I[0] + dx, I[1] + dy, I[2] + dz};
        bool ok = true;
        for (int d = 0; d < D; ++d) ok = ok && J[d] >= 0 && J[d] < a.ntiles[d];
        if (!ok) continue;
        int64_t l = J[D - 1];
        for (int d = D - 2; d >= 0; --d) l = l * a.ntiles[d] + J[d];
        nb_flag[toff + l] = 1u;
      }
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

template <typename T, int D>
__device__ __forceinline__ void zero_node(const Args<T, D>& a, int64_t i) {
  a.gm(i) = T(0);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    a.gmv(i, c) = T(0);
    a.gf(i, c) = T(0);
    a.gvn(i, c) = T(0);
    a.gv(i, c) = T(0);
  }
}

template <typename T, int D>
__global__ __launch_bounds__(256) void k_zero_grid(Args<T, D> a, int64_t nnodes) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nnodes) zero_node(a, i);
}

// Every node of every scene of a batch (blockIdx.y = scene).
template <typename T, int D>
__global__ __launch_bounds__(256) void k_zero_grid_batch(Args<T, D> a) {
  const SceneRec<T, D>& r = bind_scene(a, int(blockIdx.y));
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < r.nnodes) zero_node(a, i);
}

// Zero the nodes of node blocks touched last call but not this call.
template <typename T, int D, bool BATCH = false>
__global__ __launch_bounds__(64) void k_zero_stale(Args<T, D> a, const uint32_t* __restrict__ prev_list,
                                                    int64_t n_prev) {
  using G = Geo<D>;
  const int64_t j = blockIdx.x;
  if (j >= n_prev) return;
  const uint32_t B = prev_list[j];
  if (a.nb_flag[B]) return;
  uint32_t poff;
  const uint32_t Bl = bind_tile<BATCH>(a, B, poff);
  int Bi[D], I[D];
  unlin<D>(Bl, a.ntiles, Bi);
  int r = threadIdx.x;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    I[d] = Bi[d] * G::BW + r % G::BW;
    r /= G::BW;
    if (I[d] >= a.mesh.dims[d]) return;
  }
  const int64_t i = grid_slot(a, I);
  if (i >= 0) zero_node(a, i);
}

template <typename T, int D, int BASIS>
__device__ __forceinline__ void axis_jets(T xd, int lo_local, const T* s_nodex, T l, T hinv, T* val, T* der) {
  constexpr int SW = support_width(BASIS);
#pragma unroll
  for (int a = 0; a < SW; ++a) {
    const T xi = (xd - s_nodex[lo_local + a]) * hinv;  // separable_nodal_basis_jet, basis_weight.jl:49-60
    piecewise_jet<T, BASIS>(xi, l, val[a], der[a]);
  }
}

// Corrected deferred weights of one particle inside a tile kernel: the axis
// jets (nodewise values at the footprint's node coordinates) become the
// corrected ones, or the general fit is prepared (src/Transfer/runtime.jl:55-80
// with the deferred states of wls.jl:131-163, kernelcorrection.jl:41-63).
// ext = window extent - 1 per axis, ccoord = window start in the footprint.
template <typename T, int D, int BASIS>
__device__ __forceinline__ void correct_jets(const Args<T, D>& a, const T* x, const int* lo, const int* ext,
                                             const int* ccoord, const T (*s_nodex)[Geo<D>::TS],
                                             Corrected<T, D, support_width(BASIS)>& cs,
                                             T (&val)[D][support_width(BASIS)], T (&der)[D][support_width(BASIS)]) {
  constexpr int SW = support_width(BASIS);
  int cnt[D];
  T xi[D][SW];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    cnt[d] = ext[d] + 1;
#pragma unroll
    for (int k = 0; k < SW; ++k) xi[d][k] = (x[d] - s_nodex[d][ccoord[d] + k]) * a.mesh.h_inv;
  }
  cs.prepare(a.corr, BASIS, false, a.filter, a.mesh.dims, lo, cnt, xi, val, der);
}

// ---- P2G -------------------------------------------------------------------

__device__ __forceinline__ uint64_t lanemask_lt64() { return (uint64_t(1) << (threadIdx.x & 63)) - 1; }

// Per-class counts of one tile from its slice of the sorted keys (class = low 6
// bits), by wave-level matching; lane c gets count and offset of class c.
// sorted slots per tile whose particle ids the transfer kernels stage in LDS
constexpr int PID_CAP = 1024;

template <int NW = 1>
// With s_pid, also stages the particle ids (perm - poff) of the tile's first
// `cap` sorted slots in LDS, and with s_key their sort keys: the transfer
// kernels then address their particle loads through LDS reads instead of
// global loads that the memory counter would drain behind earlier traffic.
__device__ __forceinline__ void tile_class_counts(const uint32_t* __restrict__ skeys, uint32_t start, int cnt,
                                                  uint32_t* s_cnt, int& ccount, int& coff, int& maxcnt,
                                                  const uint32_t* __restrict__ perm = nullptr, uint32_t poff = 0,
                                                  uint32_t* s_pid = nullptr, uint32_t* s_key = nullptr,
                                                  int cap = 0) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));  // uniform: scalar branches
  // one chunk of 64 sorted slots: stage ids and keys, count the classes
  auto chunk = [&](int i, uint32_t key, uint32_t pid) {
    const bool valid = i < cnt;
    const uint32_t c = key & 63u;
    if (s_pid && valid && i < cap) {
      s_pid[i] = pid - poff;
      if (s_key) s_key[i] = key;
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const bool bit = (c >> b) & 1u;
      const uint64_t m = __ballot(valid && bit);
      peers &= bit ? m : ~m;
    }
    if (valid && (peers & lanemask_lt64()) == 0) atomicAdd(&s_cnt[c], uint32_t(__popcll(peers)));
  };
  // the keys and ids of a wave's first CC chunks are loaded together, before
  // the barrier (it waits only on LDS operations), so one memory round trip
  // serves them (and whatever the caller has in flight)
  constexpr int CC = 4;
  uint32_t kk[CC], pp[CC];
#pragma unroll
  for (int r = 0; r < CC; ++r) {
    const int i = wave * 64 + r * 64 * NW + lane;
    kk[r] = i < cnt ? skeys[start + i] : 0u;
    pp[r] = (s_pid && i < cnt && i < cap) ? perm[start + i] : 0u;
  }
  if (wave == 0) s_cnt[lane] = 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < CC; ++r) {
    const int i = wave * 64 + r * 64 * NW + lane;
    if (wave * 64 + r * 64 * NW < cnt) chunk(i, kk[r], pp[r]);  // uniform
  }
  for (int base = wave * 64 + CC * 64 * NW; base < cnt; base += 64 * NW) {
    const int i = base + lane;
    const bool valid = i < cnt;
    chunk(i, valid ? skeys[start + i] : 0u, (s_pid && valid && i < cap) ? perm[start + i] : 0u);
  }
  __syncthreads();
  ccount = int(s_cnt[lane]);
  int incl = ccount;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(incl, off, 64);
    if (lane >= off) incl += u;
  }
  coff = incl - ccount;
  maxcnt = ccount;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) maxcnt = max(maxcnt, __shfl_xor(maxcnt, off, 64));
}

// Layered-order contiguity of one layer (the estimator of
// block_ordered_particle_contiguity, src/Parallel/cpu_block_strategy.jl:222-246):
// a lane holding a particle (bit set in `has`) is paired with the next lower
// such lane, the lowest one with the last particle of the previous layer
// (`last`, updated here); returns 1 where the pair is not consecutive in
// storage.  Called by every lane of the wave; the tile's first particle has
// no pair.
__device__ __forceinline__ uint32_t broken_pair(uint64_t has, int k, uint32_t pid, uint32_t& last) {
  const uint64_t below = has & lanemask_lt64();
  const uint32_t pl = __shfl(pid, below ? 63 - __clzll(below) : 0, 64);
  const bool mine = (has >> (threadIdx.x & 63)) & 1u;
  const uint32_t r = (mine && (below || k > 0) && pid != (below ? pl : last) + 1u) ? 1u : 0u;
  last = __shfl(pid, 63 - __clzll(has), 64);
  return r;
}

// One wave per tile.  The sort key orders each tile's particles by class
// (window start relative to the footprint origin), so lane c owns class c: it
// walks its particles, accumulates their contributions to the sw^dim nodes of
// the class window in registers, and adds them into the LDS tile.  Lanes hold
// distinct classes, so every LDS add instruction touches each node at most once
// and the per-node order of the adds is the fixed program order.
template <typename T, int D, int BASIS, bool STORED, int NPASS, bool CORR = false>
__global__ __launch_bounds__(64) void k_p2g_tile(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                  uint32_t* __restrict__ misplaced) {
  using G = Geo<D>;
  constexpr int SW = support_width(BASIS);
  constexpr int NV = 1 + 2 * D;
  constexpr int NOFF = Pow<SW, D>::v;
  constexpr int NSL = NOFF / NPASS;  // offsets per pass (NPASS is 1 or SW)
  const int lane = threadIdx.x;
  const int64_t j = xcd_remap(blockIdx.x, a.n_active);
  if (j >= a.n_active) return;
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  int tI[D], origin[D];
  unlin<D>(tile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;

  __shared__ __align__(16) T s_tile[G::NFOOT][8];
  __shared__ T s_nodex[D][G::TS];
  __shared__ uint32_t s_cnt[64];
  for (int q = lane; q < 8 * G::NFOOT; q += 64) (&s_tile[0][0])[q] = T(0);
  for (int q = lane; q < D * G::TS; q += 64) {
    const int d = q / G::TS, k = q % G::TS;
    s_nodex[d][k] = a.mesh.node(d, origin[d] + k);
  }
  int ccount, coff, maxcnt;
  tile_class_counts(skeys, start, cnt, s_cnt, ccount, coff, maxcnt);
  int ccoord[D];
  {
    int r = lane;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ccoord[d] = r % G::BW;
      r /= G::BW;
    }
  }
  check_sparse_support<T, D, SW>(a, origin, ccoord, ccount);
  // neighbours in the layered order (k-th particle of every class, classes
  // ascending, then the next k) whose storage indices are not consecutive
  if (misplaced) {
    uint32_t bad = 0, last = 0;
    for (int k = 0; k < maxcnt; ++k) {
      const uint64_t has = __ballot(k < ccount);
      const uint32_t pid = k < ccount ? a.perm[start + coff + k] : 0u;
      bad += broken_pair(has, k, pid, last);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) bad += __shfl_xor(bad, off, 64);
    if (lane == 0) misplaced[j] = bad;  // per tile, summed later
  }
  const T mp = a.mp, hinv = a.mesh.h_inv;
  for (int pass = 0; pass < NPASS; ++pass) {
    T acc[NSL][NV];
#pragma unroll
    for (int s2 = 0; s2 < NSL; ++s2)
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[s2][v] = T(0);
    for (int k = 0; k < maxcnt; ++k) {
      if (k >= ccount) continue;
      const uint32_t pid = a.perm[start + coff + k];
      T x[D], l;
      int lo[D], hi[D];
      particle_window<T, D, BASIS, STORED>(a, pid, x, l, lo, hi, nullptr);
      int ext[D];
#pragma unroll
      for (int d = 0; d < D; ++d) ext[d] = hi[d] - lo[d];
      // hoisted particle factors (p2g.jl:272-307): mₚ*v[p] and (-V[p])*σ[p]
      T mv[D], A[D * D];
#pragma unroll
      for (int d = 0; d < D; ++d) mv[d] = mp * a.v(pid, d);
      const T negV = -a.V(pid);
#pragma unroll
      for (int q = 0; q < D * D; ++q) A[q] = negV * a.sigma(pid, q);
      T val[D][SW], der[D][SW];
      if constexpr (!STORED) {
#pragma unroll
        for (int d = 0; d < D; ++d) axis_jets<T, D, BASIS>(x[d], ccoord[d], s_nodex[d], l, hinv, val[d], der[d]);
      }
      Corrected<T, D, SW> cs;
      if constexpr (CORR) correct_jets<T, D, BASIS>(a, x, lo, ext, ccoord, s_nodex, cs, val, der);
#pragma unroll
      for (int s2 = 0; s2 < NSL; ++s2) {
        const int o = pass * NSL + s2;
        int aa[D];
        int r = o;
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          aa[d] = r % SW;
          r /= SW;
          ok = ok && aa[d] <= ext[d];
        }
        if (!ok) continue;
        T w, g[D];
        if constexpr (STORED) {
          const int64_t e = int64_t(o) + int64_t(NOFF) * pid;
          w = a.w(e);
#pragma unroll
          for (int c = 0; c < D; ++c) g[c] = a.gw(e, c);
        } else if (CORR && cs.mode == CORR_GENERAL) {
          cs.node(val, aa, filter_pass<D>(a.filter, a.mesh.dims, lo, aa), hinv, w, g);
        } else {
          // w = Π_d N_d (left-assoc), ∇w_c = (N'_c Π_{d≠c} N_d) h⁻¹
          w = val[0][aa[0]];
#pragma unroll
          for (int d = 1; d < D; ++d) w = w * val[d][aa[d]];
#pragma unroll
          for (int c = 0; c < D; ++c) {
            T gg = (c == 0) ? der[0][aa[0]] : val[0][aa[0]];
#pragma unroll
            for (int d = 1; d < D; ++d) gg = gg * ((d == c) ? der[d][aa[d]] : val[d][aa[d]]);
            g[c] = gg * hinv;
          }
        }
        acc[s2][0] += w * mp;
#pragma unroll
        for (int c = 0; c < D; ++c) acc[s2][1 + c] += w * mv[c];
#pragma unroll
        for (int r2 = 0; r2 < D; ++r2) {
          T f = A[r2] * g[0];
#pragma unroll
          for (int c = 1; c < D; ++c) f = f + A[r2 + D * c] * g[c];
          acc[s2][1 + D + r2] += f;
        }
      }
    }
    if (ccount > 0) {
#pragma unroll
      for (int s2 = 0; s2 < NSL; ++s2) {
        int r = pass * NSL + s2, q = 0, mul = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          q += (ccoord[d] + r % SW) * mul;
          r /= SW;
          mul *= G::TS;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) atomicAdd(&s_tile[q][v], acc[s2][v]);
      }
    }
  }
  __syncthreads();
  using PT = Part<D, SW>;
  T* out = a.partials + int64_t(j) * PT::SIZE;
  for (int e = lane; e < PT::SIZE; e += 64) {
    const int n = e / NV, v = e % NV;
    const int x = n % PT::E, y = (n / PT::E) % PT::E, z = D == 3 ? n / (PT::E * PT::E) : 0;
    out[e] = s_tile[x + G::TS * (y + G::TS * z)][v];
  }
}

template <typename T, int D>
__device__ __forceinline__ double det_d(const T* F) {
  if constexpr (D == 2) {
    return double(F[0]) * double(F[3]) - double(F[2]) * double(F[1]);
  } else {
    const double f11 = F[0], f21 = F[1], f31 = F[2], f12 = F[3], f22 = F[4], f32 = F[5], f13 = F[6], f23 = F[7],
                 f33 = F[8];
    return f11 * (f22 * f33 - f23 * f32) - f12 * (f21 * f33 - f23 * f31) + f13 * (f21 * f32 - f22 * f31);
  }
}

// Neo-Hookean Cauchy stress σ = (μ(FFᵀ-I) + λ lnJ I)/J and V = V0 J of the
// updated F (README.md:84-86), stiffness scaled by exp(logE) when has_logE.
// J = det F (cofactor expansion), ln J and 1/J are Float64 for both storage
// types, so a small or large J keeps its relative accuracy.  Float64 storage
// forms FFᵀ - I as written.  Float32 storage forms it from H = F - I as
// H + Hᵀ + HHᵀ in Float32: the difference of FFᵀ and I would lose the digits
// of a small strain, H itself is exact, and nothing in the sum cancels for F
// near I.  σ is written as sig[r + D*b].
template <typename T, int D>
__device__ __forceinline__ void neo_hookean(const T* Fn, double lam0, double mu0, double V0, bool has_logE, T logE,
                                            T* sig, T& Vout) {
  const double J = det_d<T, D>(Fn);
  double lam = lam0, mu = mu0;
  if (has_logE) {
    const double sc = exp(double(logE));
    lam *= sc;
    mu *= sc;
  }
  const double lnJ = log(J);
  const double invJ = 1.0 / J;
  if constexpr (sizeof(T) == 4) {
    const float muJ = float(mu * invJ);
    const float lamJ = float(lam * lnJ * invJ);
    float H[D * D];
#pragma unroll
    for (int b = 0; b < D; ++b)
#pragma unroll
      for (int r = 0; r < D; ++r) H[r + D * b] = Fn[r + D * b] - (r == b ? 1.0f : 0.0f);
#pragma unroll
    for (int b = 0; b < D; ++b)
#pragma unroll
      for (int r = 0; r < D; ++r) {
        float hh = H[r] * H[b];  // (HHᵀ)_rb = Σ_k H_rk H_bk
#pragma unroll
        for (int k = 1; k < D; ++k) hh = fmaf(H[r + D * k], H[b + D * k], hh);
        const float bp = (H[r + D * b] + H[b + D * r]) + hh;
        sig[r + D * b] = r == b ? fmaf(muJ, bp, lamJ) : muJ * bp;
      }
  } else {
#pragma unroll
    for (int b = 0; b < D; ++b)
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double bb = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) bb += double(Fn[r + D * k]) * double(Fn[b + D * k]);
        const double sg = mu * (bb - (r == b ? 1.0 : 0.0)) + (r == b ? lam * lnJ : 0.0);
        sig[r + D * b] = T(sg * invJ);
      }
  }
  Vout = T(V0 * J);
}

// ---- deferred-weight fast paths ------------------------------------------------
//
// Both transfer kernels below evaluate the basis on the fly.  Along an axis
// whose window is not truncated they use the reference's full-support
// polynomial (bspline.jl:90-100,146-170); truncated axes and uGIMP use the
// nodewise piecewise jet (basis_weight.jl:49-60,624-632).  Tensor-product sums
// are factored by axis (products of the y/z jets are formed once per (y,z)
// offset and reused along x); the results equal the reference's up to
// floating-point rounding.

// Loads component c of element e: Julia-native AoS fields (component stride 1,
// element stride NC) get compile-time strides so neighbouring components
// merge into vector loads.
template <typename T, int NC, bool AOS>
__device__ __forceinline__ T fld(const DField<T>& f, int64_t e, int c) {
  if constexpr (AOS) return f.p[e * NC + c];
  else return f(e, c);
}
template <typename T, int NC, bool AOS>
__device__ __forceinline__ T& fldw(const DField<T>& f, int64_t e, int c) {
  if constexpr (AOS) return f.p[e * NC + c];
  else return f(e, c);
}

// Cache policy of the G2P particle stream (x, v, F read; v, ∇v, x, F, σ, V
// written; gigabytes per launch, far beyond the L2 and the Infinity Cache):
// 0 plain accesses, 1 non-temporal stores, 2 non-temporal loads and stores.
#ifndef TESS_G2P_STREAM
#define TESS_G2P_STREAM 1
#endif
template <typename T>
__device__ __forceinline__ void st_g2p(T& dst, T v) {
  if constexpr (TESS_G2P_STREAM >= 1) __builtin_nontemporal_store(v, &dst);
  else dst = v;
}
template <typename T>
__device__ __forceinline__ T ld_g2p(T& src) {
  if constexpr (TESS_G2P_STREAM >= 2) return __builtin_nontemporal_load(&src);
  else return src;
}
// P2G's particle reads (x, v, V, σ): 0 plain, 1 non-temporal
#ifndef TESS_P2G_STREAM
#define TESS_P2G_STREAM 0
#endif
template <typename T>
__device__ __forceinline__ T ld_p2g(T& src) {
  if constexpr (TESS_P2G_STREAM >= 1) return __builtin_nontemporal_load(&src);
  else return src;
}

// 1-D jets at the window nodes lo..lo+SW-1 of one axis; entries past the
// (truncated) window are zero.  Derivatives are returned times h⁻¹.
template <typename T, int BASIS>
__device__ __forceinline__ void axis_jets_mixed(T xd, T xmin, int lo_local, int ext, const T* s_nodex, T l, T hinv,
                                                T* val, T* der) {
  constexpr int SW = support_width(BASIS);
  bool done = false;
  if constexpr (BASIS != TESS_UGIMP) {
    if (ext == SW - 1) {
      const T xi = (xd - xmin) * hinv;
      const T y = (degree_of(BASIS) % 2 == 1) ? xi : xi - T(0.5);
      bspline_local_jet<T, BASIS>(y - floor(y), val, der);
      done = true;
    }
  }
  if (!done) {
#pragma unroll
    for (int q = 0; q < SW; ++q) {
      T v0 = T(0), d0 = T(0);
      if (q <= ext) piecewise_jet<T, BASIS>((xd - s_nodex[lo_local + q]) * hinv, l, v0, d0);
      val[q] = v0;
      der[q] = d0;
    }
  }
#pragma unroll
  for (int q = 0; q < SW; ++q) der[q] *= hinv;
}

// Separable corrected bases in the fast kernels: KernelCorrection or WLS with
// Polynomial(MultiLinear()) over a quadratic or cubic B-spline and no filter
// (the reference's SeparableKernelCorrection and the MultiLinear WLS,
// kernelcorrection.jl:53-63, wls.jl:79-99, 166-189).  Every particle is either
// the plain kernel or a product of per-axis 1-D fits, so the fast kernels only
// swap the axis jets.
__host__ __device__ constexpr bool separable_correction(int corr, int basis, bool has_filter) {
  return (corr & TESS_POLY_MULTILINEAR) && (basis == TESS_BSPLINE_QUADRATIC || basis == TESS_BSPLINE_CUBIC) &&
         !has_filter;
}

// Whether a particle's axis jets are corrected: always for WLS, off a whole
// support for KernelCorrection (kernelcorrection.jl:55-63).
template <typename T, int D, int BASIS>
__device__ __forceinline__ bool sep_corrects(const Args<T, D>& a, const T* x) {
  if (!(a.corr & TESS_KERNEL_CORRECTION)) return true;
  bool full = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int lo, hi;
    axis_window<T, BASIS>((x[d] - a.mesh.xmin[d]) * a.mesh.h_inv, T(0), a.mesh.dims[d], lo, hi);
    full = full && hi - lo == support_width(BASIS) - 1;
  }
  return !full;
}

// Axis jets of a separable corrected basis: the plain ones, or the 1-D fit of
// the nodewise values (der in world units, as axis_jets_mixed returns it).
template <typename T, int BASIS>
__device__ __forceinline__ void axis_jets_sep(T xd, T xmin, int lo_local, int ext, const T* s_nodex, T hinv,
                                              bool correct, T* val, T* der) {
  constexpr int SW = support_width(BASIS);
  if (!correct) {
    axis_jets_mixed<T, BASIS>(xd, xmin, lo_local, ext, s_nodex, T(0), hinv, val, der);
    return;
  }
  T dh[SW];
#pragma unroll
  for (int q = 0; q < SW; ++q) {
    const T xi = (xd - s_nodex[lo_local + q]) * hinv;
    T v0 = T(0), d0 = T(0);
    if (q <= ext) piecewise_jet<T, BASIS>(xi, T(0), v0, d0);
    val[q] = v0;
    dh[q] = -xi;
  }
  wls_axis<T, SW>(ext + 1, degree_of(BASIS), dh, val, der);
#pragma unroll
  for (int q = 0; q < SW; ++q) der[q] *= hinv;
}

// One particle's P2G inputs, loaded one iteration ahead of their use.
template <typename T, int D>
struct P2GIn {
  uint32_t pid;
  T x[D], v[D], V, S[D * D], l;
};

template <typename T, int D, int BASIS, bool AOS>
__device__ __forceinline__ void p2g_load(const Args<T, D>& a, uint32_t pid, P2GIn<T, D>& in) {
  in.pid = pid;
#pragma unroll
  for (int d = 0; d < D; ++d) in.x[d] = ld_p2g(fldw<T, D, AOS>(a.x, pid, d));
#pragma unroll
  for (int d = 0; d < D; ++d) in.v[d] = ld_p2g(fldw<T, D, AOS>(a.v, pid, d));
  in.V = ld_p2g(a.V(pid));
#pragma unroll
  for (int q = 0; q < D * D; ++q) in.S[q] = ld_p2g(fldw<T, D * D, AOS>(a.sigma, pid, q));
  in.l = T(0);
  if constexpr (BASIS == TESS_UGIMP) in.l = a.l(pid) / a.mesh.h;
}

// One layer of class lanes into the stencil offsets a wave owns, in
// registers: the KN offsets along the last axis from vl/dl (that axis's jets
// there), all SW along the others from val/der (axes 0..D-2).  mpe is the
// particle's mass (zero on a lane without a particle, whose jets are finite),
// mpv = mpe v and A = -V σ (column-major).  Offset o = ix + SW*(jy + SW*kz).
template <typename T, int D, int SW, int KN>
__device__ __forceinline__ void p2g_accumulate(const T (*val)[SW], const T (*der)[SW], const T* vl, const T* dl,
                                               const T mpe, const T* mpv, const T* A, T* am,
                                               T (*amv)[Pow<SW, D - 1>::v * KN], T (*af)[Pow<SW, D - 1>::v * KN]) {
  if constexpr (D == 3) {
#pragma unroll
    for (int kz = 0; kz < KN; ++kz) {
#pragma unroll
      for (int jy = 0; jy < SW; ++jy) {
        const T YZ = val[1][jy] * vl[kz];
        const T YdZ = der[1][jy] * vl[kz];
        const T YZd = val[1][jy] * dl[kz];
        const T cm = YZ * mpe;
        T cmv[3], U[3], Wv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          cmv[c] = YZ * mpv[c];
          U[c] = A[c] * YZ;
          Wv[c] = A[c + 3] * YdZ + A[c + 6] * YZd;
        }
#pragma unroll
        for (int ix = 0; ix < SW; ++ix) {
          const int o = ix + SW * (jy + SW * kz);
          am[o] += val[0][ix] * cm;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            amv[c][o] += val[0][ix] * cmv[c];
            af[c][o] = fma(val[0][ix], Wv[c], fma(der[0][ix], U[c], af[c][o]));
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int jy = 0; jy < KN; ++jy) {
      const T Y = vl[jy], Yd = dl[jy];
      const T cm = Y * mpe;
      T cmv[2], U[2], Wv[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        cmv[c] = Y * mpv[c];
        U[c] = A[c] * Y;
        Wv[c] = A[c + 2] * Yd;
      }
#pragma unroll
      for (int ix = 0; ix < SW; ++ix) {
        const int o = ix + SW * jy;
        am[o] += val[0][ix] * cm;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          amv[c][o] += val[0][ix] * cmv[c];
          af[c][o] = fma(val[0][ix], Wv[c], fma(der[0][ix], U[c], af[c][o]));
        }
      }
    }
  }
}

// Adds one wave's stencil (offsets pass*KN .. pass*KN+KN-1 along the last
// axis) into the tile's node sums in LDS.  Lanes are distinct classes, so for
// one offset their target nodes are distinct; offsets are applied one after
// another (read, add, write: the LDS executes one wave's accesses in order).
template <typename T, int D, int SW, int KN>
__device__ __forceinline__ void p2g_stencil_to_lds(T (*s_tile)[Geo<D>::PSZ], const int* ccoord, int pass,
                                                   const T* am, const T (*amv)[Pow<SW, D - 1>::v * KN],
                                                   const T (*af)[Pow<SW, D - 1>::v * KN]) {
  constexpr int NV = 1 + 2 * D;
  constexpr int NSL = Pow<SW, D - 1>::v * KN;
#pragma unroll
  for (int o = 0; o < NSL; ++o) {
    int r = o, nI[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int off = (d == D - 1) ? pass * KN + r : r % SW;
      nI[d] = ccoord[d] + off;
      r /= SW;
    }
    const int q = pad_node<D>(nI[0], nI[1], nI[2]);
    T t[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) t[v] = s_tile[v][q];
    t[0] += am[o];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      t[1 + c] += amv[c][o];
      t[1 + D + c] += af[c][o];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) s_tile[v][q] = t[v];
    __asm__ volatile("" ::: "memory");  // keep the offsets in program order
  }
}

// The tile's compact partial (Part) from its node sums in LDS: the NT threads
// of the workgroup write one node's NV values each.
template <typename T, int D, int SW, int NT>
__device__ __forceinline__ void p2g_write_partial(const Args<T, D>& a, const int64_t j, T (*s_tile)[Geo<D>::PSZ]) {
  using PT = Part<D, SW>;
  constexpr int NV = 1 + 2 * D;
  T* out = a.partials + int64_t(j) * PT::SIZE;
  for (int n = int(threadIdx.x); n < PT::NODES; n += NT) {
    const int q = pad_node<D>(n % PT::E, (n / PT::E) % PT::E, D == 3 ? n / (PT::E * PT::E) : 0);
    T* o = out + int64_t(n) * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) o[v] = s_tile[v][q];
  }
}

// With several waves per tile, each owning KN offsets along the last axis,
// wave `pass` adds its stencil into a slab of its own: the PL = KN + BW - 1
// planes (last-axis node coordinates pass*KN ..) that its offsets reach, in
// the padded node order of Geo, one array of PL planes per component.  The
// waves fill their slabs at the same time; the partial is then the sum of the
// slabs in wave order (fixed per-node order, so the sums are deterministic).
template <typename T, int D, int SW, int KN>
struct Slab {
  static constexpr int PL = KN + Geo<D>::BW - 1;
  static constexpr int PITCH = D == 3 ? Geo<D>::PZ : Geo<D>::PY;  // words per plane
  static constexpr int VSTRIDE = PL * PITCH;                     // words per component
  static constexpr int SIZE = (1 + 2 * D) * VSTRIDE;             // words per wave
};

template <typename T, int D, int SW, int KN>
__device__ __forceinline__ void p2g_stencil_to_slab(T* slab, const int* ccoord, int pass, const T* am,
                                                    const T (*amv)[Pow<SW, D - 1>::v * KN],
                                                    const T (*af)[Pow<SW, D - 1>::v * KN]) {
  using SL = Slab<T, D, SW, KN>;
  constexpr int NV = 1 + 2 * D;
  constexpr int NSL = Pow<SW, D - 1>::v * KN;
#pragma unroll
  for (int o = 0; o < NSL; ++o) {
    int r = o, nI[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      nI[d] = ccoord[d] + ((d == D - 1) ? r : r % SW);  // the last axis relative to the slab's first plane
      r /= SW;
    }
    const int q = pad_node<D>(nI[0], nI[1], nI[2]);
    T t[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) t[v] = slab[v * SL::VSTRIDE + q];
    t[0] += am[o];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      t[1 + c] += amv[c][o];
      t[1 + D + c] += af[c][o];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) slab[v * SL::VSTRIDE + q] = t[v];
    __asm__ volatile("" ::: "memory");  // keep the offsets in program order
  }
}

// The tile's compact partial from the NPASS waves' slabs, summed in wave order;
// NT threads, one node each.
template <typename T, int D, int SW, int KN, int NPASS, int NT>
__device__ __forceinline__ void p2g_write_partial_slabs(const Args<T, D>& a, const int64_t j, const T* slabs) {
  using PT = Part<D, SW>;
  using SL = Slab<T, D, SW, KN>;
  constexpr int NV = 1 + 2 * D;
  T* out = a.partials + int64_t(j) * PT::SIZE;
  for (int n = int(threadIdx.x); n < PT::NODES; n += NT) {
    const int x = n % PT::E, y = (n / PT::E) % PT::E, z = D == 3 ? n / (PT::E * PT::E) : 0;
    const int last = D == 3 ? z : y;
    T sum[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) sum[v] = T(0);
#pragma unroll
    for (int w = 0; w < NPASS; ++w) {
      const int pl = last - w * KN;
      if (pl >= 0 && pl < SL::PL) {
        const int q = D == 3 ? pad_node<3>(x, y, pl) : pad_node<2>(x, pl, 0);
#pragma unroll
        for (int v = 0; v < NV; ++v) sum[v] += slabs[w * SL::SIZE + v * SL::VSTRIDE + q];
      }
    }
    T* o = out + int64_t(n) * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) o[v] = sum[v];
  }
}

// NPASS waves per tile: wave w accumulates the offsets w*KN .. w*KN+KN-1
// along the last axis for every class (the particles are read by each wave;
// the registers of one wave hold 1/NPASS of the stencil, which keeps several
// waves per SIMD resident).  The waves add their stencils into the LDS tile
// one after another, so the per-node order of the additions is fixed.
template <typename T, int D, int BASIS, int NPASS, bool AOS, bool BATCH, bool SEP>
__device__ __forceinline__ void k_p2g_fast_body(Args<T, D>& a, const uint32_t* __restrict__ skeys,
                                                          uint32_t* __restrict__ misplaced, const int64_t j) {
  using G = Geo<D>;
  using PT = Part<D, support_width(BASIS)>;
  constexpr int SW = support_width(BASIS);
  constexpr int NV = 1 + 2 * D;
  constexpr int KN = SW / NPASS;                       // last-axis offsets per pass
  constexpr int NSL = Pow<SW, D - 1>::v * KN;          // offsets per pass
  constexpr int NT = 64 * NPASS;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));  // uniform: scalar branches
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  uint32_t poff;  // a batch's scene: local tile, first particle
  const uint32_t ltile = bind_tile<BATCH>(a, tile, poff);
  int tI[D], origin[D];
  unlin<D>(ltile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;

  // The tile's node sums m, mv[D], f[D].  One wave: one array per component
  // over the padded footprint (bank-conflict-free for the class lanes).
  // Several: one slab per wave (Slab), in LDS that the per-layer jet exchange
  // and the staged particle ids use before.
  constexpr bool SHJ = NPASS > 1;
  using SL = Slab<T, D, SW, KN>;
  constexpr int JET_WORDS = SHJ ? 2 * D * 2 * SW * 64 : 0;  // [layer parity][axis][val SW, der SW][lane]
  constexpr int PID_WORDS = PID_CAP * int(sizeof(uint32_t)) / int(sizeof(T));
  constexpr int SUM_WORDS = SHJ ? NPASS * SL::SIZE : NV * G::PSZ;
  constexpr int RAW_WORDS = SHJ ? (SUM_WORDS > JET_WORDS + PID_WORDS ? SUM_WORDS : JET_WORDS + PID_WORDS)
                                : SUM_WORDS + JET_WORDS + PID_WORDS;
  __shared__ __align__(16) T s_raw[RAW_WORDS];
  T* const s_sums = s_raw;
  T(*s_tile)[G::PSZ] = reinterpret_cast<T(*)[G::PSZ]>(s_raw);  // one wave
  T(*s_jet)[D][2 * SW][64] = reinterpret_cast<T(*)[D][2 * SW][64]>(s_raw + (SHJ ? 0 : SUM_WORDS));
  uint32_t* const s_pid = reinterpret_cast<uint32_t*>(s_raw + (SHJ ? JET_WORDS : SUM_WORDS + JET_WORDS));
  __shared__ T s_nodex[D][G::TS];
  __shared__ uint32_t s_cnt[64];
  static_assert(NV * G::PSZ * sizeof(T) % 16 == 0, "LDS tile in 16-byte words");
  static_assert(JET_WORDS * sizeof(T) % 16 == 0 && SUM_WORDS * sizeof(T) % 16 == 0, "16-byte LDS sections");
  if constexpr (!SHJ) {
    for (int q = int(threadIdx.x); q < int(NV * G::PSZ * sizeof(T) / 16); q += NT)
      reinterpret_cast<uint4*>(&s_tile[0][0])[q] = make_uint4(0u, 0u, 0u, 0u);
  }
  for (int q = int(threadIdx.x); q < D * G::TS; q += NT) {
    const int d = q / G::TS, k = q % G::TS;
    s_nodex[d][k] = a.mesh.node(d, origin[d] + k);
  }
  int ccount, coff, maxcnt;
  tile_class_counts<NPASS>(skeys, start, cnt, s_cnt, ccount, coff, maxcnt, a.perm, poff, s_pid, nullptr,
                           PID_CAP);  // barrier after the staging
  int ccoord[D];
  {
    int r = lane;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ccoord[d] = r % G::BW;
      r /= G::BW;
    }
  }
  if (wave == 0) check_sparse_support<T, D, SW>(a, origin, ccoord, ccount);
  const T mp = a.mp, hinv = a.mesh.h_inv;
  uint32_t bad = 0, last = 0;
  {
    const int pass = wave;
    T am[NSL], amv[D][NSL], af[D][NSL];
#pragma unroll
    for (int o = 0; o < NSL; ++o) {
      am[o] = T(0);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        amv[c][o] = T(0);
        af[c][o] = T(0);
      }
    }
    // two register sets that alternate between the layer being accumulated and
    // the one being loaded.  Lanes past their class's count compute on the
    // finite data they hold (zeros, or an earlier particle) with zero mass and
    // stress factors, so their contributions are exact zeros.
    P2GIn<T, D> pa, pb;
    std::memset(static_cast<void*>(&pa), 0, sizeof(pa));
    std::memset(static_cast<void*>(&pb), 0, sizeof(pb));
    // particle ids come from LDS (staged with the class counts), a layer ahead
    // of the loads they address
    auto pid_of = [&](int k, auto staged) {
      const int li = coff + k;
      if (k >= ccount) return 0u;
      if constexpr (decltype(staged)::value)
        return s_pid[li];
      else
        return a.perm[start + li] - poff;
    };
    // one layer: the accumulation runs on every lane (a branch around it made
    // the compiler copy all accumulators at each iteration)
    auto layer = [&](int k, const P2GIn<T, D>& cur, P2GIn<T, D>& nxt, uint32_t q_next, uint32_t& q_after,
                     auto staged) {
      q_after = pid_of(k + 2, staged);
      if (k + 1 < ccount) p2g_load<T, D, BASIS, AOS>(a, q_next, nxt);
      const uint64_t has = __ballot(k < ccount);
      const bool act = k < ccount;
      if (misplaced && pass == 0) bad += broken_pair(has, k, act ? cur.pid : 0u, last);
      T x[D], mpv[D], A[D * D];
      const T mpe = act ? mp : T(0);
#pragma unroll
      for (int d = 0; d < D; ++d) x[d] = cur.x[d];
#pragma unroll
      for (int d = 0; d < D; ++d) mpv[d] = mpe * cur.v[d];  // hoisted mₚ*v[p]
      const T negV = act ? -cur.V : T(0);
#pragma unroll
      for (int q = 0; q < D * D; ++q) A[q] = negV * cur.S[q];  // hoisted (-V[p])*σ[p]
      const T l = cur.l;
      // jets of axes 0..D-2 at all SW offsets, of the last axis at this
      // wave's KN offsets (vl, dl)
      T val[D][SW], der[D][SW], vl[KN], dl[KN];
      auto axis_jets = [&](auto dc, T* jv, T* jd) {
        constexpr int d = decltype(dc)::value;
        const T xi = (x[d] - a.mesh.xmin[d]) * hinv;
        int lo, hi;
        axis_window<T, BASIS>(xi, l, a.mesh.dims[d], lo, hi);
        if constexpr (SEP)
          // inactive lanes carry zero mass and stress: they keep the plain jets, whose values are finite
          // whatever their position (a fit over a window that is not theirs can be singular)
          axis_jets_sep<T, BASIS>(x[d], a.mesh.xmin[d], ccoord[d], hi - lo, s_nodex[d], hinv,
                                  act && sep_corrects<T, D, BASIS>(a, x), jv, jd);
        else
          axis_jets_mixed<T, BASIS>(x[d], a.mesh.xmin[d], ccoord[d], hi - lo, s_nodex[d], l, hinv, jv, jd);
      };
      if constexpr (NPASS == 1) {
        axis_jets(std::integral_constant<int, 0>{}, val[0], der[0]);
        axis_jets(std::integral_constant<int, 1>{}, val[1], der[1]);
        if constexpr (D == 3) axis_jets(std::integral_constant<int, D - 1>{}, val[D - 1], der[D - 1]);
#pragma unroll
        for (int q = 0; q < KN; ++q) {
          vl[q] = val[D - 1][q];
          dl[q] = der[D - 1][q];
        }
      } else {
        // the waves hold the same particles: wave d % NPASS evaluates axis d
        // and shares it through LDS (double-buffered by layer, one barrier)
        const int buf = k & 1;
        auto share = [&](auto dc) {
          T jv[SW], jd[SW];
          axis_jets(dc, jv, jd);
#pragma unroll
          for (int q = 0; q < SW; ++q) {
            s_jet[buf][decltype(dc)::value][q][lane] = jv[q];
            s_jet[buf][decltype(dc)::value][SW + q][lane] = jd[q];
          }
        };
        if (wave == 0 % NPASS) share(std::integral_constant<int, 0>{});
        if (wave == 1 % NPASS) share(std::integral_constant<int, 1>{});
        if constexpr (D == 3) {
          if (wave == 2 % NPASS) share(std::integral_constant<int, D - 1>{});
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < D - 1; ++d)
#pragma unroll
          for (int q = 0; q < SW; ++q) {
            val[d][q] = s_jet[buf][d][q][lane];
            der[d][q] = s_jet[buf][d][SW + q][lane];
          }
#pragma unroll
        for (int q = 0; q < KN; ++q) {
          vl[q] = s_jet[buf][D - 1][pass * KN + q][lane];
          dl[q] = s_jet[buf][D - 1][SW + pass * KN + q][lane];
        }
      }
      p2g_accumulate<T, D, SW, KN>(val, der, vl, dl, mpe, mpv, A, am, amv, af);
    };
    auto walk = [&](auto staged) {
      uint32_t qa = pid_of(1, staged), qb = 0u;
      if (ccount > 0) p2g_load<T, D, BASIS, AOS>(a, pid_of(0, staged), pa);
      for (int k = 0; k < maxcnt; k += 2) {
        layer(k, pa, pb, qa, qb, staged);
        if (k + 1 < maxcnt) layer(k + 1, pb, pa, qb, qa, staged);
      }
    };
    if (cnt <= PID_CAP)
      walk(std::true_type{});
    else
      walk(std::false_type{});
    if constexpr (SHJ) {
      // the slabs take the LDS of the jet exchange and the ids: after every
      // wave's last layer.  Each wave zeroes and fills its own slab (the LDS
      // executes one wave's accesses in order)
      __syncthreads();
      T* const slab = s_sums + pass * SL::SIZE;
      static_assert(SL::SIZE * sizeof(T) % 16 == 0, "slabs in 16-byte words");
      for (int q = lane; q < int(SL::SIZE * sizeof(T) / 16); q += 64)
        reinterpret_cast<uint4*>(slab)[q] = make_uint4(0u, 0u, 0u, 0u);
      if (ccount > 0) p2g_stencil_to_slab<T, D, SW, KN>(slab, ccoord, pass, am, amv, af);
      __syncthreads();
    } else {
      // lanes are distinct classes, so for one offset their target nodes are
      // distinct; offsets are applied one after another (read, add, write: the
      // LDS executes one wave's accesses in order)
      if (ccount > 0) p2g_stencil_to_lds<T, D, SW, KN>(s_tile, ccoord, pass, am, amv, af);
      __syncthreads();
    }
  }
  if (misplaced && wave == 0) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) bad += __shfl_xor(bad, off, 64);
    if (lane == 0) misplaced[j] = bad;  // per tile, summed later
  }
  if constexpr (SHJ)
    p2g_write_partial_slabs<T, D, SW, KN, NPASS, NT>(a, j, s_sums);
  else
    p2g_write_partial<T, D, SW, NT>(a, j, s_tile);
}

template <typename T, int D, int BASIS, int NPASS, bool AOS, bool BATCH = false, bool SEP = false>
__global__ __launch_bounds__(64 * NPASS) void k_p2g_fast(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                          uint32_t* __restrict__ misplaced) {
  const int64_t j = item_of<BATCH>(blockIdx.x, a.n_active);
  if (j < a.n_active) k_p2g_fast_body<T, D, BASIS, NPASS, AOS, BATCH, SEP>(a, skeys, misplaced, j);
}

// Forms whose item count is read on the device at `count` (launch sizes fixed
// in advance): one workgroup per possible item, those past the count ending
// at once (_grid), or a persistent grid sized to the device walking the items
// (_persist; a multiple of 8 workgroups keeps each XCD on a contiguous run of
// them, as xcd_remap does for one workgroup per item).
template <typename T, int D, int BASIS, int NPASS, bool AOS, bool BATCH = false, bool SEP = false>
__global__ __launch_bounds__(64 * NPASS) void k_p2g_fast_grid(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                          uint32_t* __restrict__ misplaced, const uint32_t* __restrict__ count) {
  const int64_t N = int64_t(*count), R = ((N + 7) >> 3) << 3;
  // xcd_remap is a bijection of [0, R) only: workgroups past R end here
  const int64_t j = item_of<BATCH>(blockIdx.x, N);
  if (int64_t(blockIdx.x) < R && j < N) k_p2g_fast_body<T, D, BASIS, NPASS, AOS, BATCH, SEP>(a, skeys, misplaced, j);
}
template <typename T, int D, int BASIS, int NPASS, bool AOS, bool BATCH = false, bool SEP = false>
__global__ __launch_bounds__(64 * NPASS) void k_p2g_fast_persist(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                          uint32_t* __restrict__ misplaced, const uint32_t* __restrict__ count) {
  const int64_t N = int64_t(*count), R = ((N + 7) >> 3) << 3;
  for (int64_t it = blockIdx.x; it < R; it += gridDim.x) {
    const int64_t j = item_of<BATCH>(it, N);
    if (j < N) k_p2g_fast_body<T, D, BASIS, NPASS, AOS, BATCH, SEP>(a, skeys, misplaced, j);
    __syncthreads();  // the next item reuses the workgroup's LDS
  }
}


// G2P with the footprint's v and v - vⁿ staged per node in LDS; the FLIP
// increment is formed from the node differences as the reference does.  NW
// waves share one tile: wave w walks the class-major order at k = w, w+NW, ...
// and prefetches its next particle while updating the current one.
template <typename T, int D, bool AOS>
struct G2PIn {
  uint32_t pid, cls, key;  // particle, window-start class in the tile, sort key at its slot
  T x[D], v[D], F[D * D], l;
};

template <typename T, int D, int BASIS, bool AOS>
__device__ __forceinline__ void g2p_load(const Args<T, D>& a, uint32_t pid, G2PIn<T, D, AOS>& in) {
  in.pid = pid;
#pragma unroll
  for (int c = 0; c < D; ++c) in.x[c] = ld_g2p(fldw<T, D, AOS>(a.x, pid, c));
#pragma unroll
  for (int c = 0; c < D; ++c) in.v[c] = ld_g2p(fldw<T, D, AOS>(a.v, pid, c));
#pragma unroll
  for (int q = 0; q < D * D; ++q) in.F[q] = ld_g2p(fldw<T, D * D, AOS>(a.F, pid, q));
  in.l = T(0);
  if constexpr (BASIS == TESS_UGIMP) in.l = a.l(pid) / a.mesh.h;
}

// Stages a tile's footprint for G2P in two halves: g2p_stage_load issues the
// grid loads of v and vⁿ (zeros outside the grid) into registers, and
// g2p_stage_store writes v and v - vⁿ per component in the padded node order
// of Geo, with the footprint's node coordinates per axis; NT threads.  The
// loads stay in flight across the barriers in between (a barrier waits only on
// LDS operations), so the caller issues the tile's key and id loads before the
// grid values are needed, and one memory round trip serves both.  The caller
// synchronises after the store.
template <typename T, int D, int NT>
struct G2PStage {
  static constexpr int NPER = (Geo<D>::NFOOT + NT - 1) / NT;
  T v[NPER][D], vn[NPER][D];
};

template <typename T, int D, int NT>
__device__ __forceinline__ void g2p_stage_load(const Args<T, D>& a, const int* origin, G2PStage<T, D, NT>& st) {
  using G = Geo<D>;
#pragma unroll
  for (int u = 0; u < G2PStage<T, D, NT>::NPER; ++u) {
    const int q = int(threadIdx.x) + u * NT;
    int I[D], r = q;
    bool in = q < G::NFOOT;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      I[d] = origin[d] + r % G::TS;
      r /= G::TS;
      in = in && I[d] >= 0 && I[d] < a.mesh.dims[d];
    }
    const int64_t i = in ? grid_slot(a, I) : -1;
    in = i >= 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      st.v[u][c] = in ? a.gv(i, c) : T(0);
      st.vn[u][c] = in ? a.gvn(i, c) : T(0);
    }
  }
}

template <typename T, int D, int NT>
__device__ __forceinline__ void g2p_stage_store(const Args<T, D>& a, const int* origin,
                                                const G2PStage<T, D, NT>& st, T (*s_n)[Geo<D>::PSZ],
                                                T (*s_nodex)[Geo<D>::TS]) {
  using G = Geo<D>;
#pragma unroll
  for (int u = 0; u < G2PStage<T, D, NT>::NPER; ++u) {
    const int q = int(threadIdx.x) + u * NT;
    if (q < G::NFOOT) {
      const int pq = pad_node<D>(q % G::TS, (q / G::TS) % G::TS, D == 3 ? q / (G::TS * G::TS) : 0);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        s_n[c][pq] = st.v[u][c];
        s_n[D + c][pq] = st.v[u][c] - st.vn[u][c];
      }
    }
  }
  for (int q = int(threadIdx.x); q < D * G::TS; q += NT) {
    const int d = q / G::TS, k = q % G::TS;
    s_nodex[d][k] = a.mesh.node(d, origin[d] + k);
  }
}

// One particle's G2P and update (README.md:81-86) from the tile's staged
// footprint (s_n: v and v - vⁿ per component; s_nodex: node coordinates): its
// axis jets at its class window, the axis-factored sums of Δv, ∇v and the x
// rate, the writes of v, ∇v, x, F, σ, V and, with nk, of the next sort key and
// the stayer flag at its sorted slot jj.  emit(val, der, v, σ, V) receives the
// jets and the updated state.
template <typename T, int D, int BASIS, bool AOS, bool SEP, typename Emit>
__device__ __forceinline__ void g2p_particle(const Args<T, D>& a, const G2PIn<T, D, AOS>& cur,
                                             const T (*s_n)[Geo<D>::PSZ], const T (*s_nodex)[Geo<D>::TS],
                                             const uint32_t jj, const uint32_t tile, const uint32_t ltile,
                                             uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                             const uint32_t ntiles, uint32_t* __restrict__ movers, Emit&& emit) {
  using G = Geo<D>;
  constexpr int SW = support_width(BASIS);
  const T hinv = a.mesh.h_inv, dt = a.dt;
  int cc[D];  // the particle's class coordinates (its window's offset in the footprint)
  {
    uint32_t r = cur.cls;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      cc[d] = int(r % G::BW);
      r /= G::BW;
    }
  }
  T val[D][SW], der[D][SW];
  bool corrects = false;
  if constexpr (SEP) corrects = sep_corrects<T, D, BASIS>(a, cur.x);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T xi = (cur.x[d] - a.mesh.xmin[d]) * hinv;
    int lo, hi;
    axis_window<T, BASIS>(xi, cur.l, a.mesh.dims[d], lo, hi);
    if constexpr (SEP)
      axis_jets_sep<T, BASIS>(cur.x[d], a.mesh.xmin[d], cc[d], hi - lo, s_nodex[d], hinv, corrects, val[d],
                              der[d]);
    else
      axis_jets_mixed<T, BASIS>(cur.x[d], a.mesh.xmin[d], cc[d], hi - lo, s_nodex[d], cur.l, hinv, val[d],
                                der[d]);
  }
  // X = Σ w v_i, DV = Σ w (v_i - vⁿ_i), L = Σ v_i ⊗ ∇w, factored by axis
  T X[D], DV[D], L[D * D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    X[c] = T(0);
    DV[c] = T(0);
  }
#pragma unroll
  for (int q = 0; q < D * D; ++q) L[q] = T(0);
  if constexpr (D == 3) {
#pragma unroll
    for (int kz = 0; kz < SW; ++kz)
#pragma unroll
      for (int jy = 0; jy < SW; ++jy) {
        T S1[3] = {T(0), T(0), T(0)}, S2[3] = {T(0), T(0), T(0)}, Sd[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int ix = 0; ix < SW; ++ix) {
          const int q = pad_node<3>(cc[0] + ix, cc[1] + jy, cc[2] + kz);
          T nd[6];
#pragma unroll
          for (int u = 0; u < 6; ++u) nd[u] = s_n[u][q];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            S1[c] += val[0][ix] * nd[c];
            S2[c] += der[0][ix] * nd[c];
            Sd[c] += val[0][ix] * nd[3 + c];
          }
        }
        const T YZ = val[1][jy] * val[2][kz];
        const T YdZ = der[1][jy] * val[2][kz];
        const T YZd = val[1][jy] * der[2][kz];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          X[c] += YZ * S1[c];
          DV[c] += YZ * Sd[c];
          L[c] += YZ * S2[c];
          L[c + 3] += YdZ * S1[c];
          L[c + 6] += YZd * S1[c];
        }
      }
  } else {
#pragma unroll
    for (int jy = 0; jy < SW; ++jy) {
      T S1[2] = {T(0), T(0)}, S2[2] = {T(0), T(0)}, Sd[2] = {T(0), T(0)};
#pragma unroll
      for (int ix = 0; ix < SW; ++ix) {
        const int q = pad_node<2>(cc[0] + ix, cc[1] + jy, 0);
        T nd[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) nd[u] = s_n[u][q];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          S1[c] += val[0][ix] * nd[c];
          S2[c] += der[0][ix] * nd[c];
          Sd[c] += val[0][ix] * nd[2 + c];
        }
      }
      const T Y = val[1][jy], Yd = der[1][jy];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        X[c] += Y * S1[c];
        DV[c] += Y * Sd[c];
        L[c] += Y * S2[c];
        L[c + 2] += Yd * S1[c];
      }
    }
  }
  const uint32_t pid = cur.pid;
  // particle equations (README.md:81-86)
  T vn[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    vn[c] = cur.v[c] + DV[c];
    st_g2p(fldw<T, D, AOS>(a.v, pid, c), vn[c]);
  }
#pragma unroll
  for (int q = 0; q < D * D; ++q) st_g2p(fldw<T, D * D, AOS>(a.gradv, pid, q), L[q]);
  T xn[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    xn[c] = cur.x[c] + X[c] * dt;
    st_g2p(fldw<T, D, AOS>(a.x, pid, c), xn[c]);
  }
  if (a.x_keep) {
#pragma unroll
    for (int c = 0; c < D; ++c) a.x_keep[int64_t(pid) * D + c] = cur.x[c];
  }
  if (nk) {
    uint32_t nl = 1;  // the scene's tiles; in a batch its keys start at its first global tile
#pragma unroll
    for (int d = 0; d < D; ++d) nl *= uint32_t(a.ntiles[d]);
    const uint32_t kl = key_of_position<T, D, BASIS>(a, xn, cur.l, nl);
    const uint32_t key = kl == nl * 64u ? ntiles * 64u : kl + (tile - ltile) * 64u;
    nk[jj] = key;
    const bool moved = key != cur.key;
    stay[jj] = moved ? 0u : 1u;
    // particles whose key changed, counted per wave by its lowest live lane
    // (no atomic in the common case of none)
    const uint64_t mv = __ballot(moved);
    if (mv && (__ballot(true) & lanemask_lt64()) == 0)
      atomicAdd(movers + (blockIdx.x % MOVER_LANES) * MOVER_STRIDE, uint32_t(__popcll(mv)));
  }
  T Fn[D * D];
#pragma unroll
  for (int b = 0; b < D; ++b)
#pragma unroll
    for (int r = 0; r < D; ++r) {
      T t = L[r] * cur.F[D * b];  // (I + Δt∇v) F = F + Δt (∇v F)
#pragma unroll
      for (int kk = 1; kk < D; ++kk) t += L[r + D * kk] * cur.F[kk + D * b];
      Fn[r + D * b] = cur.F[r + D * b] + dt * t;
    }
#pragma unroll
  for (int q = 0; q < D * D; ++q) st_g2p(fldw<T, D * D, AOS>(a.F, pid, q), Fn[q]);
  T sg[D * D], Vn;
  neo_hookean<T, D>(Fn, a.lambda, a.mu, a.V0, a.logE.p != nullptr, a.logE.p ? a.logE(pid) : T(0), sg, Vn);
#pragma unroll
  for (int q = 0; q < D * D; ++q) st_g2p(fldw<T, D * D, AOS>(a.sigma, pid, q), sg[q]);
  st_g2p(a.V(pid), Vn);
  emit(val, der, vn, sg, Vn);
}

// With nk != nullptr the kernel also writes, at each particle's sorted position
// j, the sort key of its new position (nk[j]) and whether it equals the
// current one (stay[j]): the next step's incremental binning starts from them.
// waves per SIMD the Float32 G2P is compiled for (VGPR budget 512 / n), and
// waves per tile of the single-scene G2P
#ifndef TESS_G2P_WAVES_PER_SIMD
#define TESS_G2P_WAVES_PER_SIMD 3
#endif
#ifndef TESS_G2P_WAVES_PER_TILE
#define TESS_G2P_WAVES_PER_TILE 4
#endif
template <typename T, int D, int BASIS, bool AOS, int NW, bool BATCH, bool SEP>
__device__ __forceinline__ void k_g2p_fast_body(Args<T, D>& a, const uint32_t* __restrict__ skeys,
                                                       uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                       uint32_t ntiles, uint32_t* __restrict__ movers, const int64_t j) {
  using G = Geo<D>;
  constexpr int SW = support_width(BASIS);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));  // uniform: scalar branches
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  uint32_t poff;  // a batch's scene: local tile, first particle
  const uint32_t ltile = bind_tile<BATCH>(a, tile, poff);
  int tI[D], origin[D];
  unlin<D>(ltile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;
  // per component (v[D], then v-vⁿ[D]) the footprint in the padded
  // bank-conflict-free node order of Geo
  __shared__ T s_n[2 * D][G::PSZ];
  __shared__ T s_nodex[D][G::TS];
  __shared__ uint32_t s_cnt[64];
  G2PStage<T, D, 64 * NW> stage;
  g2p_stage_load<T, D, 64 * NW>(a, origin, stage);
  int ccount, coff, maxcnt;
  __shared__ uint32_t s_pid[PID_CAP], s_key[PID_CAP];
  tile_class_counts<NW>(skeys, start, cnt, s_cnt, ccount, coff, maxcnt, a.perm, poff, s_pid, s_key,
                        PID_CAP);  // barrier after the staging of ids and keys
  g2p_stage_store<T, D, 64 * NW>(a, origin, stage, s_n, s_nodex);
  __syncthreads();
  // Two walks of the tile's particles.  Class order (lane = class, iteration k =
  // the k-th particle of every class) reads the layered storage contiguously
  // and takes maxcnt iterations.  Sorted order (iteration k = sorted positions
  // 64k .. 64k+63) takes ceil(cnt/64); a tile whose classes are uneven or
  // sparse uses it when that saves a third of the iterations.
  const int n_sorted = (cnt + 63) / 64;
  const bool by_sorted = 3 * n_sorted < 2 * maxcnt;
  const int kend = by_sorted ? n_sorted : maxcnt;
  auto live = [&](int k) { return by_sorted ? k * 64 + lane < cnt : k < ccount; };
  auto slot = [&](int k) { return start + uint32_t(by_sorted ? k * 64 + lane : coff + k); };  // sorted position
  // a particle's id and sort key come from LDS (staged with the class
  // counts), an iteration ahead of the loads they address: those loads do not
  // wait behind the previous particle's stores, as a global load of the id
  // would (the memory counter drains in issue order)
  struct Ahead {
    uint32_t pid, key;
  };
  // staged: the tile's ids are all in LDS (cnt <= PID_CAP); the walk is
  // compiled for each case so that neither waits on the other's loads
  auto ahead = [&](int k, auto staged) {
    Ahead r{0u, 0u};
    if (live(k)) {
      const uint32_t li = slot(k) - start;
      if constexpr (decltype(staged)::value) {
        r.pid = s_pid[li];
        r.key = s_key[li];
      } else {
        r.pid = a.perm[start + li] - poff;
        r.key = skeys[start + li];
      }
    }
    return r;
  };
  auto fetch = [&](int k, const Ahead& q, G2PIn<T, D, AOS>& in) {
    if (!live(k)) return;
    g2p_load<T, D, BASIS, AOS>(a, q.pid, in);
    in.key = q.key;
    in.cls = by_sorted ? q.key & 63u : uint32_t(lane);
  };
  // two register sets alternate between the particle being updated and the
  // one being loaded (wave w walks k = w, w+NW, ...)
  G2PIn<T, D, AOS> pa, pb;
  std::memset(static_cast<void*>(&pa), 0, sizeof(pa));
  std::memset(static_cast<void*>(&pb), 0, sizeof(pb));
  auto particle = [&](int k, const G2PIn<T, D, AOS>& cur, G2PIn<T, D, AOS>& nxt, const Ahead& q_next,
                      Ahead& q_after, auto staged) {
    const bool act = live(k);
    q_after = ahead(k + 2 * NW, staged);
    fetch(k + NW, q_next, nxt);
    if (!act) return;
    g2p_particle<T, D, BASIS, AOS, SEP>(a, cur, s_n, s_nodex, nk ? slot(k) : 0u, tile, ltile, nk, stay, ntiles,
                                        movers, [](const auto&...) {});
  };
  auto walk = [&](auto staged) {
    Ahead qa = ahead(wave + NW, staged), qb{0u, 0u};
    fetch(wave, ahead(wave, staged), pa);
    for (int k = wave; k < kend; k += 2 * NW) {
      particle(k, pa, pb, qa, qb, staged);
      if (k + NW < kend) particle(k + NW, pb, pa, qb, qa, staged);
    }
  };
  if (cnt <= PID_CAP)
    walk(std::true_type{});
  else
    walk(std::false_type{});
}

template <typename T, int D, int BASIS, bool AOS, int NW, bool BATCH = false, bool SEP = false>
__global__ __launch_bounds__(64 * NW, sizeof(T) == 4 ? TESS_G2P_WAVES_PER_SIMD : 1) void k_g2p_fast(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                       uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                       uint32_t ntiles, uint32_t* __restrict__ movers) {
  const int64_t j = item_of<BATCH>(blockIdx.x, a.n_active);
  if (j < a.n_active) k_g2p_fast_body<T, D, BASIS, AOS, NW, BATCH, SEP>(a, skeys, nk, stay, ntiles, movers, j);
}

// Forms whose item count is read on the device at `count` (launch sizes fixed
// in advance): one workgroup per possible item, those past the count ending
// at once (_grid), or a persistent grid sized to the device walking the items
// (_persist; a multiple of 8 workgroups keeps each XCD on a contiguous run of
// them, as xcd_remap does for one workgroup per item).
template <typename T, int D, int BASIS, bool AOS, int NW, bool BATCH = false, bool SEP = false>
__global__ __launch_bounds__(64 * NW, sizeof(T) == 4 ? TESS_G2P_WAVES_PER_SIMD : 1) void k_g2p_fast_grid(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                       uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                       uint32_t ntiles, uint32_t* __restrict__ movers, const uint32_t* __restrict__ count) {
  const int64_t N = int64_t(*count), R = ((N + 7) >> 3) << 3;
  // xcd_remap is a bijection of [0, R) only: workgroups past R end here
  const int64_t j = item_of<BATCH>(blockIdx.x, N);
  if (int64_t(blockIdx.x) < R && j < N) k_g2p_fast_body<T, D, BASIS, AOS, NW, BATCH, SEP>(a, skeys, nk, stay, ntiles, movers, j);
}
template <typename T, int D, int BASIS, bool AOS, int NW, bool BATCH = false, bool SEP = false>
__global__ __launch_bounds__(64 * NW, sizeof(T) == 4 ? TESS_G2P_WAVES_PER_SIMD : 1) void k_g2p_fast_persist(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                       uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                       uint32_t ntiles, uint32_t* __restrict__ movers, const uint32_t* __restrict__ count) {
  const int64_t N = int64_t(*count), R = ((N + 7) >> 3) << 3;
  for (int64_t it = blockIdx.x; it < R; it += gridDim.x) {
    const int64_t j = item_of<BATCH>(it, N);
    if (j < N) k_g2p_fast_body<T, D, BASIS, AOS, NW, BATCH, SEP>(a, skeys, nk, stay, ntiles, movers, j);
    __syncthreads();  // the next item reuses the workgroup's LDS
  }
}


// ---- gather + grid-node half ------------------------------------------------
// ---- single-sweep @G2P2G -------------------------------------------------------
//
// One pass over a tile's particles runs both halves of a @G2P2G program
// (src/Transfer/g2p2g.jl:100-159) with one set of weights, those of the
// positions before the pass.  NP producer waves run the G2P half, one layer of
// class lanes each per round (the k-th particle of every class, as k_g2p_fast
// walks them: gather, update, writes).  They hand each particle's position
// before the update, mass, mₚv and -Vσ to NC consumer waves through LDS.  The
// consumers evaluate the jets of that position, own the stencil offsets of
// k_p2g_fast's waves and accumulate those layers one round behind the
// producers (double-buffered exchange, one barrier per round).
// They see the layers in k_p2g_fast's order and run its arithmetic, so the
// tile's partial is the one that the G2P kernel followed by the P2G kernel at
// the kept positions produces.  The particles are read once per pass instead
// of twice, and the G2P half's memory waits overlap the P2G half's arithmetic.
//
// σ is exchanged as its D(D+1)/2 distinct entries: neo_hookean forms it
// symmetric to the last bit (both triangles from the same products).

// index of σ's entry (r, b), r >= b, among the distinct entries
__host__ __device__ constexpr int sym_index(int r, int b, int D) { return b * D - b * (b - 1) / 2 + (r - b); }

// producer waves of the single-sweep kernel (0: as many as the LDS budget allows)
#ifndef TESS_G2P2G_PRODUCERS
#define TESS_G2P2G_PRODUCERS 0
#endif
template <typename T, int D, int BASIS>
struct G2P2GShape {
  static constexpr int SW = support_width(BASIS);
  static constexpr int NV = 1 + 2 * D;
  static constexpr int NS = D * (D + 1) / 2;
  static constexpr bool HAS_L = BASIS == TESS_UGIMP;
  // exchange rows per particle: x[D] (and l) before the update, mass, mₚv[D], -Vσ[NS]
  static constexpr int XL = D, XM = D + (HAS_L ? 1 : 0), XV = XM + 1, XA = XV + D, XR = XA + NS;
  // consumer waves: those of k_p2g_fast (3D: one per last-axis offset; 2D: one
  // while the stencil's accumulators fit)
  static constexpr int NC = (D == 3 || Pow<SW, D>::v * NV * int(sizeof(T)) / 4 > 96) ? SW : 1;
  static constexpr int lds(int np) {
    return int(sizeof(T)) * (2 * np * XR * 64 + 2 * D * Geo<D>::PSZ + D * Geo<D>::TS) + 4 * (2 * PID_CAP + 64);
  }
  // producer waves: eight waves per workgroup, two workgroups per CU at 128
  // VGPRs (Float32)
  static constexpr int NP_AUTO = 8 - NC;
  static constexpr int NP = TESS_G2P2G_PRODUCERS > 0 ? TESS_G2P2G_PRODUCERS : NP_AUTO;
  // waves per SIMD to compile for: as many workgroups per CU as the LDS holds
  // (at most two), at most 4 waves per SIMD; Float64 1
  static constexpr int wps(int np) {
    const int wg = lds(np) <= 80 * 1024 ? 2 : 1;
    const int w = (wg * (np + NC) + 3) / 4;
    return sizeof(T) != 4 ? 1 : w < 1 ? 1 : w > 4 ? 4 : w;
  }
};

template <typename T, int D, int BASIS, bool AOS, int NP, int NC>
__device__ __forceinline__ void k_g2p2g_fast_body(Args<T, D>& a, const uint32_t* __restrict__ skeys,
                                                  uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
                                                  uint32_t ntiles, uint32_t* __restrict__ movers,
                                                  uint32_t* __restrict__ misplaced, const int64_t j) {
  using G = Geo<D>;
  using SH = G2P2GShape<T, D, BASIS>;
  constexpr int SW = SH::SW, NV = SH::NV, XR = SH::XR, XM = SH::XM, XV = SH::XV, XA = SH::XA;
  constexpr int KN = SW / NC;                   // last-axis offsets per consumer
  constexpr int NSL = Pow<SW, D - 1>::v * KN;  // offsets per consumer
  constexpr int NT = 64 * (NP + NC);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));  // uniform: scalar branches
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  uint32_t poff;
  const uint32_t ltile = bind_tile<false>(a, tile, poff);
  int tI[D], origin[D];
  unlin<D>(ltile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;
  __shared__ T s_n[2 * D][G::PSZ];
  __shared__ T s_nodex[D][G::TS];
  __shared__ uint32_t s_cnt[64];
  __shared__ uint32_t s_pid[PID_CAP], s_key[PID_CAP];
  // [round parity][producer][row][class lane]; after the last round the
  // tile's node sums take its place: the consumers' slabs (several consumers)
  // or one array per component over the footprint (one)
  using SL = Slab<T, D, SW, KN>;
  constexpr int XWORDS = 2 * NP * XR * 64;
  constexpr int SUMW = NC > 1 ? NC * SL::SIZE : NV * G::PSZ;
  __shared__ __align__(16) T s_xraw[XWORDS > SUMW ? XWORDS : SUMW];
  T(*s_x)[NP][XR][64] = reinterpret_cast<T(*)[NP][XR][64]>(s_xraw);
  T(*s_tile)[G::PSZ] = reinterpret_cast<T(*)[G::PSZ]>(s_xraw);
  G2PStage<T, D, NT> stage;
  g2p_stage_load<T, D, NT>(a, origin, stage);
  int ccount, coff, maxcnt;
  tile_class_counts<NP + NC>(skeys, start, cnt, s_cnt, ccount, coff, maxcnt, a.perm, poff, s_pid, s_key,
                             PID_CAP);  // barrier after the staging of ids and keys
  g2p_stage_store<T, D, NT>(a, origin, stage, s_n, s_nodex);
  __syncthreads();
  const int nround = (maxcnt + NP - 1) / NP;  // the same in every wave
  auto zero_tile = [&]() {
    static_assert(NV * G::PSZ * sizeof(T) % 16 == 0, "node sums in 16-byte words");
    for (int q = int(threadIdx.x); q < int(NV * G::PSZ * sizeof(T) / 16); q += NT)
      reinterpret_cast<uint4*>(&s_tile[0][0])[q] = make_uint4(0u, 0u, 0u, 0u);
  };
  // Both roles pass the same barriers: nround + 1 in the rounds, then one after
  // the consumers fill their slabs (several consumers), or one after the node
  // sums are zeroed and one as the consumer adds its stencil (one).
  if (wave < NP) {
    // producer: layers k = wave, wave + NP, ... (round r takes k = r*NP + wave)
    auto ahead = [&](int k, auto staged) {
      uint32_t pid = 0u, key = 0u;
      if (k < ccount) {
        if constexpr (decltype(staged)::value) {
          pid = s_pid[coff + k];
          key = s_key[coff + k];
        } else {
          pid = a.perm[start + coff + k] - poff;
          key = skeys[start + coff + k];
        }
      }
      return uint2{pid, key};
    };
    auto fetch = [&](int k, const uint2& q, G2PIn<T, D, AOS>& in) {
      if (k >= ccount) return;
      g2p_load<T, D, BASIS, AOS>(a, q.x, in);
      in.key = q.y;
      in.cls = uint32_t(lane);
    };
    G2PIn<T, D, AOS> pa, pb;
    std::memset(static_cast<void*>(&pa), 0, sizeof(pa));
    std::memset(static_cast<void*>(&pb), 0, sizeof(pb));
    auto produce = [&](int r, const G2PIn<T, D, AOS>& cur, G2PIn<T, D, AOS>& nxt, const uint2& q_next, uint2& q_after,
                     auto staged) {
      if (r < nround) {
        const int k = r * NP + wave;
        q_after = ahead(k + 2 * NP, staged);
        fetch(k + NP, q_next, nxt);
        T(*xr)[64] = s_x[r & 1][wave];
        if (k < ccount) {
          g2p_particle<T, D, BASIS, AOS, false>(
              a, cur, s_n, s_nodex, start + uint32_t(coff + k), tile, ltile, nk, stay, ntiles, movers,
              [&](const auto&, const auto&, const auto& vn, const auto& sg, const T Vn) {
#pragma unroll
                for (int d = 0; d < D; ++d) xr[d][lane] = cur.x[d];
                if constexpr (SH::HAS_L) xr[SH::XL][lane] = cur.l;
                const T mp = a.mp;
                xr[XM][lane] = mp;
#pragma unroll
                for (int c = 0; c < D; ++c) xr[XV + c][lane] = mp * vn[c];
                const T negV = -Vn;
#pragma unroll
                for (int b = 0; b < D; ++b)
#pragma unroll
                  for (int rr = b; rr < D; ++rr) xr[XA + sym_index(rr, b, D)][lane] = negV * sg[rr + D * b];
              });
        } else {
          // no particle: a finite position, zero mass and stress (exact zero contributions)
#pragma unroll
          for (int q = 0; q < XR; ++q) xr[q][lane] = T(0);
        }
      }
      __syncthreads();
    };
    auto walk = [&](auto staged) {
      uint2 qa = ahead(wave + NP, staged), qb{0u, 0u};
      fetch(wave, ahead(wave, staged), pa);
      for (int r = 0; r <= nround; r += 2) {
        produce(r, pa, pb, qa, qb, staged);
        if (r + 1 <= nround) produce(r + 1, pb, pa, qb, qa, staged);
      }
    };
    if (cnt <= PID_CAP)
      walk(std::true_type{});
    else
      walk(std::false_type{});
    if constexpr (NC > 1) {
      __syncthreads();
    } else {
      zero_tile();
      __syncthreads();
      __syncthreads();
    }
  } else {
    // consumer `pass`: the stencil offsets with last-axis offset pass*KN .. pass*KN+KN-1
    const int pass = wave - NP;
    int ccoord[D];
    {
      int r = lane;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        ccoord[d] = r % G::BW;
        r /= G::BW;
      }
    }
    T am[NSL], amv[D][NSL], af[D][NSL];
#pragma unroll
    for (int o = 0; o < NSL; ++o) {
      am[o] = T(0);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        amv[c][o] = T(0);
        af[c][o] = T(0);
      }
    }
    const T hinv = a.mesh.h_inv;
    uint32_t bad = 0, last = 0;
    for (int r = 0; r <= nround; ++r) {
      if (r > 0) {
        for (int p = 0; p < NP; ++p) {
          const int k = (r - 1) * NP + p;
          if (k >= maxcnt) break;  // uniform
          const T(*xr)[64] = s_x[(r - 1) & 1][p];
          // the jets of the position before the update, as k_p2g_fast evaluates them
          T x[D], l = T(0);
#pragma unroll
          for (int d = 0; d < D; ++d) x[d] = xr[d][lane];
          if constexpr (SH::HAS_L) l = xr[SH::XL][lane];
          T val[D][SW], der[D][SW], vl[KN], dl[KN], mpv[D], A[D * D];
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const T xi = (x[d] - a.mesh.xmin[d]) * hinv;
            int lo, hi;
            axis_window<T, BASIS>(xi, l, a.mesh.dims[d], lo, hi);
            axis_jets_mixed<T, BASIS>(x[d], a.mesh.xmin[d], ccoord[d], hi - lo, s_nodex[d], l, hinv, val[d], der[d]);
          }
#pragma unroll
          for (int q = 0; q < KN; ++q) {
            vl[q] = val[D - 1][pass * KN + q];
            dl[q] = der[D - 1][pass * KN + q];
          }
          const T mpe = xr[XM][lane];
#pragma unroll
          for (int c = 0; c < D; ++c) mpv[c] = xr[XV + c][lane];
#pragma unroll
          for (int b = 0; b < D; ++b)
#pragma unroll
            for (int rr = 0; rr < D; ++rr)
              A[rr + D * b] = xr[XA + (rr >= b ? sym_index(rr, b, D) : sym_index(b, rr, D))][lane];
          p2g_accumulate<T, D, SW, KN>(val, der, vl, dl, mpe, mpv, A, am, amv, af);
          if (misplaced && pass == 0) {
            const bool act = k < ccount;
            uint32_t pid = 0u;
            if (act) pid = cnt <= PID_CAP ? s_pid[coff + k] : a.perm[start + coff + k] - poff;
            bad += broken_pair(__ballot(act), k, pid, last);
          }
        }
      }
      __syncthreads();
    }
    if constexpr (NC > 1) {
      // each consumer zeroes and fills its own slab (the exchange is read out:
      // the rounds ended with a barrier)
      T* const slab = s_xraw + pass * SL::SIZE;
      for (int q = lane; q < int(SL::SIZE * sizeof(T) / 16); q += 64)
        reinterpret_cast<uint4*>(slab)[q] = make_uint4(0u, 0u, 0u, 0u);
      if (ccount > 0) p2g_stencil_to_slab<T, D, SW, KN>(slab, ccoord, pass, am, amv, af);
      __syncthreads();
    } else {
      zero_tile();
      __syncthreads();
      if (ccount > 0) p2g_stencil_to_lds<T, D, SW, KN>(s_tile, ccoord, pass, am, amv, af);
      __syncthreads();
    }
    if (misplaced && pass == 0) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) bad += __shfl_xor(bad, off, 64);
      if (lane == 0) misplaced[j] = bad;  // per tile, summed later
    }
  }
  if constexpr (NC > 1)
    p2g_write_partial_slabs<T, D, SW, KN, NC, NT>(a, j, s_xraw);
  else
    p2g_write_partial<T, D, SW, NT>(a, j, s_tile);
}

template <typename T, int D, int BASIS, bool AOS, int NP, int NC>
__global__ __launch_bounds__(64 * (NP + NC), (G2P2GShape<T, D, BASIS>::wps(NP))) void k_g2p2g_fast(
    Args<T, D> a, const uint32_t* __restrict__ skeys, uint32_t* __restrict__ nk, uint32_t* __restrict__ stay,
    uint32_t ntiles, uint32_t* __restrict__ movers, uint32_t* __restrict__ misplaced) {
  const int64_t j = item_of<false>(blockIdx.x, a.n_active);
  if (j < a.n_active) k_g2p2g_fast_body<T, D, BASIS, AOS, NP, NC>(a, skeys, nk, stay, ntiles, movers, misplaced, j);
}

template <typename T, int D, int SW, bool GAOS, bool BATCH>
__device__ __forceinline__ void k_gather_grid_body(Args<T, D>& a, uint32_t* __restrict__ nodes_per_block, const int64_t j) {
  using G = Geo<D>;
  using PT = Part<D, SW>;
  constexpr int NV = 1 + 2 * D;
  constexpr int NB = (D == 3) ? 27 : 9;
  const uint32_t B = a.nb_list[j];
  uint32_t poff;  // a batch's scene: local node block; tiles of the scene start at toff
  const uint32_t Bl = bind_tile<BATCH>(a, B, poff);
  const uint32_t toff = B - Bl;
  int Bi[D], I[D], rr[D];
  unlin<D>(Bl, a.ntiles, Bi);
  // slots of the 3^dim neighbour tiles (-1: no partial)
  __shared__ int s_slot[NB];
  if (threadIdx.x < NB) {
    int r = int(threadIdx.x), t = 0, mul = 1;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int b = Bi[d] + (r % 3) - 1;
      r /= 3;
      ok = ok && b >= 0 && b < a.ntiles[d];
      t += b * mul;
      mul *= a.ntiles[d];
    }
    s_slot[threadIdx.x] = (ok && a.tile_flag[toff + t]) ? int(a.tile_slot[toff + t]) : -1;
  }
  __syncthreads();
  {
    int r = threadIdx.x;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      rr[d] = r % G::BW;
      r /= G::BW;
      I[d] = Bi[d] * G::BW + rr[d];
      if (I[d] >= a.mesh.dims[d]) return;
    }
  }
  // Tiles covering this node, per axis: B-1 when the node lies in the
  // footprint of the tile below (r < SW-2), B, and B+1 when r == BW-1; at
  // most two per axis.  All their partial values are loaded first, then
  // summed in a fixed order (last axis outermost, ascending tiles).
  int lo[3] = {0, 0, 0}, two[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = rr[d] < SW - 2 ? -1 : 0;
    two[d] = (lo[d] < 0 || rr[d] == G::BW - 1) ? 1 : 0;
  }
  constexpr int NC = 1 << D;
  T pv[NC][NV];
  bool use[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    bool ok = true;
    int nb = 0, q = 0, mulb = 1, mulq = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = (c >> d) & 1;
      ok = ok && k <= two[d];
      const int o = lo[d] + k;
      nb += (o + 1) * mulb;
      mulb *= 3;
      q += (rr[d] + 1 - o * G::BW) * mulq;  // node index inside that tile's partial
      mulq *= PT::E;
    }
    const int slot = ok ? s_slot[nb] : -1;
    use[c] = slot >= 0;
    // a skipped candidate reads tile slot 0's first node (a valid address) and is not added
    const T* P = a.partials + (use[c] ? int64_t(slot) * PT::SIZE + int64_t(q) * NV : 0);
#pragma unroll
    for (int v = 0; v < NV; ++v) pv[c][v] = P[v];
  }
  T acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = T(0);
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = use[c] ? acc[v] + pv[c][v] : acc[v];
  const int64_t i = grid_slot(a, I);
  // grid-node half with the guarded Float64 division (test/transfer.jl:102-104)
  const double md = double(acc[0]);
  const bool has = md > 0.0;
  {
    // nodes with m > 0 in this block (one plain store per block, summed later)
    const uint64_t m = __ballot(has);
    if ((threadIdx.x & 63) == __ffsll((unsigned long long)__ballot(true)) - 1)
      nodes_per_block[j] = uint32_t(__popcll(m));
  }
  const double minv = has ? 1.0 / md : 0.0;
  if (i < 0) return;  // SpGrid node of an inactive block: no storage (and no support, checked by P2G)
  a.gm(i) = acc[0];
  T mvo[D], fo[D], vno[D], vo[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const T mvc = acc[1 + c], fc = acc[1 + D + c];
    const T vn = T(double(mvc) * minv);
    const T accel = T(double(fc) * minv) + (has ? a.gravity[c] : T(0));
    T vc = vn + accel * a.dt;
    const bool wall = (((a.slip >> (2 * c)) & 1) && I[c] == 0) ||
                      (((a.slip >> (2 * c + 1)) & 1) && I[c] == a.mesh.dims[c] - 1);
    if (wall) vc = T(0);
    mvo[c] = mvc;
    fo[c] = fc;
    vno[c] = vn;
    vo[c] = vc;
  }
#pragma unroll
  for (int c = 0; c < D; ++c) {
    fldw<T, D, GAOS>(a.gmv, i, c) = mvo[c];
    fldw<T, D, GAOS>(a.gf, i, c) = fo[c];
    fldw<T, D, GAOS>(a.gvn, i, c) = vno[c];
    fldw<T, D, GAOS>(a.gv, i, c) = vo[c];
  }
}

template <typename T, int D, int SW, bool GAOS, bool BATCH = false>
__global__ __launch_bounds__(64) void k_gather_grid(Args<T, D> a, uint32_t* __restrict__ nodes_per_block) {
  const int64_t j = item_of<BATCH>(blockIdx.x, a.n_nb);
  if (j < a.n_nb) k_gather_grid_body<T, D, SW, GAOS, BATCH>(a, nodes_per_block, j);
}

// Forms whose item count is read on the device at `count` (launch sizes fixed
// in advance): one workgroup per possible item, those past the count ending
// at once (_grid), or a persistent grid sized to the device walking the items
// (_persist; a multiple of 8 workgroups keeps each XCD on a contiguous run of
// them, as xcd_remap does for one workgroup per item).
template <typename T, int D, int SW, bool GAOS, bool BATCH = false>
__global__ __launch_bounds__(64) void k_gather_grid_grid(Args<T, D> a, uint32_t* __restrict__ nodes_per_block, const uint32_t* __restrict__ count) {
  const int64_t N = int64_t(*count), R = ((N + 7) >> 3) << 3;
  // xcd_remap is a bijection of [0, R) only: workgroups past R end here
  const int64_t j = item_of<BATCH>(blockIdx.x, N);
  if (int64_t(blockIdx.x) < R && j < N) k_gather_grid_body<T, D, SW, GAOS, BATCH>(a, nodes_per_block, j);
}
template <typename T, int D, int SW, bool GAOS, bool BATCH = false>
__global__ __launch_bounds__(64) void k_gather_grid_persist(Args<T, D> a, uint32_t* __restrict__ nodes_per_block, const uint32_t* __restrict__ count) {
  const int64_t N = int64_t(*count), R = ((N + 7) >> 3) << 3;
  for (int64_t it = blockIdx.x; it < R; it += gridDim.x) {
    const int64_t j = item_of<BATCH>(it, N);
    if (j < N) k_gather_grid_body<T, D, SW, GAOS, BATCH>(a, nodes_per_block, j);
    __syncthreads();  // the next item reuses the workgroup's LDS
  }
}


// ---- G2P -------------------------------------------------------------------

// Particle equations of the README @G2P block (README.md:81-86) after the sums.
template <typename T, int D>
__device__ __forceinline__ void particle_update(const Args<T, D>& a, int64_t p, const T* dv, const T* L,
                                                const T* dx) {
  const T dt = a.dt;
#pragma unroll
  for (int c = 0; c < D; ++c) a.v(p, c) += dv[c];
#pragma unroll
  for (int k = 0; k < D * D; ++k) a.gradv(p, k) = L[k];
  if (a.x_keep) {
#pragma unroll
    for (int c = 0; c < D; ++c) a.x_keep[p * D + c] = a.x(p, c);
  }
#pragma unroll
  for (int c = 0; c < D; ++c) a.x(p, c) += dx[c];
  T Fo[D * D], M[D * D], Fn[D * D];
#pragma unroll
  for (int k = 0; k < D * D; ++k) Fo[k] = a.F(p, k);
#pragma unroll
  for (int b = 0; b < D; ++b)
#pragma unroll
    for (int r = 0; r < D; ++r) M[r + D * b] = (r == b ? T(1) : T(0)) + L[r + D * b] * dt;
#pragma unroll
  for (int b = 0; b < D; ++b)
#pragma unroll
    for (int r = 0; r < D; ++r) {
      T s = M[r] * Fo[D * b];
#pragma unroll
      for (int k = 1; k < D; ++k) s = s + M[r + D * k] * Fo[k + D * b];
      Fn[r + D * b] = s;
    }
#pragma unroll
  for (int k = 0; k < D * D; ++k) a.F(p, k) = Fn[k];
  T sg[D * D], Vn;
  neo_hookean<T, D>(Fn, a.lambda, a.mu, a.V0, a.logE.p != nullptr, a.logE.p ? a.logE(p) : T(0), sg, Vn);
#pragma unroll
  for (int k = 0; k < D * D; ++k) a.sigma(p, k) = sg[k];
  a.V(p) = Vn;
}

// One wave per tile, lanes own the tile's classes as in P2G (the walk is
// coalesced when particles are stored in the layered order).  The footprint's
// vⁿ and v are staged in LDS once per tile.
template <typename T, int D, int BASIS, bool STORED, bool CORR = false>
__global__ __launch_bounds__(64) void k_g2p_tile(Args<T, D> a, const uint32_t* __restrict__ skeys) {
  using G = Geo<D>;
  constexpr int SW = support_width(BASIS);
  const int lane = threadIdx.x;
  const int64_t j = xcd_remap(blockIdx.x, a.n_active);
  if (j >= a.n_active) return;
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  int tI[D], origin[D];
  unlin<D>(tile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;
  __shared__ T s_vn[D][G::NFOOT];
  __shared__ T s_v[D][G::NFOOT];
  __shared__ T s_nodex[D][G::TS];
  __shared__ uint32_t s_cnt[64];
  for (int q = lane; q < G::NFOOT; q += 64) {
    int I[D], r = q;
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      I[d] = origin[d] + r % G::TS;
      r /= G::TS;
      in = in && I[d] >= 0 && I[d] < a.mesh.dims[d];
    }
    const int64_t i = in ? grid_slot(a, I) : -1;
    in = i >= 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      s_vn[c][q] = in ? a.gvn(i, c) : T(0);
      s_v[c][q] = in ? a.gv(i, c) : T(0);
    }
  }
  for (int q = lane; q < D * G::TS; q += 64) {
    const int d = q / G::TS, k = q % G::TS;
    s_nodex[d][k] = a.mesh.node(d, origin[d] + k);
  }
  int ccount, coff, maxcnt;
  tile_class_counts(skeys, start, cnt, s_cnt, ccount, coff, maxcnt);  // contains the barrier after staging
  int ccoord[D];
  {
    int r = lane;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ccoord[d] = r % G::BW;
      r /= G::BW;
    }
  }
  const T hinv = a.mesh.h_inv, dt = a.dt;
  for (int k = 0; k < maxcnt; ++k) {
    if (k >= ccount) continue;
    const uint32_t pid = a.perm[start + coff + k];
    T x[D], l;
    int lo[D], hi[D];
    particle_window<T, D, BASIS, STORED>(a, pid, x, l, lo, hi, nullptr);
    int ext[D];
#pragma unroll
    for (int d = 0; d < D; ++d) ext[d] = hi[d] - lo[d];
    T val[D][SW], der[D][SW];
    if constexpr (!STORED) {
#pragma unroll
      for (int d = 0; d < D; ++d) axis_jets<T, D, BASIS>(x[d], ccoord[d], s_nodex[d], l, hinv, val[d], der[d]);
    }
    Corrected<T, D, SW> cs;
    if constexpr (CORR) correct_jets<T, D, BASIS>(a, x, lo, ext, ccoord, s_nodex, cs, val, der);
    T dv[D], L[D * D], dx[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dv[c] = T(0);
      dx[c] = T(0);
    }
#pragma unroll
    for (int q = 0; q < D * D; ++q) L[q] = T(0);
    // column-major walk of the window (eachindex(window))
#pragma unroll
    for (int o = 0; o < Pow<SW, D>::v; ++o) {
      int aa[D], r = o;
      bool ok = true;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        aa[d] = r % SW;
        r /= SW;
        ok = ok && aa[d] <= ext[d];
      }
      if (!ok) continue;
      T w, g[D];
      if constexpr (STORED) {
        const int64_t e = int64_t(o) + int64_t(Pow<SW, D>::v) * pid;
        w = a.w(e);
#pragma unroll
        for (int c = 0; c < D; ++c) g[c] = a.gw(e, c);
      } else if (CORR && cs.mode == CORR_GENERAL) {
        cs.node(val, aa, filter_pass<D>(a.filter, a.mesh.dims, lo, aa), hinv, w, g);
      } else {
        w = val[0][aa[0]];
#pragma unroll
        for (int d = 1; d < D; ++d) w = w * val[d][aa[d]];
#pragma unroll
        for (int c = 0; c < D; ++c) {
          T gg = (c == 0) ? der[0][aa[0]] : val[0][aa[0]];
#pragma unroll
          for (int d = 1; d < D; ++d) gg = gg * ((d == c) ? der[d][aa[d]] : val[d][aa[d]]);
          g[c] = gg * hinv;
        }
      }
      int q = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        q += (ccoord[d] + aa[d]) * mul;
        mul *= G::TS;
      }
      T vi[D];
#pragma unroll
      for (int c = 0; c < D; ++c) vi[c] = s_v[c][q];
#pragma unroll
      for (int c = 0; c < D; ++c) dv[c] += w * (vi[c] - s_vn[c][q]);
#pragma unroll
      for (int b = 0; b < D; ++b)
#pragma unroll
        for (int r2 = 0; r2 < D; ++r2) L[r2 + D * b] += vi[r2] * g[b];
#pragma unroll
      for (int c = 0; c < D; ++c) dx[c] += (w * vi[c]) * dt;
    }
    particle_update<T, D>(a, pid, dv, L, dx);
  }
}

// Particles outside the mesh have an empty window: the sums are zero and only
// the particle equations run (g2p.jl:123-153).
template <typename T, int D, bool BATCH = false>
__global__ __launch_bounds__(256) void k_g2p_outside(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                     uint32_t* __restrict__ nk, uint32_t* __restrict__ stay) {
  const int64_t j = a.n_inside + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= a.n) return;
  uint32_t pid = a.perm[j];
  if constexpr (BATCH) pid -= uint32_t(bind_scene(a, scene_of(a.scene_poff, a.nscenes, pid)).poff);
  if (nk) {  // the position does not change: the key stays
    nk[j] = skeys[j];
    stay[j] = 1u;
  }
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

// ---- host side ----------------------------------------------------------------

template <typename T, int D>
Args<T, D> make_args(const tess_mesh& mesh, const tess_grid* G, const tess_particles& P, const tess_weights* W,
                     const tess_params* prm) {
  Args<T, D> a;
  std::memset(static_cast<void*>(&a), 0, sizeof(a));
  a.mesh = dmesh<T, D>(mesh);
  a.x = dfield<T>(P.x);
  a.v = dfield<T>(P.v);
  a.V = dfield<T>(P.V);
  a.gradv = dfield<T>(P.grad_v);
  a.F = dfield<T>(P.F);
  a.sigma = dfield<T>(P.sigma);
  a.l = dfield<T>(P.l);
  a.logE = dfield<T>(P.logE);
  if (G) {
    a.gm = dfield<T>(G->m);
    a.gmv = dfield<T>(G->mv);
    a.gf = dfield<T>(G->f);
    a.gvn = dfield<T>(G->vn);
    a.gv = dfield<T>(G->v);
  }
  if (W) {
    a.window = W->window;
    a.w = dfield<T>(W->w);
    a.gw = dfield<T>(W->grad_w);
  }
  a.n = P.n;
  if (prm) {
    a.mp = T(prm->mp);
    a.dt = T(prm->dt);
    a.lambda = prm->lambda;
    a.mu = prm->mu;
    a.V0 = prm->V0;
    for (int d = 0; d < D; ++d) a.gravity[d] = T(prm->gravity[d]);
    a.slip = prm->slip_walls;
  }
  for (int d = 0; d < D; ++d) a.ntiles[d] = int32_t(((mesh.dims[d] - 1) >> Geo<D>::LT) + 1);
  return a;
}

template <typename T, int D>
void set_sparse(Args<T, D>& a, const tess_mesh& mesh, const tess_sparsity* sp) {
  if (!sp) return;
  a.numbering = sp->blocknumbering;
  a.spL = mesh.block_size_log2;
  for (int d = 0; d < D; ++d) a.nblk[d] = int32_t(((mesh.dims[d] - 1) >> mesh.block_size_log2) + 1);
}

template <typename T, int D>
void fill_binning_ptrs(tess_ctx* ctx, Args<T, D>& a) {
  const Binning& b = ctx->bin;
  a.perm = static_cast<const uint32_t*>(ctx->vals.ptr);
  a.tile_start = static_cast<const uint32_t*>(ctx->tile_start.ptr);
  a.tile_end = static_cast<const uint32_t*>(ctx->tile_end.ptr);
  a.active = static_cast<const uint32_t*>(ctx->active_tiles.ptr);
  a.tile_flag = static_cast<const uint32_t*>(ctx->tile_flag.ptr);
  a.tile_slot = static_cast<const uint32_t*>(ctx->tile_slot.ptr);
  a.nb_list = static_cast<const uint32_t*>(ctx->nb_list.ptr);
  a.nb_flag = static_cast<const uint32_t*>(ctx->nb_flag.ptr);
  a.n_inside = b.n_inside;
  a.n_active = b.n_active;
  a.n_nb = b.n_nodeblocks;
}

// The corrected basis and node filter of the current call (deferred weights
// only: stored weights already hold the corrected values).
template <typename T, int D, bool STORED>
void set_correction(tess_ctx* ctx, Args<T, D>& a) {
  a.corr = STORED ? 0 : ctx->corr;
  a.filter = STORED ? nullptr : ctx->corr_filter;
}

// True when every particle field is Julia-native AoS (components contiguous,
// element stride = component count), the layout the vectorised kernels assume.
inline bool julia_layout(const tess_particles& P, int D) {
  auto ok = [](const tess_field& f, int nc) { return !f.ptr || (f.comp_stride == 1 && f.elem_stride == nc) ||
                                                      (nc == 1 && f.elem_stride == 1); };
  return ok(P.x, D) && ok(P.v, D) && ok(P.V, 1) && ok(P.grad_v, D * D) && ok(P.F, D * D) && ok(P.sigma, D * D);
}

inline bool grid_julia_layout(const tess_grid& G, int D) {
  auto ok = [D](const tess_field& f) { return f.comp_stride == 1 && f.elem_stride == D; };
  return ok(G.mv) && ok(G.f) && ok(G.vn) && ok(G.v);
}

inline unsigned blocks_for(int64_t n, int threads) { return unsigned(std::max<int64_t>(1, ceil_div(n, threads))); }

int key_bits_for(uint64_t maxkey) {
  int b = 0;
  while (b < 32 && (uint64_t(1) << b) <= maxkey) ++b;
  return b;
}

// Incremental re-binning with M > 0 movers: nk[j] is the new key of the
// particle at sorted position j, spos[j] its stayer rank (exclusive scan of the
// stayer flags, spos[n] = stayers).  Few movers are sorted among themselves
// and merged into the stayers' previous order; many are sorted with the
// stayers.  The merged order replaces (keys_sorted, vals).
inline void merge_movers(tess_ctx* ctx, const uint32_t* nk, const uint32_t* spos, int64_t n, int64_t M, int kbits) {
  uint32_t* skeys = static_cast<uint32_t*>(ctx->keys_sorted.ptr);
  uint32_t* perm = static_cast<uint32_t*>(ctx->vals.ptr);
  uint32_t* skeys2 = ctx->skeys2.as<uint32_t>(size_t(n));
  uint32_t* perm2 = ctx->perm2.as<uint32_t>(size_t(n));
  if (M <= n / 8) {
    uint32_t* mkey = ctx->inc_mkey.as<uint32_t>(size_t(M));
    uint32_t* mj = ctx->inc_mj.as<uint32_t>(size_t(M));
    uint32_t* mks = ctx->inc_mks.as<uint32_t>(size_t(M));
    uint32_t* mr = ctx->inc_mr.as<uint32_t>(size_t(M));
    uint32_t* mjs = ctx->inc_mjs.as<uint32_t>(size_t(M));
    hipLaunchKernelGGL(k_compact_movers, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, nk, spos, n, mkey, mj);
    TESS_HIP(hipGetLastError());
    radix_sort_pairs(ctx, mkey, nullptr, mks, mr, M, kbits);
    hipLaunchKernelGGL(k_place_movers, dim3(blocks_for(M, 256)), dim3(256), 0, ctx->stream, mks, mr, mj, skeys, perm,
                       spos, n, M, mjs, skeys2, perm2, static_cast<const uint32_t*>(nullptr));
    TESS_HIP(hipGetLastError());
    const int64_t nsb = blocks_for(n, STAY_BLOCK);
    int64_t* ranges = ctx->inc_rng.as<int64_t>(size_t(2 * nsb));
    hipLaunchKernelGGL(k_stay_ranges, dim3(blocks_for(nsb, 256)), dim3(256), 0, ctx->stream, skeys, n, mks, mjs, M,
                       nsb, ranges, static_cast<const uint32_t*>(nullptr));
    TESS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_place_stayers, dim3(unsigned(nsb)), dim3(STAY_BLOCK), 0, ctx->stream, skeys, perm, spos, n,
                       mks, mjs, ranges, skeys2, perm2);
    TESS_HIP(hipGetLastError());
  } else {
    radix_sort_pairs(ctx, nk, perm, skeys2, perm2, n, kbits);
  }
  std::swap(ctx->keys_sorted, ctx->skeys2);
  std::swap(ctx->vals, ctx->perm2);
}

// Debug builds (-DTESS_DEBUG=1, `make debug`): after every binning the host
// reads the binning back and checks its invariants (the permutation, sorted
// keys, each tile's run of keys, the active-tile list), raising
// TESS_ERR_STATE on a violation instead of letting a later kernel index
// through it.
#ifndef TESS_DEBUG
#define TESS_DEBUG 0
#endif
inline void debug_check_binning(hipStream_t st, const uint32_t* skeys, const uint32_t* perm, const uint32_t* tstart,
                                const uint32_t* tend, const uint32_t* active, int64_t n, int64_t ntiles,
                                int64_t n_active, int64_t n_inside, const char* where) {
  if (!TESS_DEBUG) return;
  auto fail = [&](const std::string& what) { raise(TESS_ERR_STATE, std::string("debug check (") + where + "): " + what); };
  TESS_HIP(hipStreamSynchronize(st));
  const size_t nn = size_t(std::max<int64_t>(n, 0)), nt = size_t(std::max<int64_t>(ntiles, 0));
  std::vector<uint32_t> sk(nn), pm(nn), ts(nt), te(nt), act(size_t(std::max<int64_t>(n_active, 0)));
  auto get = [&](std::vector<uint32_t>& v, const uint32_t* src) {
    if (!v.empty()) TESS_HIP(hipMemcpy(v.data(), src, v.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  };
  get(sk, skeys);
  get(pm, perm);
  get(ts, tstart);
  get(te, tend);
  get(act, active);
  const uint64_t outside = uint64_t(ntiles) * 64u;
  std::vector<uint8_t> seen(sk.size(), 0);
  int64_t inside = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (pm[size_t(j)] >= uint64_t(n) || seen[pm[size_t(j)]]) fail("the permutation repeats or leaves [0, n)");
    seen[pm[size_t(j)]] = 1;
    if (sk[size_t(j)] > outside) fail("a sort key past the tiles");
    if (j > 0 && sk[size_t(j)] < sk[size_t(j - 1)]) fail("sort keys out of order at " + std::to_string(j));
    inside += sk[size_t(j)] < outside ? 1 : 0;
  }
  if (n_inside >= 0 && inside != n_inside) fail("inside count " + std::to_string(n_inside) + " != " + std::to_string(inside));
  int64_t nact = 0, j = 0;
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t a = j;
    while (j < inside && int64_t(sk[size_t(j)] >> 6) == t) ++j;
    if (j > a) {
      if (ts[size_t(t)] != uint64_t(a) || te[size_t(t)] != uint64_t(j)) fail("tile run of " + std::to_string(t));
      if (nact >= n_active || act[size_t(nact)] != uint64_t(t)) fail("active-tile list at " + std::to_string(nact));
      ++nact;
    } else if (te[size_t(t)] > ts[size_t(t)]) {
      fail("a particle run for the empty tile " + std::to_string(t));
    }
  }
  if (nact != n_active) fail("active tiles " + std::to_string(n_active) + " != " + std::to_string(nact));
}

template <typename T, int D, int BASIS, bool STORED>
void bin_particles_run(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, const tess_weights* W,
                       bool keys_from_g2p);

template <typename T, int D, int BASIS, bool STORED>
void bin_particles(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, const tess_weights* W,
                   bool keys_from_g2p = false) {
  bin_particles_run<T, D, BASIS, STORED>(ctx, mesh, P, W, keys_from_g2p);
  if (TESS_DEBUG && P.n > 0) {
    const Binning& b = ctx->bin;
    debug_check_binning(ctx->stream, static_cast<const uint32_t*>(ctx->keys_sorted.ptr),
                        static_cast<const uint32_t*>(ctx->vals.ptr), static_cast<const uint32_t*>(ctx->tile_start.ptr),
                        static_cast<const uint32_t*>(ctx->tile_end.ptr),
                        static_cast<const uint32_t*>(ctx->active_tiles.ptr), P.n, b.ntiles, b.n_active, b.n_inside,
                        "bin_particles");
  }
}

template <typename T, int D, int BASIS, bool STORED>
void bin_particles_run(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, const tess_weights* W,
                       bool keys_from_g2p) {
  Args<T, D> a = make_args<T, D>(mesh, nullptr, P, W, nullptr);
  const int64_t n = P.n;
  uint64_t ntiles = 1;
  for (int d = 0; d < D; ++d) ntiles *= uint64_t(a.ntiles[d]);
  if (ntiles * 64 >= uint64_t(UINT32_MAX) || n >= int64_t(UINT32_MAX))
    raise(TESS_ERR_ARGUMENT, "mesh or particle count exceeds the 32-bit binning range");
  Binning& b = ctx->bin;
  b.valid = false;
  uint32_t* counters = ctx->counters.as<uint32_t>(CNT_SLOTS);
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
  const int tok = prof_begin(ctx, "binning", 0.0);
  const int kbits = key_bits_for(ntiles * 64);
  const bool incremental = b.state_valid && b.state_n == n && b.state_ntiles == int64_t(ntiles) &&
                           b.state_dim == D && n > 0;
  const bool precomputed = keys_from_g2p && b.keys_ready;
  b.keys_ready = false;
  if (n > 0 && incremental && precomputed && b.x_ptr == P.x.ptr && b.tile_log2 == Geo<D>::LT &&
      b.basis == BASIS && b.dtype == mesh.dtype) {
    // The previous G2P counted the particles whose key changed.  With none,
    // the sorted keys, the order and every tile and node-block list of the
    // previous binning are what binning again would produce: keep them.
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
    TESS_HIP(hipMemcpyAsync(hc, counters, CNT_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    movers_readback(ctx);
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    if (movers_total(ctx) == 0) {
      ctx->last_movers = 0;
      ctx->last_active_nodes = int64_t(hc[CNT_NODES]);
      ctx->last_misplaced = (ctx->misplaced_x_ptr == P.x.ptr && ctx->misplaced_n == n) ? int64_t(hc[CNT_MISPLACED]) : -1;
      ctx->kept_binnings += 1;
      b.valid = true;
      prof_end(ctx, tok);
      return;
    }
  }
  TESS_HIP(hipMemsetAsync(counters, 0, CNT_BIN_SLOTS * sizeof(uint32_t), ctx->stream));
  if (n > 0 && incremental) {
    // keys in the previous sorted order, stayer flags -> stayer ranks; the
    // previous G2P of the same step loop may have written them already
    uint32_t* nk = ctx->inc_nk.as<uint32_t>(size_t(n));
    uint32_t* spos = ctx->inc_spos.as<uint32_t>(size_t(n) + 1);
    if (precomputed) {
      TESS_HIP(hipMemsetAsync(spos + n, 0, sizeof(uint32_t), ctx->stream));
    } else {
      hipLaunchKernelGGL((k_keys_sorted<T, D, BASIS, STORED>), dim3(blocks_for(n + 1, 256)), dim3(256), 0,
                         ctx->stream, a, perm, skeys, uint32_t(ntiles), nk, spos, counters);
      TESS_HIP(hipGetLastError());
    }
    exclusive_scan_u32(ctx, spos, n + 1, counters + CNT_STAY);
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
    TESS_HIP(hipMemcpyAsync(hc, counters, CNT_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t M = n - int64_t(hc[CNT_STAY]);
    ctx->last_movers = M;
    if (M == 0 && b.tile_log2 == Geo<D>::LT && b.basis == BASIS && b.dtype == mesh.dtype &&
        !(hc[CNT_ERR] & ERR_GIMP_LENGTH)) {
      // no particle changed its key: the previous order and tile lists stand
      ctx->last_active_nodes = int64_t(hc[CNT_NODES]);
      ctx->last_misplaced = (ctx->misplaced_x_ptr == P.x.ptr && ctx->misplaced_n == n) ? int64_t(hc[CNT_MISPLACED]) : -1;
      ctx->kept_binnings += 1;
      b.x_ptr = P.x.ptr;
      b.valid = true;
      prof_end(ctx, tok);
      return;
    }
    if (M > 0) {
      merge_movers(ctx, nk, spos, n, M, kbits);
      skeys = static_cast<uint32_t*>(ctx->keys_sorted.ptr);
      perm = static_cast<uint32_t*>(ctx->vals.ptr);
    }
  } else if (n > 0) {
    hipLaunchKernelGGL((k_bin_keys<T, D, BASIS, STORED>), dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a,
                       keys, uint32_t(ntiles), counters);
    TESS_HIP(hipGetLastError());
    radix_sort_pairs(ctx, keys, nullptr, skeys, perm, n, kbits);
    ctx->last_movers = -1;
  }
  TESS_HIP(hipMemsetAsync(tstart, 0, ntiles * sizeof(uint32_t), ctx->stream));
  TESS_HIP(hipMemsetAsync(tend, 0, ntiles * sizeof(uint32_t), ctx->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_tile_runs, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, skeys, n,
                       uint32_t(ntiles), tstart, tend, counters);
    TESS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_tile_flags, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, tstart, tend,
                     uint32_t(ntiles), tflag, tslot);
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(ctx, tslot, int64_t(ntiles), counters + CNT_ACTIVE);
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, tflag, tslot,
                     uint32_t(ntiles), active);
  TESS_HIP(hipGetLastError());
  TESS_HIP(hipMemsetAsync(nbflag, 0, ntiles * sizeof(uint32_t), ctx->stream));
  const int3 nt = make_int3(a.ntiles[0], a.ntiles[1], D == 3 ? a.ntiles[D == 3 ? 2 : 0] : 1);
  hipLaunchKernelGGL(k_nb_mark<D>, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, tflag,
                     uint32_t(ntiles), nt, nbflag);
  TESS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, nbflag, nbslot,
                     int64_t(ntiles));
  TESS_HIP(hipGetLastError());
  exclusive_scan_u32(ctx, nbslot, int64_t(ntiles), counters + CNT_NB);
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(ntiles, 256)), dim3(256), 0, ctx->stream, nbflag, nbslot,
                     uint32_t(ntiles), nblist);
  TESS_HIP(hipGetLastError());
  prof_end(ctx, tok);
  // the one host readback of the binning: totals for the launch sizes
  uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
  TESS_HIP(hipMemcpyAsync(hc, counters, CNT_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
  if (hc[CNT_ERR] & ERR_GIMP_LENGTH) raise(TESS_ERR_ARGUMENT, "uGIMP requires 0 ≤ pt.l ≤ spacing(mesh)");
  if (hc[CNT_ERR] & ERR_WINDOW)
    raise(TESS_ERR_DIMENSION, "stored support window is larger than the basis support width");
  b.n_inside = n > 0 ? int64_t(hc[CNT_INSIDE]) : 0;
  ctx->last_active_nodes = int64_t(hc[CNT_NODES]);
  // layout score of the previous P2G on this particle set (consumed by the step driver)
  ctx->last_misplaced = (ctx->misplaced_x_ptr == P.x.ptr && ctx->misplaced_n == n) ? int64_t(hc[CNT_MISPLACED]) : -1;
  b.n_active = int64_t(hc[CNT_ACTIVE]);
  b.n_nodeblocks = int64_t(hc[CNT_NB]);
  b.x_ptr = P.x.ptr;
  b.n = n;
  b.dim = D;
  b.dtype = mesh.dtype;
  b.basis = BASIS;
  b.tile_log2 = Geo<D>::LT;
  b.ntiles = int64_t(ntiles);
  for (int d = 0; d < 3; ++d) b.ntiles_axis[d] = d < D ? a.ntiles[d] : 1;
  b.valid = true;
  b.state_valid = true;
  b.state_n = n;
  b.state_ntiles = int64_t(ntiles);
  b.state_dim = D;
}

// P2G after binning: zeroing of what this call does not overwrite, the tile
// kernel and the gather.  count_misplaced asks the tile kernel to count
// particles not stored in the layered order (for the step driver's reorder).
// With `sweep`, the tile partials come from sweep(a, sorted keys, misplaced)
// in place of the P2G kernel (the single-sweep @G2P2G kernel, which reads the
// grid's v and vⁿ: the grid is zeroed after it).
template <typename T, int D>
using SweepFn = std::function<void(Args<T, D>&, const uint32_t*, uint32_t*)>;

template <typename T, int D, int BASIS, bool STORED>
void p2g_launch(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                const tess_weights* W, const tess_params& prm, bool count_misplaced, const tess_sparsity* sp,
                const SweepFn<T, D>* sweep = nullptr) {
  using Gm = Geo<D>;
  constexpr int NV = 1 + 2 * D;
  constexpr int SW = support_width(BASIS);
  // all sw^dim offsets in one pass while the accumulators fit ~200 VGPRs,
  // else one pass per offset along the last axis
  constexpr int NPASS = (Pow<SW, D>::v * NV * int(sizeof(T)) / 4 <= 100) ? 1 : SW;
  Args<T, D> a = make_args<T, D>(mesh, &G, P, W, &prm);
  fill_binning_ptrs<T, D>(ctx, a);
  set_sparse<T, D>(a, mesh, sp);
  set_correction<T, D, STORED>(ctx, a);
  uint32_t* counters = static_cast<uint32_t*>(ctx->counters.ptr);
  int64_t nnodes = 1;
  for (int d = 0; d < D; ++d) nnodes *= mesh.dims[d];
  // zero what this call will not overwrite: everything on first use of a
  // grid, else the node blocks that were written last time and are not now.
  // An SpGrid's storage is zeroed whole (its numbering may have changed).
  const bool fresh = ctx->last_grid_m != G.m.ptr || ctx->prev_tile_log2 != Gm::LT ||
                     ctx->prev_ntiles != ctx->bin.ntiles || ctx->prev_nnodes != nnodes;
  if (sp) {
    a.p2g_err = counters + CNT_P2GERR;
    TESS_HIP(hipMemsetAsync(a.p2g_err, 0, sizeof(uint32_t), ctx->stream));
  }
  auto zero_grid = [&]() {
    if (sp) {
      if (sp->storage_len > 0) {
        hipLaunchKernelGGL((k_zero_grid<T, D>), dim3(blocks_for(sp->storage_len, 256)), dim3(256), 0, ctx->stream,
                           a, sp->storage_len);
        TESS_HIP(hipGetLastError());
      }
    } else if (fresh) {
      hipLaunchKernelGGL((k_zero_grid<T, D>), dim3(blocks_for(nnodes, 256)), dim3(256), 0, ctx->stream, a, nnodes);
      TESS_HIP(hipGetLastError());
    } else if (ctx->prev_nodeblocks > 0) {
      hipLaunchKernelGGL((k_zero_stale<T, D>), dim3(unsigned(ctx->prev_nodeblocks)), dim3(64), 0, ctx->stream, a,
                         static_cast<const uint32_t*>(ctx->nb_list_prev.ptr), ctx->prev_nodeblocks);
      TESS_HIP(hipGetLastError());
    }
  };
  if (!sweep) zero_grid();
  const Binning& b = ctx->bin;
  uint32_t* misplaced = nullptr;
  uint32_t* per_tile = ctx->per_tile.as<uint32_t>(size_t(std::max<int64_t>(b.n_active, 1)));
  uint32_t* per_nb = ctx->per_nodeblock.as<uint32_t>(size_t(std::max<int64_t>(b.n_nodeblocks, 1)));
  if (count_misplaced) {
    misplaced = per_tile;
    ctx->misplaced_x_ptr = P.x.ptr;
    ctx->misplaced_n = P.n;
    ctx->misplaced_pairs = b.n_inside - b.n_active;
  } else {
    ctx->misplaced_x_ptr = nullptr;
  }
  if (b.n_active > 0 && sweep) {
    a.partials = ctx->partials.as<T>(size_t(b.n_active) * Part<D, SW>::SIZE);
    (*sweep)(a, static_cast<const uint32_t*>(ctx->keys_sorted.ptr), misplaced);
    if (misplaced) sum_u32(ctx, misplaced, b.n_active, counters + CNT_MISPLACED);
  } else if (b.n_active > 0) {
    a.partials = ctx->partials.as<T>(size_t(b.n_active) * Part<D, SW>::SIZE);
    // algorithmic bytes: x, v, V, σ per particle
    const int tok = prof_begin(ctx, "k_p2g_tile", double(b.n_inside) * sizeof(T) * (2 * D + 1 + D * D));
    const uint32_t* sk = static_cast<const uint32_t*>(ctx->keys_sorted.ptr);
    if constexpr (STORED) {
      hipLaunchKernelGGL((k_p2g_tile<T, D, BASIS, STORED, NPASS>), dim3(xcd_grid(b.n_active)), dim3(64), 0,
                         ctx->stream, a, sk, misplaced);
    } else if (a.corr && !separable_correction(a.corr, BASIS, a.filter != nullptr)) {
      hipLaunchKernelGGL((k_p2g_tile<T, D, BASIS, false, NPASS, true>), dim3(xcd_grid(b.n_active)), dim3(64), 0,
                         ctx->stream, a, sk, misplaced);
    } else {
      // 3D: one wave per offset along z; 2D: one wave while the stencil fits
      constexpr int NPF = (D == 3 || Pow<SW, D>::v * NV * int(sizeof(T)) / 4 > 96) ? SW : 1;
      const bool aos = julia_layout(P, D);
      bool done = false;
      if constexpr (BASIS == TESS_BSPLINE_QUADRATIC || BASIS == TESS_BSPLINE_CUBIC) {
        if (a.corr) {  // separable corrected basis: the same kernel with corrected axis jets
          if (aos)
            hipLaunchKernelGGL((k_p2g_fast<T, D, BASIS, NPF, true, false, true>), dim3(xcd_grid(b.n_active)),
                               dim3(64 * NPF), 0, ctx->stream, a, sk, misplaced);
          else
            hipLaunchKernelGGL((k_p2g_fast<T, D, BASIS, NPF, false, false, true>), dim3(xcd_grid(b.n_active)),
                               dim3(64 * NPF), 0, ctx->stream, a, sk, misplaced);
          done = true;
        }
      }
      if (!done) {
        if (aos)
          hipLaunchKernelGGL((k_p2g_fast<T, D, BASIS, NPF, true>), dim3(xcd_grid(b.n_active)), dim3(64 * NPF), 0,
                             ctx->stream, a, sk, misplaced);
        else
          hipLaunchKernelGGL((k_p2g_fast<T, D, BASIS, NPF, false>), dim3(xcd_grid(b.n_active)), dim3(64 * NPF), 0,
                             ctx->stream, a, sk, misplaced);
      }
    }
    TESS_HIP(hipGetLastError());
    prof_end(ctx, tok);
    if (misplaced) {
      sum_u32(ctx, misplaced, b.n_active, counters + CNT_MISPLACED);
    }
  }
  if (sweep) zero_grid();
  TESS_HIP(hipMemsetAsync(counters + CNT_NODES, 0, sizeof(uint32_t), ctx->stream));
  if (b.n_nodeblocks > 0) {
    // algorithmic bytes per active node: m, mv, f written, read back by the
    // grid-node half, vⁿ and v written (the node count of the previous call)
    const double nodes = double(std::max<int64_t>(ctx->last_active_nodes, 0));
    const int tok = prof_begin(ctx, "k_gather_grid", nodes * sizeof(T) * (2 * (2 * D + 1) + 2 * D));
    if (grid_julia_layout(G, D))
      hipLaunchKernelGGL((k_gather_grid<T, D, SW, true>), dim3(xcd_grid(b.n_nodeblocks)), dim3(64), 0, ctx->stream, a,
                         per_nb);
    else
      hipLaunchKernelGGL((k_gather_grid<T, D, SW, false>), dim3(xcd_grid(b.n_nodeblocks)), dim3(64), 0, ctx->stream, a,
                         per_nb);
    TESS_HIP(hipGetLastError());
    prof_end(ctx, tok);
    sum_u32(ctx, per_nb, b.n_nodeblocks, counters + CNT_NODES);
  }
  // remember this call's node blocks for the next call's stale zeroing
  uint32_t* prev = ctx->nb_list_prev.as<uint32_t>(size_t(std::max<int64_t>(b.ntiles, 1)));
  if (b.n_nodeblocks > 0)
    TESS_HIP(hipMemcpyAsync(prev, ctx->nb_list.ptr, size_t(b.n_nodeblocks) * sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, ctx->stream));
  ctx->prev_nodeblocks = b.n_nodeblocks;
  ctx->last_grid_m = sp ? nullptr : G.m.ptr;  // a dense grid after an SpGrid call starts from a full zero
  ctx->prev_tile_log2 = Gm::LT;
  ctx->prev_ntiles = b.ntiles;
  ctx->prev_nnodes = nnodes;
  if (sp) {
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx->counters_host);
    TESS_HIP(hipMemcpyAsync(hc + CNT_P2GERR, counters + CNT_P2GERR, sizeof(uint32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
    TESS_HIP(hipStreamSynchronize(ctx->stream));
    if (hc[CNT_P2GERR] & ERR_INACTIVE)
      raise(TESS_ERR_STATE, "@P2G: inactive SpGrid support node. Call update_sparsity! before @P2G.");
  }
}

template <typename T, int D, int BASIS, bool STORED>
void g2p_launch(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                const tess_weights* W, const tess_params& prm, const tess_sparsity* sp, bool write_keys = false,
                T* x_keep = nullptr) {
  const Binning& b = ctx->bin;
  Args<T, D> a = make_args<T, D>(mesh, &G, P, W, &prm);
  a.x_keep = x_keep;
  fill_binning_ptrs<T, D>(ctx, a);
  set_sparse<T, D>(a, mesh, sp);
  set_correction<T, D, STORED>(ctx, a);
  // next-step keys at the sorted positions (deferred weights in the fast kernel)
  const bool generic = a.corr && !separable_correction(a.corr, BASIS, a.filter != nullptr);
  uint32_t* nk = nullptr;
  uint32_t* stay = nullptr;
  uint32_t* mv = nullptr;  // movers: particles whose key changes
  if (write_keys && !STORED && !generic && P.n > 0) {
    nk = ctx->inc_nk.as<uint32_t>(size_t(P.n));
    stay = ctx->inc_spos.as<uint32_t>(size_t(P.n) + 1);
    mv = movers_counters(ctx);
    movers_zero(ctx);
  }
  const uint32_t* skeys_now = static_cast<const uint32_t*>(ctx->keys_sorted.ptr);
  const bool aos = julia_layout(P, D);
  if (b.n_active > 0) {
    // algorithmic bytes: x, v, F read and v, ∇v, x, F, σ, V written per
    // particle; vⁿ, v read per active node
    const double nodes = double(std::max<int64_t>(ctx->last_active_nodes, 0));
    const int per_p = 2 * D + D * D + 2 * D + 3 * D * D + 1;
    const int tok = prof_begin(ctx, "k_g2p_tile",
                               double(b.n_inside) * sizeof(T) * per_p + nodes * sizeof(T) * 2 * D);
    const uint32_t* sk = static_cast<const uint32_t*>(ctx->keys_sorted.ptr);
    if constexpr (STORED) {
      hipLaunchKernelGGL((k_g2p_tile<T, D, BASIS, STORED>), dim3(xcd_grid(b.n_active)), dim3(64), 0, ctx->stream, a,
                         sk);
    } else if (generic) {
      hipLaunchKernelGGL((k_g2p_tile<T, D, BASIS, false, true>), dim3(xcd_grid(b.n_active)), dim3(64), 0,
                         ctx->stream, a, sk);
    } else {
      constexpr int NW = TESS_G2P_WAVES_PER_TILE;
      const uint32_t nt = uint32_t(b.ntiles);
      bool done = false;
      if constexpr (BASIS == TESS_BSPLINE_QUADRATIC || BASIS == TESS_BSPLINE_CUBIC) {
        if (a.corr) {  // separable corrected basis
          if (aos)
            hipLaunchKernelGGL((k_g2p_fast<T, D, BASIS, true, NW, false, true>), dim3(xcd_grid(b.n_active)),
                               dim3(64 * NW), 0, ctx->stream, a, sk, nk, stay, nt, mv);
          else
            hipLaunchKernelGGL((k_g2p_fast<T, D, BASIS, false, NW, false, true>), dim3(xcd_grid(b.n_active)),
                               dim3(64 * NW), 0, ctx->stream, a, sk, nk, stay, nt, mv);
          done = true;
        }
      }
      if (!done) {
        if (aos)
          hipLaunchKernelGGL((k_g2p_fast<T, D, BASIS, true, NW>), dim3(xcd_grid(b.n_active)), dim3(64 * NW), 0,
                             ctx->stream, a, sk, nk, stay, nt, mv);
        else
          hipLaunchKernelGGL((k_g2p_fast<T, D, BASIS, false, NW>), dim3(xcd_grid(b.n_active)), dim3(64 * NW), 0,
                             ctx->stream, a, sk, nk, stay, nt, mv);
      }
    }
    TESS_HIP(hipGetLastError());
    prof_end(ctx, tok);
  }
  if (P.n > b.n_inside) {
    hipLaunchKernelGGL((k_g2p_outside<T, D>), dim3(blocks_for(P.n - b.n_inside, 256)), dim3(256), 0, ctx->stream,
                       a, skeys_now, nk, stay);
    TESS_HIP(hipGetLastError());
  }
  ctx->bin.valid = false;  // positions moved
  ctx->bin.keys_ready = nk != nullptr;
}

// Layered storage order of the binned particles: inside each tile the k-th
// particle of every class, classes ascending, then the next k.  Writes the
// gather order (new storage slot -> old particle) and the sorted-slot -> new
// storage index map.
template <typename T, int D>
__global__ __launch_bounds__(64) void k_layered_order(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                       uint32_t* __restrict__ order, uint32_t* __restrict__ perm_new,
                                                       uint32_t* __restrict__ broken = nullptr) {
  const int lane = threadIdx.x;
  const int64_t j = blockIdx.x;
  if (j >= a.n_active) return;
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  __shared__ uint32_t s_cnt[64];
  int ccount, coff, maxcnt;
  tile_class_counts(skeys, start, cnt, s_cnt, ccount, coff, maxcnt);
  uint32_t layer_off = 0, bad = 0, last = 0;
  for (int k = 0; k < maxcnt; ++k) {
    const uint64_t has = __ballot(k < ccount);
    const uint32_t pid = k < ccount ? a.perm[start + coff + k] : 0u;
    if (k < ccount) {
      const uint32_t dest = start + layer_off + uint32_t(__popcll(has & lanemask_lt64()));
      order[dest] = pid;
      perm_new[start + coff + k] = dest;
    }
    if (broken) bad += broken_pair(has, k, pid, last);
    layer_off += uint32_t(__popcll(has));
  }
  if (broken) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) bad += __shfl_xor(bad, off, 64);
    if (lane == 0) broken[j] = bad;
  }
}

__global__ void k_tail_identity(const uint32_t* __restrict__ perm, int64_t from, int64_t n,
                                uint32_t* __restrict__ order, uint32_t* __restrict__ perm_new) {
  const int64_t j = from + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  order[j] = perm[j];
  perm_new[j] = uint32_t(j);
}

// fraction of layered neighbour pairs (particles inside the mesh minus active
// tiles) that are consecutive in storage; 1 with no pairs
inline double layered_contiguity(int64_t broken, int64_t pairs) {
  return pairs > 0 ? 1.0 - double(broken) / double(pairs) : 1.0;
}

template <typename T, int D>
void layered_reorder(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P) {
  const Binning& b = ctx->bin;
  const int64_t n = P.n;
  if (n == 0) return;
  Args<T, D> a = make_args<T, D>(mesh, nullptr, P, nullptr, nullptr);
  fill_binning_ptrs<T, D>(ctx, a);
  uint32_t* order = ctx->keys.as<uint32_t>(size_t(n));        // raw keys are dead after the sort
  uint32_t* perm_new = ctx->keys_alt.as<uint32_t>(size_t(n));
  if (b.n_active > 0) {
    hipLaunchKernelGGL((k_layered_order<T, D>), dim3(unsigned(b.n_active)), dim3(64), 0, ctx->stream, a,
                       static_cast<const uint32_t*>(ctx->keys_sorted.ptr), order, perm_new);
    TESS_HIP(hipGetLastError());
  }
  if (n > b.n_inside) {
    hipLaunchKernelGGL(k_tail_identity, dim3(blocks_for(n - b.n_inside, 256)), dim3(256), 0, ctx->stream,
                       static_cast<const uint32_t*>(ctx->vals.ptr), b.n_inside, n, order, perm_new);
    TESS_HIP(hipGetLastError());
  }
  permute_particles(ctx, mesh, P, order);
  TESS_HIP(hipMemcpyAsync(ctx->vals.ptr, perm_new, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                          ctx->stream));
}

template <typename T, int D, int BASIS, bool STORED>
void p2g_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
              const tess_weights* W, const tess_params& prm, const tess_sparsity* sp) {
  bin_particles<T, D, BASIS, STORED>(ctx, mesh, P, W);
  p2g_launch<T, D, BASIS, STORED>(ctx, mesh, G, P, W, prm, false, sp);
}

template <typename T, int D, int BASIS, bool STORED>
void g2p_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
              const tess_weights* W, const tess_params& prm, const tess_sparsity* sp) {
  // a standalone G2P bins again: device addresses alone cannot prove that the
  // particle set is unchanged since the last P2G
  bin_particles<T, D, BASIS, STORED>(ctx, mesh, P, W);
  g2p_launch<T, D, BASIS, STORED>(ctx, mesh, G, P, W, prm, sp);
}

// A @G2P2G pass as one particle sweep (k_g2p2g_fast) where it is asked for
// (TESS_G2P2G_SWEEP=1 in the environment) and applies (deferred weights
// without kernel correction).  The particles were binned at their current
// positions.  The sweep runs both halves and writes the tile partials;
// p2g_launch then zeroes the grid and runs the grid-node half.  With
// write_keys the sweep also writes the next binning's keys.  Returns false
// (nothing enqueued) where the caller composes the G2P and P2G kernels, the
// default: on the 13.4 M-particle block drop the sweep takes 1.23 ms against
// 0.72 + 0.43 ms for the two kernels (its producer waves, 10 per CU against
// the G2P kernel's 12, wait on memory, and the per-round barrier couples them
// to the consumers).
inline bool g2p2g_sweep_enabled() {
  const char* e = std::getenv("TESS_G2P2G_SWEEP");
  return e && e[0] == '1' && e[1] == '\0';
}

template <typename T, int D, int BASIS>
bool g2p2g_sweep(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                 const tess_params& prm, bool write_keys, bool count_misplaced) {
  if (ctx->corr != 0 || !g2p2g_sweep_enabled()) return false;
  using SH = G2P2GShape<T, D, BASIS>;
  const Binning& b = ctx->bin;
  uint32_t* nk = nullptr;
  uint32_t* stay = nullptr;
  uint32_t* mv = nullptr;  // movers: particles whose key changes
  if (write_keys && P.n > 0) {
    nk = ctx->inc_nk.as<uint32_t>(size_t(P.n));
    stay = ctx->inc_spos.as<uint32_t>(size_t(P.n) + 1);
    mv = movers_counters(ctx);
    movers_zero(ctx);
  }
  const bool aos = julia_layout(P, D);
  const uint32_t nt = uint32_t(b.ntiles);
  const SweepFn<T, D> sweep = [&](Args<T, D>& a, const uint32_t* sk, uint32_t* misplaced) {
    // algorithmic bytes: those of the G2P kernel (x, v, F read; v, ∇v, x, F,
    // σ, V written; vⁿ, v read per active node); the P2G half reads nothing
    const double nodes = double(std::max<int64_t>(ctx->last_active_nodes, 0));
    const int per_p = 2 * D + D * D + 2 * D + 3 * D * D + 1;
    const int tok = prof_begin(ctx, "k_g2p2g_tile", double(b.n_inside) * sizeof(T) * per_p + nodes * sizeof(T) * 2 * D);
    const dim3 grid(xcd_grid(b.n_active)), block(64 * (SH::NP + SH::NC));
    if (aos)
      hipLaunchKernelGGL((k_g2p2g_fast<T, D, BASIS, true, SH::NP, SH::NC>), grid, block, 0, ctx->stream, a, sk, nk,
                         stay, nt, mv, misplaced);
    else
      hipLaunchKernelGGL((k_g2p2g_fast<T, D, BASIS, false, SH::NP, SH::NC>), grid, block, 0, ctx->stream, a, sk, nk,
                         stay, nt, mv, misplaced);
    TESS_HIP(hipGetLastError());
    prof_end(ctx, tok);
  };
  p2g_launch<T, D, BASIS, false>(ctx, mesh, G, P, nullptr, prm, count_misplaced, nullptr, &sweep);
  if (P.n > b.n_inside) {  // particles outside the mesh: the G2P half's update with zero sums
    Args<T, D> a = make_args<T, D>(mesh, &G, P, nullptr, &prm);
    fill_binning_ptrs<T, D>(ctx, a);
    hipLaunchKernelGGL((k_g2p_outside<T, D>), dim3(blocks_for(P.n - b.n_inside, 256)), dim3(256), 0, ctx->stream,
                       a, static_cast<const uint32_t*>(ctx->keys_sorted.ptr), nk, stay);
    TESS_HIP(hipGetLastError());
  }
  ctx->bin.valid = false;  // positions moved
  ctx->bin.keys_ready = nk != nullptr;
  return true;
}

// @G2P2G of the README program (src/Transfer/g2p2g.jl:100-159): one set of
// weights, at the positions before the pass, serves the G2P half and the P2G
// half.  One binning at those positions orders both halves: the G2P kernel
// updates the particles, then the P2G kernels scatter the updated v, V, σ with
// the weights of the old positions (stored windows, or the positions the G2P
// kernel writes aside as it moves the particles), and the grid-node half follows.
template <typename T, int D, int BASIS, bool STORED>
void g2p2g_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                const tess_weights* W, const tess_params& prm) {
  bin_particles<T, D, BASIS, STORED>(ctx, mesh, P, W);
  if constexpr (!STORED) {
    if (g2p2g_sweep<T, D, BASIS>(ctx, mesh, G, P, prm, false, false)) return;
  }
  // the G2P half keeps each particle's position before its update aside
  // (deferred weights), for the P2G half's weights
  tess_particles Pold = P;
  T* xo = nullptr;
  if (!STORED && P.n > 0) {
    xo = ctx->g2p2g_x.as<T>(size_t(P.n) * D);
    Pold.x = tess_field{xo, D, 1};
  }
  g2p_launch<T, D, BASIS, STORED>(ctx, mesh, G, P, W, prm, nullptr, false, xo);
  ctx->bin.valid = true;  // the binning describes the old positions the P2G half reads
  p2g_launch<T, D, BASIS, STORED>(ctx, mesh, G, Pold, W, prm, false, nullptr);
  ctx->bin.valid = false;  // the particles have moved
}

// nsteps @G2P2G passes of the README program (deferred weights), the time loop
// of a program written as one @G2P2G block: each pass bins at the positions
// before it (from the second pass on with the keys the previous pass's G2P
// wrote), keeps those positions for its P2G half, and reorders the particles
// into the layered order at the threshold, as step_impl does.
template <typename T, int D, int BASIS>
void g2p2g_steps_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                      const tess_params& prm, int nsteps, double reorder_threshold) {
  for (int s = 0; s < nsteps; ++s) {
    bin_particles<T, D, BASIS, false>(ctx, mesh, P, nullptr, s > 0);
    const Binning& b = ctx->bin;
    if (reorder_threshold > 0.0 && b.n_inside > 0) {
      bool go = reorder_threshold >= 1.0;
      if (!go && ctx->last_misplaced >= 0) {
        const double score = layered_contiguity(ctx->last_misplaced, ctx->misplaced_pairs);
        go = score < reorder_threshold;
      }
      if (go) {
        layered_reorder<T, D>(ctx, mesh, P);
        ctx->reorders += 1;
      }
    }
    if (!g2p2g_sweep<T, D, BASIS>(ctx, mesh, G, P, prm, s + 1 < nsteps, reorder_threshold > 0.0)) {
      tess_particles Pold = P;
      T* xo = nullptr;
      if (P.n > 0) {
        xo = ctx->g2p2g_x.as<T>(size_t(P.n) * D);
        Pold.x = tess_field{xo, D, 1};
      }
      // the G2P half writes the positions before its update to xo
      g2p_launch<T, D, BASIS, false>(ctx, mesh, G, P, nullptr, prm, nullptr, s + 1 < nsteps, xo);
      ctx->bin.valid = true;  // the binning describes the old positions the P2G half reads
      p2g_launch<T, D, BASIS, false>(ctx, mesh, G, Pold, nullptr, prm, reorder_threshold > 0.0, nullptr);
      ctx->bin.valid = false;
    }
    // the storage order the P2G half measured is that of the particles themselves
    if (reorder_threshold > 0.0) ctx->misplaced_x_ptr = P.x.ptr;
  }
}

// nsteps of the README loop body (deferred weights).  With reorder_threshold
// in (0, 1], particles are moved into the layered order when the contiguity
// score measured by the previous step's P2G (the fraction of neighbours in the
// layered order whose storage indices are consecutive, the estimator of
// block_ordered_particle_contiguity) is below the threshold (1 reorders every
// step), the role of reorder_particles!(particles, partition; threshold) in
// the reference loop.
//
// *done (may be NULL) counts the steps whose transfers were enqueued; an SpGrid
// step that finds its storage too small raises before any of its own work, so
// after the error *done steps have run and the caller resizes and runs the
// remaining nsteps - *done.
//
// hook (may be NULL) is called as (*hook)(0, s) before step s touches the
// particles and (*hook)(1, s) once its binning and P2G grid are in place,
// before its G2P (the adjoint's replays copy states and snapshots there).
template <typename T, int D, int BASIS, bool STORED>
void step_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
               const tess_params& prm, int nsteps, double reorder_threshold, tess_sparsity* sp,
               int* done = nullptr, const std::function<void(int, int)>* hook = nullptr) {
  for (int s = 0; s < nsteps; ++s) {
    if (done) *done = s;
    if (hook) (*hook)(0, s);
    if (sp) {
      // update_sparsity!(grid, xₚ) at the top of the SpGrid loop (src/Grids/grid.jl:97-101)
      run_sparsity_update(ctx, mesh, P.x, P.n, *sp, nullptr);
      if (sp->storage_len > sp->capacity)
        raise(TESS_ERR_STATE, "SpGrid storage holds " + std::to_string(sp->capacity) + " nodes but the active blocks need " +
                                  std::to_string(sp->storage_len) + "; resize the grid arrays and call again");
    }
    // from the second step on, the previous G2P wrote the keys (same call, so
    // nothing else moved the particles in between)
    bin_particles<T, D, BASIS, false>(ctx, mesh, P, nullptr, s > 0);
    const Binning& b = ctx->bin;
    if (reorder_threshold > 0.0 && b.n_inside > 0) {
      bool go = reorder_threshold >= 1.0;
      if (!go && ctx->last_misplaced >= 0) {
        const double score = layered_contiguity(ctx->last_misplaced, ctx->misplaced_pairs);
        go = score < reorder_threshold;
      }
      if (go) {
        layered_reorder<T, D>(ctx, mesh, P);
        ctx->reorders += 1;
      }
    }
    p2g_launch<T, D, BASIS, false>(ctx, mesh, G, P, nullptr, prm, reorder_threshold > 0.0, sp);
    if (hook) (*hook)(1, s);
    g2p_launch<T, D, BASIS, false>(ctx, mesh, G, P, nullptr, prm, sp, s + 1 < nsteps);
  }
  if (done) *done = nsteps;
}

template <template <typename, int, int, bool> class Op, typename... Args_>
void dispatch(int basis, const tess_mesh& mesh, bool stored, Args_&&... args) {
#define TESS_CASE(T_, D_, B_)                                     \
  if (mesh.dim == D_ && basis == B_) {                            \
    if (stored)                                                   \
      Op<T_, D_, B_, true>::run(std::forward<Args_>(args)...);    \
    else                                                          \
      Op<T_, D_, B_, false>::run(std::forward<Args_>(args)...);   \
    return;                                                       \
  }
#define TESS_CASES(T_, D_)                        \
  TESS_CASE(T_, D_, TESS_BSPLINE_LINEAR)          \
  TESS_CASE(T_, D_, TESS_BSPLINE_QUADRATIC)       \
  TESS_CASE(T_, D_, TESS_BSPLINE_CUBIC)           \
  TESS_CASE(T_, D_, TESS_UGIMP)
  if (mesh.dtype == TESS_FLOAT32) {
    TESS_CASES(float, 2)
    TESS_CASES(float, 3)
  } else {
    TESS_CASES(double, 2)
    TESS_CASES(double, 3)
  }
#undef TESS_CASES
#undef TESS_CASE
  raise(TESS_ERR_UNSUPPORTED, "unsupported (dim, dtype, basis) combination");
}

template <typename T, int D, int B, bool S>
struct P2GOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_weights* W, const tess_params& prm, const tess_sparsity* sp) {
    p2g_impl<T, D, B, S>(ctx, m, G, P, W, prm, sp);
  }
};
template <typename T, int D, int B, bool S>
struct G2POp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_weights* W, const tess_params& prm, const tess_sparsity* sp) {
    g2p_impl<T, D, B, S>(ctx, m, G, P, W, prm, sp);
  }
};
template <typename T, int D, int B, bool S>
struct G2P2GOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_weights* W, const tess_params& prm) {
    g2p2g_impl<T, D, B, S>(ctx, m, G, P, W, prm);
  }
};
template <typename T, int D, int B, bool S>
struct G2P2GStepsOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_params& prm, int nsteps, double thr) {
    if constexpr (!S) g2p2g_steps_impl<T, D, B>(ctx, m, G, P, prm, nsteps, thr);
  }
};
template <typename T, int D, int B, bool S>
struct StepOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_params& prm, int nsteps, double thr, tess_sparsity* sp, int* done) {
    if constexpr (!S) step_impl<T, D, B, false>(ctx, m, G, P, prm, nsteps, thr, sp, done);
  }
};

}  // namespace
}  // namespace tess
