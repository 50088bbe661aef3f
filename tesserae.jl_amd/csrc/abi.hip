// C ABI of libtesserae_hip: argument checks (mirroring the host-side checks of
// src/Transfer/runtime.jl:177-229 and src/Basis/basis_weight.jl:714-752),
// context management, and conversion of internal errors into tess_status +
// a thread-local message.  No exception crosses this boundary.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <new>
#include <string>

#include "tess_internal.hpp"

namespace tess {

namespace {
thread_local std::string g_last_error;
// failure injection (tess_debug_fail_after): the n-th checked HIP call from
// now on reports a failure; 0 is off
std::atomic<int64_t> g_fail_after{0};
// bumped by every call that returns an error: contexts and batches that saw
// an older value drop their cached binnings and zeroing bookkeeping
std::atomic<uint64_t> g_error_epoch{0};
}  // namespace

void raise(tess_status code, const std::string& msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) raise(TESS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  if (g_fail_after.load(std::memory_order_relaxed) > 0 && g_fail_after.fetch_sub(1) == 1)
    raise(TESS_ERR_DEVICE, std::string(what) + ": injected failure");
}

void check_hip_exact(hipError_t e, const char* what) {
  if (e != hipSuccess) raise(TESS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

uint64_t error_epoch() { return g_error_epoch.load(std::memory_order_relaxed); }

void* Buffer::get(size_t need) {
  if (need == 0) need = 1;
  if (need <= bytes) return ptr;
  if (ptr) {
    // the previous contents are dropped: callers treat scratch as uninitialised
    TESS_HIP(hipDeviceSynchronize());
    const hipError_t e = hipFree(ptr);
    ptr = nullptr;  // forgotten before any error is raised
    bytes = 0;
    TESS_HIP(e);
  }
  const size_t cap = need + need / 8;
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, cap);
  if (e == hipSuccess) {
    ptr = p;
    bytes = cap;
  }
  TESS_HIP(e);
  return ptr;
}

void Buffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int prof_begin(tess_ctx* ctx, const char* name, double bytes) {
  Profiler& P = ctx->prof;
  if (!P.on) return -1;
  int e = -1;
  for (size_t i = 0; i < P.entries.size(); ++i)
    if (P.entries[i].name == name) e = int(i);
  if (e < 0) {
    P.entries.push_back(ProfEntry{name});
    e = int(P.entries.size() - 1);
  }
  P.entries[e].bytes += bytes;
  P.entries[e].calls += 1;
  while (P.pool.size() < P.next_event + 2) {
    hipEvent_t ev = nullptr;
    const hipError_t err = hipEventCreate(&ev);
    if (err == hipSuccess) P.pool.push_back(ev);  // owned before any error is raised
    TESS_HIP(err);
  }
  hipEvent_t a = P.pool[P.next_event], b = P.pool[P.next_event + 1];
  P.next_event += 2;
  TESS_HIP(hipEventRecord(a, ctx->stream));
  P.pending.push_back(Profiler::Pending{e, a, b});
  return int(P.pending.size() - 1);
}

void prof_end(tess_ctx* ctx, int token) {
  if (token < 0) return;
  TESS_HIP(hipEventRecord(ctx->prof.pending[size_t(token)].b, ctx->stream));
}

void prof_flush(tess_ctx* ctx) {
  Profiler& P = ctx->prof;
  if (P.pending.empty()) return;
  TESS_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& q : P.pending) {
    float ms = 0.f;
    TESS_HIP(hipEventElapsedTime(&ms, q.a, q.b));
    P.entries[size_t(q.entry)].total_ms += ms;
  }
  P.pending.clear();
  P.next_event = 0;
}

namespace {

template <typename F>
tess_status guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return TESS_OK;
  } catch (const Error& e) {
    g_last_error = e.msg;
    g_error_epoch.fetch_add(1);
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    g_error_epoch.fetch_add(1);
    return TESS_ERR_DEVICE;
  } catch (...) {
    g_last_error = "unexpected internal error";
    g_error_epoch.fetch_add(1);
    return TESS_ERR_STATE;
  }
}

// A context's caches after a failed call somewhere: the binning, the grid's
// zeroing bookkeeping and the rollout's kept segment may describe work that a
// failure cut short, so the next call rebuilds them (full binning, whole-grid
// zero), after its streams have drained.
void ctx_drop_caches(tess_ctx* c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  else (void)hipDeviceSynchronize();
  c->bin.valid = false;
  c->bin.state_valid = false;
  c->bin.keys_ready = false;
  c->last_grid_m = nullptr;
  c->prev_nodeblocks = 0;
  c->misplaced_x_ptr = nullptr;
  c->last_misplaced = -1;
  c->rollout.last_kept = false;
}

void check_ctx(const tess_ctx* cctx) {
  if (!cctx) raise(TESS_ERR_ARGUMENT, "context is NULL");
  TESS_HIP(hipSetDevice(cctx->device));
  tess_ctx* ctx = const_cast<tess_ctx*>(cctx);
  const uint64_t ep = error_epoch();
  if (ctx->seen_epoch != ep) {
    ctx_drop_caches(ctx);
    if (ctx->replay) ctx_drop_caches(ctx->replay);
    ctx->seen_epoch = ep;
  }
}

void check_mesh(const tess_mesh* m) {
  if (!m) raise(TESS_ERR_ARGUMENT, "mesh is NULL");
  if (m->dim != 2 && m->dim != 3) raise(TESS_ERR_UNSUPPORTED, "mesh dimension must be 2 or 3");
  if (m->dtype != TESS_FLOAT32 && m->dtype != TESS_FLOAT64) raise(TESS_ERR_ARGUMENT, "mesh dtype must be Float32 or Float64");
  if (m->block_size_log2 < 0)
    raise(TESS_ERR_ARGUMENT, "block_size_log2 must be non-negative, got " + std::to_string(m->block_size_log2));
  for (int d = 0; d < m->dim; ++d)
    if (m->dims[d] < 2) raise(TESS_ERR_ARGUMENT, "a CartesianMesh needs at least 2 nodes per axis");
  if (!(m->h > 0) || !std::isfinite(m->h)) raise(TESS_ERR_ARGUMENT, "mesh spacing must be positive");
  if (!(m->h_inv > 0) || !std::isfinite(m->h_inv)) raise(TESS_ERR_ARGUMENT, "mesh inverse spacing must be positive");
}

// A kernel, optionally wrapped in one of WLS / KernelCorrection (the
// polynomial flag only with a wrapper).
void check_basis(int32_t basis) {
  const int32_t k = basis & 0xff, flags = basis & ~0xff;
  if (k != TESS_BSPLINE_LINEAR && k != TESS_BSPLINE_QUADRATIC && k != TESS_BSPLINE_CUBIC && k != TESS_UGIMP)
    raise(TESS_ERR_UNSUPPORTED, "basis is not one of BSpline(Linear/Quadratic/Cubic) or uGIMP");
  const int32_t wrap = flags & (TESS_WLS | TESS_KERNEL_CORRECTION);
  if ((flags & ~(TESS_WLS | TESS_KERNEL_CORRECTION | TESS_POLY_MULTILINEAR)) != 0 ||
      wrap == (TESS_WLS | TESS_KERNEL_CORRECTION) || (wrap == 0 && flags != 0))
    raise(TESS_ERR_UNSUPPORTED, "basis flags: one of TESS_WLS / TESS_KERNEL_CORRECTION, with an optional "
                                "TESS_POLY_MULTILINEAR");
}

void check_plain_basis(int32_t basis, const char* what) {
  check_basis(basis);
  if (basis & ~0xff) raise(TESS_ERR_UNSUPPORTED, std::string(what) + " supports the plain kernels only");
}

void need(const tess_field& f, const char* what) {
  if (!f.ptr) raise(TESS_ERR_ARGUMENT, std::string("missing field ") + what);
}

void check_particles(const tess_particles* P, int32_t basis, bool transfer) {
  if (!P) raise(TESS_ERR_ARGUMENT, "particles is NULL");
  if (P->n < 0) raise(TESS_ERR_DIMENSION, "negative particle count");
  need(P->x, "x");
  if (transfer) {
    need(P->v, "v");
    need(P->V, "V");
    need(P->grad_v, "∇v");
    need(P->F, "F");
    need(P->sigma, "σ");
  }
  if ((basis & 0xff) == TESS_UGIMP && !P->l.ptr)
    raise(TESS_ERR_ARGUMENT, "uGIMP requires the initial particle length `l` in the particle property");
}

void check_grid(const tess_grid* G) {
  if (!G) raise(TESS_ERR_ARGUMENT, "grid is NULL");
  need(G->m, "m");
  need(G->mv, "mv");
  need(G->f, "f");
  need(G->vn, "vⁿ");
  need(G->v, "v");
}

void check_weights(const tess_weights* W) {
  if (!W || !W->w.ptr) return;  // deferred (a weights struct without storage carries only a filter)
  if (!W->window) raise(TESS_ERR_ARGUMENT, "stored weights need the support-window array");
  need(W->w, "w");
  need(W->grad_w, "∇w");
}

void check_params(const tess_params* p) {
  if (!p) raise(TESS_ERR_ARGUMENT, "params is NULL");
  if (!std::isfinite(p->dt)) raise(TESS_ERR_ARGUMENT, "Δt must be finite");
}

// SpIndices over the mesh's blocks; transfers additionally need the storage
// to hold the active blocks.
void check_sparsity(const tess_mesh* m, const tess_sparsity* sp, bool transfer) {
  if (!sp) raise(TESS_ERR_ARGUMENT, "sparsity is NULL");
  if (!sp->blocknumbering) raise(TESS_ERR_ARGUMENT, "blocknumbering is NULL");
  if (m->block_size_log2 < 1 || m->block_size_log2 * m->dim > 30)
    raise(TESS_ERR_UNSUPPORTED, "SpGrid blocks need 1 <= block_size_log2 and 2^(block_size_log2*dim) <= 2^30");
  if (transfer) {
    if (sp->storage_len < 0 || sp->capacity < 0) raise(TESS_ERR_DIMENSION, "negative SpArray storage length");
    if (sp->storage_len > sp->capacity)
      raise(TESS_ERR_DIMENSION, "SpArray storage (" + std::to_string(sp->capacity) +
                                    " nodes) is smaller than the active blocks (" + std::to_string(sp->storage_len) +
                                    " nodes)");
  }
}

}  // namespace
}  // namespace tess

using namespace tess;

namespace tess {

tess_ctx* ctx_new(int device, hipStream_t stream) {
  tess_ctx* c = new tess_ctx();
  c->device = device;
  c->stream = stream;
  c->seen_epoch = error_epoch();
  void* pinned = nullptr;
  // pinned readback slots: counters (words 0-63), mover counts (words 64-319)
  hipError_t e = hipHostMalloc(&pinned, 512 * sizeof(int64_t), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete c;
    check_hip(e, "hipHostMalloc");
  }
  c->counters_host = static_cast<int64_t*>(pinned);
  try {
    // device counters start at zero (the binning reads slots written by the previous P2G)
    TESS_HIP(hipMemset(c->counters.get(64 * sizeof(uint32_t)), 0, 64 * sizeof(uint32_t)));
  } catch (...) {
    ctx_free(c);
    throw;
  }
  return c;
}

void ctx_free(tess_ctx* ctx) {
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  if (ctx->replay) {
    hipStream_t rs = ctx->replay->stream;
    ctx_free(ctx->replay);
    if (rs) (void)hipStreamDestroy(rs);
    ctx->replay = nullptr;
  }
  tess::Buffer* bufs[] = {&ctx->keys,      &ctx->keys_sorted, &ctx->vals,         &ctx->keys_alt,
                          &ctx->vals_alt,  &ctx->hist,        &ctx->scan_tmp,     &ctx->tile_start,
                          &ctx->tile_end,  &ctx->tile_flag,   &ctx->tile_slot,    &ctx->active_tiles,
                          &ctx->nb_flag,   &ctx->nb_slot,     &ctx->nb_list,      &ctx->nb_list_prev,
                          &ctx->partials,  &ctx->counters,    &ctx->host_pinned_dummy, &ctx->reorder_tmp, &ctx->reorder_cnt,
                          &ctx->per_tile,  &ctx->per_nodeblock, &ctx->skeys2, &ctx->perm2,
                          &ctx->inc_nk,    &ctx->inc_spos,  &ctx->inc_mkey, &ctx->inc_mj,
                          &ctx->inc_mks,   &ctx->inc_mr,    &ctx->inc_mjs, &ctx->inc_rng,
                          &ctx->sp_occ,    &ctx->sp_act,    &ctx->sp_scan,
                          &ctx->adj_partials, &ctx->adj_nodes, &ctx->ro_checkpoints, &ctx->ro_segment,
                          &ctx->ro_gradv,  &ctx->ro_snapshots, &ctx->ro_grid, &ctx->movers, &ctx->g2p2g_x,
                          &ctx->export_buf};
  for (auto* b : bufs) b->release();
  for (auto ev : ctx->prof.pool) (void)hipEventDestroy(ev);
  if (ctx->counters_host) (void)hipHostFree(ctx->counters_host);
  delete ctx;
}

}  // namespace tess

extern "C" {

int32_t tess_abi_version(void) { return TESS_ABI_VERSION; }

const char* tess_last_error(void) { return g_last_error.c_str(); }

#ifndef TESS_DEBUG
#define TESS_DEBUG 0
#endif
int32_t tess_build_flags(void) { return TESS_DEBUG ? TESS_BUILD_DEBUG_CHECKS : 0; }

tess_status tess_debug_fail_after(int64_t n) {
  return guarded([&] {
    if (n < 0) raise(TESS_ERR_ARGUMENT, "n must be non-negative");
    g_fail_after.store(n);
  });
}

tess_status tess_ctx_create(int32_t device, void* stream, tess_ctx** out) {
  return guarded([&] {
    if (!out) raise(TESS_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    int count = 0;
    TESS_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) raise(TESS_ERR_ARGUMENT, "no such HIP device");
    TESS_HIP(hipSetDevice(device));
    tess_ctx* c = ctx_new(device, static_cast<hipStream_t>(stream));
    try {
      // the allocation-free partition's work area on this device
      noalloc_scratch_resolve();
    } catch (...) {
      ctx_free(c);
      throw;
    }
    *out = c;
  });
}

tess_status tess_ctx_destroy(tess_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx_free(ctx);
  });
}

tess_status tess_sync(tess_ctx* ctx) {
  return guarded([&] {
    check_ctx(ctx);
    TESS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

tess_status tess_update_weights(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_particles* particles,
                                const tess_weights* weights) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_particles(particles, basis, false);
    if (!weights) raise(TESS_ERR_ARGUMENT, "weights is NULL");
    if (!weights->w.ptr) raise(TESS_ERR_ARGUMENT, "update! of stored weights needs the w array");
    check_weights(weights);
    if (weights->filter && !(basis & (TESS_WLS | TESS_KERNEL_CORRECTION)))  // require_filtered_update_support
      raise(TESS_ERR_ARGUMENT, "the basis does not support filtered updates");
    launch_update_weights(ctx, basis, *mesh, *particles, *weights);
  });
}

tess_status tess_partition_update(tess_ctx* ctx, const tess_mesh* mesh, tess_field x, int64_t n,
                                  int32_t* particleindices, int32_t* starts, int32_t* stops, int64_t* nassigned) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    if (n < 0) raise(TESS_ERR_DIMENSION, "negative particle count");
    if (!x.ptr && n > 0) raise(TESS_ERR_ARGUMENT, "missing field x");
    if (!starts || !stops || !nassigned || (!particleindices && n > 0))
      raise(TESS_ERR_ARGUMENT, "partition output arrays must not be NULL");
    run_partition_update(ctx, *mesh, x, n, particleindices, starts, stops, nassigned);
  });
}

tess_status tess_partition_update_noalloc(const tess_mesh* mesh, tess_field x, int64_t n, int32_t* particleindices,
                                          int32_t* starts, int32_t* stops, int64_t* nassigned, void* stream) {
  return guarded([&] {
    check_mesh(mesh);
    if (n < 0) raise(TESS_ERR_DIMENSION, "negative particle count");
    if (n > partition_noalloc_max_particles())
      raise(TESS_ERR_ARGUMENT, "tess_partition_update_noalloc handles at most " +
                                   std::to_string(partition_noalloc_max_particles()) + " particles, got " +
                                   std::to_string(n));
    if (!x.ptr && n > 0) raise(TESS_ERR_ARGUMENT, "missing field x");
    if (!starts || !stops || !nassigned || (!particleindices && n > 0))
      raise(TESS_ERR_ARGUMENT, "partition output arrays must not be NULL");
    run_partition_update_noalloc(*mesh, x, n, particleindices, starts, stops, nassigned,
                                 static_cast<hipStream_t>(stream));
  });
}

tess_status tess_reorder_particles(tess_ctx* ctx, const tess_mesh* mesh, const tess_particles* particles,
                                   int32_t* particleindices, int64_t nassigned, double threshold,
                                   int32_t* reordered) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    if (!particles) raise(TESS_ERR_ARGUMENT, "particles is NULL");
    if (!reordered) raise(TESS_ERR_ARGUMENT, "reordered is NULL");
    if (nassigned < 0) raise(TESS_ERR_DIMENSION, "negative nassigned");
    if (nassigned > 0 && !particleindices) raise(TESS_ERR_ARGUMENT, "particleindices is NULL");
    run_reorder(ctx, *mesh, *particles, particleindices, nassigned, threshold, reordered);
  });
}

tess_status tess_p2g(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                     const tess_particles* particles, const tess_weights* weights, const tess_params* params) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_weights(weights);
    check_params(params);
    run_p2g(ctx, basis, *mesh, *grid, *particles, weights, *params);
  });
}

tess_status tess_g2p2g(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                       const tess_particles* particles, const tess_weights* weights, const tess_params* params) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_weights(weights);
    check_params(params);
    run_g2p2g(ctx, basis, *mesh, *grid, *particles, weights, *params);
  });
}

tess_status tess_g2p(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                     const tess_particles* particles, const tess_weights* weights, const tess_params* params) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_weights(weights);
    check_params(params);
    run_g2p(ctx, basis, *mesh, *grid, *particles, weights, *params);
  });
}

tess_status tess_g2p2g_steps(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                             const tess_particles* particles, const tess_params* params, int32_t nsteps,
                             double reorder_threshold) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_params(params);
    if (nsteps < 0) raise(TESS_ERR_ARGUMENT, "nsteps must be non-negative");
    if (!(reorder_threshold >= 0.0 && reorder_threshold <= 1.0))
      raise(TESS_ERR_ARGUMENT, "threshold must be in [0, 1].");
    run_g2p2g_steps(ctx, basis, *mesh, *grid, *particles, *params, nsteps, reorder_threshold);
  });
}

tess_status tess_step(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                      const tess_particles* particles, const tess_params* params, int32_t nsteps,
                      double reorder_threshold) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_params(params);
    if (nsteps < 0) raise(TESS_ERR_ARGUMENT, "nsteps must be non-negative");
    if (!(reorder_threshold >= 0.0 && reorder_threshold <= 1.0))
      raise(TESS_ERR_ARGUMENT, "threshold must be in [0, 1].");
    run_step(ctx, basis, *mesh, *grid, *particles, *params, nsteps, reorder_threshold);
  });
}

tess_status tess_sparsity_update(tess_ctx* ctx, const tess_mesh* mesh, tess_field x, int64_t n, tess_sparsity* sp,
                                 int32_t* changed) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_sparsity(mesh, sp, false);
    if (n < 0) raise(TESS_ERR_DIMENSION, "negative particle count");
    if (!x.ptr && n > 0) raise(TESS_ERR_ARGUMENT, "missing field x");
    run_sparsity_update(ctx, *mesh, x, n, *sp, changed);
  });
}

tess_status tess_p2g_sparse(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                            const tess_sparsity* sp, const tess_particles* particles, const tess_weights* weights,
                            const tess_params* params) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_sparsity(mesh, sp, true);
    check_particles(particles, basis, true);
    check_weights(weights);
    check_params(params);
    run_p2g_sparse(ctx, basis, *mesh, *grid, *sp, *particles, weights, *params);
  });
}

tess_status tess_g2p_sparse(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                            const tess_sparsity* sp, const tess_particles* particles, const tess_weights* weights,
                            const tess_params* params) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_sparsity(mesh, sp, true);
    check_particles(particles, basis, true);
    check_weights(weights);
    check_params(params);
    run_g2p_sparse(ctx, basis, *mesh, *grid, *sp, *particles, weights, *params);
  });
}

tess_status tess_step_sparse(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                             tess_sparsity* sp, const tess_particles* particles, const tess_params* params,
                             int32_t nsteps, double reorder_threshold, int32_t* steps_done) {
  int done = 0;
  if (steps_done) *steps_done = 0;
  const tess_status st = guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_basis(basis);
    check_grid(grid);
    check_sparsity(mesh, sp, false);
    check_particles(particles, basis, true);
    check_params(params);
    if (nsteps < 0) raise(TESS_ERR_ARGUMENT, "nsteps must be non-negative");
    if (!(reorder_threshold >= 0.0 && reorder_threshold <= 1.0))
      raise(TESS_ERR_ARGUMENT, "threshold must be in [0, 1].");
    run_step_sparse(ctx, basis, *mesh, *grid, *sp, *particles, *params, nsteps, reorder_threshold, &done);
  });
  if (steps_done) *steps_done = done;
  return st;
}

namespace tess {
namespace {
void check_adjoint_basis(int32_t basis) {
  check_plain_basis(basis, "the adjoint");
  if (basis == TESS_UGIMP) raise(TESS_ERR_UNSUPPORTED, "the adjoint supports the B-spline bases");
}

void check_adjoint(const tess_adjoint* adj, bool need_state) {
  if (!adj) raise(TESS_ERR_ARGUMENT, "adjoint is NULL");
  if (need_state) {
    need(adj->x, "x̄");
    need(adj->v, "v̄");
    need(adj->F, "F̄");
  }
}
}  // namespace
}  // namespace tess

tess_status tess_step_adjoint(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                              const tess_particles* particles, const tess_params* params, tess_adjoint* adj) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_adjoint_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_params(params);
    check_adjoint(adj, true);
    run_step_adjoint(ctx, basis, *mesh, *grid, *particles, *params, *adj);
  });
}

tess_status tess_rollout(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                         const tess_particles* particles, const tess_params* params, int32_t nsteps,
                         int32_t checkpoint_every) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_adjoint_basis(basis);
    check_grid(grid);
    check_particles(particles, basis, true);
    check_params(params);
    if (nsteps < 1) raise(TESS_ERR_ARGUMENT, "a rollout needs at least one step");
    if (checkpoint_every < 1) raise(TESS_ERR_ARGUMENT, "checkpoint_every must be positive");
    run_rollout(ctx, basis, *mesh, *grid, *particles, *params, nsteps, checkpoint_every);
  });
}

tess_status tess_rollout_adjoint(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                                 const tess_params* params, tess_adjoint* adj) {
  return guarded([&] {
    check_ctx(ctx);
    check_mesh(mesh);
    check_adjoint_basis(basis);
    check_grid(grid);
    check_params(params);
    check_adjoint(adj, true);
    run_rollout_adjoint(ctx, basis, *mesh, *grid, *params, *adj);
  });
}

tess_status tess_batch_create(tess_ctx* ctx, int32_t basis, int32_t nscenes, const tess_scene* scenes,
                              tess_batch** out) {
  return guarded([&] {
    check_ctx(ctx);
    check_plain_basis(basis, "a batch");
    if (!out) raise(TESS_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (nscenes < 1 || !scenes) raise(TESS_ERR_ARGUMENT, "a batch needs at least one scene");
    for (int32_t s = 0; s < nscenes; ++s) {
      check_mesh(&scenes[s].mesh);
      check_grid(&scenes[s].grid);
      check_particles(&scenes[s].particles, basis, true);
      check_params(&scenes[s].params);
    }
    run_batch_create(ctx, basis, nscenes, scenes, out);
  });
}

tess_status tess_batch_set_reorder(tess_batch* batch, double threshold) {
  return guarded([&] {
    if (!batch) raise(TESS_ERR_ARGUMENT, "batch is NULL");
    if (!(threshold >= 0.0 && threshold <= 1.0)) raise(TESS_ERR_ARGUMENT, "threshold must be in [0, 1].");
    run_batch_set_reorder(batch, threshold);
  });
}

tess_status tess_batch_step(tess_ctx* ctx, tess_batch* batch, int32_t nsteps) {
  return guarded([&] {
    check_ctx(ctx);
    if (!batch) raise(TESS_ERR_ARGUMENT, "batch is NULL");
    if (nsteps < 0) raise(TESS_ERR_ARGUMENT, "nsteps must be non-negative");
    run_batch_step(ctx, batch, nsteps);
  });
}

tess_status tess_batch_autotune(tess_ctx* ctx, tess_batch* batch, int32_t steps_per_variant, int32_t* variant) {
  return guarded([&] {
    check_ctx(ctx);
    if (!batch) raise(TESS_ERR_ARGUMENT, "batch is NULL");
    if (steps_per_variant < 1) raise(TESS_ERR_ARGUMENT, "steps_per_variant must be positive");
    const int dim = batch->dim;
    // 2D: P2G split (bit 0) × G2P with 4, 2 or 1 waves per tile; 3D: the P2G
    // split is fixed, G2P with 4 or 2 waves; each launched eagerly, as
    // captured graph replays (bit 3) with one workgroup per possible tile, and
    // as graph replays of persistent launches (bits 3 and 4)
    const int cand2[18] = {0, 1, 2, 3, 4, 5, 8, 9, 10, 11, 12, 13, 24, 25, 26, 27, 28, 29};
    const int cand3[6] = {0, 2, 8, 10, 24, 26};
    const int ncand = dim == 2 ? 18 : 6;
    // the caller's reorder threshold and variant come back and the events are
    // released on every exit, a failed step included
    struct Restore {
      tess_batch* b;
      double thr;
      int variant;
      hipEvent_t e[2] = {nullptr, nullptr};
      bool done = false;
      ~Restore() {
        for (hipEvent_t ev : e)
          if (ev) (void)hipEventDestroy(ev);
        if (!done) {
          run_batch_set_variant(b, variant);
          run_batch_set_reorder(b, thr);
        }
      }
    } guard{batch, batch->reorder_threshold, run_batch_variant(batch)};
    TESS_HIP(hipEventCreate(&guard.e[0]));
    TESS_HIP(hipEventCreate(&guard.e[1]));
    hipEvent_t e0 = guard.e[0], e1 = guard.e[1];
    float best = 0.0f;
    int best_v = guard.variant;
    // with a reorder threshold set, every variant is timed on a freshly
    // reordered layout and without reorders in its timed steps
    const double thr = guard.thr;
    for (int k = 0; k < ncand; ++k) {
      const int v = dim == 2 ? cand2[k] : cand3[k];
      run_batch_set_variant(batch, v);
      if (thr > 0.0) run_batch_set_reorder(batch, 1.0);
      run_batch_step(ctx, batch, 1);  // untimed: code loading, first-use allocations, the reorder
      run_batch_set_reorder(batch, 0.0);
      TESS_HIP(hipEventRecord(e0, ctx->stream));
      run_batch_step(ctx, batch, steps_per_variant);
      TESS_HIP(hipEventRecord(e1, ctx->stream));
      TESS_HIP(hipEventSynchronize(e1));
      run_batch_set_reorder(batch, thr);
      float ms = 0.0f;
      TESS_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (k == 0 || ms < best) {
        best = ms;
        best_v = v;
      }
    }
    guard.done = true;
    run_batch_set_variant(batch, best_v);
    if (variant) *variant = best_v;
  });
}

tess_status tess_batch_set_variant(tess_batch* batch, int32_t variant) {
  return guarded([&] {
    if (!batch) raise(TESS_ERR_ARGUMENT, "batch is NULL");
    // bit 3: graph mode; bit 4 (with bit 3): its persistent launches
    const int v = variant & ~24;
    const bool ok = variant >= 0 && ((variant & 24) != 16) && (batch->dim == 2 ? v <= 5 : (v == 0 || v == 2));
    if (!ok) raise(TESS_ERR_ARGUMENT, "no launch variant " + std::to_string(variant) + " for a " +
                                          std::to_string(batch->dim) + "D batch");
    run_batch_set_variant(batch, variant);
  });
}

tess_status tess_batch_stats(const tess_batch* batch, int64_t* out, int32_t n) {
  return guarded([&] {
    if (!batch) raise(TESS_ERR_ARGUMENT, "batch is NULL");
    if (n < 0 || (n > 0 && !out)) raise(TESS_ERR_ARGUMENT, "out is NULL");
    run_batch_stats(batch, out, n);
  });
}

tess_status tess_batch_destroy(tess_batch* batch) {
  return guarded([&] {
    if (batch) run_batch_destroy(batch);
  });
}

tess_status tess_export_field(tess_ctx* ctx, int32_t dtype, int32_t dim, tess_field field, int64_t n, int32_t kind,
                              float* dst) {
  return guarded([&] {
    check_ctx(ctx);
    if (dtype != TESS_FLOAT32 && dtype != TESS_FLOAT64) raise(TESS_ERR_ARGUMENT, "dtype");
    if (dim != 2 && dim != 3) raise(TESS_ERR_UNSUPPORTED, "dim must be 2 or 3");
    if (n < 0) raise(TESS_ERR_DIMENSION, "negative element count");
    if (n > 0 && (!field.ptr || !dst)) raise(TESS_ERR_ARGUMENT, "field or destination is NULL");
    run_export_field(ctx, dtype, dim, field, n, kind, dst);
  });
}

tess_status tess_vtk_write_points(const char* path, int64_t n, const float* xyz, int32_t narrays,
                                  const tess_vtk_array* arrays, int32_t compress) {
  return guarded([&] {
    if (!path) raise(TESS_ERR_ARGUMENT, "path is NULL");
    if (n < 0) raise(TESS_ERR_DIMENSION, "negative point count");
    if (n > 0 && !xyz) raise(TESS_ERR_ARGUMENT, "point coordinates are NULL");
    run_vtk_write_points(path, n, xyz, narrays, arrays, compress);
  });
}

tess_status tess_vtk_write_mesh(const char* path, const tess_mesh* mesh, int32_t narrays,
                                const tess_vtk_array* arrays, int32_t compress) {
  return guarded([&] {
    if (!path) raise(TESS_ERR_ARGUMENT, "path is NULL");
    check_mesh(mesh);
    run_vtk_write_mesh(path, *mesh, narrays, arrays, compress);
  });
}

tess_status tess_vtm_write(const char* path, int32_t nblocks, const char* const* files) {
  return guarded([&] {
    if (!path) raise(TESS_ERR_ARGUMENT, "path is NULL");
    run_vtm_write(path, nblocks, files);
  });
}

tess_status tess_pvd_write(const char* path, int32_t n, const double* times, const char* const* files) {
  return guarded([&] {
    if (!path) raise(TESS_ERR_ARGUMENT, "path is NULL");
    run_pvd_write(path, n, times, files);
  });
}

tess_status tess_ply_write(const char* name, int64_t n, const float* xyz) {
  return guarded([&] {
    if (!name) raise(TESS_ERR_ARGUMENT, "name is NULL");
    run_ply_write(name, n, xyz);
  });
}

tess_status tess_profile_enable(tess_ctx* ctx, int32_t on) {
  return guarded([&] {
    check_ctx(ctx);
    prof_flush(ctx);
    ctx->prof.on = on != 0;
    ctx->prof.entries.clear();
  });
}

tess_status tess_profile_read(tess_ctx* ctx, int32_t index, char* name, int32_t name_len, double* total_ms,
                              int64_t* calls, double* bytes) {
  return guarded([&] {
    check_ctx(ctx);
    prof_flush(ctx);
    const auto& E = ctx->prof.entries;
    if (index < 0 || size_t(index) >= E.size()) raise(TESS_ERR_ARGUMENT, "no such profile entry");
    const ProfEntry& e = E[size_t(index)];
    if (name && name_len > 0) {
      const size_t n = std::min(size_t(name_len - 1), e.name.size());
      e.name.copy(name, n);
      name[n] = '\0';
    }
    if (total_ms) *total_ms = e.total_ms;
    if (calls) *calls = e.calls;
    if (bytes) *bytes = e.bytes;
  });
}

tess_status tess_device_copy(tess_ctx* ctx, const void* src, void* dst, int64_t bytes) {
  return guarded([&] {
    check_ctx(ctx);
    if (bytes < 0 || bytes % 16) raise(TESS_ERR_ARGUMENT, "bytes must be a non-negative multiple of 16");
    if (bytes > 0 && (!src || !dst)) raise(TESS_ERR_ARGUMENT, "src or dst is NULL");
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15)
      raise(TESS_ERR_ARGUMENT, "src and dst must be 16-byte aligned");
    run_device_copy(ctx, src, dst, bytes);
  });
}

tess_status tess_stats(tess_ctx* ctx, int64_t* out, int32_t count) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out || count < 0) raise(TESS_ERR_ARGUMENT, "out is NULL");
    const int64_t vals[] = {ctx->bin.n_inside, ctx->bin.n_active, ctx->bin.n_nodeblocks, ctx->last_active_nodes,
                            ctx->reorders, ctx->last_misplaced, ctx->last_movers, ctx->kept_binnings};
    for (int32_t i = 0; i < count && i < int32_t(sizeof(vals) / sizeof(vals[0])); ++i) out[i] = vals[i];
  });
}

}  // extern "C"
