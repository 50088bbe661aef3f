/*
 * tesserae_hip.h — C ABI of libtesserae_hip, the MI355X (gfx950) engine for the
 * explicit material-point-method step of Tesserae.jl v0.7.5.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns a
 * tess_status, never throws or aborts across the boundary, and enqueues its GPU
 * work asynchronously on the stream bound to the context.  Only calls that say
 * so synchronise (they read a small total back to the host, as the reference's
 * GPU partition does at src/Parallel/gpu_block_strategy.jl:161-165).
 *
 * Conventions shared by all calls
 *  - Arrays are caller-owned device memory (Julia-GC-owned CuArray/ROCArray on
 *    the reference side).  The library never frees or reallocates them; its own
 *    scratch (sort buffers, tile partials) belongs to the tess_ctx.
 *  - A field is described by tess_field: component c of element e lives at
 *    ptr + (e*elem_stride + c*comp_stride) scalars.  Julia-native layouts
 *    (Vec{d,T} arrays: elem_stride=d, comp_stride=1; Mat{d,d,T}: elem_stride=d*d,
 *    comp_stride=1, column-major components) and pure SoA (elem_stride=1,
 *    comp_stride=n) are both accepted.
 *  - Grid nodes are numbered x-fastest (Julia column-major Array{T,dim}).
 *  - Scalars are the mesh's dtype (TESS_FLOAT32 / TESS_FLOAT64).
 *
 * Error classes mirror the exceptions the reference raises on the host before
 * any launch (src/Transfer/runtime.jl:177-229, src/Basis/basis_weight.jl:714-752,
 * src/Parallel/gpu_block_strategy.jl:147-148): the Julia binding rethrows
 * ArgumentError / DimensionMismatch / AssertionError / ErrorException.
 */
#ifndef TESSERAE_HIP_H
#define TESSERAE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TESS_ABI_VERSION 4

typedef int32_t tess_status;
enum {
  TESS_OK = 0,
  TESS_ERR_ARGUMENT = 1,    /* Julia ArgumentError */
  TESS_ERR_DIMENSION = 2,   /* Julia DimensionMismatch */
  TESS_ERR_ASSERTION = 3,   /* Julia AssertionError (the reference's @assert) */
  TESS_ERR_STATE = 4,       /* Julia ErrorException (the reference's error(...)) */
  TESS_ERR_DEVICE = 5,      /* HIP runtime failure (launch, allocation, copy) */
  TESS_ERR_UNSUPPORTED = 6  /* program/basis/dimension outside the library's fixed set */
};

enum { TESS_FLOAT32 = 0, TESS_FLOAT64 = 1 };

/* Bases: BSpline(Linear()), BSpline(Quadratic()), BSpline(Cubic()) and uGIMP()
 * (src/Basis/bspline.jl:15-40, src/Basis/gimp.jl:16-40). */
enum {
  TESS_BSPLINE_LINEAR = 1,
  TESS_BSPLINE_QUADRATIC = 2,
  TESS_BSPLINE_CUBIC = 3,
  TESS_UGIMP = 16
};

/* Corrected bases: one of the kernel codes above OR'd with a wrapper and,
 * optionally, the polynomial flag.
 *   WLS(kernel[, poly]) (src/Basis/wls.jl:1-17): the weighted least-squares fit
 *     of poly over the kernel's support; poly defaults to Polynomial(Linear()).
 *   KernelCorrection(kernel[, poly]) (src/Basis/kernelcorrection.jl:1-24): the
 *     plain kernel on a whole support, WLS(kernel, poly) where the mesh (or the
 *     node filter) cuts it; poly defaults to Polynomial(MultiLinear()).
 * TESS_POLY_MULTILINEAR selects Polynomial(MultiLinear()) (1, x, y, xy[, z, yz,
 * zx, xyz], src/Basis/polynomials.jl:42-46); without it the polynomial is
 * Polynomial(Linear()) (1, x[, y, z]).  Corrected weights carry w and ∇w. */
enum {
  TESS_WLS = 0x100,
  TESS_KERNEL_CORRECTION = 0x200,
  TESS_POLY_MULTILINEAR = 0x400
};

/* CartesianMesh{dim,T} (src/Mesh/cartesian.jl:38-42, ctor :89-103).
 * xmin and h are the Float64 values the mesh was built from; the library uses
 * T(xmin), T(h), T(h_inv) exactly as the device copy of the mesh holds them
 * (src/adapt.jl:53-57), and node coordinates T(xmin + i*h). */
typedef struct tess_mesh {
  int32_t dim;              /* 2 or 3 */
  int32_t dtype;            /* TESS_FLOAT32 / TESS_FLOAT64 */
  int32_t block_size_log2;  /* cartesian.jl:11, blocks of 2^L cells per axis */
  int32_t reserved;
  int64_t dims[3];          /* nodes per axis, size(mesh) */
  double xmin[3];
  double h;                 /* spacing(mesh) */
  double h_inv;             /* spacing_inv(mesh) = inv(h) */
} tess_mesh;

typedef struct tess_field {
  void* ptr;
  int64_t elem_stride;      /* in scalars */
  int64_t comp_stride;      /* in scalars */
} tess_field;

/* Particle StructArray of the README program (README.md:53-54):
 * {x::Vec, V::T, v::Vec, ∇v::Mat, F::Mat, σ::Mat}; l is the uGIMP particle
 * length (src/Basis/gimp.jl:25); logE is an optional per-particle stiffness
 * exponent, E_p = E*exp(logE_p) (NULL ptr: E_p = E). */
typedef struct tess_particles {
  int64_t n;
  tess_field x, v, V, grad_v, F, sigma;
  tess_field l;
  tess_field logE;
} tess_particles;

/* Grid StructArray of the README program: {m::T, mv::Vec, f::Vec, vⁿ::Vec, v::Vec}. */
typedef struct tess_grid {
  tess_field m, mv, f, vn, v;
} tess_grid;

/* BasisWeightArray in stored mode (src/Basis/basis_weight.jl:345-358):
 * w is Array{T,dim+1} of size (sw,...,sw,n), grad_w Array{Vec{dim,T},dim+1},
 * window the per-particle CartesianIndices as dim (start, stop) Int64 pairs,
 * 1-based inclusive, (1,0) when empty.  Entry ip of particle p is element
 * ip + sw^dim * p, ip being the column-major position of the node inside the
 * (possibly truncated) window placed at the origin of the sw^dim storage.
 * filter is the node mask of update!(weights, particles, mesh, filter)
 * (basis_weight.jl:714-752): one byte per mesh node in column-major node
 * order, nonzero = the node takes part in the corrected fit; NULL = Trues.
 * Only corrected bases accept a filter (supports_filtered_updates).
 * A tess_weights whose w.ptr is NULL stands for deferred weights that
 * recorded a filter (update!(...; deferred=true), basis_weight.jl:398,716-721):
 * transfers evaluate the basis and read only filter. */
typedef struct tess_weights {
  tess_field w;
  tess_field grad_w;
  int64_t* window;
  const uint8_t* filter;
} tess_weights;

/* Constants of the README program (README.md:46-49, 71-87) plus the config
 * extras: gravity (applied as an acceleration on nodes with m>0) and slip walls
 * (bit 2*d: low face of axis d, bit 2*d+1: high face; the normal component of
 * the grid velocity v is zeroed on that face's nodes, docs/src/gpu.md:116-129). */
typedef struct tess_params {
  double dt;
  double lambda;            /* Lamé λ = Eν/((1+ν)(1-2ν)) */
  double mu;                /* Lamé μ = E/(2(1+ν)) */
  double mp;                /* particle mass mₚ = ρ⁰V⁰ₚ */
  double V0;                /* initial particle volume V⁰ₚ */
  double gravity[3];
  int32_t slip_walls;
  int32_t reserved;
} tess_params;

typedef struct tess_ctx tess_ctx;

/* ---- context, errors ---------------------------------------------------- */

/* Replaces the backend glue of ext/TesseraeCUDAExt.jl:11-39 (backend <-> device,
 * workgroup sizing, temporaries).  stream may be NULL (default stream). */
tess_status tess_ctx_create(int32_t device, void* stream, tess_ctx** out);
tess_status tess_ctx_destroy(tess_ctx* ctx);
/* Blocks until the context's stream is idle (KernelAbstractions.synchronize). */
tess_status tess_sync(tess_ctx* ctx);
/* Message of the last failed call on this host thread ("" if none). */
const char* tess_last_error(void);
int32_t tess_abi_version(void);
/* Failure injection for tests of the error paths (no counterpart in the
 * reference, whose tests inject `error("boom")` into equations and phases,
 * test/transfer.jl:587-604, test/parallel_region.jl:76-170): the n-th HIP
 * call the library checks from now on, in this process, reports
 * TESS_ERR_DEVICE ("injected failure") after the call itself went through; 0
 * turns it off, negative n is TESS_ERR_ARGUMENT.  Any failed call makes every
 * context and batch rebuild its cached binning and zero its grids on its next
 * call, so a failure never leaves a later call on stale state. */
tess_status tess_debug_fail_after(int64_t n);
/* How this build was made: TESS_BUILD_DEBUG_CHECKS for the debug build
 * (`make debug`, libtesserae_hip_debug.so), whose every binning is read back
 * and its invariants checked on the host (permutation, sorted keys, tile runs,
 * active-tile list), failing with TESS_ERR_STATE; the counterpart of the
 * reference's debug_mode preference (src/Core/debug.jl:5-15). */
#define TESS_BUILD_DEBUG_CHECKS 1
int32_t tess_build_flags(void);

/* ---- basis weights ------------------------------------------------------ */

/* update!(weights, particles, mesh[, filter]) in stored mode
 * (src/Basis/basis_weight.jl:714-752; per particle :634-654, B-spline fast path
 * src/Basis/bspline.jl:102-170, truncated windows basis_weight.jl:624-632;
 * corrected bases src/Basis/wls.jl:19-128, kernelcorrection.jl:26-35).
 * Writes w, grad_w and window for particles 0..n-1.  A weights->filter on a
 * basis that is not corrected fails with TESS_ERR_ARGUMENT. */
tess_status tess_update_weights(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh,
                                const tess_particles* particles, const tess_weights* weights);

/* ---- partition ---------------------------------------------------------- */

/* update!(partition, x) with the CPU block strategy's result
 * (src/Parallel/cpu_block_strategy.jl:62-219): particleindices = inside-mesh
 * particle ids (1-based) stably sorted by (block colour, linear block id);
 * starts/stops (1-based, 0 for empty blocks) per block in linear block order;
 * *nassigned (host) = number of assigned particles.  Synchronises. */
tess_status tess_partition_update(tess_ctx* ctx, const tess_mesh* mesh, tess_field x, int64_t n,
                                  int32_t* particleindices, int32_t* starts, int32_t* stops,
                                  int64_t* nassigned);

/* The same partition without allocating: no counterpart in the reference.  It
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Results equal
 * tess_partition_update's; *nassigned is DEVICE memory (one int64) and nothing
 * is read back, so the call does not synchronise.  Work runs on `stream` (the
 * current device's; NULL = its default stream) over a fixed work area of the
 * library (about 21 MB per GPU, held until the process ends) sized for at most
 * TESS_PARTITION_NOALLOC_MAX particles; larger n is TESS_ERR_ARGUMENT.  All
 * calls on one GPU share that area: calls that overlap on a device race, so
 * serialise them (one stream, or events).  The first call on a device resolves
 * the area's address (tess_ctx_create does it for its device). */
#define TESS_PARTITION_NOALLOC_MAX 1048576
tess_status tess_partition_update_noalloc(const tess_mesh* mesh, tess_field x, int64_t n, int32_t* particleindices,
                                          int32_t* starts, int32_t* stops, int64_t* nassigned, void* stream);

/* reorder_particles!(particles, partition; threshold)
 * (src/Parallel/cpu_block_strategy.jl:248-355): permutes every particle field
 * into the partition's block order when the block-ordered contiguity is below
 * threshold (threshold 1 always reorders, 0 never).  particleindices/nassigned
 * are those of tess_partition_update and are rewritten to the identity.
 * *reordered (host) receives 1 if the particles were moved.  Synchronises. */
tess_status tess_reorder_particles(tess_ctx* ctx, const tess_mesh* mesh, const tess_particles* particles,
                                   int32_t* particleindices, int64_t nassigned, double threshold,
                                   int32_t* reordered);

/* ---- transfers ---------------------------------------------------------- */

/* @P2G of the README program (README.md:71-77; lowering src/Transfer/p2g.jl:145-178,
 * grid-node half :377-395): m = Σ w mₚ, mv = Σ w (mₚ v), f = Σ ((-V σ)·∇w), then on
 * every node m⁻¹ = m>0 ? 1/m : 0, vⁿ = mv m⁻¹, v = vⁿ + Δt (f m⁻¹ + g[m>0]), slip
 * walls.  weights == NULL (or weights->w.ptr == NULL, carrying a filter)
 * evaluates the basis inside the transfer (deferred weights,
 * src/Transfer/runtime.jl:21-24,55-130).  Deterministic: no float atomics; the
 * same inputs give bitwise-identical grids.  Corrected bases (TESS_WLS,
 * TESS_KERNEL_CORRECTION) are accepted by every forward entry point except the
 * batch ones. */
tess_status tess_p2g(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                     const tess_particles* particles, const tess_weights* weights,
                     const tess_params* params);

/* @G2P of the README program (README.md:80-87; lowering src/Transfer/g2p.jl:110-157):
 * v += Σ w (v_i - vⁿ_i); ∇v = Σ v_i ⊗ ∇w; x += Σ w v_i Δt; F = (I + Δt∇v)F;
 * J = det F; σ = (μ(FFᵀ - I) + λ ln J I)/J; V = V⁰J (J, ln J, 1/J in Float64). */
tess_status tess_g2p(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                     const tess_particles* particles, const tess_weights* weights,
                     const tess_params* params);

/* @G2P2G of the README program (src/Transfer/g2p2g.jl:100-159): one pass
 * doing the G2P half above, then the P2G half (m, mv, f of the updated v, V, σ)
 * and the grid-node half of tess_p2g.  Both halves use the weights at the
 * positions the particles have when the call starts: the macro binds them once
 * and the P2G half does not rebind them after x is written
 * (g2p2g.jl:107-114).  weights == NULL evaluates them inside the pass. */
tess_status tess_g2p2g(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                       const tess_particles* particles, const tess_weights* weights,
                       const tess_params* params);

/* nsteps passes of tess_g2p2g with deferred weights: the time loop of a
 * program written as one @G2P2G block (g2p2g.jl:100-159).  Within the call
 * each pass bins with the keys the previous pass's G2P half wrote, and the
 * particles are reordered into the engine's tile order at reorder_threshold
 * exactly as tess_step does (the role of reorder_particles! around the
 * reference's loop).  Reordering permutes the caller's particle arrays. */
tess_status tess_g2p2g_steps(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                             const tess_particles* particles, const tess_params* params, int32_t nsteps,
                             double reorder_threshold);

/* One README step with deferred weights: update!(wgts,…) + @P2G + @G2P
 * (README.md:66-87), repeated nsteps times on the context's stream.
 * reorder_threshold plays the role of the reference loop's
 * reorder_particles!(particles, partition; threshold)
 * (src/Parallel/cpu_block_strategy.jl:222-278): before a step's P2G, the
 * particle fields are permuted into the engine's tile order when the
 * contiguity score measured by the previous step (the fraction of neighbours
 * in that order whose storage indices are consecutive, the meaning of
 * block_ordered_particle_contiguity) is below the threshold; 0 never
 * reorders, 1 reorders every step.  Reordering permutes the
 * caller's particle arrays, exactly like reorder_particles!. */
tess_status tess_step(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                      const tess_particles* particles, const tess_params* params, int32_t nsteps,
                      double reorder_threshold);

/* ---- sparse grid (SpArray / SpGrid) ------------------------------------- */

/* SpIndices of a CartesianMesh (src/Grids/spindices.jl:57-140): nodes grouped
 * in blocks of 2^L per axis (L = mesh->block_size_log2), nblocks(mesh) =
 * ((dims-1)>>L)+1 per axis.  blocknumbering holds one int32 per block in
 * column-major block order: 0 = inactive, k = the k-th active block in linear
 * order (src/Grids/sparsity.jl:53-60).  A grid field of an SpGrid is the
 * compact SpArray data vector: node I of an active block lives at element
 * (k-1)*2^(L*dim) + local, local = column-major index of I & (2^L-1) inside the
 * block (spindices.jl:89,96-102).  storage_len = active blocks * 2^(L*dim);
 * capacity = elements the caller's grid arrays hold. */
typedef struct tess_sparsity {
  int32_t* blocknumbering;
  int64_t storage_len;
  int64_t capacity;
} tess_sparsity;

/* update_sparsity!(spinds, xₚ, mesh) (src/Grids/sparsity.jl:23-40,197-238):
 * blocks holding a particle (findblock, cartesian.jl:277-285) are occupied,
 * every block within ±1 block of an occupied one is active, and the active
 * blocks are numbered in linear order.  Writes sp->blocknumbering and
 * sp->storage_len; *changed (host, may be NULL) = 1 if the active set differs
 * from the numbering passed in (the reference returns `nothing` otherwise).
 * Synchronises (one readback of the active-block count). */
tess_status tess_sparsity_update(tess_ctx* ctx, const tess_mesh* mesh, tess_field x, int64_t n,
                                 tess_sparsity* sp, int32_t* changed);

/* tess_p2g / tess_g2p / tess_step on an SpGrid: the grid fields are SpArray
 * data vectors addressed through sp (src/Transfer/runtime.jl:132-170,
 * basis_weight.jl:270-278).  P2G zeroes the whole storage [0, storage_len)
 * before writing (the `=` targets, p2g.jl:159) and fails with TESS_ERR_STATE
 * when a particle's support reaches an inactive block (the reference's
 * "inactive SpGrid support node" error, runtime.jl:155-161).  G2P reads
 * inactive nodes as zero.  tess_step_sparse runs update_sparsity! before each
 * step's P2G, as the SpGrid loop does (src/Grids/grid.jl:97-101), and fails with TESS_ERR_STATE when the
 * new storage_len exceeds sp->capacity.  That check comes before any of the step's own work, so
 * *steps_done (host, may be NULL) then holds the number of steps that did run, and the caller resizes the
 * SpArray data vectors to sp->storage_len and calls again for the remaining nsteps - *steps_done steps.
 * On success *steps_done = nsteps; on any other error it is the number of steps enqueued before it. */
tess_status tess_p2g_sparse(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                            const tess_sparsity* sp, const tess_particles* particles, const tess_weights* weights,
                            const tess_params* params);
tess_status tess_g2p_sparse(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                            const tess_sparsity* sp, const tess_particles* particles, const tess_weights* weights,
                            const tess_params* params);
tess_status tess_step_sparse(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                             tess_sparsity* sp, const tess_particles* particles, const tess_params* params,
                             int32_t nsteps, double reorder_threshold, int32_t* steps_done);

/* ---- differentiation (no counterpart in the reference) ------------------- */

/* Adjoint of the README step's state (x, v, F) and of the per-particle
 * stiffness exponent logE (E_p = E exp(logE_p)); Float64 fields whatever the
 * mesh dtype.  The step treats V = V0 J and σ = (μ(FFᵀ-I) + λ lnJ I)/J as
 * functions of F (the README particle update), so a particle state must hold
 * σ and V consistent with F (as every state after a step does, and F = I,
 * σ = 0, V = V0 does).  logE.ptr may be NULL when its gradient is not wanted. */
typedef struct tess_adjoint {
  tess_field x, v, F;
  tess_field logE;
} tess_adjoint;

/* Reverse of one tess_step with deferred weights and no reordering (B-spline
 * bases): particles is the state at the START of the step (not modified); on
 * entry adj holds the adjoint of the state after the step, on return the
 * adjoint of the state before it, and adj->logE accumulates (+=) the step's
 * contribution.  The grid receives the step's forward P2G (scratch).
 * Deterministic (no float atomics outside LDS with a fixed order). */
tess_status tess_step_adjoint(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                              const tess_particles* particles, const tess_params* params, tess_adjoint* adj);

/* Forward rollout of nsteps README steps (deferred weights, no reordering) on
 * the caller's particles, recording a checkpoint of (x, v, F, V, σ) every
 * checkpoint_every steps in the context (the logE field is referenced, not
 * copied: keep it alive and unchanged until tess_rollout_adjoint). */
tess_status tess_rollout(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                         const tess_particles* particles, const tess_params* params, int32_t nsteps,
                         int32_t checkpoint_every);

/* Pullback of the context's last tess_rollout: adj holds the adjoint of the
 * final state on entry and of the initial state on return; adj->logE
 * accumulates ∂/∂logE over all steps.  Each segment is replayed from its
 * checkpoint, then reversed step by step.  The grid is scratch. */
tess_status tess_rollout_adjoint(tess_ctx* ctx, int32_t basis, const tess_mesh* mesh, const tess_grid* grid,
                                 const tess_params* params, tess_adjoint* adj);

/* ---- batches of scenes (no counterpart in the reference) ----------------- */

/* One scene of a batch: its own mesh, grid, particles and constants (Δt
 * included).  Fields use the Julia-native layout. */
typedef struct tess_scene {
  tess_mesh mesh;
  tess_grid grid;
  tess_particles particles;
  tess_params params;
} tess_scene;

typedef struct tess_batch tess_batch;

/* Many independent scenes of one dimension, dtype and basis, stepped together:
 * the scenes' tiles and particles share one index space, so every kernel of a
 * step runs once for the whole batch.  The scene arrays are referenced, not
 * copied (keep them alive and in place).  Deferred weights, no reordering. */
tess_status tess_batch_create(tess_ctx* ctx, int32_t basis, int32_t nscenes, const tess_scene* scenes,
                              tess_batch** out);
/* nsteps README steps of every scene of the batch (each with its own Δt). */
tess_status tess_batch_step(tess_ctx* ctx, tess_batch* batch, int32_t nsteps);
/* Reorders each scene's particles into the engine's tile order before a
 * step's P2G when the batch's contiguity score (the fraction of neighbours in
 * that order whose storage indices are consecutive) is below threshold, or
 * every step at threshold >= 1 (reorder_particles! and
 * block_ordered_particle_contiguity, src/Parallel/cpu_block_strategy.jl:222-278); particles
 * outside their mesh follow, and every particle stays in its scene's arrays.
 * 0 (the default) never reorders.  Reordering permutes the caller's particle
 * arrays, exactly like reorder_particles!. */
tess_status tess_batch_set_reorder(tess_batch* batch, double threshold);
/* Times each launch variant (waves per tile in P2G and G2P, eager or graph mode) over
 * steps_per_variant steps of the batch after one untimed step, keeps the
 * fastest and writes its number to *variant.  The tuning steps are steps of
 * the simulation: the scenes advance by (variants * (steps_per_variant + 1))
 * steps.  With a reorder threshold set, each untimed step reorders and the
 * timed steps do not, so the variants are compared on the same freshly
 * ordered layout; the threshold is restored.  Synchronises. */
tess_status tess_batch_autotune(tess_ctx* ctx, tess_batch* batch, int32_t steps_per_variant, int32_t* variant);
/* Sets the launch variant, e.g. a number tess_batch_autotune returned and the
 * caller cached for this workload shape: in 2D bit 0 splits P2G into one wave
 * per last-axis offset, and G2P runs 4 waves per tile, 2 with bit 1, 1 with
 * bit 2 (0..5); in 3D 0 or 2.  Bit 3 (+8) selects graph mode: the batch steps
 * on a private stream and scratch, each step one replay of a captured hipGraph
 * whose binning sorts every particle and whose tile and node-block kernels
 * read their counts on the device, launched with one workgroup per possible
 * tile or, with bit 4 as well (+24), as persistent grids sized to the device
 * (no host readback inside a step; the reorder test and one error-flag
 * readback run per call and per 16 replays).  Other numbers are
 * TESS_ERR_ARGUMENT.  No counterpart in the reference, which has no launch
 * variants. */
tess_status tess_batch_set_variant(tess_batch* batch, int32_t variant);
/* Counters of the batch, host values, the first n of: [0] reorders done, [1] of
 * those, reorders decided from the broken-pair count of the previous step's
 * P2G (not from a layered-order pass on a call's first step), [2] binnings that
 * kept the previous tile layout, [3] the launch variant, [4] graph replays,
 * [5] graph captures.  No counterpart in the reference (measurement). */
tess_status tess_batch_stats(const tess_batch* batch, int64_t* out, int32_t n);
tess_status tess_batch_destroy(tess_batch* batch);

/* ---- output (src/IO/export.jl) ------------------------------------------ */

/* Component layouts of export.jl's vtk_format (:73-85) for one field. */
enum {
  TESS_VTK_SCALAR = 1,     /* one component */
  TESS_VTK_VECTOR = 2,     /* dim components (a Vec) */
  TESS_VTK_POINTS = 3,     /* dim components padded with zeros to 3 (coordinates, resize(x, 3)) */
  TESS_VTK_TENSOR = 4,     /* dim*dim components, column-major (a SecondOrderTensor) */
  TESS_VTK_SYMMETRIC = 5   /* (xx, yy, zz, xy, yz, zx) of a symmetric tensor held in full dim*dim storage;
                              in 2D (xx, yy, NaN, xy, NaN, NaN) (export.jl:79-81) */
};

/* f32(vtk_format(field)) (export.jl:66-85, 104-107) on the device: n elements
 * of a particle or grid field of type dtype (any strides) become Float32
 * components in the layout `kind`, written to dst (host or device memory,
 * n * components floats) with one copy.  Synchronises. */
tess_status tess_export_field(tess_ctx* ctx, int32_t dtype, int32_t dim, tess_field field, int64_t n, int32_t kind,
                              float* dst);

/* One named point-data array of the writers below (vtk[name] = data):
 * n x ncomp Float32, the components of one point contiguous. */
typedef struct tess_vtk_array {
  const char* name;
  const float* data;
  int32_t ncomp;
  int32_t reserved;
} tess_vtk_array;

/* openvtk(path, x::AbstractVector{<:Vec}) with point data, then close
 * (export.jl:47-52, 89-99): "path.vtu", an UnstructuredGrid of n VTK_VERTEX
 * cells at xyz (n x 3 Float32).  Data are appended raw, zlib-compressed when
 * compress != 0 (vtkZLibDataCompressor, UInt64 headers).  Host only. */
tess_status tess_vtk_write_points(const char* path, int64_t n, const float* xyz, int32_t narrays,
                                  const tess_vtk_array* arrays, int32_t compress);
/* openvtk(path, mesh::CartesianMesh) (export.jl:53-55): "path.vts", a
 * StructuredGrid of the mesh nodes as Float32 points, point data over the
 * nodes in column-major order.  Host only. */
tess_status tess_vtk_write_mesh(const char* path, const tess_mesh* mesh, int32_t narrays,
                                const tess_vtk_array* arrays, int32_t compress);
/* openvtm (export.jl:101): "path.vtm", a vtkMultiBlockDataSet of the given
 * block files (paths as they are to be written in the file).  Host only. */
tess_status tess_vtm_write(const char* path, int32_t nblocks, const char* const* files);
/* openpvd + pvd[t] = file (export.jl:102): "path.pvd", a ParaView collection
 * of (time, file) entries.  Host only. */
tess_status tess_pvd_write(const char* path, int32_t n, const double* times, const char* const* files);
/* write_ply(name, xs) (export.jl:115-132): "name.ply" with the header lines and
 * float x, y, z per vertex, binary little-endian (xyz: n x 3).  Host only. */
tess_status tess_ply_write(const char* name, int64_t n, const float* xyz);

/* ---- measurement (no counterpart in the reference) ----------------------- */

/* Per-kernel timing with HIP events recorded on the context's stream around
 * each profiled launch (binning, k_p2g_tile, k_gather_grid, k_g2p_tile).
 * Enabling (or disabling) clears the accumulated entries. */
tess_status tess_profile_enable(tess_ctx* ctx, int32_t on);
/* Entry `index` (0-based, in first-launch order): name, summed duration,
 * launch count and summed algorithmic bytes.  TESS_ERR_ARGUMENT past the last
 * entry.  Synchronises. */
tess_status tess_profile_read(tess_ctx* ctx, int32_t index, char* name, int32_t name_len, double* total_ms,
                              int64_t* calls, double* bytes);
/* Host copies of the last binning and step: out[0] particles inside the mesh,
 * out[1] active tiles, out[2] node blocks written, out[3] grid nodes with m>0
 * written by the previous P2G, out[4] reorders done by tess_step, out[5]
 * neighbours in the layered order (the tile kernels' read order) whose
 * storage indices are not consecutive, counted by the previous P2G (-1
 * unknown; the contiguity score is 1 - out[5] / (out[0] - out[1])),
 * out[6] particles whose tile or window class changed at the last incremental
 * re-binning (-1 after a full sort), out[7] binnings that kept the previous
 * tile layout because no particle changed its key. */
tess_status tess_stats(tess_ctx* ctx, int64_t* out, int32_t count);
/* A plain device copy of `bytes` (a multiple of 16, both pointers 16-byte
 * aligned) on the context's stream, one 16-byte load and store per element: the
 * library's own measure of the copy bandwidth the box reaches.  No counterpart
 * in the reference (measurement).  Does not synchronise. */
tess_status tess_device_copy(tess_ctx* ctx, const void* src, void* dst, int64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* TESSERAE_HIP_H */
