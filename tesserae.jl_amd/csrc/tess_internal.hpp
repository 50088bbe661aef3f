// Host-side internals of libtesserae_hip: the context (stream + scratch owned by
// the library), error reporting across the C ABI, and the launchers each
// kernel file provides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/tesserae_hip.h"

namespace tess {

// Error raised inside the library; converted to a tess_status + message at the
// C boundary (nothing propagates across it).
struct Error {
  tess_status code;
  std::string msg;
};

[[noreturn]] void raise(tess_status code, const std::string& msg);
void check_hip(hipError_t e, const char* what);
#define TESS_HIP(call) ::tess::check_hip((call), #call)
// the same check, never subject to failure injection (stream-capture control)
void check_hip_exact(hipError_t e, const char* what);
#define TESS_HIP_EXACT(call) ::tess::check_hip_exact((call), #call)
// incremented by every C-ABI call that fails (see check_ctx)
uint64_t error_epoch();

// A device buffer owned by the context; grows on demand, never shrinks.
struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* get(size_t need);  // (re)allocates when need > bytes; contents undefined
  void release();
  template <typename U>
  U* as(size_t count) { return static_cast<U*>(get(count * sizeof(U))); }
};

// Binning of one particle set into the P2G/G2P tiles, valid until the
// particles move.  All arrays live in ctx scratch.
struct Binning {
  bool valid = false;
  const void* x_ptr = nullptr;  // identity of the particle set it was built for
  int64_t n = 0;
  int32_t dim = 0, dtype = 0, basis = 0;
  int32_t tile_log2 = 0;
  int64_t ntiles_axis[3] = {1, 1, 1};
  int64_t ntiles = 0;
  int64_t n_inside = 0;     // particles with a non-empty window (host copy)
  int64_t n_active = 0;     // tiles holding >= 1 particle (host copy)
  int64_t n_nodeblocks = 0; // node blocks touched by some tile footprint (host copy)
  // the sorted (key, perm) arrays are a valid previous order for incremental re-binning
  bool state_valid = false;
  int64_t state_n = -1, state_ntiles = -1;
  int32_t state_dim = 0;
  // the last G2P wrote the next keys and stay flags at the sorted positions
  bool keys_ready = false;
};

// Per-kernel timing with HIP events on the context's stream.  Disabled by
// default; begin/end cost two event records per profiled launch when enabled.
struct ProfEntry {
  std::string name;
  double total_ms = 0.0;
  int64_t calls = 0;
  double bytes = 0.0;  // algorithmic bytes summed over the calls
};
struct Profiler {
  bool on = false;
  std::vector<ProfEntry> entries;
  std::vector<hipEvent_t> pool;
  struct Pending {
    int entry;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  size_t next_event = 0;
};

// The last tess_rollout of a context: checkpoints of (x, v, F, V, σ) every
// `every` steps, replayed segment by segment by the adjoint.
struct Rollout {
  bool valid = false;
  int nsteps = 0, every = 1;
  int64_t n = 0;
  int32_t dim = 0, dtype = 0, basis = 0;
  int64_t dims[3] = {0, 0, 0};
  tess_field logE{nullptr, 0, 0};
  // the last segment's states, binnings and grids as the rollout ran it, in
  // the pullback's slots (so the pullback need not replay that segment), and
  // the buffers they were written to
  bool last_kept = false;
  std::vector<Binning> last_bins;
  const void* last_seg = nullptr;
  const void* last_snaps = nullptr;
};

}  // namespace tess

struct tess_ctx {
  tess::Profiler prof;
  int device = 0;
  hipStream_t stream = nullptr;
  // sort + binning: keys (unsorted), keys_sorted + vals (= perm) sorted output,
  // *_alt ping-pong buffers of the radix sort
  tess::Buffer keys, keys_sorted, vals, keys_alt, vals_alt, hist, scan_tmp;
  tess::Buffer tile_start, tile_end, tile_flag, tile_slot, active_tiles;
  tess::Buffer nb_flag, nb_slot, nb_list, nb_list_prev;
  tess::Buffer partials;
  tess::Buffer per_tile, per_nodeblock;  // per-workgroup counts summed by a reduction
  // incremental re-binning scratch
  tess::Buffer skeys2, perm2, inc_nk, inc_spos, inc_mkey, inc_mj, inc_mks, inc_mr, inc_mjs, inc_rng;
  int64_t last_movers = -1;
  int64_t kept_binnings = 0;  // binnings that kept the previous layout (no particle changed its key)
  tess::Buffer counters;   // small device scalars
  tess::Buffer g2p2g_x;    // positions before a G2P2G pass (its P2G half's weights)
  tess::Buffer export_buf; // Float32 staging of tess_export_field
  // corrected basis of the current call (TESS_WLS / TESS_KERNEL_CORRECTION /
  // TESS_POLY_MULTILINEAR bits, 0 for a plain kernel) and its node filter
  int32_t corr = 0;
  const uint8_t* corr_filter = nullptr;
  tess::Buffer movers;     // G2P's changed-key counts (spread over cache lines)
  tess::Buffer host_pinned_dummy;
  int64_t* counters_host = nullptr;  // pinned readback slots
  tess::Binning bin;
  // zeroing bookkeeping for the grid of the previous P2G
  const void* last_grid_m = nullptr;
  int64_t prev_nodeblocks = 0;
  int32_t prev_tile_log2 = -1;
  int64_t prev_ntiles = -1;
  int64_t prev_nnodes = -1;
  // reorder scratch and the layout score of the last step's P2G
  tess::Buffer reorder_tmp, reorder_cnt;
  const void* misplaced_x_ptr = nullptr;
  int64_t misplaced_n = 0;
  int64_t misplaced_pairs = 0;  // layered neighbour pairs of that P2G's binning
  int64_t last_misplaced = -1;  // of them, pairs not consecutive in storage
  int64_t reorders = 0;
  // nodes with m > 0 written by the last gather (read back with the next binning)
  int64_t last_active_nodes = -1;
  // SpGrid block activation scratch (occupied flags, active flags, scan)
  tess::Buffer sp_occ, sp_act, sp_scan;
  // adjoint: per-tile Float64 partials, node adjoints, rollout checkpoints and replay
  tess::Buffer adj_partials, adj_nodes, ro_checkpoints, ro_segment, ro_gradv, ro_snapshots;
  tess::Rollout rollout;
  // the pullback replays segments on a second context (own stream and
  // scratch, private grid ro_grid) while this context reverses the previous one
  tess_ctx* replay = nullptr;
  tess::Buffer ro_grid;
  // the failure epoch this context's caches belong to
  uint64_t seen_epoch = 0;
};

namespace tess {
// A context's state without the C-ABI checks (tess_ctx_create/destroy use them).
tess_ctx* ctx_new(int device, hipStream_t stream);
void ctx_free(tess_ctx* ctx);
}  // namespace tess

// A batch of scenes (batch.hip): the scenes as given, where each scene's tiles
// and particles start in the batch-wide index spaces, and the device copies.
struct tess_batch {
  int32_t basis = 0, dim = 0, dtype = 0, nscenes = 0;
  int64_t ntot = 0, tiles_total = 0, max_nodes = 0;
  std::vector<tess_scene> scenes;
  std::vector<int64_t> toff, poff;
  tess::Buffer recs, d_toff, d_poff, prev_list;
  int64_t prev_nb = 0;
  bool zeroed = false;
  // the binning in the context is this batch's, and the last G2P wrote every
  // particle's next key (stepping within one call)
  bool binned = false, keys_ready = false;
  int64_t kept_binnings = 0;
  // bit 0: P2G one wave per last-axis offset (2D); G2P waves per tile: 4, 2 with bit 1, 1 with bit 2
  int variant = 0;
  // layered reorder of the scenes' particles when the contiguity score falls
  // below the threshold (0: never, 1: every step); reorders done so far
  double reorder_threshold = 0.0;
  int64_t reorders = 0;
  // of those, reorders decided from the broken-pair count of the previous step's P2G
  int64_t reorders_by_count = 0;
  // the last P2G of this call counted the broken layered pairs (the sum sits in
  // the counters), over this many pairs
  bool counted = false;
  int64_t counted_pairs = 0;
  // graph mode (variant bit 3): a private context on its own stream steps the
  // batch as replays of captured hipGraphs, one per parity of the binning's
  // double buffer (graph[p] reads binning set p and writes set 1-p); the
  // previous step's node-block list and count live on the device, and the
  // buffers the graphs point at are recorded to detect a reallocation
  tess_ctx* gctx = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  hipEvent_t graph_ev[2] = {nullptr, nullptr};
  int captured_variant = -1;
  std::vector<std::pair<const tess::Buffer*, void*>> graph_deps;
  tess::Buffer g_prev_list, g_prev_count;
  unsigned grid_p2g = 0, grid_g2p = 0, grid_gather = 0, grid_zero = 0, grid_out = 0;
  // the grids were last written by graph-mode steps (the eager bookkeeping does not know them)
  bool last_graph = false;
  // replays between two reorder tests (each test bins on the host side)
  int graph_chunk = 16;
  int64_t graph_replays = 0, captures = 0;
  // the failure epoch the batch's caches belong to
  uint64_t seen_epoch = 0;
};

namespace tess {

// ---- scans and sorts (sort.hip) ----
// Plain copy of bytes (a multiple of 16) on the context's stream.
void run_device_copy(tess_ctx* ctx, const void* src, void* dst, int64_t bytes);
// Exclusive prefix sum of n uint32 values in place; returns nothing, total is
// written to *total_dev (device) when non-null.
void exclusive_scan_u32(tess_ctx* ctx, uint32_t* data, int64_t n, uint32_t* total_dev);
// *out (device) = sum of v[0..n), one workgroup (for per-workgroup counts).
void sum_u32(tess_ctx* ctx, const uint32_t* v, int64_t n, uint32_t* out);
// Stable LSD radix sort of (key, value) pairs on the low key_bits of key.
// vals_in == nullptr means values 0..n-1.  Result lands in (keys_out, vals_out).
// With n_dev the element count is read on the device (at most n, which sizes
// the launches), so the sort can be captured in a graph.
void radix_sort_pairs(tess_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                      uint32_t* vals_out, int64_t n, int key_bits, const uint32_t* n_dev = nullptr);
// The same sort over caller-provided work buffers on a stream (no allocation):
// hist holds sort_hist_words(n) words, keys_alt/vals_alt n words each, scan_tmp
// scan_work_bytes(sort_hist_words(n)) bytes.
struct SortWork {
  hipStream_t stream = nullptr;
  uint32_t* hist = nullptr;
  uint32_t* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  uint8_t* scan_tmp = nullptr;
};
size_t sort_hist_words(int64_t n);
size_t scan_work_bytes(int64_t n);
void radix_sort_pairs_ws(const SortWork& w, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                         uint32_t* vals_out, int64_t n, int key_bits, const uint32_t* n_dev = nullptr);

// ---- reference-layout weights and partition (partition.hip) ----
void launch_update_weights(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_particles& P,
                           const tess_weights& W);
void run_partition_update(tess_ctx* ctx, const tess_mesh& mesh, tess_field x, int64_t n, int32_t* pidx,
                          int32_t* starts, int32_t* stops, int64_t* nassigned);
// The partition over the library's fixed device work area (no allocation);
// *nassigned is device memory.  Raises TESS_ERR_ARGUMENT beyond
// partition_noalloc_max_particles().
void run_partition_update_noalloc(const tess_mesh& mesh, tess_field x, int64_t n, int32_t* pidx, int32_t* starts,
                                  int32_t* stops, int64_t* nassigned, hipStream_t stream);
int64_t partition_noalloc_max_particles();
// Resolves the work area's address on the current device (once per device).
void noalloc_scratch_resolve();
void run_reorder(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, int32_t* pidx,
                 int64_t nassigned, double threshold, int32_t* reordered);

// ---- SpGrid block numbering (sparse.hip) ----
// update_sparsity!(spinds, xₚ, mesh): writes sp.blocknumbering and
// sp.storage_len, *changed (host, optional) = active set changed.  Synchronises.
void run_sparsity_update(tess_ctx* ctx, const tess_mesh& mesh, tess_field x, int64_t n, tess_sparsity& sp,
                         int32_t* changed);

// ---- transfers (transfer.hip) ----
void run_p2g(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
             const tess_weights* W, const tess_params& prm);
void run_g2p(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
             const tess_weights* W, const tess_params& prm);
void run_g2p2g(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
               const tess_weights* W, const tess_params& prm);
void run_g2p2g_steps(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                     const tess_params& prm, int nsteps, double reorder_threshold);
void run_step(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
              const tess_params& prm, int nsteps, double reorder_threshold);
// The same on an SpGrid: grid fields are SpArray data vectors addressed through sp.
void run_p2g_sparse(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_sparsity& sp,
                    const tess_particles& P, const tess_weights* W, const tess_params& prm);
void run_g2p_sparse(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_sparsity& sp,
                    const tess_particles& P, const tess_weights* W, const tess_params& prm);
// *done: steps run before a TESS_ERR_STATE (storage too small), nsteps on success.
void run_step_sparse(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, tess_sparsity& sp,
                     const tess_particles& P, const tess_params& prm, int nsteps, double reorder_threshold,
                     int* done);
// ---- adjoint (adjoint.hip) ----
void run_step_adjoint(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                      const tess_params& prm, const tess_adjoint& adj);
void run_rollout(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                 const tess_params& prm, int nsteps, int every);
void run_rollout_adjoint(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G,
                         const tess_params& prm, const tess_adjoint& adj);
// ---- batches (batch.hip) ----
void run_export_field(tess_ctx* ctx, int32_t dtype, int32_t dim, tess_field field, int64_t n, int32_t kind,
                      float* dst);
void run_vtk_write_points(const char* path, int64_t n, const float* xyz, int32_t narrays,
                          const tess_vtk_array* arrays, int32_t compress);
void run_vtk_write_mesh(const char* path, const tess_mesh& mesh, int32_t narrays, const tess_vtk_array* arrays,
                        int32_t compress);
void run_vtm_write(const char* path, int32_t nblocks, const char* const* files);
void run_pvd_write(const char* path, int32_t n, const double* times, const char* const* files);
void run_ply_write(const char* name, int64_t n, const float* xyz);
void run_batch_create(tess_ctx* ctx, int basis, int nscenes, const tess_scene* scenes, tess_batch** out);
void run_batch_step(tess_ctx* ctx, tess_batch* b, int nsteps);
void run_batch_set_variant(tess_batch* b, int variant);
void run_batch_set_reorder(tess_batch* b, double threshold);
int run_batch_variant(const tess_batch* b);
// out[0..n): reorders, reorders decided from the previous P2G's count, kept binnings, variant
void run_batch_stats(const tess_batch* b, int64_t* out, int n);
void run_batch_destroy(tess_batch* b);
// Moves every particle field so that slot k receives particle order[k].
void permute_particles(tess_ctx* ctx, const tess_mesh& mesh, const tess_particles& P, const uint32_t* order);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Profiling hooks (abi.hip): prof_begin returns a token (-1 when disabled).
int prof_begin(tess_ctx* ctx, const char* name, double bytes);
void prof_end(tess_ctx* ctx, int token);
void prof_flush(tess_ctx* ctx);

}  // namespace tess
