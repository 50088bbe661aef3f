// Forward entry points of the README step (tess_p2g / tess_g2p / tess_g2p2g /
// tess_step and their SpGrid forms); the kernels and launchers live in
// transfer_impl.hpp.
#include "transfer_impl.hpp"

namespace tess {

// A corrected basis runs on its kernel's instantiation: the wrapper and
// polynomial bits and the node filter travel in the context.  Returns the
// stored weights, or NULL for deferred ones (weights without storage carry
// only the filter their update! recorded).
const tess_weights* bind_basis(tess_ctx* ctx, int& basis, const tess_weights* W) {
  ctx->corr = basis & ~0xff;
  ctx->corr_filter = W ? W->filter : nullptr;
  basis &= 0xff;
  return (W && W->w.ptr) ? W : nullptr;
}

void run_p2g(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
             const tess_weights* W, const tess_params& prm) {
  W = bind_basis(ctx, basis, W);
  dispatch<P2GOp>(basis, mesh, W != nullptr, ctx, mesh, G, P, W, prm, static_cast<const tess_sparsity*>(nullptr));
}

void run_g2p(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
             const tess_weights* W, const tess_params& prm) {
  W = bind_basis(ctx, basis, W);
  dispatch<G2POp>(basis, mesh, W != nullptr, ctx, mesh, G, P, W, prm, static_cast<const tess_sparsity*>(nullptr));
}

void run_g2p2g(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
               const tess_weights* W, const tess_params& prm) {
  W = bind_basis(ctx, basis, W);
  dispatch<G2P2GOp>(basis, mesh, W != nullptr, ctx, mesh, G, P, W, prm);
}

void run_g2p2g_steps(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                     const tess_params& prm, int nsteps, double reorder_threshold) {
  bind_basis(ctx, basis, nullptr);
  dispatch<G2P2GStepsOp>(basis, mesh, false, ctx, mesh, G, P, prm, nsteps, reorder_threshold);
}

void run_step(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
              const tess_params& prm, int nsteps, double reorder_threshold) {
  bind_basis(ctx, basis, nullptr);
  dispatch<StepOp>(basis, mesh, false, ctx, mesh, G, P, prm, nsteps, reorder_threshold,
                   static_cast<tess_sparsity*>(nullptr), static_cast<int*>(nullptr));
}

void run_p2g_sparse(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_sparsity& sp,
                    const tess_particles& P, const tess_weights* W, const tess_params& prm) {
  W = bind_basis(ctx, basis, W);
  dispatch<P2GOp>(basis, mesh, W != nullptr, ctx, mesh, G, P, W, prm, &sp);
}

void run_g2p_sparse(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_sparsity& sp,
                    const tess_particles& P, const tess_weights* W, const tess_params& prm) {
  W = bind_basis(ctx, basis, W);
  dispatch<G2POp>(basis, mesh, W != nullptr, ctx, mesh, G, P, W, prm, &sp);
}

void run_step_sparse(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, tess_sparsity& sp,
                     const tess_particles& P, const tess_params& prm, int nsteps, double reorder_threshold,
                     int* done) {
  bind_basis(ctx, basis, nullptr);
  dispatch<StepOp>(basis, mesh, false, ctx, mesh, G, P, prm, nsteps, reorder_threshold, &sp, done);
}

}  // namespace tess
