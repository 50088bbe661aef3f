// Reverse mode of the README step (the engine's differentiation extension; no
// counterpart in the reference).  State of a step: (x, v, F); V = V0 J and
// σ = (μ(FFᵀ-I) + λ lnJ I)/J follow from F and the per-particle stiffness
// E_p = E exp(logE_p), so P = V σ = V0 (μ(FFᵀ-I) + λ lnJ I).
//
// One reversed step, given the state at its start and the adjoint (x̄', v̄', F̄')
// of the state after it (Float64 throughout):
//   bin + forward P2G      the step's grid (m, mv, f, vⁿ, v), recomputed
//   k_adj_scatter          particles -> nodes through the same sorted tiles and
//                          class-owner lanes as P2G (deterministic):
//                          Δv̄_i = Σ w v̄',  v̄_i = Σ (L̄ ∇w + Δt w x̄'),  L̄ = Δt F̄' Fᵀ
//   k_adj_grid             per node block: the grid update reversed (m>0 guard,
//                          slip walls): f̄, q̄ (of mv), m̄
//   k_adj_gather           per particle over its window, with ∇w and ∇∇w:
//                          x̄ = x̄' + Σ ∇w (v̄'·Δv + Δt x̄'·v + mₚ m̄ + mₚ q̄·v_p)
//                                  + Σ ∇∇w (L̄ᵀ v - Pᵀ f̄)
//                          v̄ = v̄' + mₚ Σ w q̄
//                          P̄ = -Σ f̄ ⊗ ∇w;  F̄ = (I+ΔtL)ᵀ F̄' + V0 (μ (P̄+P̄ᵀ) F + λ tr(P̄) F⁻ᵀ)
//                          ∂/∂logE += P̄ : P
// Rollouts keep checkpoints every k steps and replay each segment before
// reversing it.
#include "transfer_impl.hpp"

// Reversed steps a segment replay keeps enqueued ahead of its own progress.
// The replay's host thread waits at each step's binning readback; the lead is
// what the reversing stream has queued meanwhile.
#ifndef TESS_ADJ_REVERSE_LEAD
#define TESS_ADJ_REVERSE_LEAD 6
#endif

namespace tess {
namespace {

// Centred B-spline at node-relative coordinate r = ξ - i with its first and
// second derivatives in r (the textbook piecewise polynomials).
template <int BASIS>
__device__ __forceinline__ void bspline_jet2(double r, double& N, double& dN, double& ddN) {
  const double a = fabs(r), sg = signbit(r) ? -1.0 : 1.0;
  N = 0.0;
  dN = 0.0;
  ddN = 0.0;
  if constexpr (BASIS == TESS_BSPLINE_LINEAR) {
    if (a < 1.0) {
      N = 1.0 - a;
      dN = -sg;
    }
  } else if constexpr (BASIS == TESS_BSPLINE_QUADRATIC) {
    if (a < 0.5) {
      N = 0.75 - r * r;
      dN = -2.0 * r;
      ddN = -2.0;
    } else if (a < 1.5) {
      const double u = 1.5 - a;
      N = 0.5 * u * u;
      dN = -sg * u;
      ddN = 1.0;
    }
  } else {
    if (a < 1.0) {
      N = 0.5 * a * a * a - r * r + 2.0 / 3.0;
      dN = 1.5 * sg * r * r - 2.0 * r;
      ddN = 3.0 * a - 2.0;
    } else if (a < 2.0) {
      const double u = 2.0 - a;
      N = u * u * u / 6.0;
      dN = -sg * 0.5 * u * u;
      ddN = u;
    }
  }
}

// Doubles per node of the node-adjoint array: f̄ (D), q̄ (D), m̄, then the
// forward v (D) and v - vⁿ (D) that the gather stages.
template <int D>
constexpr int NODE_ADJ = 4 * D + 1;

// Jets of node lo+k along axis d (zero past hi), derivatives in x.
template <typename T, int D, int BASIS>
__device__ __forceinline__ void node_jet2(const Args<T, D>& a, T x, int d, int node, int hi, double& n0, double& n1,
                                          double& n2) {
  n0 = n1 = n2 = 0.0;
  if (node <= hi) {
    const double hinv = a.mesh.hinv_d;
    const double r = (double(x) - (a.mesh.xmin_d[d] + double(node) * a.mesh.h_d)) * hinv;
    bspline_jet2<BASIS>(r, n0, n1, n2);
    n1 *= hinv;
    n2 *= hinv * hinv;
  }
}

// Per-axis jets of a particle's window: node lo[d]+k for k < SW (zero past hi).
template <typename T, int D, int BASIS, bool SECOND>
__device__ __forceinline__ void window_jets2(const Args<T, D>& a, const T* x, const int* lo, const int* hi,
                                             double (*N)[support_width(BASIS)], double (*dN)[support_width(BASIS)],
                                             double (*ddN)[support_width(BASIS)]) {
  constexpr int SW = support_width(BASIS);
  const double hinv = a.mesh.hinv_d;
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      double n0 = 0.0, n1 = 0.0, n2 = 0.0;
      if (lo[d] + k <= hi[d]) {
        const double r = (double(x[d]) - (a.mesh.xmin_d[d] + double(lo[d] + k) * a.mesh.h_d)) * hinv;
        bspline_jet2<BASIS>(r, n0, n1, n2);
      }
      N[d][k] = n0;
      dN[d][k] = n1 * hinv;
      if constexpr (SECOND) ddN[d][k] = n2 * hinv * hinv;
    }
}

// Particle adjoints -> per-tile node partials [tile][NFOOT][2D]: Δv̄ (D) then v̄ (D).
// SW waves per tile, lane = window-start class, as in P2G: wave w accumulates
// the offsets w along the last axis (the Float64 accumulators of one offset
// plane fit in registers); the waves add into the LDS tile one after another.
template <typename T, int D, int BASIS>
__global__ __launch_bounds__(64 * support_width(BASIS)) void k_adj_scatter(Args<T, D> a, const uint32_t* __restrict__ skeys,
                                                    DField<double> ax, DField<double> av, DField<double> aF,
                                                    double* __restrict__ partials) {
  using G = Geo<D>;
  constexpr int SW = support_width(BASIS);
  constexpr int NA = 2 * D;
  constexpr int NSL = Pow<SW, D - 1>::v;
  constexpr int NT = 64 * SW;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int64_t j = xcd_remap(blockIdx.x, a.n_active);
  if (j >= a.n_active) return;
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  int tI[D], origin[D];
  unlin<D>(tile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;
  __shared__ double s_tile[G::NFOOT][NA];
  __shared__ uint32_t s_cnt[64];
  for (int q = int(threadIdx.x); q < NA * G::NFOOT; q += NT) (&s_tile[0][0])[q] = 0.0;
  int ccount, coff, maxcnt;
  __shared__ uint32_t s_pid[PID_CAP];
  tile_class_counts<SW>(skeys, start, cnt, s_cnt, ccount, coff, maxcnt, a.perm, 0u, s_pid, nullptr, PID_CAP);
  int ccoord[D];
  {
    int r = lane;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ccoord[d] = r % G::BW;
      r /= G::BW;
    }
  }
  const double dt = double(a.dt);
  {
    const int pass = wave;
    double acc[NSL][NA];
#pragma unroll
    for (int s = 0; s < NSL; ++s)
#pragma unroll
      for (int v = 0; v < NA; ++v) acc[s][v] = 0.0;
    // the inputs of a lane's next particle are loaded while its current one
    // is processed (the loads of one particle would otherwise stall each step)
    struct In {
      T x[D];
      double vb[D], xb[D], aF[D * D];
      T F[D * D];
    };
    auto load = [&](int k, In& in) {
      const int li = coff + k;
      const uint32_t pid = li < PID_CAP ? s_pid[li] : a.perm[start + li];
#pragma unroll
      for (int d = 0; d < D; ++d) in.x[d] = a.x(pid, d);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        in.vb[c] = av(pid, c);
        in.xb[c] = dt * ax(pid, c);
      }
#pragma unroll
      for (int q = 0; q < D * D; ++q) {
        in.aF[q] = aF(pid, q);
        in.F[q] = a.F(pid, q);
      }
    };
    In cur, nxt;
    if (ccount > 0) load(0, cur);
    for (int k = 0; k < maxcnt; ++k) {
      if (k + 1 < ccount) load(k + 1, nxt);
      if (k < ccount) {
      const T* x = cur.x;
      int lo[D], hi[D];  // the window (a binned particle lies inside the mesh)
      {
        T xi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xi[d] = (x[d] - a.mesh.xmin[d]) * a.mesh.h_inv;
#pragma unroll
        for (int d = 0; d < D; ++d) axis_window<T, BASIS>(xi[d], T(0), a.mesh.dims[d], lo[d], hi[d]);
      }
      double N[D][SW], dN[D][SW];
      window_jets2<T, D, BASIS, false>(a, x, lo, hi, N, dN, nullptr);
      const double* vb = cur.vb;
      const double* xb = cur.xb;
      double Lb[D * D];
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < D; ++q) s += cur.aF[r + D * q] * double(cur.F[c + D * q]);
          Lb[r + D * c] = dt * s;  // L̄ = Δt F̄' Fᵀ
        }
      // rows of this wave's offsets: the last axis at `pass`, the middle axis
      // (3D) enumerated, the nodes along x inside a row.  With w = N_x R and
      // ∇w = (N_x' R, N_x R_c) per node along x:  Δv̄ += N_x γ,
      // v̄ += N_x α + N_x' β,  γ = R v̄',  α = R Δt x̄' + Σ_{c>0} L̄_{:c} R_c,  β = L̄_{:0} R
      constexpr int NROW = NSL / SW;
#pragma unroll
      for (int row = 0; row < NROW; ++row) {
        int o[D];
        o[0] = 0;
        if constexpr (D == 3) o[1] = row;
        o[D - 1] = pass;
        double R = 1.0, Rc[D];
#pragma unroll
        for (int d = 1; d < D; ++d) R *= N[d][o[d]];
#pragma unroll
        for (int c = 1; c < D; ++c) {
          double t = dN[c][o[c]];
#pragma unroll
          for (int d = 1; d < D; ++d)
            if (d != c) t *= N[d][o[d]];
          Rc[c] = t;
        }
        double al[D], be[D], ga[D];
#pragma unroll
        for (int r2 = 0; r2 < D; ++r2) {
          ga[r2] = R * vb[r2];
          double t = R * xb[r2];
#pragma unroll
          for (int c = 1; c < D; ++c) t += Lb[r2 + D * c] * Rc[c];
          al[r2] = t;
          be[r2] = Lb[r2] * R;
        }
#pragma unroll
        for (int ox = 0; ox < SW; ++ox) {
          const int sl = ox + SW * row;
#pragma unroll
          for (int r2 = 0; r2 < D; ++r2) {
            acc[sl][r2] += N[0][ox] * ga[r2];
            acc[sl][D + r2] += N[0][ox] * al[r2] + dN[0][ox] * be[r2];
          }
        }
      }
      }
      cur = nxt;
    }
    // lanes own distinct classes: one add instruction never hits a node twice;
    // the waves add one after another, so the per-node order is fixed
    for (int w = 0; w < SW; ++w) {
      if (w == wave && ccount > 0) {
#pragma unroll
        for (int s = 0; s < NSL; ++s) {
          int r = s, q = 0, mul = 1;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const int od = (d == D - 1) ? pass : r % SW;
            if (d < D - 1) r /= SW;
            q += (ccoord[d] + od) * mul;
            mul *= G::TS;
          }
#pragma unroll
          for (int v = 0; v < NA; ++v) atomicAdd(&s_tile[q][v], acc[s][v]);
        }
      }
      __syncthreads();
    }
  }
  double* out = partials + int64_t(j) * NA * G::NFOOT;
  for (int q = int(threadIdx.x); q < NA * G::NFOOT; q += NT) out[q] = (&s_tile[0][0])[q];
}

// Node adjoints [node][2D+1] = f̄ (D), q̄ (D), m̄ of every node of the touched
// node blocks: the covering tiles' partials summed in a fixed order, then the
// grid update reversed.
template <typename T, int D>
__global__ __launch_bounds__(64) void k_adj_grid(Args<T, D> a, const double* __restrict__ partials,
                                                 double* __restrict__ nodeadj, const T* __restrict__ snap_grid) {
  using G = Geo<D>;
  constexpr int NA = 2 * D;
  constexpr int NB = (D == 3) ? 27 : 9;
  const int64_t j = xcd_remap(blockIdx.x, a.n_nb);
  if (j >= a.n_nb) return;
  const uint32_t B = a.nb_list[j];
  int Bi[D], I[D], rr[D];
  unlin<D>(B, a.ntiles, Bi);
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
    s_slot[threadIdx.x] = (ok && a.tile_flag[t]) ? int(a.tile_slot[t]) : -1;
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
  double acc[NA];
#pragma unroll
  for (int v = 0; v < NA; ++v) acc[v] = 0.0;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = rr[d] <= 1 ? -1 : 0;
    hi[d] = rr[d] == G::BW - 1 ? 1 : 0;
  }
  int o[3] = {0, 0, 0};
  for (o[2] = lo[2]; o[2] <= hi[2]; ++o[2])
    for (o[1] = lo[1]; o[1] <= hi[1]; ++o[1])
      for (o[0] = lo[0]; o[0] <= hi[0]; ++o[0]) {
        int nb = 0, q = 0, mulb = 1, mulq = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          nb += (o[d] + 1) * mulb;
          mulb *= 3;
          q += (rr[d] + 1 - o[d] * G::BW) * mulq;
          mulq *= G::TS;
        }
        const int slot = s_slot[nb];
        if (slot < 0) continue;
        const double* P = partials + (int64_t(slot) * G::NFOOT + q) * NA;
#pragma unroll
        for (int v = 0; v < NA; ++v) acc[v] += P[v];
      }
  const int64_t i = a.mesh.lin(I);
  double* out = nodeadj + i * NODE_ADJ<D>;
  // the step's forward grid at this node: the grid itself, or node r of the
  // j-th listed block of a compact snapshot (m, mv, f, vⁿ, v)
  const T* sg = snap_grid ? snap_grid + (j * int64_t(blockDim.x) + threadIdx.x) * (1 + 4 * D) : nullptr;
  const double m = sg ? double(sg[0]) : double(a.gm(i));
  // v and v - vⁿ for the gather's footprint (zero where m = 0, as the forward grid is)
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const double vv = sg ? double(sg[1 + 3 * D + c]) : double(a.gv(i, c));
    const double vn = sg ? double(sg[1 + 2 * D + c]) : double(a.gvn(i, c));
    out[2 * D + 1 + c] = vv;
    out[3 * D + 1 + c] = vv - vn;
  }
  if (!(m > 0.0)) {
#pragma unroll
    for (int v = 0; v < 2 * D + 1; ++v) out[v] = 0.0;
    return;
  }
  const double dt = double(a.dt), im = 1.0 / m;
  double mb = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const bool wall = (((a.slip >> (2 * c)) & 1) && I[c] == 0) ||
                      (((a.slip >> (2 * c + 1)) & 1) && I[c] == a.mesh.dims[c] - 1);
    const double dvb = acc[c], vgb = acc[D + c];
    const double s = wall ? 0.0 : vgb + dvb;  // adjoint reaching v through its update
    const double vnb = s - dvb;               // vⁿ enters v and Δv = v - vⁿ
    out[c] = s * dt * im;                     // f̄
    out[D + c] = vnb * im;                    // q̄ (of mv)
    const double fc = sg ? double(sg[1 + D + c]) : double(a.gf(i, c));
    const double mvc = sg ? double(sg[1 + c]) : double(a.gmv(i, c));
    mb -= (s * dt * fc + vnb * mvc) * im * im;
  }
  out[2 * D] = mb;
}

// Nodes -> particle adjoints: the footprint's forward v, Δv and node adjoints
// are staged in LDS; each thread reverses one particle of the tile (512
// threads: a tile's particles at 8 per cell in one pass).
#ifndef TESS_ADJ_GATHER_THREADS
#define TESS_ADJ_GATHER_THREADS 512
#endif
constexpr int ADJ_GATHER_THREADS = TESS_ADJ_GATHER_THREADS;
template <typename T, int D, int BASIS>
__global__ __launch_bounds__(ADJ_GATHER_THREADS) void k_adj_gather(Args<T, D> a, const double* __restrict__ nodeadj,
                                                    DField<double> ax, DField<double> av, DField<double> aF,
                                                    DField<double> ath) {
  using G = Geo<D>;
  constexpr int SW = support_width(BASIS);
  constexpr int NS = 4 * D + 1;  // v (D), Δv (D), f̄ (D), q̄ (D), m̄
  const int64_t j = xcd_remap(blockIdx.x, a.n_active);
  if (j >= a.n_active) return;
  const uint32_t tile = a.active[j];
  const uint32_t start = a.tile_start[tile];
  const int cnt = int(a.tile_end[tile] - start);
  int tI[D], origin[D];
  unlin<D>(tile, a.ntiles, tI);
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = tI[d] * G::BW - 1;
  __shared__ double s_n[G::NFOOT][NS];
  for (int q = int(threadIdx.x); q < G::NFOOT; q += int(blockDim.x)) {
    int I[D], r = q;
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      I[d] = origin[d] + r % G::TS;
      r /= G::TS;
      in = in && I[d] >= 0 && I[d] < a.mesh.dims[d];
    }
    const int64_t i = in ? a.mesh.lin(I) : 0;
    const double* na = nodeadj + i * NODE_ADJ<D>;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      s_n[q][c] = in ? na[2 * D + 1 + c] : 0.0;       // v
      s_n[q][D + c] = in ? na[3 * D + 1 + c] : 0.0;   // v - vⁿ
    }
#pragma unroll
    for (int v = 0; v < 2 * D + 1; ++v) s_n[q][2 * D + v] = in ? na[v] : 0.0;
  }
  __syncthreads();
  const double dt = double(a.dt), mp = double(a.mp);
  // rows of the window: offsets along y (unrolled, jets in registers) inside
  // offsets along z (3D; that axis' jet evaluated per row plane and pass)
  constexpr int NZ = (D == 3) ? SW : 1;
  for (int t = int(threadIdx.x); t < cnt; t += int(blockDim.x)) {
    const uint32_t p = a.perm[start + t];
    T x[D], l;
    int lo[D], hi[D];
    particle_window<T, D, BASIS, false>(a, p, x, l, lo, hi, nullptr);
    // x- and y-axis jets (value, first, second derivative per window node)
    double N[1][SW], dN[1][SW], ddN[1][SW], J1[SW][3];
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      node_jet2<T, D, BASIS>(a, x[0], 0, lo[0] + k, hi[0], N[0][k], dN[0][k], ddN[0][k]);
      node_jet2<T, D, BASIS>(a, x[1], 1, lo[1] + k, hi[1], J1[k][0], J1[k][1], J1[k][2]);
    }
    int base[D];
#pragma unroll
    for (int d = 0; d < D; ++d) base[d] = lo[d] - origin[d];  // the class coordinate, <= BW-1
    // Sums are factored by axis: for each offset of the other axes (rest), x-sums
    // of node data weighted by N_x (A), N_x' (B), N_x'' (C), combined with the
    // rest's jets.  With R = Π_{d>0} N_d, R_c = N_c' Π_{d>0,d!=c} N_d,
    // R_cc = N_c'' Π.., R_ce = N_c' N_e' Π..:  w = N_x R, ∇w = (N_x' R, N_x R_c),
    // ∇∇w = [[N_x'' R, N_x' R_c], [N_x' R_c, N_x R_cc | N_x R_ce]].
    // Two passes keep the Float64 working set in registers: the ∇∇w terms of x̄
    // (with L̄ and P), then the ∇w / w terms (x̄, v̄, P̄, L).
    double gx[D];
#pragma unroll
    for (int c = 0; c < D; ++c) gx[c] = 0.0;
    {
      double Lb[D * D], Pm[D * D];
      const double Vp = double(a.V(p));
#pragma unroll
      for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double s2 = 0.0;
#pragma unroll
          for (int q = 0; q < D; ++q) s2 += aF(p, r2 + D * q) * double(a.F(p, c + D * q));
          Lb[r2 + D * c] = dt * s2;  // L̄ = Δt F̄' Fᵀ
          Pm[r2 + D * c] = Vp * double(a.sigma(p, r2 + D * c));
        }
#pragma unroll 1
      for (int oz = 0; oz < NZ; ++oz) {
        double J2[3] = {1.0, 0.0, 0.0};
        int qz = 0;
        if constexpr (D == 3) {
          node_jet2<T, D, BASIS>(a, x[D - 1], D - 1, lo[D - 1] + oz, hi[D - 1], J2[0], J2[1], J2[2]);
          qz = (base[D - 1] + oz) * G::TS * G::TS;
        }
#pragma unroll
      for (int oy = 0; oy < SW; ++oy) {
        const int qrow = qz + (base[1] + oy) * G::TS;
        double rn[D], rd[D], rdd[D];
        rn[1] = J1[oy][0];
        rd[1] = J1[oy][1];
        rdd[1] = J1[oy][2];
        if constexpr (D == 3) {
          rn[D - 1] = J2[0];
          rd[D - 1] = J2[1];
          rdd[D - 1] = J2[2];
        }
        // A/B x-sums of v and f̄, and C of u0 = (L̄ᵀ v - Pᵀ f̄)_0
        double Av[D], Bv[D], Af[D], Bf[D], Cu0 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) Av[c] = Bv[c] = Af[c] = Bf[c] = 0.0;
#pragma unroll
        for (int ix = 0; ix < SW; ++ix) {
          const double* nd = s_n[qrow + base[0] + ix];
          const double n0 = N[0][ix], n1 = dN[0][ix], n2 = ddN[0][ix];
          double u0 = 0.0;
#pragma unroll
          for (int c = 0; c < D; ++c) {
            const double vg = nd[c], fb = nd[2 * D + c];
            u0 += Lb[c] * vg - Pm[c] * fb;
            Av[c] += n0 * vg;
            Bv[c] += n1 * vg;
            Af[c] += n0 * fb;
            Bf[c] += n1 * fb;
          }
          Cu0 += n2 * u0;
        }
        double UA[D], UB[D];  // U(S)_e = Σ_r L̄_re S[v]_r - P_re S[f̄]_r
#pragma unroll
        for (int e = 0; e < D; ++e) {
          double ua = 0.0, ub = 0.0;
#pragma unroll
          for (int r2 = 0; r2 < D; ++r2) {
            ua += Lb[r2 + D * e] * Av[r2] - Pm[r2 + D * e] * Af[r2];
            ub += Lb[r2 + D * e] * Bv[r2] - Pm[r2 + D * e] * Bf[r2];
          }
          UA[e] = ua;
          UB[e] = ub;
        }
        double R = 1.0;
#pragma unroll
        for (int d = 1; d < D; ++d) R *= rn[d];
        gx[0] += R * Cu0;
#pragma unroll
        for (int c = 1; c < D; ++c) {
          double Rc = rd[c], Rcc = rdd[c];
#pragma unroll
          for (int d = 1; d < D; ++d)
            if (d != c) {
              Rc *= rn[d];
              Rcc *= rn[d];
            }
          gx[0] += Rc * UB[c];
          double tt = Rc * UB[0] + Rcc * UA[c];
#pragma unroll
          for (int e = 1; e < D; ++e)
            if (e != c) {
              double pce = rd[c] * rd[e];
#pragma unroll
              for (int d = 1; d < D; ++d)
                if (d != c && d != e) pce *= rn[d];
              tt += pce * UA[e];
            }
          gx[c] += tt;
        }
      }
      }
    }
    double gq[D], Pb[D * D], L[D * D];
#pragma unroll
    for (int c = 0; c < D; ++c) gq[c] = 0.0;
#pragma unroll
    for (int k = 0; k < D * D; ++k) Pb[k] = L[k] = 0.0;
    {
      double vb[D], dtxb[D], mpvp[D];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        vb[c] = av(p, c);
        dtxb[c] = dt * ax(p, c);
        mpvp[c] = mp * double(a.v(p, c));
      }
#pragma unroll 1
      for (int oz = 0; oz < NZ; ++oz) {
        double J2[2] = {1.0, 0.0}, unused;
        int qz = 0;
        if constexpr (D == 3) {
          node_jet2<T, D, BASIS>(a, x[D - 1], D - 1, lo[D - 1] + oz, hi[D - 1], J2[0], J2[1], unused);
          qz = (base[D - 1] + oz) * G::TS * G::TS;
        }
#pragma unroll
      for (int oy = 0; oy < SW; ++oy) {
        const int qrow = qz + (base[1] + oy) * G::TS;
        double rn[D], rd[D];
        rn[1] = J1[oy][0];
        rd[1] = J1[oy][1];
        if constexpr (D == 3) {
          rn[D - 1] = J2[0];
          rd[D - 1] = J2[1];
        }
        // A/B x-sums of v, f̄ and s = v̄'·Δv + Δt x̄'·v + mₚ m̄ + mₚ q̄·v_p, A of q̄
        double Av[D], Bv[D], Af[D], Bf[D], Aq[D], As = 0.0, Bs = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) Av[c] = Bv[c] = Af[c] = Bf[c] = Aq[c] = 0.0;
#pragma unroll
        for (int ix = 0; ix < SW; ++ix) {
          const double* nd = s_n[qrow + base[0] + ix];
          const double n0 = N[0][ix], n1 = dN[0][ix];
          double sv = mp * nd[4 * D];
#pragma unroll
          for (int c = 0; c < D; ++c) {
            const double vg = nd[c], fb = nd[2 * D + c], qb = nd[3 * D + c];
            sv += vb[c] * nd[D + c] + dtxb[c] * vg + mpvp[c] * qb;
            Av[c] += n0 * vg;
            Bv[c] += n1 * vg;
            Af[c] += n0 * fb;
            Bf[c] += n1 * fb;
            Aq[c] += n0 * qb;
          }
          As += n0 * sv;
          Bs += n1 * sv;
        }
        double R = 1.0;
#pragma unroll
        for (int d = 1; d < D; ++d) R *= rn[d];
        gx[0] += R * Bs;
#pragma unroll
        for (int r2 = 0; r2 < D; ++r2) {
          gq[r2] += R * Aq[r2];
          L[r2] += R * Bv[r2];
          Pb[r2] -= R * Bf[r2];
        }
#pragma unroll
        for (int c = 1; c < D; ++c) {
          double Rc = rd[c];
#pragma unroll
          for (int d = 1; d < D; ++d)
            if (d != c) Rc *= rn[d];
          gx[c] += Rc * As;
#pragma unroll
          for (int r2 = 0; r2 < D; ++r2) {
            L[r2 + D * c] += Rc * Av[r2];
            Pb[r2 + D * c] -= Rc * Af[r2];
          }
        }
      }
      }
    }
    // particle side: x̄, v̄, F̄ = (I+ΔtL)ᵀ F̄' + V0 (μ (P̄+P̄ᵀ) F + λ tr(P̄) F⁻ᵀ), ∂/∂logE += P̄ : P
#pragma unroll
    for (int c = 0; c < D; ++c) {
      ax(p, c) = ax(p, c) + gx[c];
      av(p, c) = av(p, c) + mp * gq[c];
    }
    double F[D * D], Fb[D * D];
#pragma unroll
    for (int k = 0; k < D * D; ++k) {
      F[k] = double(a.F(p, k));
      Fb[k] = aF(p, k);
    }
    double lam = a.lambda, mu = a.mu;
    if (a.logE.p) {
      const double e = exp(double(a.logE(p)));
      lam *= e;
      mu *= e;
    }
    double Fi[D * D];  // F⁻¹ by the adjugate
    if constexpr (D == 2) {
      const double det = F[0] * F[3] - F[2] * F[1];
      Fi[0] = F[3] / det;
      Fi[1] = -F[1] / det;
      Fi[2] = -F[2] / det;
      Fi[3] = F[0] / det;
    } else {
      double cof[9];
#pragma unroll
      for (int r2 = 0; r2 < 3; ++r2)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int r1 = (r2 + 1) % 3, r3 = (r2 + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
          cof[r2 + 3 * c] = F[r1 + 3 * c1] * F[r3 + 3 * c2] - F[r1 + 3 * c2] * F[r3 + 3 * c1];
        }
      const double det = F[0] * cof[0] + F[3] * cof[3] + F[6] * cof[6];
#pragma unroll
      for (int r2 = 0; r2 < 3; ++r2)
#pragma unroll
        for (int c = 0; c < 3; ++c) Fi[r2 + 3 * c] = cof[c + 3 * r2] / det;
    }
    const double Vp = double(a.V(p));
    double trPb = 0.0, PbP = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) trPb += Pb[c + D * c];
#pragma unroll
    for (int k = 0; k < D * D; ++k) PbP += Pb[k] * (Vp * double(a.sigma(p, k)));
    double Fn[D * D];
#pragma unroll
    for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double acc = Fb[r2 + D * c];
#pragma unroll
        for (int k = 0; k < D; ++k) acc += dt * L[k + D * r2] * Fb[k + D * c];
        double sym = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) sym += (Pb[r2 + D * k] + Pb[k + D * r2]) * F[k + D * c];
        Fn[r2 + D * c] = acc + a.V0 * (mu * sym + lam * trPb * Fi[c + D * r2]);
      }
#pragma unroll
    for (int k = 0; k < D * D; ++k) aF(p, k) = Fn[k];
    if (ath.p) ath(p) += PbP;
  }
}

inline DField<double> dfield_d(const tess_field& f) {
  return DField<double>{static_cast<double*>(f.ptr), f.elem_stride, f.comp_stride};
}

// Layout of a snapshot (binning words, then the grid of the listed node
// blocks): perm and sorted keys (n words each), then tile_start, tile_end, the
// active tiles, tile flags and slots, the node-block list and flags (ntiles
// words each).
inline int64_t snapshot_bin_words(int64_t n, int64_t ntiles) { return (2 * n + 7 * ntiles + 1) & ~int64_t(1); }
inline void snapshot_views(const uint32_t* snap, int64_t n, int64_t ntiles, const uint32_t* w[9]) {
  const int64_t lens[9] = {n, n, ntiles, ntiles, ntiles, ntiles, ntiles, ntiles, ntiles};
  int64_t off = 0;
  for (int i = 0; i < 9; ++i) {
    w[i] = snap + off;
    off += lens[i];
  }
}

// The reversed kernels of one step, after the binning of S and its forward P2G
// grid are in place, or read from the step's snapshot.
template <typename T, int D, int BASIS>
void adjoint_kernels(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& S,
                     const tess_params& prm, const tess_adjoint& adj, const uint32_t* snap = nullptr) {
  using Gm = Geo<D>;
  const Binning& b = ctx->bin;
  if (b.n_active == 0) return;  // every particle outside: the step is the identity on the state
  Args<T, D> a = make_args<T, D>(mesh, &G, S, nullptr, &prm);
  fill_binning_ptrs<T, D>(ctx, a);
  const T* snap_grid = nullptr;
  const uint32_t* sk = static_cast<const uint32_t*>(ctx->keys_sorted.ptr);
  if (snap) {
    // the step's binning arrays and forward grid straight from its snapshot
    const uint32_t* w[9];
    snapshot_views(snap, b.n, b.ntiles, w);
    a.perm = w[0];
    sk = w[1];
    a.tile_start = w[2];
    a.tile_end = w[3];
    a.active = w[4];
    a.tile_flag = w[5];
    a.tile_slot = w[6];
    a.nb_list = w[7];
    a.nb_flag = w[8];
    snap_grid = reinterpret_cast<const T*>(snap + snapshot_bin_words(b.n, b.ntiles));
  }
  int64_t nnodes = 1;
  for (int d = 0; d < D; ++d) nnodes *= mesh.dims[d];
  double* partials = ctx->adj_partials.as<double>(size_t(b.n_active) * 2 * D * Gm::NFOOT);
  double* nodeadj = ctx->adj_nodes.as<double>(size_t(nnodes) * NODE_ADJ<D>);
  const DField<double> ax = dfield_d(adj.x), av = dfield_d(adj.v), aF = dfield_d(adj.F), ath = dfield_d(adj.logE);
  int tok = prof_begin(ctx, "k_adj_scatter", 0.0);
  hipLaunchKernelGGL((k_adj_scatter<T, D, BASIS>), dim3(xcd_grid(b.n_active)), dim3(64 * support_width(BASIS)), 0,
                     ctx->stream, a, sk, ax,
                     av, aF, partials);
  TESS_HIP(hipGetLastError());
  prof_end(ctx, tok);
  tok = prof_begin(ctx, "k_adj_grid", 0.0);
  hipLaunchKernelGGL((k_adj_grid<T, D>), dim3(xcd_grid(b.n_nodeblocks)), dim3(64), 0, ctx->stream, a, partials,
                     nodeadj, snap_grid);
  TESS_HIP(hipGetLastError());
  prof_end(ctx, tok);
  tok = prof_begin(ctx, "k_adj_gather", 0.0);
  hipLaunchKernelGGL((k_adj_gather<T, D, BASIS>), dim3(xcd_grid(b.n_active)), dim3(ADJ_GATHER_THREADS), 0,
                     ctx->stream, a,
                     nodeadj, ax, av, aF, ath);
  TESS_HIP(hipGetLastError());
  prof_end(ctx, tok);
}

// One reversed step: S is the state at the start of the step.
template <typename T, int D, int BASIS>
void step_adjoint_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& S,
                       const tess_params& prm, const tess_adjoint& adj) {
  bin_particles<T, D, BASIS, false>(ctx, mesh, S, nullptr);
  p2g_launch<T, D, BASIS, false>(ctx, mesh, G, S, nullptr, prm, false, nullptr);
  adjoint_kernels<T, D, BASIS>(ctx, mesh, G, S, prm, adj);
}

// ---- snapshots of a step's binning and forward grid ------------------------------
//
// While a segment is replayed, each step's binning arrays and P2G grid are
// copied aside, so that the reversed step restores them instead of binning and
// scattering again.  One launch copies the binning arrays (4-byte words); the
// grid is copied only on the node blocks of the step's list (nb_list), the
// only nodes the reversed step reads: tile footprints lie inside them, and the
// next forward step zeroes the whole grid before scattering.

struct CopyTask {
  const uint32_t* src;
  uint32_t* dst;
  int64_t words;
};
struct CopyTable {
  CopyTask t[16];
};

// Task blockIdx.y of the table; 16-byte words where both ends are 16-byte aligned.
__global__ __launch_bounds__(256) void k_copy_multi(CopyTable tab) {
  const CopyTask c = tab.t[blockIdx.y];
  const int64_t i0 = int64_t(blockIdx.x) * 256 + threadIdx.x, stride = int64_t(gridDim.x) * 256;
  int64_t done = 0;
  if (((reinterpret_cast<uintptr_t>(c.src) | reinterpret_cast<uintptr_t>(c.dst)) & 15) == 0) {
    const int64_t n4 = c.words / 4;
    const uint4* s4 = reinterpret_cast<const uint4*>(c.src);
    uint4* d4 = reinterpret_cast<uint4*>(c.dst);
    for (int64_t i = i0; i < n4; i += stride) d4[i] = s4[i];
    done = n4 * 4;
  }
  for (int64_t i = done + i0; i < c.words; i += stride) c.dst[i] = c.src[i];
}

// words of a field of nc components over n elements, whatever its strides
template <typename T>
int64_t field_words(const tess_field& f, int64_t n, int nc) {
  if (!f.ptr || n == 0) return 0;
  const int64_t span = (n - 1) * f.elem_stride + int64_t(nc - 1) * f.comp_stride + 1;
  return span * int64_t(sizeof(T)) / 4;
}

// The arrays of one snapshot: binning (perm, sorted keys, tile runs and lists)
// and the grid fields m, mv, f, vⁿ, v.
int snapshot_arrays(tess_ctx* ctx, int64_t n, int64_t ntiles, void* ptrs[9], int64_t words[9]) {
  tess::Buffer* bufs[] = {&ctx->vals,      &ctx->keys_sorted, &ctx->tile_start, &ctx->tile_end, &ctx->active_tiles,
                          &ctx->tile_flag, &ctx->tile_slot,   &ctx->nb_list,    &ctx->nb_flag};
  const int64_t lens[] = {n, n, ntiles, ntiles, ntiles, ntiles, ntiles, ntiles, ntiles};
  for (int i = 0; i < 9; ++i) {
    ptrs[i] = bufs[i]->ptr;
    words[i] = lens[i];
  }
  return 9;
}


// node values of one snapshot's grid part: m, mv, f, vⁿ, v of up to every node
// block (BW^D nodes each)
template <typename T, int D>
int64_t snapshot_grid_words(int64_t ntiles) {
  return ntiles * Pow<Geo<D>::BW, D>::v * (1 + 4 * D) * int64_t(sizeof(T) / 4);
}

template <typename T, int D>
int64_t snapshot_words(int64_t n, int64_t ntiles) {
  return snapshot_bin_words(n, ntiles) + snapshot_grid_words<T, D>(ntiles);
}

// Grid values of the node blocks in nb_list to (save) or from (restore) a
// compact store: node r of the j-th listed block at (j*BW^D + r)*(1+4D).
template <typename T, int D>
__global__ __launch_bounds__(64) void k_snap_grid(Args<T, D> a, T* __restrict__ store, int to_store) {
  using G = Geo<D>;
  constexpr int NV = 1 + 4 * D;
  const int64_t j = blockIdx.x;
  if (j >= a.n_nb) return;
  int Bi[D], I[D];
  unlin<D>(a.nb_list[j], a.ntiles, Bi);
  int r = int(threadIdx.x);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    I[d] = Bi[d] * G::BW + r % G::BW;
    r /= G::BW;
    if (I[d] >= a.mesh.dims[d]) return;
  }
  const int64_t i = a.mesh.lin(I);
  T* s = store + (j * int64_t(blockDim.x) + threadIdx.x) * NV;
  if (to_store) {
    s[0] = a.gm(i);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      s[1 + c] = a.gmv(i, c);
      s[1 + D + c] = a.gf(i, c);
      s[1 + 2 * D + c] = a.gvn(i, c);
      s[1 + 3 * D + c] = a.gv(i, c);
    }
  } else {
    a.gm(i) = s[0];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      a.gmv(i, c) = s[1 + c];
      a.gf(i, c) = s[1 + D + c];
      a.gvn(i, c) = s[1 + 2 * D + c];
      a.gv(i, c) = s[1 + 3 * D + c];
    }
  }
}

// save (to_store) or restore the snapshot at store; ctx->bin describes the
// step whose binning is saved or restored.  A further copy of xwords 32-bit
// words from xsrc to xdst rides along in the same launch.
template <typename T, int D>
void snapshot_copy(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, int64_t n, int64_t ntiles,
                   uint32_t* store, bool to_store, const void* xsrc = nullptr, void* xdst = nullptr,
                   int64_t xwords = 0) {
  void* p[9];
  int64_t w[9];
  const int k = snapshot_arrays(ctx, n, ntiles, p, w);
  CopyTable tab;
  std::memset(static_cast<void*>(&tab), 0, sizeof(tab));
  int64_t off = 0, most = 0;
  for (int i = 0; i < k; ++i) {
    uint32_t* arr = static_cast<uint32_t*>(p[i]);
    tab.t[i].src = to_store ? arr : store + off;
    tab.t[i].dst = to_store ? store + off : arr;
    tab.t[i].words = (arr && w[i] > 0) ? w[i] : 0;
    off += w[i];
    most = std::max(most, w[i]);
  }
  int ntask = k;
  if (xwords > 0) {
    tab.t[ntask].src = static_cast<const uint32_t*>(xsrc);
    tab.t[ntask].dst = static_cast<uint32_t*>(xdst);
    tab.t[ntask].words = xwords;
    ++ntask;
    most = std::max(most, xwords / 4);
  }
  if (most > 0) {
    const unsigned gx = unsigned(std::min<int64_t>(blocks_for(most, 256), 2048));
    hipLaunchKernelGGL(k_copy_multi, dim3(gx, unsigned(ntask)), dim3(256), 0, ctx->stream, tab);
    TESS_HIP(hipGetLastError());
  }
  const Binning& b = ctx->bin;
  if (b.n_nodeblocks > 0) {  // after the binning part: a restore reads the restored nb_list
    tess_particles none;
    std::memset(static_cast<void*>(&none), 0, sizeof(none));
    Args<T, D> a = make_args<T, D>(mesh, &G, none, nullptr, nullptr);
    fill_binning_ptrs<T, D>(ctx, a);
    T* gstore = reinterpret_cast<T*>(store + snapshot_bin_words(n, ntiles));
    hipLaunchKernelGGL((k_snap_grid<T, D>), dim3(unsigned(b.n_nodeblocks)), dim3(Pow<Geo<D>::BW, D>::v), 0,
                       ctx->stream, a, gstore, to_store ? 1 : 0);
    TESS_HIP(hipGetLastError());
  }
}

// ---- checkpointed rollouts -----------------------------------------------------

// A particle state stored by the context: x, v, F, V, σ packed field after field
// in the Julia layout.
template <typename T, int D>
constexpr int64_t state_scalars() {
  return 2 * D + 2 * D * D + 1;
}

template <typename T, int D>
tess_particles state_view(T* base, int64_t n, const Rollout& ro, T* gradv) {
  tess_particles P;
  std::memset(static_cast<void*>(&P), 0, sizeof(P));
  P.n = n;
  auto fld = [](void* p, int64_t es) {
    tess_field f;
    f.ptr = p;
    f.elem_stride = es;
    f.comp_stride = 1;
    return f;
  };
  T* q = base;
  P.x = fld(q, D);
  q += n * D;
  P.v = fld(q, D);
  q += n * D;
  P.F = fld(q, D * D);
  q += n * D * D;
  P.V = fld(q, 1);
  q += n;
  P.sigma = fld(q, D * D);
  P.grad_v = fld(gradv, D * D);
  P.logE = ro.logE;
  return P;
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_field(DField<T> src, DField<T> dst, int64_t n, int nc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  for (int c = 0; c < nc; ++c) dst(e, c) = src(e, c);
}

template <typename T, int D>
void copy_state(tess_ctx* ctx, const tess_particles& src, const tess_particles& dst) {
  const int64_t n = src.n;
  if (n == 0) return;
  const tess_field* s[] = {&src.x, &src.v, &src.F, &src.V, &src.sigma};
  const tess_field* t[] = {&dst.x, &dst.v, &dst.F, &dst.V, &dst.sigma};
  const int nc[] = {D, D, D * D, 1, D * D};
  for (int k = 0; k < 5; ++k) {
    hipLaunchKernelGGL(k_copy_field<T>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, dfield<T>(*s[k]),
                       dfield<T>(*t[k]), n, nc[k]);
    TESS_HIP(hipGetLastError());
  }
}

// The pullback's kept layout: whether two segments of snapshots fit the 8 GiB
// budget, the words of one snapshot, and the state slots of one buffer set
// (every + 1: the last one the replay's working state).
struct KeptLayout {
  bool kept;
  int64_t snap_words, seg_set;
};
template <typename T, int D>
KeptLayout kept_layout(const tess_mesh& mesh, int64_t n, int every) {
  constexpr int64_t kSnapBytes = int64_t(8) << 30;
  int64_t ntiles = 1;
  for (int d = 0; d < D; ++d) ntiles *= ((mesh.dims[d] - 1) >> Geo<D>::LT) + 1;
  KeptLayout k;
  k.snap_words = snapshot_words<T, D>(n, ntiles);
  k.kept = k.snap_words * 4 * int64_t(every) * 2 <= kSnapBytes;
  k.seg_set = state_scalars<T, D>() * n * (every + 1);
  return k;
}

template <typename T, int D, int BASIS>
void rollout_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                  const tess_params& prm, int nsteps, int every) {
  Rollout& ro = ctx->rollout;
  ro.valid = false;
  ro.last_kept = false;
  const int64_t n = P.n;
  const int ncheck = (nsteps + every - 1) / every;
  const int64_t per = state_scalars<T, D>() * n;
  T* ck = ctx->ro_checkpoints.as<T>(size_t(std::max<int64_t>(per * ncheck, 1)));
  ro.logE = P.logE;
  // The last segment is the first the pullback reverses: its states, binnings
  // and P2G grids are kept as this loop makes them, in the slots of the
  // pullback's buffer set for that segment, so that the pullback starts
  // reversing at once instead of replaying it alone.
  const KeptLayout kl = kept_layout<T, D>(mesh, n, every);
  const int last = ncheck - 1, lset = last & 1;
  T* seg = nullptr;
  uint32_t* snaps = nullptr;
  if (kl.kept && ncheck > 0) {
    seg = ctx->ro_segment.as<T>(size_t(std::max<int64_t>(2 * kl.seg_set, 1)));
    snaps = ctx->ro_snapshots.as<uint32_t>(size_t(kl.snap_words) * size_t(every) * 2);
    ro.last_bins.assign(size_t(every), Binning{});
  }
  // one step loop per segment between checkpoints
  for (int c = 0; c < ncheck; ++c) {
    const int len = std::min(every, nsteps - c * every);
    copy_state<T, D>(ctx, P, state_view<T, D>(ck + per * c, n, ro, nullptr));
    if (c < last || !seg) {
      step_impl<T, D, BASIS, false>(ctx, mesh, G, P, prm, len, 0.0, nullptr);
      continue;
    }
    // the last segment runs on the set's working slot (packed, as the
    // pullback's replays do); after each step's P2G, which leaves the
    // particles as at the step's start, one launch copies the state into the
    // step's slot and the binning and grid aside.  The final state goes back
    // to the caller's arrays; ∇v is written to them directly.
    T* sg = seg + kl.seg_set * lset;
    T* work = sg + per * every;
    copy_state<T, D>(ctx, P, state_view<T, D>(work, n, ro, nullptr));
    tess_particles W = state_view<T, D>(work, n, ro, nullptr);
    W.grad_v = P.grad_v;
    std::function<void(int, int)> hook = [&](int phase, int s) {
      if (phase == 0) return;
      snapshot_copy<T, D>(ctx, mesh, G, n, ctx->bin.ntiles, snaps + kl.snap_words * (int64_t(every) * lset + s),
                          true, work, sg + per * s, per * int64_t(sizeof(T)) / 4);
      ro.last_bins[size_t(s)] = ctx->bin;
    };
    step_impl<T, D, BASIS, false>(ctx, mesh, G, W, prm, len, 0.0, nullptr, nullptr, &hook);
    copy_state<T, D>(ctx, W, P);
  }
  if (seg) {
    ro.last_kept = true;
    ro.last_seg = seg;
    ro.last_snaps = snaps;
  }
  ro.valid = true;
  ro.nsteps = nsteps;
  ro.every = every;
  ro.n = n;
  ro.dim = D;
  ro.dtype = mesh.dtype;
  ro.basis = BASIS;
  for (int d = 0; d < 3; ++d) ro.dims[d] = mesh.dims[d];
}

// The adjoint's per-tile and per-node buffers of a context at their largest
// sizes for ntiles tiles, so that the reversed steps never grow (and free) a
// buffer while the other stream runs.
template <typename T, int D>
void reserve_reverse_buffers(tess_ctx* ctx, int64_t ntiles, int64_t nnodes) {
  using Gm = Geo<D>;
  ctx->adj_partials.as<double>(size_t(std::max<int64_t>(ntiles, 1)) * 2 * D * Gm::NFOOT);
  ctx->adj_nodes.as<double>(size_t(nnodes) * NODE_ADJ<D>);
}

// The replay context of ctx: a second context on the same device with a
// non-blocking stream of default priority.
tess_ctx* replay_context(tess_ctx* ctx) {
  if (!ctx->replay) {
    hipStream_t rs = nullptr;
    const hipError_t err = hipStreamCreateWithFlags(&rs, hipStreamNonBlocking);
    try {
      TESS_HIP(err);
      ctx->replay = ctx_new(ctx->device, rs);
    } catch (...) {
      if (rs) (void)hipStreamDestroy(rs);
      throw;
    }
  }
  return ctx->replay;
}

template <typename T, int D, int BASIS>
void rollout_adjoint_impl(tess_ctx* ctx, const tess_mesh& mesh, const tess_grid& G, const tess_params& prm,
                          const tess_adjoint& adj) {
  Rollout& ro = ctx->rollout;
  const int64_t n = ro.n;
  const int64_t per = state_scalars<T, D>() * n;
  const int ncheck = (ro.nsteps + ro.every - 1) / ro.every;
  T* ck = static_cast<T*>(ctx->ro_checkpoints.ptr);
  int64_t nnodes = 1, ntiles = 1;
  for (int d = 0; d < D; ++d) {
    nnodes *= mesh.dims[d];
    ntiles *= ((mesh.dims[d] - 1) >> Geo<D>::LT) + 1;
  }
  // Each reversed step k reads the binning and P2G grid of state k from the
  // snapshot the replay of its segment kept: the replay advances a copy of
  // every state of the segment by one step (the last one included) and copies
  // its context's binning arrays and the grid on the step's node blocks.  The
  // replay of segment c-1 runs on a second context and stream while this
  // context reverses segment c: both are latency-bound at this size, and they
  // use separate state, snapshot and scratch buffers (two sets, alternating by
  // segment), ordered by events.  When two segments of snapshots exceed 8 GiB,
  // each reversed step bins and scatters again instead, on this stream alone.
  const KeptLayout kl = kept_layout<T, D>(mesh, n, ro.every);
  const int64_t snap_words = kl.snap_words;
  const int every = ro.every;
  const bool kept = kl.kept;
  if (!kept) {
    T* seg = ctx->ro_segment.as<T>(size_t(std::max<int64_t>(per * every, 1)));
    T* gradv = ctx->ro_gradv.as<T>(size_t(std::max<int64_t>(n * D * D, 1)));
    for (int c = ncheck - 1; c >= 0; --c) {
      const int s0 = c * every;
      const int len = std::min(every, ro.nsteps - s0);
      TESS_HIP(hipMemcpyAsync(seg, ck + per * c, size_t(per) * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
      for (int k = 1; k < len; ++k) {
        TESS_HIP(hipMemcpyAsync(seg + per * k, seg + per * (k - 1), size_t(per) * sizeof(T),
                                hipMemcpyDeviceToDevice, ctx->stream));
        step_impl<T, D, BASIS, false>(ctx, mesh, G, state_view<T, D>(seg + per * k, n, ro, gradv), prm, 1, 0.0,
                                      nullptr);
      }
      for (int k = len - 1; k >= 0; --k)
        step_adjoint_impl<T, D, BASIS>(ctx, mesh, G, state_view<T, D>(seg + per * k, n, ro, gradv), prm, adj);
    }
    return;
  }
  tess_ctx* rc = replay_context(ctx);
  // per set: every + 1 state slots (the last one a scratch copy stepped for the
  // segment's last binning), every snapshots
  const int64_t seg_set = kl.seg_set;
  T* seg = ctx->ro_segment.as<T>(size_t(std::max<int64_t>(2 * seg_set, 1)));
  T* gradv = ctx->ro_gradv.as<T>(size_t(std::max<int64_t>(2 * n * D * D, 1)));  // replay, reverse
  uint32_t* snaps = ctx->ro_snapshots.as<uint32_t>(size_t(snap_words) * size_t(every) * 2);
  // the replay's private grid, Julia layout: m, then mv, f, vⁿ, v with D components
  T* gr = ctx->ro_grid.as<T>(size_t(nnodes) * (1 + 4 * D));
  tess_grid Gr;
  Gr.m = tess_field{gr, 1, 1};
  Gr.mv = tess_field{gr + nnodes, D, 1};
  Gr.f = tess_field{gr + nnodes * (1 + D), D, 1};
  Gr.vn = tess_field{gr + nnodes * (1 + 2 * D), D, 1};
  Gr.v = tess_field{gr + nnodes * (1 + 3 * D), D, 1};
  rc->last_grid_m = nullptr;  // the replay's first P2G zeroes its grid whole
  reserve_reverse_buffers<T, D>(ctx, ntiles, nnodes);
  std::vector<Binning> bins[2];
  bins[0].assign(size_t(every), Binning{});
  bins[1].assign(size_t(every), Binning{});
  struct Events {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};  // replay done (set 0, 1), reverse done (set 0, 1)
    ~Events() {
      for (hipEvent_t ev : e)
        if (ev) (void)hipEventDestroy(ev);
    }
  } ev;
  for (auto& e : ev.e) TESS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  bool reversed[2] = {false, false};
  // the caller's work on ctx's stream (the rollout, the adjoint's inputs) comes before the first replay
  TESS_HIP(hipEventRecord(ev.e[2], ctx->stream));
  TESS_HIP(hipStreamWaitEvent(rc->stream, ev.e[2], 0));
  // one reversed step k of the segment in buffer set `set` (enqueued only)
  auto reverse_step = [&](int set, int k) {
    const tess_particles S = state_view<T, D>(seg + seg_set * set + per * k, n, ro, gradv + n * D * D);
    Binning& b = ctx->bin;
    b = bins[set][size_t(k)];
    b.valid = false;  // the context's own arrays do not hold this binning
    b.state_valid = false;
    adjoint_kernels<T, D, BASIS>(ctx, mesh, G, S, prm, adj, snaps + snap_words * (int64_t(every) * set + k));
  };
  auto seg_len = [&](int c) { return std::min(every, ro.nsteps - c * every); };
  // Segment c is replayed as one step loop (the binning carried from step to
  // step as in the forward) on a working copy of its checkpoint, the last slot
  // of its buffer set: after each step's P2G the working state is copied to
  // the step's slot and the binning and grid aside.  While segment c-1 is
  // replayed, segment c is reversed: each replay step enqueues reversed steps
  // on this context's stream before its G2P, so the two streams stay in step
  // without either queueing far ahead.
  // the segment being reversed during a replay: its buffer set, next step to
  // enqueue, length; and the replay's length
  int rev_set = -1, rev_k = -1, rev_len = 0, rep_len = 1;
  std::function<void(int, int)> hook;
  int cur_set = 0;
  hook = [&](int phase, int s) {
    if (phase == 0) return;
    // after the step's P2G (which leaves the particles as they were at the
    // step's start): the working state into the step's slot, with the binning
    // and grid, in one launch
    T* sg = seg + seg_set * cur_set;
    snapshot_copy<T, D>(rc, mesh, Gr, n, rc->bin.ntiles, snaps + snap_words * (int64_t(every) * cur_set + s), true,
                        sg + per * every, sg + per * s, per * int64_t(sizeof(T)) / 4);
    bins[cur_set][size_t(s)] = rc->bin;
    {
      // keep the reversal TESS_ADJ_REVERSE_LEAD steps ahead of its share of the replay
      const int target = std::min(rev_len, (s + 1) * rev_len / rep_len + TESS_ADJ_REVERSE_LEAD);
      while (rev_k >= 0 && rev_len - 1 - rev_k < target) reverse_step(rev_set, rev_k--);
    }
  };
  auto replay = [&](int c) {
    const int set = c & 1;
    if (reversed[set]) TESS_HIP(hipStreamWaitEvent(rc->stream, ev.e[2 + set], 0));  // its buffers are free again
    cur_set = set;
    T* work = seg + seg_set * set + per * every;
    TESS_HIP(hipMemcpyAsync(work, ck + per * c, size_t(per) * sizeof(T), hipMemcpyDeviceToDevice, rc->stream));
    step_impl<T, D, BASIS, false>(rc, mesh, Gr, state_view<T, D>(work, n, ro, gradv), prm, seg_len(c), 0.0, nullptr,
                                  nullptr, &hook);
    TESS_HIP(hipEventRecord(ev.e[set], rc->stream));
  };
  // the last segment as the rollout ran it, when it kept it in these buffers;
  // else its replay runs first, alone
  const bool last_from_rollout = ro.last_kept && ro.last_seg == seg && ro.last_snaps == snaps &&
                                 ro.last_bins.size() == size_t(every);
  if (last_from_rollout) {
    bins[(ncheck - 1) & 1] = ro.last_bins;
    TESS_HIP(hipEventRecord(ev.e[(ncheck - 1) & 1], ctx->stream));
  } else {
    replay(ncheck - 1);
  }
  // the pullback's replays overwrite those slots
  ro.last_kept = false;
  for (int c = ncheck - 1; c >= 0; --c) {
    const int set = c & 1;
    TESS_HIP(hipStreamWaitEvent(ctx->stream, ev.e[set], 0));
    rev_set = set;
    rev_len = seg_len(c);
    rev_k = rev_len - 1;
    if (c > 0) {
      rep_len = seg_len(c - 1);
      replay(c - 1);
    }
    while (rev_k >= 0) reverse_step(set, rev_k--);
    TESS_HIP(hipEventRecord(ev.e[2 + set], ctx->stream));
    reversed[set] = true;
  }
  // nothing of the replay context is left running when the call returns
  TESS_HIP(hipStreamWaitEvent(ctx->stream, ev.e[0], 0));
  TESS_HIP(hipStreamWaitEvent(ctx->stream, ev.e[1], 0));
  TESS_HIP(hipStreamSynchronize(rc->stream));
  ctx->bin.valid = false;
  ctx->bin.state_valid = false;
}

template <typename T, int D, int B, bool S>
struct AdjointOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_params& prm, const tess_adjoint& adj) {
    if constexpr (B == TESS_UGIMP || S) raise(TESS_ERR_UNSUPPORTED, "the adjoint supports the B-spline bases");
    else step_adjoint_impl<T, D, B>(ctx, m, G, P, prm, adj);
  }
};
template <typename T, int D, int B, bool S>
struct RolloutOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_particles& P,
                  const tess_params& prm, int nsteps, int every) {
    if constexpr (B == TESS_UGIMP || S) raise(TESS_ERR_UNSUPPORTED, "the adjoint supports the B-spline bases");
    else rollout_impl<T, D, B>(ctx, m, G, P, prm, nsteps, every);
  }
};
template <typename T, int D, int B, bool S>
struct RolloutAdjointOp {
  static void run(tess_ctx* ctx, const tess_mesh& m, const tess_grid& G, const tess_params& prm,
                  const tess_adjoint& adj) {
    if constexpr (B == TESS_UGIMP || S) raise(TESS_ERR_UNSUPPORTED, "the adjoint supports the B-spline bases");
    else {
      try {
        rollout_adjoint_impl<T, D, B>(ctx, m, G, prm, adj);
      } catch (...) {
        // nothing of a failed pullback keeps running on the replay's stream
        if (ctx->replay) (void)hipStreamSynchronize(ctx->replay->stream);
        throw;
      }
    }
  }
};

}  // namespace

void run_step_adjoint(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                      const tess_params& prm, const tess_adjoint& adj) {
  ctx->corr = 0;  // plain kernels only
  ctx->corr_filter = nullptr;
  dispatch<AdjointOp>(basis, mesh, false, ctx, mesh, G, P, prm, adj);
}

void run_rollout(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G, const tess_particles& P,
                 const tess_params& prm, int nsteps, int every) {
  ctx->corr = 0;
  ctx->corr_filter = nullptr;
  dispatch<RolloutOp>(basis, mesh, false, ctx, mesh, G, P, prm, nsteps, every);
}

void run_rollout_adjoint(tess_ctx* ctx, int basis, const tess_mesh& mesh, const tess_grid& G,
                         const tess_params& prm, const tess_adjoint& adj) {
  const Rollout& ro = ctx->rollout;
  if (!ro.valid) raise(TESS_ERR_STATE, "no recorded rollout on this context: call tess_rollout first");
  bool same = ro.dim == mesh.dim && ro.dtype == mesh.dtype && ro.basis == basis;
  for (int d = 0; d < mesh.dim; ++d) same = same && ro.dims[d] == mesh.dims[d];
  if (!same) raise(TESS_ERR_ARGUMENT, "mesh, dtype or basis differ from the recorded rollout");
  ctx->corr = 0;
  ctx->corr_filter = nullptr;
  dispatch<RolloutAdjointOp>(basis, mesh, false, ctx, mesh, G, prm, adj);
}

}  // namespace tess
