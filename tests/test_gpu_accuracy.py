"""Accuracy bars of the five workloads in their configured modes, against the Float64 oracle.

* block drop (dx=1/128, 1,671,168 particles, gravity, slip walls, the headline's
  reorder at 0.85), Float32 engine, 100 steps: relative error on x ≤ 1e-5
  (max|a-b|/max|b| and ||a-b||/||b||); v and F within the bounds written below;
* sparse-grid sphere (dx=1/64), cubic B-spline and uGIMP, SpArray 8³ blocks with
  the reorder at 0.85: Float64 engine within 1e-10 of the oracle after 10 steps
  (at least one reorder happened); Float32 engine after 50 steps within the
  bounds written below;
* 256 uneven 2D scenes in one batch, Float32, 200 steps: the checked scenes
  within the bounds below;
* inverse cube (dx=1/64, 132,651 particles, 200 steps, checkpoints every 20):
  Float64 rollout gradients along 64 seeded directions of v(0) within 1e-6 of
  Float64 central differences (h = 1e-6); the Float32-forward / Float64-adjoint
  gradient within the bound below of the Float64 one.

What sets the size of a Float32 error: the oracle, which restates the
reference's CPU algorithm, runs in Float32 as well as Float64, and its own
Float32 error is the error of the reference's algorithm in that precision.
Where a scene is sensitive (the batch's particles with independent random
velocities, the sparse fringe nodes of uGIMP supports, whose mass falls off
quadratically while the force terms fall off linearly), that error is large,
and the engine's Float32 run is held to at most twice it (it is usually
smaller: J, ln J and 1/J are Float64 in the engine, and its strain term has no
cancellation).

Particles are matched by position (nearest neighbour, checked to be a
bijection) where the engine reorders them.
"""
import os

import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree

import tesserae_hip as th
from gpu_util import particles_np, rel_err, to_gpu
from oracle_bind import HostState
from tesserae_hip.scenes import batched_scenes, block_drop, inverse_cube, sphere_spin

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _errors(got, st, names=("x", "v", "F")):
    out = {}
    for name in names:
        a = np.asarray(got[name], np.float64).ravel()
        b = np.asarray(getattr(st, name), np.float64).ravel()
        out[name] = (float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)), rel_err(a, b))
    return out


def _match(x_got, x_ref):
    """Index j with x_got[i] ≈ x_ref[j[i]], a bijection (lattice spacing ≫ the errors)."""
    d, j = cKDTree(x_ref).query(x_got)
    assert np.unique(j).size == len(j), "position matching is not one to one"
    return j


def _reordered(got, j):
    """The engine's particle fields put into the oracle's particle order."""
    inv = np.empty_like(j)
    inv[j] = np.arange(len(j))
    return {k: v[inv] for k, v in got.items()}


def _oracle_run(oracle, sc, nsteps, chunk=10):
    st = HostState(sc)
    W, done = None, 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        W = oracle.step_threaded(st, sc.basis, sc.params, THREADS, k, W)
        done += k
        print(f"  oracle {done}/{nsteps}", flush=True)
    return st


def _oracle_fp32(oracle, sc, nsteps, chunk=10):
    """The reference algorithm (the oracle's threaded restatement) run in Float32."""
    st = HostState(sc, dtype=np.float32)
    W, done = None, 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        W = oracle.step_threaded(st, sc.basis, sc.params, THREADS, k, W)
        done += k
    return st


def _within_reference_fp32(e, r, names=("x", "v", "F"), factor=2.0, floor=1e-9):
    """The engine's Float32 error e is at most factor times the reference algorithm's own Float32 error r."""
    for name in names:
        for i, norm in enumerate(("L2", "Linf")):
            assert e[name][i] <= factor * r[name][i] + floor, (name, norm, e[name][i], r[name][i])


def _report(tag, errs):
    print(f"[{tag}] " + ", ".join(f"{k}: L2 {a:.2e} Linf {b:.2e}" for k, (a, b) in errs.items()), flush=True)


# ------------------------------------------------------------------ block drop

def test_block_drop_fp32_100_steps_vs_fp64_oracle(oracle):
    sc = block_drop(dtype=torch.float64, dx=1 / 128, steps=100)
    print(f"[block drop] n={sc.n} 100 steps FP32, reorder 0.85", flush=True)
    mesh, P, G = to_gpu(sc, dtype=torch.float32)
    ctx = th.Context()
    th.step(G, P, sc.basis, sc.params, nsteps=100, reorder_threshold=0.85, ctx=ctx)
    torch.cuda.synchronize()
    assert ctx.stats()["reorders"] >= 1
    got = particles_np(P)
    ctx.close()
    del P, G
    st = _oracle_run(oracle, sc, 100)
    got = _reordered(got, _match(got["x"], st.x))
    e = _errors(got, st)
    _report("block drop FP32 vs FP64 oracle", e)
    assert max(e["x"]) <= 1e-5
    assert max(e["v"]) <= 1e-5 and max(e["F"]) <= 1e-7


# ------------------------------------------------------------------ sparse sphere

SPHERE_BASES = [th.BSpline(th.Cubic()), th.uGIMP()]


@pytest.mark.parametrize("basis", SPHERE_BASES, ids=["cubic", "ugimp"])
def test_sphere_sparse_fp64_with_reorder_vs_oracle(oracle, basis):
    sc = sphere_spin(basis, dtype=torch.float64, dx=1 / 64, steps=10)
    print(f"[sphere fp64] n={sc.n} basis={basis} 10 steps, SpArray, reorder 0.85", flush=True)
    mesh, P, _ = to_gpu(sc)
    G = th.generate_grid(th.SpArray, mesh, device="cuda")
    ctx = th.Context()
    th.step(G, P, basis, sc.params, nsteps=10, reorder_threshold=0.85, ctx=ctx)
    torch.cuda.synchronize()
    reorders = ctx.stats()["reorders"]
    got = particles_np(P)
    ctx.close()
    assert reorders >= 1
    st = _oracle_run(oracle, sc, 10)
    j = _match(got["x"], st.x)
    assert np.any(j != np.arange(len(j)))  # the arrays were permuted
    got = _reordered(got, j)
    for name in ("x", "v", "F", "sigma", "V", "grad_v"):
        assert rel_err(got[name], getattr(st, name)) <= 1e-10, name


# Float32 bounds after 50 steps at dx=1/64 (relative L2, relative Linf), per basis
SPHERE_FP32 = {
    "cubic": {"x": (1e-6, 1e-5), "v": (1e-4, 1e-3), "F": (1e-6, 1e-5)},
    "ugimp": {"x": (1e-6, 1e-5), "v": (1e-4, 1e-3), "F": (1e-6, 2e-4)},
}


@pytest.mark.parametrize("basis", SPHERE_BASES, ids=["cubic", "ugimp"])
def test_sphere_sparse_fp32_vs_fp64_oracle(oracle, basis):
    sc = sphere_spin(basis, dtype=torch.float64, dx=1 / 64, steps=50)
    key = "cubic" if basis.kernel != 16 else "ugimp"
    print(f"[sphere fp32] n={sc.n} basis={basis} 50 steps, SpArray, reorder 0.85", flush=True)
    mesh, P, _ = to_gpu(sc, dtype=torch.float32)
    G = th.generate_grid(th.SpArray, mesh, device="cuda")
    ctx = th.Context()
    th.step(G, P, basis, sc.params, nsteps=50, reorder_threshold=0.85, ctx=ctx)
    torch.cuda.synchronize()
    got = particles_np(P)
    ctx.close()
    st = _oracle_run(oracle, sc, 50)
    got = _reordered(got, _match(got["x"], st.x))
    e = _errors(got, st)
    _report(f"sphere {key} FP32 vs FP64 oracle", e)
    r = _errors({k: getattr(_oracle_fp32(oracle, sc, 50), k) for k in ("x", "v", "F")}, st)
    _report(f"sphere {key} reference algorithm FP32 vs FP64", r)
    for name, (l2, linf) in SPHERE_FP32[key].items():
        assert e[name][0] <= l2 and e[name][1] <= linf, (name, e[name])
    _within_reference_fp32(e, r)


# ------------------------------------------------------------------ batch

BATCH_CHECKED = (0, 1, 2, 3, 4, 5, 6, 7)


def test_batch_fp32_200_steps_vs_fp64_oracle(oracle):
    scs = batched_scenes(256, dtype=torch.float64)
    members, gpu = [], []
    for sc in scs:
        mesh, P, G = to_gpu(sc, dtype=torch.float32)
        members.append((G, P, sc.params))
        gpu.append(P)
    print(f"[batch fp32] 256 scenes, {sum(s.n for s in scs)} particles, 200 steps", flush=True)
    ctx = th.Context()
    batch = th.Batch(members, scs[0].basis, ctx=ctx)
    batch.step(200)
    ctx.sync()
    got = {k: particles_np(gpu[k]) for k in BATCH_CHECKED}
    batch.close()
    ctx.close()
    del members, gpu
    for k in BATCH_CHECKED:
        st = _oracle_run(oracle, scs[k], 200, chunk=200)
        e = _errors(got[k], st)
        r = _errors({n: getattr(_oracle_fp32(oracle, scs[k], 200, chunk=200), n) for n in ("x", "v", "F")}, st)
        _report(f"batch scene {k} n={scs[k].n} E={scs[k].params.E:.3g} FP32 vs FP64", e)
        _report(f"batch scene {k} reference algorithm FP32 vs FP64", r)
        # positions to 1e-5 in both norms; v, F within 2e-3 Linf (the worst scenes are soft, E ~ 1e3, and
        # their velocity field is random per particle)
        assert max(e["x"]) <= 1e-5, (k, e["x"])
        assert e["v"][1] <= 2e-3 and e["F"][1] <= 2e-3, (k, e["v"], e["F"])
        _within_reference_fp32(e, r)


# ------------------------------------------------------------------ adjoint, config 5

def _final_x(sc, ctx, nsteps=200):
    mesh, P, G = to_gpu(sc)
    th.step(G, P, sc.basis, sc.params, nsteps=nsteps, ctx=ctx)
    return P.get("x").double().cpu().numpy()


def _central_difference(sc, name, d, h, target, ctx, nsteps=200, restate=None):
    """(L(θ+hd) - L(θ-hd)) / 2h for L = mean ||x(T) - target||², formed as
    Σ (x₊ - x₋)·(x₊ + x₋ - 2 target) / (2hn): the same number, without subtracting two
    nearly equal losses (at h = 1e-6 they share their first ten digits).  restate(scene)
    makes a perturbed state consistent (σ and V follow logE)."""
    plus, minus = sc.copy(), sc.copy()
    setattr(plus, name, getattr(sc, name) + h * d)
    setattr(minus, name, getattr(sc, name) - h * d)
    if restate is not None:
        plus, minus = restate(plus), restate(minus)
    xp, xm = _final_x(plus, ctx, nsteps), _final_x(minus, ctx, nsteps)
    return float(((xp - xm) * (xp + xm - 2 * target)).sum() / (2 * h * sc.n))


def _check_directions(tag, sc, grads, target, ctx, ndir, seed, restate=None):
    """Directional derivatives of the adjoint against central differences at h = 1e-6 and h = 1e-5.

    A central difference has errors of its own: at h = 1e-6 the rounding of two 200-step Float64 trajectories,
    about 1e-8 of the derivative in most directions and up to a few 1e-6 in some; at h = 1e-5 ten times less
    rounding, but truncation where a particle's support crosses a cell face between the two runs (the
    quadratic B-spline's second derivative jumps there).  The two differences' disagreement bounds their error.
    Errors are relative to the derivative, or to the median derivative over the directions when a direction's
    own derivative is smaller (a random direction can partly cancel the gradient, and the rounding of the
    adjoint and of the differences, about 1e-12 here, does not shrink with it).  Every direction must agree to 1e-6 with at least
    one of the two differences and with the h = 1e-6 one to within twice their disagreement plus 1e-6; the
    median error at h = 1e-6 must stay below 1e-7."""
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(ndir):
        name, grad = grads[k % len(grads)]
        d = rng.standard_normal(grad.shape)
        rs = restate if name == "logE" else None
        fd6 = _central_difference(sc, name, d, 1e-6, target, ctx, restate=rs)
        fd5 = _central_difference(sc, name, d, 1e-5, target, ctx, restate=rs)
        rows.append((name, float((grad * d).sum()), fd6, fd5))
    scale = {nm: np.median([abs(r[3]) for r in rows if r[0] == nm]) for nm, _ in grads}
    errs6, errs5 = [], []
    for k, (name, ad, fd6, fd5) in enumerate(rows):
        den = max(abs(fd5), scale[name])
        e6, e5, noise = abs(ad - fd6) / den, abs(ad - fd5) / den, abs(fd6 - fd5) / den
        errs6.append(e6)
        errs5.append(e5)
        print(f"[fd {tag} {name} {k}] adjoint {ad:.12e} fd(1e-6) {fd6:.12e} rel {e6:.2e}; fd(1e-5) rel {e5:.2e}; "
              f"fd(1e-6) vs fd(1e-5) {noise:.2e}", flush=True)
        assert min(e6, e5) <= 1e-6, (k, name, e6, e5)
        assert e6 <= 2 * noise + 1e-6, (k, name, e6, noise)
    print(f"[fd {tag}] {ndir} directions: h=1e-6 median {np.median(errs6):.2e} worst {max(errs6):.2e}; "
          f"h=1e-5 median {np.median(errs5):.2e} worst {max(errs5):.2e}", flush=True)
    assert np.median(errs6) <= 1e-7
    return errs6, errs5


def _grad(sc, target, ctx, dtype=None, nsteps=200):
    mesh, P, G = to_gpu(sc, dtype=dtype)
    th.rollout(G, P, sc.basis, sc.params, nsteps, checkpoint_every=20, ctx=ctx)
    x = P.get("x").double()
    t = torch.from_numpy(target).cuda()
    adj = th.Adjoint(sc.n, 3)
    adj.x.copy_(2.0 * (x - t) / sc.n)
    th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
    torch.cuda.synchronize()
    return adj.v.cpu().numpy(), adj.logE.cpu().numpy()


def test_inverse_cube_gradients_64_directions():
    """Config 5 as configured: dx=1/64 (132,651 particles), 200 steps, checkpoints every 20, loss = mean
    squared distance of x(T) to x(0) + (0.1, 0, 0), Float64.  ∂L/∂v(0) along 64 seeded directions against
    central differences (h = 1e-6, and 1e-5; see _check_directions).

    ∂L/∂logE of this loss is about 1e-6 of ∂L/∂v(0) along a direction of the same norm: changing a
    particle's stiffness moves no mass centre (internal forces conserve momentum), and the target is a
    translation, so only the spread of the particles about their centre feels it.  A logE perturbation of
    1e-6 moves the positions by about 1e-10, near the rounding of the 200-step Float64 trajectory itself,
    so central differences resolve it to a few percent only; the logE directions are checked at h = 1e-4
    to 5e-3 here, and to 1e-6 under a loss that feels the stiffness at first order (next test)."""
    sc = inverse_cube(dtype=torch.float64)
    target = sc.x + np.array([0.1, 0.0, 0.0])
    ctx = th.Context()
    gv, gth = _grad(sc, target, ctx)
    print(f"[adjoint] n={sc.n} |dL/dv0|={np.linalg.norm(gv):.3e} |dL/dlogE|={np.linalg.norm(gth):.3e}", flush=True)
    _check_directions("v(0)", sc, [("v", gv)], target, ctx, 64, 2024)
    rng = np.random.default_rng(7)
    for k in range(4):
        d = rng.standard_normal(gth.shape)
        fd = _central_difference(sc, "logE", d, 1e-4, target, ctx)
        ad = float((gth * d).sum())
        print(f"[fd logE {k}] adjoint {ad:.12e} fd(h=1e-4) {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) <= 5e-3 * abs(fd)
    # Float32 forward, Float64 adjoint accumulation
    gv32, gth32 = _grad(sc, target, ctx, dtype=torch.float32)
    ev = np.linalg.norm(gv32 - gv) / np.linalg.norm(gv)
    et = np.linalg.norm(gth32 - gth) / np.linalg.norm(gth)
    print(f"[adjoint] FP32-forward gradient vs FP64: v {ev:.2e}, logE {et:.2e} (relative L2)", flush=True)
    assert ev <= 1e-5 and et <= 5e-2
    ctx.close()


def test_inverse_cube_stiffness_gradient_200_steps():
    """The config-5 cube, mesh, 200 steps and checkpoints every 20, from a pre-strained state (F = 0.97 I +
    0.02 N(0,1), logE = 0.3 N(0,1), σ and V consistent) towards the cube contracted by 10% about its
    centre: the stress-driven expansion makes the loss depend on every particle's stiffness at first order.
    ∂L/∂logE and ∂L/∂v(0) along 16 seeded directions each against central differences (_check_directions)."""
    from test_gpu_adjoint import _consistent_stress, _prestrained
    sc = _prestrained(inverse_cube(dtype=torch.float64), 17)
    c = sc.x.mean(axis=0)
    target = c + 0.9 * (sc.x - c)
    ctx = th.Context()
    gv, gth = _grad(sc, target, ctx)
    _check_directions("prestrained", sc, [("logE", gth), ("v", gv)], target, ctx, 32, 64, restate=_consistent_stress)
    ctx.close()
