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
  Float64 rollout gradients along 64 seeded directions within 1e-6 of Float64
  central differences (h = 1e-6); the Float32-forward / Float64-adjoint
  gradient within the bound below of the Float64 one.

Errors of a Float32 run have two sources: the rounding of the inputs to Float32
and the rounding of the arithmetic during the run.  The dynamics amplify both
the same way.  Where a scene is sensitive (particles with independent random
velocities, sparse fringes of uGIMP support), the tests also run the oracle
from the Float32-rounded inputs: the distance between the two oracle runs is
the scene's own sensitivity to a Float32-sized perturbation, and the engine's
Float32 error is bounded by a small multiple of it.

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


def _rounded_inputs(sc):
    """The scene with its particle inputs rounded to Float32 (what a Float32 run starts from)."""
    r = sc.copy()
    for name in ("x", "v", "V", "F", "sigma", "grad_v", "l", "logE"):
        a = getattr(r, name)
        if a is not None:
            setattr(r, name, a.astype(np.float32).astype(np.float64))
    return r


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
    "ugimp": {"x": (1e-6, 1e-5), "v": (1e-4, 1e-2), "F": (1e-6, 1e-3)},
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
    st_r = _oracle_run(oracle, _rounded_inputs(sc), 50)
    s = _errors({k: getattr(st_r, k) for k in ("x", "v", "F")}, st)
    _report(f"sphere {key} FP64 oracle from FP32-rounded inputs vs FP64 oracle", s)
    for name, (l2, linf) in SPHERE_FP32[key].items():
        assert e[name][0] <= l2 and e[name][1] <= linf, (name, e[name])


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
    worst_ratio = 0.0
    for k in BATCH_CHECKED:
        st = _oracle_run(oracle, scs[k], 200, chunk=200)
        e = _errors(got[k], st)
        st_r = _oracle_run(oracle, _rounded_inputs(scs[k]), 200, chunk=200)
        s = _errors({n: getattr(st_r, n) for n in ("x", "v", "F")}, st)
        _report(f"batch scene {k} n={scs[k].n} E={scs[k].params.E:.3g} FP32 vs FP64", e)
        _report(f"batch scene {k} FP64 from FP32-rounded inputs vs FP64", s)
        assert max(e["x"]) <= 1e-4, (k, e["x"])
        for name in ("v", "F"):
            worst_ratio = max(worst_ratio, e[name][0] / max(s[name][0], 1e-12))
    print(f"[batch] worst L2 ratio engine-FP32 error / input-rounding sensitivity: {worst_ratio:.2f}", flush=True)
    assert worst_ratio <= 100.0


# ------------------------------------------------------------------ adjoint, config 5

def _loss(sc, target, nsteps, ctx, dtype=None):
    mesh, P, G = to_gpu(sc, dtype=dtype)
    th.step(G, P, sc.basis, sc.params, nsteps=nsteps, ctx=ctx)
    x = P.get("x").double()
    return float(((x - torch.from_numpy(target).cuda()) ** 2).sum(dim=1).mean())


def _grad(sc, target, ctx, dtype=None):
    mesh, P, G = to_gpu(sc, dtype=dtype)
    th.rollout(G, P, sc.basis, sc.params, 200, checkpoint_every=20, ctx=ctx)
    x = P.get("x").double()
    t = torch.from_numpy(target).cuda()
    adj = th.Adjoint(sc.n, 3)
    adj.x.copy_(2.0 * (x - t) / sc.n)
    th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
    torch.cuda.synchronize()
    return adj.v.cpu().numpy(), adj.logE.cpu().numpy()


def test_inverse_cube_gradients_64_directions():
    """Config 5 as configured: dx=1/64 (132,651 particles), 200 steps, checkpoints every 20, loss = mean
    squared distance of x(T) to x(0) + (0.1, 0, 0); gradients with respect to v(0) and logE."""
    sc = inverse_cube(dtype=torch.float64)
    target = sc.x + np.array([0.1, 0.0, 0.0])
    ctx = th.Context()
    gv, gth = _grad(sc, target, ctx)
    print(f"[adjoint] n={sc.n} |dL/dv0|={np.linalg.norm(gv):.3e} |dL/dlogE|={np.linalg.norm(gth):.3e}", flush=True)
    rng = np.random.default_rng(2024)
    h = 1e-6
    worst = 0.0
    for k in range(64):
        name, grad = (("v", gv), ("logE", gth))[k % 2]
        d = rng.standard_normal(grad.shape)
        plus, minus = sc.copy(), sc.copy()
        setattr(plus, name, getattr(sc, name) + h * d)
        setattr(minus, name, getattr(sc, name) - h * d)
        fd = (_loss(plus, target, 200, ctx) - _loss(minus, target, 200, ctx)) / (2 * h)
        ad = float((grad * d).sum())
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"[fd {k}] {name}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
    print(f"[adjoint] worst relative FD error over 64 directions: {worst:.2e}", flush=True)
    assert worst <= 1e-6
    # Float32 forward, Float64 adjoint accumulation
    gv32, gth32 = _grad(sc, target, ctx, dtype=torch.float32)
    ev = np.linalg.norm(gv32 - gv) / np.linalg.norm(gv)
    et = np.linalg.norm(gth32 - gth) / np.linalg.norm(gth)
    print(f"[adjoint] FP32-forward gradient vs FP64: v {ev:.2e}, logE {et:.2e} (relative L2)", flush=True)
    assert ev <= 1e-3 and et <= 1e-2
    ctx.close()
