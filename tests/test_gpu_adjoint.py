"""GPU adjoint (differentiation extension) against the oracle's adjoint step and
against Float64 central differences of GPU rollouts.

Bars: one reversed step within 1e-10 of the oracle (Float64, different but
fixed summation orders); rollout gradients of the inverse-cube loss along
seeded random directions within 1e-6 of central differences (h = 1e-6);
run-to-run bitwise determinism.
"""
import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import particles_np, rel_err, to_gpu
from oracle_bind import HostState
from tesserae_hip.scenes import inverse_cube

pytestmark = pytest.mark.gpu


def _advance_on_oracle(oracle, sc, k):
    """A consistent state after k oracle steps (σ, V follow from F)."""
    st = HostState(sc)
    oracle.step(st, sc.basis, sc.params, k)
    out = sc.copy()
    for name in ("x", "v", "F", "sigma", "V", "grad_v"):
        setattr(out, name, getattr(st, name).copy())
    return out


@pytest.mark.parametrize("basis", [th.BSpline(th.Quadratic()), th.BSpline(th.Cubic())], ids=["quadratic", "cubic"])
def test_step_adjoint_matches_oracle(oracle, basis):
    sc = inverse_cube(dtype=torch.float64, dx=1 / 16)
    sc.basis = basis
    sc.logE = 0.3 * np.random.default_rng(1).standard_normal(sc.n)
    sc = _advance_on_oracle(oracle, sc, 3)
    rng = np.random.default_rng(2)
    a, b, C = rng.standard_normal((sc.n, 3)), rng.standard_normal((sc.n, 3)), rng.standard_normal((sc.n, 9))
    # oracle
    hs = HostState(sc)
    ax, av, aF, ath = a.copy(), b.copy(), C.copy(), np.zeros(sc.n)
    oracle.step_adjoint(hs, basis, sc.params, ax, av, aF, ath)
    # GPU
    mesh, P, G = to_gpu(sc)
    adj = th.Adjoint(sc.n, 3)
    adj.x.copy_(torch.from_numpy(a))
    adj.v.copy_(torch.from_numpy(b))
    adj.F.copy_(torch.from_numpy(C))
    th.step_adjoint(G, P, basis, sc.params, adj)
    torch.cuda.synchronize()
    for name, got, ref in (("x", adj.x, ax), ("v", adj.v, av), ("F", adj.F, aF), ("logE", adj.logE, ath)):
        assert rel_err(got.cpu().numpy(), ref) <= 1e-10, name
    # the start state is not modified
    np.testing.assert_array_equal(particles_np(P)["x"], sc.x)


def _loss_and_grad(sc, target, nsteps, every, ctx):
    mesh, P, G = to_gpu(sc)
    th.rollout(G, P, sc.basis, sc.params, nsteps, checkpoint_every=every, ctx=ctx)
    x = P.get("x").double()
    t = torch.from_numpy(target).cuda()
    loss = float(((x - t) ** 2).sum(dim=1).mean())
    adj = th.Adjoint(sc.n, 3)
    adj.x.copy_(2.0 * (x - t) / sc.n)
    th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
    torch.cuda.synchronize()
    return loss, adj


def _loss(sc, target, nsteps, ctx):
    mesh, P, G = to_gpu(sc)
    th.step(G, P, sc.basis, sc.params, nsteps=nsteps, ctx=ctx)
    x = P.get("x").double().cpu().numpy()
    return float(((x - target) ** 2).sum(axis=1).mean())


def _consistent_stress(sc):
    """σ = (μ(FFᵀ-I) + λ lnJ I)/J and V = V0 J from F and logE (column-major 3x3 storage)."""
    n = sc.n
    Fm = sc.F.reshape(n, 3, 3).transpose(0, 2, 1)
    J = np.linalg.det(Fm)
    lam = sc.params.lam * np.exp(sc.logE)
    mu = sc.params.mu * np.exp(sc.logE)
    B = Fm @ Fm.transpose(0, 2, 1)
    S = (mu[:, None, None] * (B - np.eye(3)) + (lam * np.log(J))[:, None, None] * np.eye(3)) / J[:, None, None]
    sc.sigma = S.transpose(0, 2, 1).reshape(n, 9)
    sc.V = sc.params.V0 * J
    return sc


def _prestrained(sc, seed):
    """F = 0.97 I + 0.02 N(0,1) and logE = 0.3 N(0,1) with a consistent σ and V: the
    compressed cube expands from the first step at a rate set by E."""
    rng = np.random.default_rng(seed)
    sc.F = 0.97 * sc.F + 0.02 * rng.standard_normal(sc.F.shape)
    sc.logE = 0.3 * rng.standard_normal(sc.n)
    return _consistent_stress(sc)


def test_rollout_gradient_matches_central_differences():
    """Inverse cube at dx=1/32 (17,576 particles), Float64, pre-strained (F, logE), 30 steps
    with checkpoints every 7; directional derivatives along seeded random directions of v(0)
    and logE against central differences with h = 1e-6.  The target is the cube contracted
    by 10% about its centre: the expansion, driven by the stress, moves the loss at first
    order (a uniform offset would not, since the internal forces conserve momentum), and
    the differences of the oracle's rollouts agree to 1e-7 between h = 1e-5 and 1e-6."""
    sc = _prestrained(inverse_cube(dtype=torch.float64, dx=1 / 32), 17)
    c = sc.x.mean(axis=0)
    target = c + 0.9 * (sc.x - c)
    ctx = th.Context()
    loss, adj = _loss_and_grad(sc, target, 30, 7, ctx)
    gv, gth = adj.v.cpu().numpy(), adj.logE.cpu().numpy()
    assert np.abs(gth).max() > 0
    rng = np.random.default_rng(64)
    h = 1e-6
    for name, grad in (("v", gv), ("logE", gth), ("v", gv), ("logE", gth)):
        d = rng.standard_normal(grad.shape)
        plus, minus = sc.copy(), sc.copy()
        setattr(plus, name, getattr(sc, name) + h * d)
        setattr(minus, name, getattr(sc, name) - h * d)
        if name == "logE":  # σ follows E
            for s in (plus, minus):
                _consistent_stress(s)
        fd = (_loss(plus, target, 30, ctx) - _loss(minus, target, 30, ctx)) / (2 * h)
        ad = float((grad * d).sum())
        print(f"[fd] {name}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd), (name, ad, fd)
    ctx.close()


def test_rollout_adjoint_checkpointing_and_determinism():
    """Checkpoint spacing does not change the gradient beyond rounding, and repeated
    pullbacks are bitwise identical."""
    sc = inverse_cube(dtype=torch.float64, dx=1 / 32)
    target = sc.x + np.array([0.1, 0.0, 0.0])
    res = []
    for every in (1, 7, 7, 30):
        ctx = th.Context()
        _, adj = _loss_and_grad(sc, target, 30, every, ctx)
        res.append((adj.v.cpu().numpy(), adj.logE.cpu().numpy(), adj.x.cpu().numpy()))
        ctx.close()
    for k in range(3):
        np.testing.assert_array_equal(res[1][k], res[2][k])
        assert rel_err(res[0][k], res[1][k]) <= 1e-12
        assert rel_err(res[3][k], res[1][k]) <= 1e-12


def test_rollout_keeps_last_segment_for_the_first_pullback():
    """The rollout keeps its last segment's states, binnings and grids for the pullback, which then reverses
    that segment without replaying it; a second pullback of the same rollout replays it.  Both give the same
    gradients (Float64, to rounding: the replay may order ties differently), with a last segment shorter than
    the checkpoint spacing (30 steps, checkpoints every 7)."""
    sc = inverse_cube(dtype=torch.float64, dx=1 / 32)
    target = sc.x + np.array([0.1, 0.0, 0.0])
    ctx = th.Context()
    mesh, P, G = to_gpu(sc)
    th.rollout(G, P, sc.basis, sc.params, 30, checkpoint_every=7, ctx=ctx)
    x = P.get("x").double()
    t = torch.from_numpy(target).cuda()
    res = []
    for _ in range(2):
        adj = th.Adjoint(sc.n, 3)
        adj.x.copy_(2.0 * (x - t) / sc.n)
        th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
        torch.cuda.synchronize()
        res.append((adj.v.cpu().numpy(), adj.logE.cpu().numpy(), adj.x.cpu().numpy()))
    for k in range(3):
        assert rel_err(res[0][k], res[1][k]) <= 1e-12
    ctx.close()


def test_adjoint_errors():
    sc = inverse_cube(dtype=torch.float64, dx=1 / 16)
    mesh, P, G = to_gpu(sc)
    ctx = th.Context()
    adj = th.Adjoint(sc.n, 3)
    with pytest.raises(th.ErrorException):
        th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)  # nothing recorded
    with pytest.raises(th.UnsupportedError):
        th.step_adjoint(G, P, th.uGIMP(), sc.params, adj, ctx=ctx)
    ctx.close()
