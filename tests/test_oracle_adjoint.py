"""The oracle's adjoint step against central finite differences of the oracle's
forward steps (Float64).

The adjoint is the engine's differentiation extension and has no counterpart in
the reference; what pins it is the mathematics: for a random linear functional
L = Σ a·x(K) + b·v(K) + C:F(K) of the state after K steps, the adjoint's
directional derivatives along random directions of x(0), v(0), F(0) and logE
must equal (L(+h) - L(-h)) / 2h.  The initial σ and V are kept consistent with
F and E (σ = (μ(FFᵀ-I) + λ lnJ I)/J, V = V0 J), as the step assumes.
"""
import numpy as np
import pytest

import tesserae_hip as th
from oracle_bind import HostState
from tesserae_hip.scenes import random_box


def consistent(sc, F, logE):
    n, d = sc.n, sc.mesh.dim
    Fm = F.reshape(n, d, d).transpose(0, 2, 1)  # column-major storage -> matrices
    J = np.linalg.det(Fm)
    lam = sc.params.lam * np.exp(logE)
    mu = sc.params.mu * np.exp(logE)
    B = Fm @ Fm.transpose(0, 2, 1)
    S = (mu[:, None, None] * (B - np.eye(d)) + (lam * np.log(J))[:, None, None] * np.eye(d)) / J[:, None, None]
    return S.transpose(0, 2, 1).reshape(n, d * d), sc.params.V0 * J


def make_scene(dim, seed, basis):
    sc = random_box(dim, 150 if dim == 2 else 120, seed, basis, h=0.1, lo=0.3, hi=0.62, vscale=0.5)
    # stiff enough (c = 10, CFL 0.1) that the stress moves the particles noticeably in K steps
    sc.params = th.MPMParams(dt=1e-3, E=1e5, nu=0.3, rho0=1e3, V0=sc.params.V0)
    rng = np.random.default_rng(seed)
    sc.F = sc.F + 0.03 * rng.standard_normal(sc.F.shape)
    sc.logE = 0.3 * rng.standard_normal(sc.n)
    sc.sigma, sc.V = consistent(sc, sc.F, sc.logE)
    return sc


def rollout(oracle, sc, K, keep=False):
    st = HostState(sc)
    states = []
    for _ in range(K):
        if keep:
            states.append(HostState(_snapshot(sc, st)))
        oracle.step(st, sc.basis, sc.params, 1)
    return st, states


def _snapshot(sc, st):
    c = sc.copy()
    for name in ("x", "v", "F", "sigma", "V", "grad_v"):
        setattr(c, name, getattr(st, name).copy())
    return c


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", [th.BSpline(th.Quadratic()), th.BSpline(th.Cubic())], ids=["quadratic", "cubic"])
def test_adjoint_matches_central_differences(oracle, dim, basis):
    K = 6
    sc = make_scene(dim, 11 + dim, basis)
    rng = np.random.default_rng(99)
    n, d = sc.n, dim
    a = rng.standard_normal((n, d))
    b = rng.standard_normal((n, d))
    C = rng.standard_normal((n, d * d))

    def loss(s):
        st, _ = rollout(oracle, s, K)
        return float((a * st.x).sum() + (b * st.v).sum() + (C * st.F).sum())

    _, states = rollout(oracle, sc, K, keep=True)
    ax, av, aF = a.copy(), b.copy(), C.copy()
    ath = np.zeros(n)
    for s in reversed(states):
        oracle.step_adjoint(s, sc.basis, sc.params, ax, av, aF, ath)

    h = 1e-6
    for name, grad in (("x", ax), ("v", av), ("F", aF), ("logE", ath)):
        direction = rng.standard_normal(grad.shape)
        plus, minus = sc.copy(), sc.copy()
        setattr(plus, name, getattr(sc, name) + h * direction)
        setattr(minus, name, getattr(sc, name) - h * direction)
        for s in (plus, minus):
            s.sigma, s.V = consistent(s, s.F, s.logE)
        fd = (loss(plus) - loss(minus)) / (2 * h)
        ad = float((grad * direction).sum())
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3), (name, ad, fd)


def test_adjoint_of_a_mistake_is_caught(oracle):
    """Planting a wrong factor in the gradient (here: dropping the stress term by
    zeroing σ in the state handed to the adjoint) makes the check fail, so the
    finite-difference test above can see a dropped term."""
    K = 4
    sc = make_scene(2, 5, th.BSpline(th.Quadratic()))
    rng = np.random.default_rng(3)
    a = rng.standard_normal((sc.n, 2))
    _, states = rollout(oracle, sc, K, keep=True)
    ax, av, aF = a.copy(), np.zeros((sc.n, 2)), np.zeros((sc.n, 4))
    ath = np.zeros(sc.n)
    for s in reversed(states):
        s.sigma[:] = 0.0
        oracle.step_adjoint(s, sc.basis, sc.params, ax, av, aF, ath)
    direction = rng.standard_normal(sc.logE.shape)
    plus, minus = sc.copy(), sc.copy()
    plus.logE = sc.logE + 1e-6 * direction
    minus.logE = sc.logE - 1e-6 * direction
    for s in (plus, minus):
        s.sigma, s.V = consistent(s, s.F, s.logE)
    fd = (float((a * rollout(oracle, plus, K)[0].x).sum()) - float((a * rollout(oracle, minus, K)[0].x).sum())) / 2e-6
    assert abs(float((ath * direction).sum()) - fd) > 1e-6 * abs(fd)
