"""Oracle partition and transfer checks.

Partition: brute-force grouping, group order and colour disjointness
(test/partition.jl:53-158).  Transfers: a dense brute-force loop over all
node/particle pairs (no support windows, closed-form kernels) on the
reference's transfer fixture geometry (test/transfer.jl:2-29,61-82), plus the
conservation and affine-reproduction invariants of MPM transfers.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle_bind import HostState, HostWeights
from tesserae_hip import BSpline, CartesianMesh, Cubic, Linear, Quadratic, uGIMP
from tesserae_hip.core import MPMParams, grid_sampling_points
from tesserae_hip.scenes import Scene, _state, random_box, two_disks


# ---------------------------------------------------------------- partition

def _brute_blocks(oracle, mesh, x):
    nb = mesh.nblocks()
    expected = {}
    na = 0
    for p, xp in enumerate(x):
        b = oracle.findblock(mesh, xp)
        if b is not None:
            lin = (b[0] - 1) + nb[0] * (b[1] - 1)
            expected.setdefault(lin, []).append(p + 1)
            na += 1
    return expected, na


def _check_partition(oracle, mesh, x):
    pidx, starts, stops, na = oracle.partition(mesh, x)
    expected, na_exp = _brute_blocks(oracle, mesh, x)
    assert na == na_exp
    nb = mesh.nblocks()
    for b in range(int(np.prod(nb))):
        got = list(pidx[starts[b] - 1:stops[b]]) if starts[b] else []
        assert got == expected.get(b, []), b
    # group order: colours ascending, blocks in linear order within a colour
    order = []
    for color in range(4):
        for b in range(int(np.prod(nb))):
            bx, by = b % nb[0], b // nb[0]
            if (bx & 1) + ((by & 1) << 1) == color and starts[b]:
                order.extend(pidx[starts[b] - 1:stops[b]])
    assert order == list(pidx)
    return pidx, starts, stops, na


def test_partition_circle_shuffled(oracle):
    mesh = CartesianMesh(0.25, (0, 4), (0, 4))
    pts = grid_sampling_points(mesh)
    pts = pts[(pts[:, 0] - 2) ** 2 + (pts[:, 1] - 2) ** 2 < 1]
    pts = pts[np.random.default_rng(1234).permutation(len(pts))]
    pidx, starts, stops, na = _check_partition(oracle, mesh, pts)
    # same-colour blocks never share a cubic support node (test/partition.jl:92-104)
    nb = mesh.nblocks()
    for color in range(4):
        seen = set()
        for b in range(int(np.prod(nb))):
            bx, by = b % nb[0], b // nb[0]
            if (bx & 1) + ((by & 1) << 1) != color or not starts[b]:
                continue
            nodes = set()
            for p in pidx[starts[b] - 1:stops[b]]:
                win, _, _ = oracle.basis_weight(mesh, BSpline(Cubic()), pts[p - 1])
                nodes |= set(itertools.product(range(win[0][0], win[0][1] + 1), range(win[1][0], win[1][1] + 1)))
            assert not (seen & nodes)
            seen |= nodes


def test_partition_moving_and_outside(oracle):
    mesh = CartesianMesh(0.25, (0, 4), (0, 4))
    _check_partition(oracle, mesh, np.array([[0.125, 0.125], [0.375, 0.375], [3.625, 3.625], [3.875, 3.875]]))
    _check_partition(oracle, mesh, np.array([[0.125, 0.125], [0.375, 0.375]]))
    pidx, starts, stops, na = _check_partition(oracle, mesh, np.array([[0.125, 0.125], [10.0, 10.0],
                                                                       [3.875, 0.125]]))
    assert na == 2 and list(pidx) == [1, 3]


# ---------------------------------------------------------------- transfers

def fixture_scene(basis=None, dtype=torch.float64):
    """Geometry and state of test/transfer.jl:2-29,61-82 with a constant particle mass."""
    mesh = CartesianMesh(1.0, (0, 2), (0, 2), dtype=dtype)
    x = grid_sampling_points(mesh)
    n = len(x)
    p = np.arange(1, n + 1, dtype=np.float64)
    st = _state(x, np.stack([0.15 + 0.02 * p, -0.25 + 0.015 * p], axis=1), 0.0, 2)
    st["V"] = 0.2 + 0.01 * p
    st["F"] = np.stack([1.0 + 0.002 * p, 0 * p, 0 * p, 1.0 - 0.001 * p], axis=1)
    a = np.stack([1.0 + 0.03 * p, 0.2 * np.ones(n)], axis=1)
    b = np.stack([0.4 * np.ones(n), 0.8 + 0.01 * p], axis=1)
    outer = a[:, :, None] * b[:, None, :]
    sym = 0.5 * (outer + outer.transpose(0, 2, 1))
    st["sigma"] = sym.transpose(0, 2, 1).reshape(n, 4)
    params = MPMParams(dt=0.01, E=10.0, nu=0.3, rho0=1.0, V0=1.0)
    return Scene("transfer_fixture", mesh, basis or BSpline(Linear()), params, **st)


def _kernel(basis, xi, l=0.0):
    a = abs(xi)
    if basis.kind == 1:
        return max(0.0, 1 - a), (-np.sign(xi) if a < 1 else 0.0)
    if basis.kind == 2:
        if a < 0.5:
            return 0.75 - a * a, -2 * xi
        if a < 1.5:
            return 0.5 * (1.5 - a) ** 2, -(1.5 - a) * np.sign(xi)
        return 0.0, 0.0
    if basis.kind == 3:
        if a < 1:
            return 0.5 * a ** 3 - a ** 2 + 2 / 3, (1.5 * a * a - 2 * a) * np.sign(xi)
        if a < 2:
            return (2 - a) ** 3 / 6, -0.5 * (2 - a) ** 2 * np.sign(xi)
        return 0.0, 0.0
    if a < l / 2:
        return 1 - (4 * a * a + l * l) / (4 * l), -2 * xi / l
    if a < 1 - l / 2:
        return 1 - a, -np.sign(xi)
    if a < 1 + l / 2:
        return (1 + l / 2 - a) ** 2 / (2 * l), -(1 + l / 2 - a) / l * np.sign(xi)
    return 0.0, 0.0


def brute_p2g(scene):
    """Dense loop over every node/particle pair with the closed-form kernel."""
    mesh, basis, prm = scene.mesh, scene.basis, scene.params
    d, N = mesh.dim, mesh.nnodes()
    axes = [mesh.axis(k).astype(np.float64) for k in range(d)]
    m = np.zeros(N)
    mv = np.zeros((N, d))
    f = np.zeros((N, d))
    for p in range(scene.n):
        x = scene.x[p]
        A = -scene.V[p] * scene.sigma[p].reshape(d, d, order="F")
        for i, I in enumerate(itertools.product(*[range(n) for n in reversed(mesh.dims)])):
            I = I[::-1]
            lin = sum(I[k] * int(np.prod(mesh.dims[:k])) for k in range(d))
            jets = [_kernel(basis, (x[k] - axes[k][I[k]]) / mesh.h,
                            (scene.l[p] / mesh.h if scene.l is not None else 0.0)) for k in range(d)]
            w = np.prod([j[0] for j in jets])
            if w == 0 and all(j[1] == 0 for j in jets):
                continue
            g = np.array([np.prod([jets[k][1] if k == c else jets[k][0] for k in range(d)]) / mesh.h
                          for c in range(d)])
            m[lin] += w * prm.mp
            mv[lin] += w * prm.mp * scene.v[p]
            f[lin] += A @ g
    return m, mv, f


@pytest.mark.parametrize("basis", [BSpline(Linear()), BSpline(Quadratic()), BSpline(Cubic())],
                         ids=["linear", "quadratic", "cubic"])
def test_p2g_matches_brute_force(oracle, basis):
    sc = fixture_scene(basis)
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params)
    m, mv, f = brute_p2g(sc)
    np.testing.assert_allclose(st.m, m, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.mv, mv, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.f, f, rtol=1e-12, atol=1e-13)
    # stored weights (fast path) agree with the deferred evaluation
    st2 = HostState(sc)
    W = oracle.update_weights(st2, basis)
    oracle.p2g(st2, basis, sc.params, W)
    np.testing.assert_allclose(st2.m, st.m, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st2.f, st.f, rtol=1e-12, atol=1e-13)


def test_p2g_brute_force_ugimp(oracle):
    sc = random_box(2, 40, 3, uGIMP(), h=0.25, lo=0.3, hi=0.7, stress=1.0, with_l=True)
    st = HostState(sc)
    oracle.p2g(st, sc.basis, sc.params)
    m, mv, f = brute_p2g(sc)
    np.testing.assert_allclose(st.m, m, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.mv, mv, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.f, f, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("deg", [1, 2, 3])
def test_p2g_conservation(oracle, dim, deg):
    basis = BSpline(deg)
    sc = random_box(dim, 300, 11 + deg, basis, h=0.1, lo=0.3, hi=0.7, stress=1.0)
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params)
    mp = sc.params.mp
    assert st.m.sum() == pytest.approx(sc.n * mp, rel=1e-12)
    np.testing.assert_allclose(st.mv.sum(axis=0), mp * sc.v.sum(axis=0), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(st.f.sum(axis=0), 0.0, atol=1e-11 * np.abs(st.f).max())


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", [BSpline(Linear()), BSpline(Quadratic()), BSpline(Cubic()), uGIMP()],
                         ids=["linear", "quadratic", "cubic", "ugimp"])
def test_g2p_affine_reproduction(oracle, dim, basis):
    sc = random_box(dim, 200, 5, basis, h=0.1, lo=0.3, hi=0.7, with_l=basis.kind == 16)
    st = HostState(sc)
    rng = np.random.default_rng(0)
    A = rng.normal(size=(dim, dim))
    b = rng.normal(size=dim)
    c = rng.normal(size=dim)
    axes = np.meshgrid(*[sc.mesh.axis(k) for k in range(dim)], indexing="ij")
    X = np.stack([a.reshape(-1, order="F") for a in axes], axis=1)
    st.gv[:] = X @ A.T + b
    st.vn[:] = st.gv - c
    # a deformed F (det > 0) so that the stress check below sees finite strains
    F0 = np.eye(dim) + 0.3 * rng.uniform(-1, 1, size=(sc.n, dim, dim))
    st.F[:] = F0.transpose(0, 2, 1).reshape(sc.n, dim * dim)
    x0, v0 = st.x.copy(), st.v.copy()
    oracle.g2p(st, basis, sc.params)
    dt = sc.params.dt
    np.testing.assert_allclose(st.grad_v.reshape(-1, dim, dim).transpose(0, 2, 1), np.broadcast_to(A, (sc.n, dim, dim)),
                               atol=1e-10)
    np.testing.assert_allclose(st.x - x0, (x0 @ A.T + b) * dt, atol=1e-12)
    np.testing.assert_allclose(st.v - v0, np.broadcast_to(c, (sc.n, dim)), atol=1e-11)
    # Neo-Hookean closed form on the updated F
    F = st.F.reshape(-1, dim, dim).transpose(0, 2, 1)
    J = np.linalg.det(F)
    lam, mu = sc.params.lam, sc.params.mu
    sig = (mu * (F @ F.transpose(0, 2, 1) - np.eye(dim)) + lam * np.log(J)[:, None, None] * np.eye(dim)) \
        / J[:, None, None]
    np.testing.assert_allclose(st.sigma.reshape(-1, dim, dim).transpose(0, 2, 1), sig, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(st.V, sc.params.V0 * J, rtol=1e-14)
    # F advanced by F <- (I + dt ∇v) F
    np.testing.assert_allclose(F, (np.eye(dim) + dt * A) @ F0, rtol=1e-12, atol=1e-12)
    # the stress is the Kirchhoff stress of the compressible Neo-Hookean energy
    # ψ(F) = μ/2 (tr FᵀF - d) - μ ln J + λ/2 (ln J)², by central differences:
    # J σ = (∂ψ/∂F) Fᵀ.  A dropped term, a wrong sign or a transposed product fails it.
    def psi(G):
        lj = np.log(np.linalg.det(G))
        return mu / 2 * ((G * G).sum(axis=(1, 2)) - dim) - mu * lj + lam / 2 * lj ** 2

    eps = 1e-5
    dpsi = np.zeros_like(F)
    for a in range(dim):
        for b in range(dim):
            E = np.zeros((dim, dim))
            E[a, b] = eps
            dpsi[:, a, b] = (psi(F + E) - psi(F - E)) / (2 * eps)
    tau = dpsi @ F.transpose(0, 2, 1)
    got = st.sigma.reshape(-1, dim, dim).transpose(0, 2, 1) * J[:, None, None]
    np.testing.assert_allclose(got, tau, rtol=1e-6, atol=1e-7 * np.abs(tau).max())


def test_threaded_baseline_matches_sequential(oracle):
    sc = two_disks(h=0.05, steps=5)
    a, b = HostState(sc), HostState(sc)
    oracle.step(a, sc.basis, sc.params, 5)
    oracle.step_threaded(b, sc.basis, sc.params, nthreads=4, nsteps=5)
    for name in ("x", "v", "F", "sigma", "V"):
        ref = getattr(a, name)
        np.testing.assert_allclose(getattr(b, name), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # the threaded colour-phased run is deterministic
    c = HostState(sc)
    oracle.step_threaded(c, sc.basis, sc.params, nthreads=3, nsteps=5)
    for name in ("x", "v", "F", "sigma", "V"):
        np.testing.assert_array_equal(getattr(b, name), getattr(c, name))
