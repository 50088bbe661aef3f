"""GPU corrected bases (WLS, KernelCorrection) against the oracle's.

Stored weights (update!) and deferred transfers, with and without a node
filter, in 2D and 3D over scenes whose particles spill past the mesh faces so
that truncated windows (the corrected branches) occur alongside whole ones.
Float64 within 1e-11 relative to each field's largest entry (the device fits in
units of h by Cholesky, the oracle in world units by Gauss-Jordan: the two agree
to rounding times the fit's conditioning); Float32 against the Float64 oracle
within 1e-5 (weights) and 1e-4 (grid fields).
"""
import itertools

import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import grid_np, particles_np, rel_err, to_gpu
from oracle_bind import HostState, HostWeights, host_filter
from tesserae_hip.scenes import random_box

pytestmark = pytest.mark.gpu

LIN, ML = th.Polynomial(th.Linear()), th.Polynomial(th.MultiLinear())
Q, C, LN, G = th.BSpline(th.Quadratic()), th.BSpline(th.Cubic()), th.BSpline(th.Linear()), th.uGIMP()
BASES = [th.KernelCorrection(Q), th.KernelCorrection(C), th.KernelCorrection(G), th.KernelCorrection(Q, LIN),
         th.WLS(Q), th.WLS(C), th.WLS(LN), th.WLS(G), th.WLS(Q, ML), th.WLS(G, ML)]
IDS = [repr(b).replace("Polynomial", "P").replace("BSpline", "B") for b in BASES]


def _scene(dim, basis, seed, dtype=torch.float64, n=None):
    n = n or (2500 if dim == 2 else 3000)
    return random_box(dim, n, seed, basis, dtype=dtype, h=0.1, lo=-0.04, hi=1.04, stress=1.0,
                      with_l=basis.kernel == 16)


def _mask(sc, hole=True):
    """Every node within the particles' supports, minus one node in the middle (a hole inside the body)."""
    mask = np.zeros(sc.mesh.size(), bool)
    inside = np.all((sc.x >= 0) & (sc.x <= 1), axis=1)
    idx = np.floor(sc.x[inside] / sc.mesh.h).astype(int)
    for off in itertools.product(*[range(-1, 3)] * sc.mesh.dim):
        J = np.clip(idx + np.array(off), 0, np.array(sc.mesh.size()) - 1)
        mask[tuple(J.T)] = True
    if hole:
        mask[tuple(s // 2 for s in sc.mesh.size())] = False
    return mask


def _check_weights(W, Wo, mesh, basis, tol):
    win = W.window.cpu().numpy()
    np.testing.assert_array_equal(win, Wo.window)
    sw, dim = basis.support_width, mesh.dim
    w, gw = W.w.double().cpu().numpy(), W.grad_w.double().cpu().numpy()
    gscale = np.abs(Wo.grad_w).max()
    for p in range(len(win)):
        ext = [win[p, 2 * d + 1] - win[p, 2 * d] + 1 for d in range(dim)]
        if min(ext) <= 0:
            continue
        for s in range(sw ** dim):
            loc = [(s // sw ** d) % sw for d in range(dim)]
            if any(loc[d] >= ext[d] for d in range(dim)):
                continue
            assert abs(w[p, s] - Wo.w[p, s]) <= tol, (p, s, w[p, s], Wo.w[p, s])
            assert np.all(np.abs(gw[p, s] - Wo.grad_w[p, s]) <= tol * gscale), (p, s)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=IDS)
def test_update_weights_fp64(oracle, dim, basis):
    print(f"[corr weights f64] dim={dim} basis={basis!r}", flush=True)
    sc = _scene(dim, basis, 31 + dim)
    mesh, P, _ = to_gpu(sc)
    W = th.generate_basis_weights(basis, mesh, sc.n)
    th.update(W, P, mesh)
    torch.cuda.synchronize()
    _check_weights(W, oracle.update_weights(HostState(sc), basis), mesh, basis, 1e-11)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", [th.KernelCorrection(Q), th.WLS(Q), th.WLS(Q, ML), th.KernelCorrection(C)],
                         ids=["KC(Q)", "WLS(Q)", "WLS(Q,ML)", "KC(C)"])
def test_update_weights_filtered_fp64(oracle, dim, basis):
    print(f"[corr weights filter] dim={dim} basis={basis!r}", flush=True)
    sc = _scene(dim, basis, 41 + dim)
    mask = _mask(sc)
    mesh, P, _ = to_gpu(sc)
    W = th.generate_basis_weights(basis, mesh, sc.n)
    th.update(W, P, mesh, mask)
    torch.cuda.synchronize()
    _check_weights(W, oracle.update_weights(HostState(sc), basis, filter=mask), mesh, basis, 1e-11)


def test_update_weights_fp32(oracle):
    basis = th.KernelCorrection(Q)
    sc = _scene(3, basis, 5)
    mesh, P, _ = to_gpu(sc, dtype=torch.float32)
    W = th.generate_basis_weights(basis, mesh, sc.n)
    th.update(W, P, mesh)
    torch.cuda.synchronize()
    _check_weights(W, oracle.update_weights(HostState(sc), basis), mesh, basis, 1e-5)


def _transfers(sc, basis, mode, dtype=None, mask=None):
    """P2G then G2P on the GPU with stored ('stored') or deferred ('deferred') weights."""
    mesh, P, Gr = to_gpu(sc, dtype=dtype)
    W = th.generate_basis_weights(basis, mesh, sc.n, deferred=mode == "deferred")
    th.update(W, P, mesh, mask)
    th.P2G(Gr, P, W, sc.params)
    grid = grid_np(Gr)
    th.update(W, P, mesh, mask)
    th.G2P(Gr, P, W, sc.params)
    torch.cuda.synchronize()
    return grid, particles_np(P)


def _oracle_transfers(oracle, sc, basis, mode, mask=None):
    st = HostState(sc)
    if mode == "stored":
        W = oracle.update_weights(st, basis, filter=mask)
    else:
        W = HostWeights(basis, sc.mesh.dim, sc.n, st.npdt, filter=host_filter(sc.mesh, mask), deferred=True) \
            if mask is not None else None
    oracle.p2g(st, basis, sc.params, W)
    grid = {"m": st.m.copy(), "mv": st.mv.copy(), "f": st.f.copy(), "vn": st.vn.copy(), "v": st.gv.copy()}
    if mode == "stored":
        W = oracle.update_weights(st, basis, filter=mask)
    oracle.g2p(st, basis, sc.params, W)
    return grid, st


@pytest.mark.parametrize("mode", ["deferred", "stored"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=IDS)
def test_transfers_fp64(oracle, mode, dim, basis):
    print(f"[corr transfers f64] mode={mode} dim={dim} basis={basis!r}", flush=True)
    sc = _scene(dim, basis, 51 + dim)
    grid, parts = _transfers(sc, basis, mode)
    og, st = _oracle_transfers(oracle, sc, basis, mode)
    for name in ("m", "mv", "f", "vn", "v"):
        assert rel_err(grid[name], og[name]) <= 1e-11, name
    for name in ("x", "v", "grad_v", "F", "sigma", "V"):
        assert rel_err(parts[name], getattr(st, name)) <= 1e-11, name


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", [th.KernelCorrection(Q), th.WLS(Q), th.KernelCorrection(C)],
                         ids=["KC(Q)", "WLS(Q)", "KC(C)"])
def test_deferred_transfers_with_filter_fp64(oracle, dim, basis):
    print(f"[corr transfers filter] dim={dim} basis={basis!r}", flush=True)
    sc = _scene(dim, basis, 61 + dim)
    mask = _mask(sc)
    grid, parts = _transfers(sc, basis, "deferred", mask=mask)
    og, st = _oracle_transfers(oracle, sc, basis, "deferred", mask=mask)
    for name in ("m", "mv", "f", "v"):
        assert rel_err(grid[name], og[name]) <= 1e-11, name
    for name in ("x", "v", "grad_v", "F"):
        assert rel_err(parts[name], getattr(st, name)) <= 1e-11, name
    # stored weights updated with the same filter give the same transfers
    grid_s, parts_s = _transfers(sc, basis, "stored", mask=mask)
    for name in ("m", "mv", "f"):
        assert rel_err(grid_s[name], grid[name]) <= 1e-11, name


@pytest.mark.parametrize("dim", [2, 3])
def test_step_fp64_and_fp32(oracle, dim):
    """Steps of the README loop with KernelCorrection(BSpline(Quadratic())) against the oracle's; Float32 twice
    bitwise."""
    basis = th.KernelCorrection(Q)
    sc = _scene(dim, basis, 71 + dim)
    st = HostState(sc)
    oracle.step(st, basis, sc.params, nsteps=3)
    mesh, P, Gr = to_gpu(sc)
    th.step(Gr, P, basis, sc.params, nsteps=3)
    torch.cuda.synchronize()
    parts = particles_np(P)
    for name in ("x", "v", "F", "sigma"):
        assert rel_err(parts[name], getattr(st, name)) <= 1e-10, name
    runs = []
    for _ in range(2):
        mesh, P, Gr = to_gpu(sc, dtype=torch.float32)
        th.step(Gr, P, basis, sc.params, nsteps=3)
        torch.cuda.synchronize()
        runs.append(particles_np(P))
    for name in runs[0]:
        np.testing.assert_array_equal(runs[0][name], runs[1][name])
    assert rel_err(runs[0]["x"], st.x) <= 1e-5


def test_sparse_grid_with_kernel_correction():
    """An SpGrid step with a corrected basis equals the dense-grid step node for node."""
    basis = th.KernelCorrection(Q)
    sc = _scene(3, basis, 81)
    mesh = th.CartesianMesh(sc.mesh.h, *[(0.0, 1.0)] * 3, block_size_log2=2)
    P = th.Particles(sc.n, 3, torch.float64, device="cuda")
    for name in ("x", "v", "V", "F", "sigma", "grad_v"):
        P.set(name, getattr(sc, name))
    Gd = th.generate_grid(mesh)
    th.step(Gd, P, basis, sc.params, nsteps=2)
    dense = particles_np(P)
    P2 = th.Particles(sc.n, 3, torch.float64, device="cuda")
    for name in ("x", "v", "V", "F", "sigma", "grad_v"):
        P2.set(name, getattr(sc, name))
    Gs = th.generate_grid(th.SpArray, mesh)
    th.step(Gs, P2, basis, sc.params, nsteps=2)
    torch.cuda.synchronize()
    sparse = particles_np(P2)
    for name in ("x", "v", "F"):
        assert rel_err(sparse[name], dense[name]) <= 1e-13, name


def test_corrected_basis_errors():
    """A plain kernel refuses a filter (require_filtered_update_support); batches and the adjoint take plain
    kernels only."""
    sc = _scene(2, Q, 3, n=200)
    mesh, P, Gr = to_gpu(sc)
    W = th.generate_basis_weights(Q, mesh, sc.n)
    with pytest.raises(th.ArgumentError, match="filtered"):
        th.update(W, P, mesh, np.ones(mesh.size(), bool))
    with pytest.raises(th.UnsupportedError):
        th.rollout(Gr, P, th.KernelCorrection(Q), sc.params, nsteps=1)


def test_taylor_rod_fp64(oracle):
    """The Taylor-impact rod (one particle per cell) with KernelCorrection(BSpline(Quadratic())): the bottom layer
    sits in truncated supports on the floor, so the separable correction runs in the fast kernels; two steps
    against the oracle's."""
    from tesserae_hip.scenes import taylor_rod
    sc = taylor_rod(dtype=torch.float64, per_cell=1)
    basis = sc.basis
    print(f"[taylor rod f64] n={sc.n}", flush=True)
    st = HostState(sc)
    oracle.step(st, basis, sc.params, nsteps=2)
    mesh, P, Gr = to_gpu(sc)
    th.step(Gr, P, basis, sc.params, nsteps=2)
    torch.cuda.synchronize()
    parts = particles_np(P)
    for name in ("x", "v", "F", "sigma"):
        assert rel_err(parts[name], getattr(st, name)) <= 1e-10, name
    # the plain kernel differs on the floor layer: the correction is in effect
    P2 = to_gpu(sc)[1]
    Gr2 = th.generate_grid(mesh)
    th.step(Gr2, P2, Q, sc.params, nsteps=2)
    torch.cuda.synchronize()
    assert rel_err(particles_np(P2)["v"], parts["v"]) > 1e-6
