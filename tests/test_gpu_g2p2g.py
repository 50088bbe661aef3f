"""GPU @G2P2G (tess_g2p2g) against the oracle's: one README P2G, then one G2P2G
pass (the G2P half, then the P2G and grid-node halves with the weights at the
positions before the pass).  Float64 within 1e-12 relative to each field's
largest entry; stored and deferred weights agree; run-to-run bitwise equal.

The library runs a pass as the G2P kernel followed by the P2G kernels at the
kept positions; with TESS_G2P2G_SWEEP=1 and deferred weights it runs one
particle sweep (k_g2p2g_fast) instead.  After one pass the two give the same
particles and grid, bit for bit.  Over several passes they stay within a few
units in the last place: the G2P kernel's two walks of a tile (class order and
sorted order) are separately compiled copies of the same arithmetic, which the
compiler may contract differently, and the two paths walk some tiles
differently."""
import contextlib
import os

import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import grid_np, particles_np, rel_err, to_gpu
from oracle_bind import HostState
from tesserae_hip.scenes import random_box

pytestmark = pytest.mark.gpu

BASES = [th.BSpline(th.Linear()), th.BSpline(th.Quadratic()), th.BSpline(th.Cubic()), th.uGIMP()]
BASIS_IDS = ["linear", "quadratic", "cubic", "ugimp"]


def _scene(dim, basis, seed, dtype=torch.float64):
    # particles spill past the mesh faces so truncated and empty windows occur
    return random_box(dim, 4000 if dim == 3 else 3000, seed, basis, dtype=dtype, h=0.1, lo=-0.05, hi=1.05,
                      stress=1.0, with_l=basis.kind == 16)


def _gpu(sc, basis, stored, dtype=None):
    mesh, P, G = to_gpu(sc, dtype=dtype)
    th.P2G(G, P, th.generate_basis_weights(basis, mesh, sc.n, deferred=True), sc.params)
    if stored:
        W = th.generate_basis_weights(basis, mesh, sc.n)
        th.update(W, P, mesh)
    else:
        W = th.generate_basis_weights(basis, mesh, sc.n, deferred=True)
    th.G2P2G(G, P, W, sc.params)
    torch.cuda.synchronize()
    return particles_np(P), grid_np(G)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=BASIS_IDS)
def test_g2p2g_fp64_matches_oracle(oracle, dim, basis):
    print(f"[g2p2g f64] dim={dim} basis={basis}", flush=True)
    sc = _scene(dim, basis, 71 + dim)
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params)
    oracle.g2p2g(st, basis, sc.params)
    parts, grid = _gpu(sc, basis, stored=False)
    for name in ("x", "v", "grad_v", "F", "sigma", "V"):
        assert rel_err(parts[name], getattr(st, name)) <= 1e-12, name
    for name, ref in (("m", st.m), ("mv", st.mv), ("f", st.f), ("vn", st.vn), ("v", st.gv)):
        assert rel_err(grid[name], ref) <= 1e-12, name
    # stored weights (updated at the positions before the pass) give the same pass
    parts_s, grid_s = _gpu(sc, basis, stored=True)
    for name in ("x", "v", "F", "sigma"):
        assert rel_err(parts_s[name], parts[name]) <= 1e-12, name
    for name in ("m", "mv", "f", "v"):
        assert rel_err(grid_s[name], grid[name]) <= 1e-12, name


def test_g2p2g_fp32_and_deterministic(oracle):
    """FP32 G2P2G of a 3D quadratic scene against the FP64 oracle, and twice bitwise."""
    basis = th.BSpline(th.Quadratic())
    sc = _scene(3, basis, 5)
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params)
    oracle.g2p2g(st, basis, sc.params)
    a = _gpu(sc, basis, stored=False, dtype=torch.float32)
    b = _gpu(sc, basis, stored=False, dtype=torch.float32)
    for name in a[0]:
        np.testing.assert_array_equal(a[0][name], b[0][name])
    for name in a[1]:
        np.testing.assert_array_equal(a[1][name], b[1][name])
    assert rel_err(a[0]["x"], st.x) <= 1e-6
    assert rel_err(a[1]["m"], st.m) <= 1e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p2g_steps_match_oracle_and_single_passes(oracle, dim):
    """tess_g2p2g_steps: four passes in one call equal the oracle's four passes (FP64) and four single-pass
    calls; with the layered reorder the particles are the same set (up to their order)."""
    basis = th.BSpline(th.Quadratic())
    sc = _scene(dim, basis, 91 + dim)
    sc.x = np.clip(sc.x, 0.02, 0.98)  # keep every particle inside so the reorder keeps them all binned
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params)
    for _ in range(4):
        oracle.g2p2g(st, basis, sc.params)
    runs = {}
    for mode in ("loop", "single", "reorder"):
        mesh, P, G = to_gpu(sc)
        th.P2G(G, P, th.generate_basis_weights(basis, mesh, sc.n, deferred=True), sc.params)
        if mode == "single":
            for _ in range(4):
                th.G2P2G(G, P, basis, sc.params)
        else:
            th.g2p2g_steps(G, P, basis, sc.params, nsteps=4, reorder_threshold=1.0 if mode == "reorder" else 0.0)
        torch.cuda.synchronize()
        runs[mode] = (particles_np(P), grid_np(G))
    for name in ("x", "v", "F", "sigma"):
        assert rel_err(runs["loop"][0][name], getattr(st, name)) <= 1e-12, name
        assert rel_err(runs["loop"][0][name], runs["single"][0][name]) <= 1e-13, name
    for name in ("m", "mv", "f"):
        assert rel_err(runs["loop"][1][name], runs["single"][1][name]) <= 1e-13, name
        assert rel_err(runs["reorder"][1][name], runs["loop"][1][name]) <= 1e-12, name
    key = lambda p: np.lexsort(p["x"].T[::-1])
    a, b = runs["loop"][0], runs["reorder"][0]
    ia, ib = key(a), key(b)
    for name in ("x", "v", "F"):
        assert rel_err(b[name][ib], a[name][ia]) <= 1e-12, name


@contextlib.contextmanager
def _sweep(on):
    """TESS_G2P2G_SWEEP=1 (one particle sweep per pass) or 0 (the G2P kernel, then the P2G kernels)."""
    old = os.environ.get("TESS_G2P2G_SWEEP")
    os.environ["TESS_G2P2G_SWEEP"] = "1" if on else "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["TESS_G2P2G_SWEEP"]
        else:
            os.environ["TESS_G2P2G_SWEEP"] = old


def _assert_same(a, b):
    for part_a, part_b in zip(a, b):
        assert part_a.keys() == part_b.keys()
        for name in part_a:
            np.testing.assert_array_equal(part_a[name], part_b[name], err_msg=name)


def _assert_close(a, b, tol):
    for part_a, part_b in zip(a, b):
        assert part_a.keys() == part_b.keys()
        for name in part_a:
            assert rel_err(part_a[name], part_b[name]) <= tol, name


def _passes(sc, basis, dtype, layout, sweep, npass):
    mesh, P, G = to_gpu(sc, layout=layout, dtype=dtype)
    W = th.generate_basis_weights(basis, mesh, sc.n, deferred=True)
    th.P2G(G, P, W, sc.params)
    with _sweep(sweep):
        for _ in range(npass):
            th.G2P2G(G, P, W, sc.params)
    torch.cuda.synchronize()
    return particles_np(P), grid_np(G)


@pytest.mark.parametrize("layout", ["julia", "soa"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=BASIS_IDS)
def test_g2p2g_sweep_equals_composed_kernels(dim, basis, dtype, layout):
    """One G2P2G pass: the single sweep and the composed kernels agree bit for bit on every particle and grid
    field (truncated windows and particles outside the mesh included); over three passes within a few units in
    the last place (1e-13 Float64, 1e-6 Float32, relative to each field's largest entry)."""
    print(f"[g2p2g sweep] dim={dim} basis={basis} dtype={dtype} layout={layout}", flush=True)
    sc = _scene(dim, basis, 131 + dim)
    _assert_same(_passes(sc, basis, dtype, layout, True, 1), _passes(sc, basis, dtype, layout, False, 1))
    tol = 1e-13 if dtype == torch.float64 else 1e-6
    _assert_close(_passes(sc, basis, dtype, layout, True, 3), _passes(sc, basis, dtype, layout, False, 3), tol)


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p2g_sweep_dense_tiles(dim):
    """Tiles holding more particles than the kernels stage in LDS (over 1024, read through the global
    permutation): one sweep equals the composed kernels bit for bit (Float32)."""
    basis = th.BSpline(th.Quadratic())
    n = 60000 if dim == 3 else 20000
    sc = random_box(dim, n, 17 + dim, basis, h=0.1, lo=0.0, hi=1.0, stress=1.0)
    runs = []
    for sweep in (True, False):
        ctx = th.Context()
        mesh, P, G = to_gpu(sc, dtype=torch.float32)
        W = th.generate_basis_weights(basis, mesh, sc.n, deferred=True)
        th.P2G(G, P, W, sc.params, ctx=ctx)
        tiles = ctx.stats()["active_tiles"]
        with _sweep(sweep):
            th.G2P2G(G, P, W, sc.params, ctx=ctx)
        ctx.sync()
        runs.append((particles_np(P), grid_np(G)))
        ctx.close()
    assert n / tiles > 1024, (n, tiles)
    _assert_same(runs[0], runs[1])


@pytest.mark.parametrize("threshold", [0.0, 0.85, 1.0])
@pytest.mark.parametrize("dim", [2, 3])
def test_g2p2g_steps_sweep_matches_composed_kernels(dim, threshold):
    """tess_g2p2g_steps with the sweep (next keys written by its G2P half, the contiguity count of its P2G
    half, reorders at the threshold) against the composed kernels: the same reorders, movers and kept
    binnings, and fields within 1e-6 (Float32, six passes)."""
    basis = th.BSpline(th.Quadratic())
    sc = _scene(dim, basis, 151 + dim)
    runs, stats = [], []
    for sweep in (True, False):
        ctx = th.Context()
        mesh, P, G = to_gpu(sc, dtype=torch.float32)
        th.P2G(G, P, th.generate_basis_weights(basis, mesh, sc.n, deferred=True), sc.params, ctx=ctx)
        with _sweep(sweep):
            th.g2p2g_steps(G, P, basis, sc.params, nsteps=6, reorder_threshold=threshold, ctx=ctx)
        ctx.sync()
        runs.append((particles_np(P), grid_np(G)))
        st = ctx.stats()
        stats.append({k: st[k] for k in ("reorders", "misplaced", "movers", "kept_binnings")})
        ctx.close()
    assert stats[0] == stats[1]
    _assert_close(runs[0], runs[1], 1e-6)
