"""GPU parity tests: libtesserae_hip against the CPU oracle on the same inputs.

Bars: bit-exact for integer work (support windows, partition); Float64
transfers within 1e-12 of the oracle relative to the field's largest entry;
Float32 transfers against the Float64 oracle within the stated tolerances.
"""
import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import grid_np, particles_np, rel_err, to_gpu
from oracle_bind import HostState
from tesserae_hip.scenes import block_drop, random_box, two_disks

pytestmark = pytest.mark.gpu

BASES = [th.BSpline(th.Linear()), th.BSpline(th.Quadratic()), th.BSpline(th.Cubic()), th.uGIMP()]
BASIS_IDS = ["linear", "quadratic", "cubic", "ugimp"]


def _scene(dim, basis, n, seed, dtype, h=0.1, lo=-0.05, hi=1.05):
    # particles spill past the mesh faces so truncated and empty windows occur
    return random_box(dim, n, seed, basis, dtype=dtype, h=h, lo=lo, hi=hi, stress=1.0, with_l=basis.kind == 16)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=BASIS_IDS)
def test_update_weights_parity(oracle, dtype, dim, basis):
    print(f"[weights] dim={dim} dtype={dtype} basis={basis}", flush=True)
    sc = _scene(dim, basis, 3000, 21, dtype)
    mesh, P, _ = to_gpu(sc)
    W = th.generate_basis_weights(basis, mesh, sc.n)
    th.update(W, P, mesh)
    torch.cuda.synchronize()
    Wo = oracle.update_weights(HostState(sc), basis)
    win = W.window.cpu().numpy()
    np.testing.assert_array_equal(win, Wo.window)
    sw = basis.support_width
    w, gw = W.w.double().cpu().numpy(), W.grad_w.double().cpu().numpy()
    tol = 1e-14 if dtype == torch.float64 else 2e-6
    for p in range(0, sc.n, 7):
        ext = [win[p, 2 * d + 1] - win[p, 2 * d] + 1 for d in range(dim)]
        if min(ext) <= 0:
            continue
        for s in range(sw ** dim):
            loc = [(s // sw ** d) % sw for d in range(dim)]
            if any(loc[d] >= ext[d] for d in range(dim)):
                continue
            assert abs(w[p, s] - Wo.w[p, s]) <= tol, (p, s)
            assert np.all(np.abs(gw[p, s] - Wo.grad_w[p, s]) <= tol * 10 / mesh.h), (p, s)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("L", [0, 2, 3])
def test_partition_bit_exact(oracle, dtype, dim, L):
    print(f"[partition] dim={dim} dtype={dtype} L={L}", flush=True)
    sc = _scene(dim, BASES[1], 20000, 5 + L, dtype, h=0.02)
    mesh = th.CartesianMesh(sc.mesh.h, *[(0.0, 1.0)] * dim, dtype=dtype, block_size_log2=L)
    _, P, _ = to_gpu(sc)
    part = th.Partition(mesh)
    th.update(part, P.x)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    pidx, starts, stops, na = oracle.partition(mesh, sc.x, dtype=npdt)
    assert part.nassigned == na
    np.testing.assert_array_equal(part.particleindices[:na].cpu().numpy(), pidx)
    np.testing.assert_array_equal(part.starts.cpu().numpy(), starts)
    np.testing.assert_array_equal(part.stops.cpu().numpy(), stops)


def test_partition_edge_cases(oracle):
    mesh = th.CartesianMesh(0.25, (0, 4), (0, 4))
    for pts in ([[0.125, 0.125], [10.0, 10.0], [3.875, 0.125]], [[10.0, 10.0]], [[0.125, 0.125]]):
        x = torch.tensor(pts, dtype=torch.float64, device="cuda")
        part = th.Partition(mesh)
        th.update(part, x)
        pidx, starts, stops, na = oracle.partition(mesh, np.array(pts))
        assert part.nassigned == na
        np.testing.assert_array_equal(part.particleindices[:na].cpu().numpy(), pidx)
        np.testing.assert_array_equal(part.starts.cpu().numpy(), starts)


def _p2g_both(oracle, sc, stored, layout="julia"):
    mesh, P, G = to_gpu(sc, layout=layout)
    if stored:
        W = th.generate_basis_weights(sc.basis, mesh, sc.n)
        th.update(W, P, mesh)
    else:
        W = th.generate_basis_weights(sc.basis, mesh, sc.n, deferred=True)
    th.P2G(G, P, W, sc.params)
    torch.cuda.synchronize()
    st = HostState(sc)
    Wo = oracle.update_weights(HostState(sc), sc.basis) if stored else None
    oracle.p2g(st, sc.basis, sc.params, Wo)
    return G, P, W, st, Wo


@pytest.mark.parametrize("stored", [False, True], ids=["deferred", "stored"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=BASIS_IDS)
def test_p2g_fp64_parity(oracle, stored, dim, basis):
    print(f"[p2g f64] dim={dim} basis={basis} stored={stored}", flush=True)
    sc = _scene(dim, basis, 4000 if dim == 3 else 3000, 31, torch.float64)
    G, _, _, st, _ = _p2g_both(oracle, sc, stored)
    g = grid_np(G)
    for name, ref in (("m", st.m), ("mv", st.mv), ("f", st.f), ("vn", st.vn), ("v", st.gv)):
        assert rel_err(g[name], ref) <= 1e-12, name


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", BASES, ids=BASIS_IDS)
def test_g2p_fp64_parity(oracle, dim, basis):
    print(f"[g2p f64] dim={dim} basis={basis}", flush=True)
    sc = _scene(dim, basis, 4000 if dim == 3 else 3000, 41, torch.float64)
    mesh, P, G = to_gpu(sc)
    st = HostState(sc)
    oracle.p2g(st, sc.basis, sc.params)
    # both G2P calls start from the oracle's grid
    G.m.copy_(torch.from_numpy(st.m))
    G.mv.copy_(torch.from_numpy(st.mv))
    G.f.copy_(torch.from_numpy(st.f))
    G.vn.copy_(torch.from_numpy(st.vn))
    G.v.copy_(torch.from_numpy(st.gv))
    th.G2P(G, P, th.generate_basis_weights(basis, mesh, sc.n, deferred=True), sc.params)
    torch.cuda.synchronize()
    oracle.g2p(st, sc.basis, sc.params)
    got = particles_np(P)
    for name in ("x", "v", "grad_v", "F", "sigma", "V"):
        assert rel_err(got[name], getattr(st, name)) <= 1e-12, name


def test_p2g_deterministic_and_layouts(oracle):
    sc = _scene(3, BASES[1], 20000, 51, torch.float32, h=0.05)
    outs = []
    for layout in ("julia", "julia", "soa"):
        G, _, _, _, _ = _p2g_both(oracle, sc, stored=False, layout=layout)
        outs.append(grid_np(G))
    for name in outs[0]:
        np.testing.assert_array_equal(outs[0][name], outs[1][name])  # run to run: bitwise
        np.testing.assert_array_equal(outs[0][name], outs[2][name])  # layout does not change the sums


def test_incremental_binning_matches_oracle_and_is_deterministic(oracle):
    """25 Float64 steps in which about 5 % of the particles change window class every
    step (incremental re-binning by merge), with reorders: two contexts give
    bitwise-identical particles, and the trajectory equals the oracle's once
    particles are matched by identity."""
    sc = block_drop(dtype=torch.float64, dx=1 / 48, steps=25)
    # a drift of a third of a cell over the run plus noise
    sc.v[:] = np.random.default_rng(3).normal(scale=2.0, size=sc.v.shape) + np.array([3.0, -2.0, 1.0])
    sc.l = np.arange(sc.n, dtype=np.float64)  # identity tag carried along by the permutation
    outs = []
    for _ in range(2):
        ctx = th.Context()
        mesh, P, G = to_gpu(sc)
        th.step(G, P, sc.basis, sc.params, nsteps=25, reorder_threshold=0.9, ctx=ctx)
        ctx.sync()
        st = ctx.stats()
        assert 0 < st["movers"] <= sc.n // 8, st
        outs.append(particles_np(P))
        ctx.close()
    for name in outs[0]:
        np.testing.assert_array_equal(outs[0][name], outs[1][name])
    got = outs[0]
    order = np.argsort(got["l"])
    assert np.array_equal(got["l"][order], np.arange(sc.n))
    hs = HostState(sc)
    oracle.step(hs, sc.basis, sc.params, 25)
    for name in ("x", "v", "F", "sigma", "V"):
        assert rel_err(got[name][order], getattr(hs, name)) <= 1e-10, name


def test_two_disks_fp64_trajectory(oracle):
    """README two-disk collision, Float64: 200 GPU steps against 200 oracle steps."""
    sc = two_disks(steps=200)
    mesh, P, G = to_gpu(sc)
    th.step(G, P, sc.basis, sc.params, nsteps=200)
    torch.cuda.synchronize()
    st = HostState(sc)
    oracle.step(st, sc.basis, sc.params, 200)
    got = particles_np(P)
    for name in ("x", "v", "F", "sigma", "V"):
        err = rel_err(got[name], getattr(st, name))
        print(f"two_disks f64 200 steps {name}: rel err {err:.3e}")
        assert err <= 1e-10, name


def test_contiguity_score_of_the_step_loop():
    """The reorder score (block_ordered_particle_contiguity's meaning, cpu_block_strategy.jl:222-246): right after
    a reorder every neighbour pair of the layered order is consecutive in storage; with the positions randomly
    permuted almost none is."""
    def score_after(sc, threshold):
        ctx = th.Context()
        mesh, P, G = to_gpu(sc)
        th.step(G, P, sc.basis, sc.params, nsteps=2, reorder_threshold=threshold, ctx=ctx)
        ctx.sync()
        st = ctx.stats()
        ctx.close()
        pairs = st["n_inside"] - st["active_tiles"]
        assert st["misplaced"] >= 0 and pairs > 1000, st
        return 1.0 - st["misplaced"] / pairs, st

    sc = two_disks(h=0.01, steps=2)
    score, st = score_after(sc, 1.0)  # reorders before each P2G: the counted P2G sees the layered order
    assert st["misplaced"] == 0 and st["reorders"] == 2, st
    sc = two_disks(h=0.01, steps=2)
    sc.x[:] = sc.x[np.random.default_rng(5).permutation(sc.n)]
    score, st = score_after(sc, 1e-9)  # never reorders (the score stays above 1e-9)
    assert st["reorders"] == 0 and score < 0.02, (score, st)


def test_step_with_reorder_matches_oracle(oracle):
    """Steps that permute particles into the engine's tile order (threshold 1: every
    step) give the oracle's trajectory once particles are matched by identity."""
    sc = two_disks(h=0.02, steps=30)
    sc.l = np.arange(sc.n, dtype=np.float64)  # identity tag carried along by the permutation
    mesh, P, G = to_gpu(sc)
    th.step(G, P, sc.basis, sc.params, nsteps=30, reorder_threshold=1.0)
    torch.cuda.synchronize()
    got = particles_np(P)
    order = np.argsort(got["l"])
    assert np.array_equal(got["l"][order], np.arange(sc.n))
    st = HostState(sc)
    oracle.step(st, sc.basis, sc.params, 30)
    for name in ("x", "v", "F", "sigma", "V"):
        assert rel_err(got[name][order], getattr(st, name)) <= 1e-10, name


@pytest.mark.parametrize("drift", [0.0, 10.0], ids=["still", "drift"])
def test_steps_in_one_call_equal_single_step_calls(drift):
    """Block drop at dx=1/64, FP32, 8 steps.  In one call of 8 steps each binning after
    the first starts from the keys the previous G2P wrote; in 8 calls of one step each
    binning recomputes the keys from the positions in the previous order.  Either way a
    binning in which no particle changed its key (the still block: gravity moves it by
    3e-6 m) keeps the previous tile layout.  The particles agree bitwise.  With a 10 m/s
    drift along x particles cross cells and the merge of movers runs instead."""
    out, kept = [], []
    for calls in (1, 8):
        ctx = th.Context()
        sc = block_drop(dx=1 / 64, steps=8)
        sc.v = sc.v + np.array([drift, 0.0, 0.0])
        mesh, P, G = to_gpu(sc, dtype=torch.float32)
        for _ in range(calls):
            th.step(G, P, sc.basis, sc.params, nsteps=8 // calls, ctx=ctx)
        torch.cuda.synchronize()
        out.append(particles_np(P))
        st = ctx.stats()
        kept.append(st["kept_binnings"])
        print(f"drift {drift} calls {calls}: kept binnings {st['kept_binnings']} movers {st['movers']}")
        ctx.close()
    for name, ref in out[1].items():
        np.testing.assert_array_equal(out[0][name], ref, err_msg=name)
    if drift == 0.0:
        assert kept == [7, 7]
    else:
        assert kept[0] < 7 and kept[1] < 7


def test_block_drop_fp32_vs_fp64_oracle(oracle):
    """3D block drop at dx=1/64 (reduced size), Float32 GPU vs Float64 oracle after 20 steps."""
    sc = block_drop(dx=1 / 64, steps=20)
    mesh, P, G = to_gpu(sc, dtype=torch.float32)
    th.step(G, P, sc.basis, sc.params, nsteps=20)
    torch.cuda.synchronize()
    sc64 = sc.copy()
    sc64.mesh = th.CartesianMesh(sc.mesh.h, (0, 1), (0, 1), (0, 1), dtype=torch.float64)
    st = HostState(sc64)
    oracle.step(st, sc.basis, sc.params, 20)
    got = particles_np(P)
    ex = rel_err(got["x"], st.x)
    print(f"block_drop f32 vs f64 oracle, 20 steps: x {ex:.3e} v {rel_err(got['v'], st.v):.3e} "
          f"F {rel_err(got['F'], st.F):.3e}")
    assert ex <= 1e-5


def test_reorder_particles(oracle):
    sc = _scene(2, BASES[1], 5000, 61, torch.float64, h=0.05, lo=0.0, hi=1.0)
    mesh, P, _ = to_gpu(sc)
    part = th.Partition(mesh)
    th.update(part, P.x)
    before = particles_np(P)
    perm = part.particleindices[:part.nassigned].cpu().numpy() - 1
    assert th.reorder_particles(P, part, threshold=1.0)
    after = particles_np(P)
    for name, a in before.items():
        np.testing.assert_array_equal(after[name][:len(perm)], a[perm])
    np.testing.assert_array_equal(part.particleindices[:part.nassigned].cpu().numpy(),
                                  np.arange(1, part.nassigned + 1))
    th.update(part, P.x)
    assert not th.reorder_particles(P, part, threshold=0.85)


def test_errors_mirror_reference():
    mesh = th.CartesianMesh(0.1, (0, 1), (0, 1))
    sc = random_box(2, 100, 1, th.uGIMP(), h=0.1, with_l=False)
    _, P, G = to_gpu(sc)
    with pytest.raises(th.ArgumentError):
        th.P2G(G, P, th.uGIMP(), sc.params)  # uGIMP needs `l`
    sc2 = random_box(2, 100, 1, th.uGIMP(), h=0.1, with_l=True)
    sc2.l[:] = 0.5  # l > h
    _, P2, G2 = to_gpu(sc2)
    with pytest.raises(th.ArgumentError):
        th.P2G(G2, P2, th.uGIMP(), sc2.params)
    with pytest.raises(th.ArgumentError):
        th.CartesianMesh(0.1, (0, 1), (0, 1), block_size_log2=-1)
