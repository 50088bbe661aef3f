"""SpGrid on the GPU: block numbering bit-exact against the oracle, and P2G / G2P /
steps into compact SpArray storage equal to the dense-grid engine node for node
(the reference's SpGrid P2G == dense check, test/sparray.jl:386-437) and within
1e-12 of the oracle in Float64."""
import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import grid_np, particles_np, rel_err, to_gpu
from oracle_bind import HostState
from tesserae_hip.scenes import random_box, sphere_spin

pytestmark = pytest.mark.gpu


def _sp_grid(mesh):
    return th.generate_grid(th.SpArray, mesh, device="cuda")


@pytest.mark.parametrize("dim,L", [(2, 2), (2, 3), (3, 2), (3, 3)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_sparsity_numbering_bit_exact(oracle, dim, L, dtype):
    print(f"[sparsity] dim={dim} L={L} dtype={dtype}", flush=True)
    mesh = th.CartesianMesh(0.02, *[(0, 1.3)] * dim, dtype=dtype, block_size_log2=L)
    rng = np.random.default_rng(dim * 10 + L)
    x = np.concatenate([rng.uniform(0.1, 0.5, size=(4000, dim)), rng.uniform(-0.2, 1.5, size=(300, dim))])
    npdt = np.float32 if dtype == torch.float32 else np.float64
    G = _sp_grid(mesh)
    xt = torch.from_numpy(x.astype(npdt)).cuda()
    n = th.update_sparsity(G, xt)
    num_ref, len_ref, _ = oracle.sparsity(mesh, x.astype(npdt), dtype=npdt)
    np.testing.assert_array_equal(G.spinds.blocknumbering.cpu().numpy(), num_ref)
    assert n == len_ref == G.spinds.storage_len
    # unchanged active set: None, same numbering; then an emptied set
    assert th.update_sparsity(G, xt) is None
    np.testing.assert_array_equal(G.spinds.blocknumbering.cpu().numpy(), num_ref)
    th.update_sparsity(G, xt[:0])
    assert G.spinds.storage_len == 0 and not G.spinds.blocknumbering.any()


@pytest.mark.parametrize("basis", [th.BSpline(th.Quadratic()), th.BSpline(th.Cubic()), th.uGIMP()],
                         ids=["quadratic", "cubic", "ugimp"])
@pytest.mark.parametrize("dim", [2, 3])
def test_sparse_p2g_g2p_equal_dense(oracle, basis, dim):
    print(f"[sparse p2g/g2p] dim={dim} basis={basis}", flush=True)
    sc = random_box(dim, 6000 if dim == 3 else 3000, 17, basis, dtype=torch.float64, h=0.02, lo=0.2, hi=0.55,
                    stress=1.0, with_l=basis.kind == 16)
    sc.mesh.block_size_log2 = 3
    mesh, P, Gd = to_gpu(sc)
    _, Ps, _ = to_gpu(sc)
    Gs = _sp_grid(mesh)
    ctx = th.Context()
    th.update_sparsity(Gs, Ps.x, ctx=ctx)
    assert Gs.spinds.storage_len < mesh.nnodes()
    th.P2G(Gd, P, basis, sc.params, ctx=ctx)
    th.P2G(Gs, Ps, basis, sc.params, ctx=ctx)
    torch.cuda.synchronize()
    dense = grid_np(Gd)
    idx = th.sparse_storage_index(mesh, Gs.spinds.numbering())
    for name in Gd.FIELDS:
        sp = Gs.tensors[name].cpu().numpy()
        dn = dense[name]
        np.testing.assert_array_equal(sp[idx[idx >= 0]], dn[idx >= 0])  # same sums, bit for bit
        assert not np.any(dn[idx < 0]), name  # nothing lands outside the active blocks
    # against the oracle (dense restatement)
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params)
    for name, ref in (("m", st.m), ("mv", st.mv), ("f", st.f), ("vn", st.vn), ("v", st.gv)):
        assert rel_err(Gs.to_dense(name), ref) <= 1e-12, name
    # G2P from the same grid state
    th.G2P(Gd, P, basis, sc.params, ctx=ctx)
    th.G2P(Gs, Ps, basis, sc.params, ctx=ctx)
    torch.cuda.synchronize()
    a, b = particles_np(P), particles_np(Ps)
    for name in a:
        np.testing.assert_array_equal(a[name], b[name])
    ctx.close()


@pytest.mark.parametrize("basis", [th.BSpline(th.Cubic()), th.uGIMP()], ids=["cubic", "ugimp"])
def test_sphere_spin_sparse_steps_equal_dense(basis):
    """The sparse-grid sphere (reduced to dx=1/48), Float32: 10 SpGrid steps (numbering,
    storage growth, zeroing) equal 10 dense-grid steps bit for bit."""
    sc = sphere_spin(basis, dtype=torch.float32, dx=1 / 48, steps=10)
    print(f"[sphere_spin] n={sc.n} basis={basis}", flush=True)
    mesh, Pd, Gd = to_gpu(sc)
    _, Ps, _ = to_gpu(sc)
    Gs = _sp_grid(mesh)
    # fresh contexts: both runs start from a full sort and re-bin incrementally alike
    cd, cs = th.Context(), th.Context()
    th.step(Gd, Pd, basis, sc.params, nsteps=10, ctx=cd)
    th.step(Gs, Ps, basis, sc.params, nsteps=10, ctx=cs)
    torch.cuda.synchronize()
    cd.close()
    cs.close()
    assert 0 < Gs.spinds.storage_len <= Gs.capacity
    a, b = particles_np(Pd), particles_np(Ps)
    for name in a:
        np.testing.assert_array_equal(a[name], b[name])
    dense = grid_np(Gd)
    for name in ("m", "v"):
        np.testing.assert_array_equal(Gs.to_dense(name), dense[name])


def test_sparse_errors():
    basis = th.BSpline(th.Quadratic())
    sc = random_box(3, 2000, 3, basis, dtype=torch.float64, h=0.05, lo=0.2, hi=0.4)
    sc.mesh.block_size_log2 = 2  # blocks of 4 nodes: the active set stops near x = 0.55
    mesh, P, _ = to_gpu(sc)
    G = _sp_grid(mesh)
    th.update_sparsity(G, P.x)
    # particles moved out of the active blocks without update_sparsity!
    P.x.add_(0.4)
    with pytest.raises(th.ErrorException, match="inactive SpGrid support node"):
        th.P2G(G, P, basis, sc.params)
    # storage smaller than the active blocks
    P.x.sub_(0.4)
    G.spinds.storage_len = G.capacity + 1
    with pytest.raises(th.DimensionMismatch):
        th.P2G(G, P, basis, sc.params)
    # block_size_log2 = 0 has no room for a quadratic window inside the ±1-block activation
    m0 = th.CartesianMesh(0.05, (0, 1), (0, 1), (0, 1), dtype=torch.float64, block_size_log2=0)
    G0 = _sp_grid(m0)
    with pytest.raises(th.UnsupportedError):
        th.update_sparsity(G0, P.x)


def test_step_sparse_storage_grows_mid_call():
    """A fast-moving box crosses into new blocks during a 12-step call whose storage fits only the first
    active set: tess_step_sparse stops before the step that needs more storage and reports the steps that
    ran; after resizing, the remaining steps complete the run, which equals 12 dense-grid steps."""
    import ctypes

    from tesserae_hip import _abi
    basis = th.BSpline(th.Quadratic())
    sc = random_box(3, 3000, 21, basis, dtype=torch.float64, h=0.02, lo=0.2, hi=0.3)
    sc.mesh.block_size_log2 = 2
    sc.v[:] = (2.0, 0.5, 0.0)
    sc.params = th.MPMParams(dt=5e-3, E=1e3, nu=0.3, rho0=1e3, V0=sc.params.V0)
    nsteps = 12
    mesh, Pd, Gd = to_gpu(sc)
    _, Ps, _ = to_gpu(sc)
    Gs = _sp_grid(mesh)
    ctx = th.Context()
    th.update_sparsity(Gs, Ps.x, ctx=ctx)
    assert Gs.capacity == Gs.spinds.storage_len
    lib = ctx.lib
    P = Ps.abi()
    done_total, calls, stops = 0, 0, []
    while done_total < nsteps:
        sp, G = Gs.sparsity_abi(), Gs.abi()
        done = ctypes.c_int32(-1)
        st = lib.tess_step_sparse(ctx.handle, basis.kind, ctypes.byref(mesh.abi()), ctypes.byref(G),
                                  ctypes.byref(sp), ctypes.byref(P), ctypes.byref(sc.params.abi()),
                                  nsteps - done_total, 0.0, ctypes.byref(done))
        calls += 1
        Gs.spinds.storage_len = int(sp.storage_len)
        done_total += done.value
        if st == _abi.TESS_ERR_STATE:
            assert 0 <= done.value < nsteps and sp.storage_len > Gs.capacity
            stops.append(done.value)
            Gs.resize_fillzero(int(sp.storage_len))
        else:
            assert st == _abi.TESS_OK and done_total == nsteps
    print(f"[storage growth] calls={calls} stopped after {stops} steps", flush=True)
    assert stops and stops[0] > 0  # the storage had to grow after steps of the same call had run
    th.step(Gd, Pd, basis, sc.params, nsteps=nsteps, ctx=th.Context())
    torch.cuda.synchronize()
    a, b = particles_np(Pd), particles_np(Ps)
    # the box moves rigidly and unstrained: ∇v and σ are rounding noise (~1e-14, ~1e-13) whose bits follow
    # the binning order (a failed call makes the next one sort from scratch), so they are bounded against
    # the flow's own scales, |v|/h and E; x, v, F and V relative to themselves
    scale = {"grad_v": 2.0 / 0.02, "sigma": 1e3}
    for name in a:
        if name in scale:
            assert np.abs(b[name] - a[name]).max() <= 1e-12 * scale[name], name
        else:
            assert rel_err(b[name], a[name]) <= 1e-12, name
    ctx.close()
