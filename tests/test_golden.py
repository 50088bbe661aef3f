"""The committed transfer fixture (tests/golden/transfer_fixture.json: the
reference's test/transfer.jl fixture through the README program, computed by an
independent numpy restatement) against the oracle (CPU) and the HIP engine (GPU)."""
import json
import os

import numpy as np
import pytest
import torch

import tesserae_hip as th
from oracle_bind import HostState
from tesserae_hip.scenes import Scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "transfer_fixture.json")


def fixture_scene(dtype=torch.float64):
    d = json.load(open(GOLDEN))
    mesh = th.CartesianMesh(d["mesh"]["h"], (0, 2), (0, 2), dtype=dtype)
    assert tuple(mesh.dims) == tuple(d["mesh"]["dims"])
    pr = d["params"]
    params = th.MPMParams(dt=pr["dt"], E=1.0, nu=0.3, rho0=1.0, V0=pr["V0"])
    params.lam, params.mu, params.mp = pr["lambda"], pr["mu"], pr["mp"]
    P = d["particles"]
    n = len(P["x"])
    sc = Scene("transfer_fixture", mesh, th.BSpline(th.Linear()), params, x=np.array(P["x"]), v=np.array(P["v"]),
               V=np.array(P["V"]), F=np.array(P["F"]), sigma=np.array(P["sigma"]), grad_v=np.zeros((n, 4)))
    return sc, d


def _check(got, ref, tol):
    ref = np.asarray(ref, dtype=float)
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.abs(np.asarray(got, dtype=float).reshape(ref.shape) - ref).max() <= tol * scale


def test_oracle_matches_transfer_fixture(oracle):
    sc, d = fixture_scene()
    st = HostState(sc)
    oracle.p2g(st, sc.basis, sc.params)
    for name, got in (("m", st.m), ("mv", st.mv), ("f", st.f), ("vn", st.vn), ("v", st.gv)):
        _check(got, d["p2g"][name], 1e-14)
    st = HostState(sc)
    st.vn[:] = np.array(d["g2p_grid"]["vn"])
    st.gv[:] = np.array(d["g2p_grid"]["v"])
    oracle.g2p(st, sc.basis, sc.params)
    for name in ("x", "v", "grad_v", "F", "sigma", "V"):
        _check(getattr(st, name), d["g2p"][name], 1e-14)


@pytest.mark.gpu
def test_engine_matches_transfer_fixture():
    from gpu_util import grid_np, particles_np, to_gpu
    sc, d = fixture_scene()
    mesh, P, G = to_gpu(sc)
    th.P2G(G, P, sc.basis, sc.params)
    torch.cuda.synchronize()
    g = grid_np(G)
    for name in ("m", "mv", "f", "vn", "v"):
        _check(g[name], d["p2g"][name], 1e-13)
    G.vn.copy_(torch.tensor(d["g2p_grid"]["vn"], dtype=torch.float64))
    G.v.copy_(torch.tensor(d["g2p_grid"]["v"], dtype=torch.float64))
    th.G2P(G, P, sc.basis, sc.params)
    torch.cuda.synchronize()
    got = particles_np(P)
    for name in ("x", "v", "grad_v", "F", "sigma", "V"):
        _check(got[name], d["g2p"][name], 1e-13)
