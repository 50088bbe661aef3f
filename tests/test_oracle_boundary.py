"""Independent pins of the oracle's gravity and slip-wall branches.

The grid-node half of the headline program adds gravity to the nodal
acceleration of every node with mass and zeroes the normal component of the
grid velocity ``v`` on wall faces, the reference's slip-floor idiom
(docs/src/gpu.md:116-129: ``v[i] = v[i] .* (true, true, false)`` on the
``[:,:,begin]`` slice, which leaves ``vⁿ`` alone).  These tests do not compare
against a second implementation of that code; they check what the mathematics
fixes:

* free fall: a stress-free block with one uniform velocity has f = 0 on every
  node (σ = 0), vⁿ = v_p exactly on every node with mass, and G2P of a
  uniform nodal field over a whole support returns it unchanged (Σw = 1,
  Σ∇w = 0), so after k steps v = v₀ + k g Δt and
  x = x₀ + k v₀ Δt + g Δt² k(k+1)/2, with F = I and σ = 0 throughout;
* walls: with one face's bit set, exactly the normal component of ``v`` on
  that face's nodes becomes zero; every other number (vⁿ, the tangential
  components, every node off the face, m, mv, f) equals the run without walls
  bit for bit.

Each check fails for a sign error in g, g applied twice or to massless nodes,
the tangential instead of the normal component zeroed, a wall on the wrong
face, or vⁿ zeroed as well.
"""
import numpy as np
import pytest
import torch

from oracle_bind import HostState
from tesserae_hip import BSpline, CartesianMesh, Quadratic
from tesserae_hip.core import MPMParams
from tesserae_hip.scenes import Scene, _state


def _block(dim, lo, hi, h, v0, gravity, walls=0, dt=1e-3, E=1e4):
    """Particles 2 per cell and axis in the box [lo, hi]^dim of a mesh on [0, 1]^dim."""
    mesh = CartesianMesh(h, *[(0.0, 1.0)] * dim, dtype=torch.float64)
    s = h / 2
    ax = np.arange(lo + s / 2, hi, s)
    g = np.meshgrid(*[ax] * dim, indexing="ij")
    x = np.stack([a.reshape(-1, order="F") for a in g], axis=1)
    v = np.tile(np.asarray(v0, float), (len(x), 1))
    V0 = s ** dim
    params = MPMParams(dt=dt, E=E, nu=0.3, rho0=1e3, V0=V0, gravity=tuple(gravity), slip_walls=walls)
    return Scene("block", mesh, BSpline(Quadratic()), params, **_state(x, v, V0, dim))


@pytest.mark.parametrize("dim", [2, 3])
def test_free_fall_closed_form(oracle, dim):
    g = (1.5, -2.0, -9.81)[:dim] if dim == 3 else (1.5, -9.81)
    v0 = (0.25, -0.5, 0.125)[:dim]
    sc = _block(dim, 0.4, 0.6, 1 / 32, v0, g)
    st = HostState(sc)
    x0 = st.x.copy()
    dt, K = sc.params.dt, 25
    for k in range(1, K + 1):
        oracle.step(st, sc.basis, sc.params, 1)
        vk = np.asarray(v0) + k * np.asarray(g) * dt
        np.testing.assert_allclose(st.v, np.broadcast_to(vk, st.v.shape), rtol=1e-12, atol=1e-13)
    xk = x0 + K * np.asarray(v0) * dt + np.asarray(g) * dt * dt * K * (K + 1) / 2
    np.testing.assert_allclose(st.x, xk, rtol=1e-12, atol=1e-14)
    eye = np.eye(dim).reshape(-1, order="F")
    np.testing.assert_allclose(st.F, np.broadcast_to(eye, st.F.shape), atol=1e-13)
    np.testing.assert_allclose(st.sigma, 0.0, atol=1e-6)  # σ = O(E·ulp)
    np.testing.assert_allclose(st.grad_v, 0.0, atol=1e-11)


@pytest.mark.parametrize("dim", [2, 3])
def test_gravity_only_where_mass(oracle, dim):
    g = (0.5, -3.0, -9.81)[:dim] if dim == 3 else (0.5, -9.81)
    sc = _block(dim, 0.4, 0.6, 1 / 16, (0.0,) * dim, g)
    st = HostState(sc)
    oracle.p2g(st, sc.basis, sc.params)
    has = st.m > 0
    assert has.any() and (~has).any()
    # v = vⁿ + Δt g on nodes with mass (f = 0 for σ = 0, vⁿ = 0 for v_p = 0), v = 0 elsewhere
    np.testing.assert_allclose(st.gv[has], np.broadcast_to(np.asarray(g) * sc.params.dt, st.gv[has].shape),
                               rtol=1e-13)
    assert np.all(st.gv[~has] == 0) and np.all(st.vn == 0)


def _face_nodes(mesh, axis, upper):
    dims = mesh.size()
    idx = np.arange(int(np.prod(dims)))
    coord = (idx // int(np.prod(dims[:axis]))) % dims[axis]
    return coord == (dims[axis] - 1 if upper else 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_slip_wall_each_face(oracle, dim):
    h = 1 / 16
    for axis in range(dim):
        for upper in (0, 1):
            # a stressed block pushed into the face: particles within a cell of it
            lo, hi = (1 - 0.25, 1.0 - 1e-9) if upper else (0.0, 0.25)
            v0 = np.full(dim, 0.3)
            v0[axis] = 1.0 if upper else -1.0
            g = np.full(dim, -0.7)
            g[axis] = 9.0 if upper else -9.0
            base = _block(dim, lo, hi, h, v0, g)
            # the block occupies the whole slab next to the face, so give it a real stress state
            base.F = base.F * 1.01
            base.sigma = np.tile(np.arange(1.0, dim * dim + 1), (base.n, 1)) * 10.0
            free = HostState(base)
            oracle.p2g(free, base.basis, base.params)
            bit = 1 << (2 * axis + upper)
            walled_sc = base.copy()
            walled_sc.params = MPMParams(dt=base.params.dt, E=base.params.E, nu=0.3, rho0=1e3, V0=base.params.V0,
                                         gravity=base.params.gravity, slip_walls=bit)
            wall = HostState(walled_sc)
            oracle.p2g(wall, walled_sc.basis, walled_sc.params)
            face = _face_nodes(base.mesh, axis, upper)
            hit = face & (free.m > 0)
            assert hit.sum() > 4, (axis, upper)
            assert np.any(free.gv[hit, axis] != 0)
            # untouched: m, mv, f, vⁿ everywhere, v off the face, tangential v on the face
            for name in ("m", "mv", "f", "vn"):
                assert np.array_equal(getattr(wall, name), getattr(free, name)), name
            assert np.array_equal(wall.gv[~face], free.gv[~face])
            tang = [c for c in range(dim) if c != axis]
            assert np.array_equal(wall.gv[face][:, tang], free.gv[face][:, tang])
            # the normal component is zero on the face
            assert np.all(wall.gv[face, axis] == 0)
            # the opposite face's bit leaves this face alone
            other = MPMParams(dt=base.params.dt, E=base.params.E, nu=0.3, rho0=1e3, V0=base.params.V0,
                              gravity=base.params.gravity, slip_walls=1 << (2 * axis + (1 - upper)))
            opp = HostState(base)
            oracle.p2g(opp, base.basis, other)
            assert np.array_equal(opp.gv[face], free.gv[face])


def test_slip_wall_floor_step_threaded_matches_sequential(oracle):
    """The threaded oracle (the CPU baseline) applies gravity and the walls like the sequential one."""
    sc = _block(3, 0.3, 0.7, 1 / 16, (0.2, -0.1, -2.0), (0.0, 0.0, -9.81), walls=0b111111)
    sc.x[:, 2] -= 0.29  # the block rests within a cell of the floor
    a, b = HostState(sc), HostState(sc)
    oracle.step(a, sc.basis, sc.params, 5)
    oracle.step_threaded(b, sc.basis, sc.params, 4, nsteps=5)
    for k in ("x", "v", "F"):
        np.testing.assert_allclose(getattr(b, k), getattr(a, k), rtol=1e-12, atol=1e-14)
    # after five steps into the floor the floor nodes carry no normal velocity
    floor = _face_nodes(sc.mesh, 2, 0)
    assert np.all(a.gv[floor, 2] == 0) and np.any(a.m[floor] > 0)
