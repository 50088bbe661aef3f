"""The oracle's @G2P2G (src/Transfer/g2p2g.jl:100-159) of the README program.

What pins it: the G2P half is the oracle's G2P with the weights at the
positions the particles have before the pass; the P2G half, checked against a
dense brute-force scatter of the updated particles' m, mv and f with the
weights at those same old positions (the macro does not rebind the weights
after x[p] is written, g2p2g.jl:107-114; test/transfer.jl:267-314 checks the
same composition by hand); mass and momentum are conserved by the P2G half; a
P2G at the updated positions (the planted mistake) differs.
"""
import numpy as np
import pytest

from oracle_bind import HostState
from test_oracle_partition_transfer import brute_p2g, fixture_scene
from tesserae_hip import BSpline, Cubic, Linear, Quadratic
from tesserae_hip.scenes import random_box

BASES = [BSpline(Linear()), BSpline(Quadratic()), BSpline(Cubic())]


def _after_p2g(oracle, sc):
    """State after one README P2G, so the grid holds v and vⁿ for the pass."""
    st = HostState(sc)
    oracle.p2g(st, sc.basis, sc.params)
    return st


def _snapshot(sc, st):
    c = sc.copy()
    for name in ("x", "v", "F", "sigma", "V", "grad_v"):
        setattr(c, name, getattr(st, name).copy())
    return c


@pytest.mark.parametrize("basis", BASES, ids=["linear", "quadratic", "cubic"])
def test_g2p2g_halves(oracle, basis):
    sc = fixture_scene(basis)
    st = _after_p2g(oracle, sc)
    x_old = st.x.copy()
    # the G2P half alone, from the same grid
    ref = HostState(_snapshot(sc, st))
    for name in ("m", "mv", "f", "vn", "gv"):
        getattr(ref, name)[...] = getattr(st, name)
    oracle.g2p(ref, basis, sc.params)
    oracle.g2p2g(st, basis, sc.params)
    for name in ("x", "v", "grad_v", "F", "sigma", "V"):
        np.testing.assert_array_equal(getattr(st, name), getattr(ref, name), err_msg=name)
    # the P2G half: updated v, V, σ scattered with the weights at the old positions
    moved = _snapshot(sc, st)
    at_old = moved.copy()
    at_old.x = x_old
    m, mv, f = brute_p2g(at_old)
    np.testing.assert_allclose(st.m, m, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.mv, mv, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.f, f, rtol=1e-12, atol=1e-13)
    # the planted mistake, weights at the new positions, is told apart
    m2, mv2, f2 = brute_p2g(moved)
    assert np.abs(f2 - st.f).max() > 1e-6 * np.abs(st.f).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p2g_conservation(oracle, dim):
    basis = BSpline(Quadratic())
    sc = random_box(dim, 400, 21 + dim, basis, h=0.1, lo=0.3, hi=0.7, stress=1.0)
    st = _after_p2g(oracle, sc)
    oracle.g2p2g(st, basis, sc.params)
    mp = sc.params.mp
    assert st.m.sum() == pytest.approx(sc.n * mp, rel=1e-12)
    np.testing.assert_allclose(st.mv.sum(axis=0), mp * st.v.sum(axis=0), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(st.f.sum(axis=0), 0.0, atol=1e-11 * np.abs(st.f).max())
