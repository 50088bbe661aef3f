"""Oracle geometry and basis checks against the reference's own known answers.

Known answers: test/mesh.jl:27-65, test/basis.jl:226-346, docstrings of
src/Mesh/cartesian.jl (findblock :247-262).  Closed forms: SURVEY a-3/a-5
(quadratic/cubic B-spline pieces, bspline.jl:176-189; uGIMP gimp.jl:28-33).
"""
import numpy as np
import pytest
import torch

from tesserae_hip import BSpline, CartesianMesh, Cubic, Linear, Quadratic, uGIMP


def test_mesh_construction_known_answers():
    mesh = CartesianMesh(0.2, (0, 3), (0, 4))
    assert mesh.size() == (16, 21)
    assert mesh.spacing() == 0.2
    assert mesh.h_inv == 1 / 0.2
    m3 = CartesianMesh(0.2, (0, 3), (0, 4), block_size_log2=3)
    assert m3.nblocks() == (2, 3)


def test_findblock_known_answers(oracle):
    m3 = CartesianMesh(0.2, (0, 3), (0, 4), block_size_log2=3)
    assert oracle.findblock(m3, (1.7, 2.1)) == (2, 2)                      # test/mesh.jl:39
    m10 = CartesianMesh(1.0, (0, 10), (0, 10))
    assert oracle.findblock(m10, (8.5, 1.5)) == (3, 1)                     # cartesian.jl:262


def test_findcell_and_isinside_known_answers(oracle):
    # findcell == findblock with 1-cell blocks (cartesian.jl:227-234, 277-285)
    mesh = CartesianMesh(0.2, (0, 3), (0, 4), block_size_log2=0)
    assert oracle.findblock(mesh, (0.1, 0.1)) == (1, 1)                    # test/mesh.jl:61
    assert oracle.findblock(mesh, (2.3, 1.1)) == (12, 6)
    assert oracle.findblock(mesh, (0.0, 0.0)) == (1, 1)
    assert oracle.findblock(mesh, (3.0, 4.0)) is None
    # isinside (test/mesh.jl:42-49)
    for x, inside in [((0.1, 0.3), True), ((0.2, 0.4), True), ((0.0, 0.0), True), ((3.0, 4.0), False),
                      ((-1.0, 3.0), False), ((1.0, -3.0), False)]:
        assert (oracle.findblock(mesh, x) is not None) == inside, x


@pytest.mark.parametrize("x,r,expected", [
    ((0.1, 0.1), 1, [(1, 2), (1, 2)]),
    ((0.3, 0.1), 2, [(1, 4), (1, 3)]),
    ((0.1, 0.3), 2, [(1, 3), (1, 4)]),
    ((0.2, 0.4), 1, [(2, 3), (3, 4)]),
    ((0.2, 0.4), 2, [(1, 4), (2, 5)]),
    ((3.0, 4.0), 2, [(1, 0), (1, 0)]),
    ((-0.1, 3.05), 3, [(1, 0), (1, 0)]),
])
def test_supportnodes_radius_known_answers(oracle, x, r, expected):
    mesh = CartesianMesh(0.2, (0, 3), (0, 4))                              # test/mesh.jl:51-58
    assert oracle.supportnodes_r(mesh, x, r) == expected


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_bspline_window_equals_generic_radius(oracle, deg):
    # bspline.jl:27 "Equivalent to supportnodes(x, (n + 1) / 2, mesh)"
    mesh = CartesianMesh(0.1, (0, 1), (0, 1))
    rng = np.random.default_rng(7)
    basis = BSpline(deg)
    for x in rng.uniform(0.0, 0.999, size=(200, 2)):
        win, _, _ = oracle.basis_weight(mesh, basis, x)
        assert [tuple(r) for r in win] == oracle.supportnodes_r(mesh, x, (deg + 1) / 2)


def test_jet_known_answers(oracle):
    assert oracle.jet1d(BSpline(Quadratic()), 0.0) == (0.75, 0.0)          # test/basis.jl:311
    assert oracle.jet1d(uGIMP(), 0.0, 1.0) == (0.75, 0.0)                  # test/basis.jl:312


def _closed_form(deg, xi):
    """Local-coordinate B-spline pieces (SURVEY a-3)."""
    if deg == 1:
        return [1 - xi, xi]
    if deg == 2:
        return [0.5 * (1 - xi) ** 2, -xi ** 2 + xi + 0.5, 0.5 * xi ** 2]
    return [(1 - xi) ** 3 / 6, (3 * xi ** 3 - 6 * xi ** 2 + 4) / 6, (-3 * xi ** 3 + 3 * xi ** 2 + 3 * xi + 1) / 6,
            xi ** 3 / 6]


@pytest.mark.parametrize("deg", [1, 2, 3])
@pytest.mark.parametrize("deferred", [False, True])
def test_weights_match_closed_forms_and_fd(oracle, deg, deferred):
    mesh = CartesianMesh(0.1, (0, 1), (0, 1))
    basis = BSpline(deg)
    sw = deg + 1
    rng = np.random.default_rng(deg)
    for x in rng.uniform(0.3, 0.7, size=(50, 2)):
        win, w, gw = oracle.basis_weight(mesh, basis, x, deferred=deferred)
        xi = x / mesh.h
        loc = xi - np.floor(xi) if deg % 2 == 1 else (xi - 0.5) - np.floor(xi - 0.5)
        N = [_closed_form(deg, t) for t in loc]
        for b in range(sw):
            for a in range(sw):
                assert w[a + sw * b] == pytest.approx(N[0][a] * N[1][b], rel=1e-12, abs=1e-15)
        # gradient by central differences of the closed form in x
        eps = 1e-6
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[c] += eps
            xm[c] -= eps
            _, wp, _ = oracle.basis_weight(mesh, basis, xp, deferred=deferred)
            _, wm, _ = oracle.basis_weight(mesh, basis, xm, deferred=deferred)
            _, _, _ = win, w, gw
            np.testing.assert_allclose(gw[:, c], (wp - wm) / (2 * eps), rtol=1e-5, atol=1e-6)


def _pu_lfr(mesh, win, w, gw, x, sw):
    d = mesh.dim
    ranges = [range(win[k][0] - 1, win[k][1]) for k in range(d)]
    sw_sum, g_sum, X_sum, XG = 0.0, np.zeros(d), np.zeros(d), np.zeros((d, d))
    axes = [mesh.axis(k).astype(np.float64) for k in range(d)]
    import itertools
    for local in itertools.product(*[range(len(r)) for r in ranges]):
        slot = sum(local[k] * sw ** k for k in range(d))
        X = np.array([axes[k][ranges[k][local[k]]] for k in range(d)])
        sw_sum += w[slot]
        g_sum += gw[slot]
        X_sum += X * w[slot]
        XG += np.outer(X, gw[slot])
    pu = np.isclose(sw_sum, 1) and np.allclose(g_sum + 1, 1)
    lfr = np.allclose(X_sum, x) and np.allclose(XG, np.eye(d))
    truncated = any(len(r) != sw for r in ranges)
    return pu, lfr, truncated


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("deg", [1, 2, 3])
def test_partition_of_unity_and_linear_reproduction(oracle, dim, deg):
    # test/basis.jl:226-270
    mesh = CartesianMesh(0.1, *[(0, 1)] * dim)
    basis = BSpline(deg)
    x = np.array([0.45 + 0.01 * (i + 1) for i in range(dim)])
    for deferred in (False, True):
        win, w, gw = oracle.basis_weight(mesh, basis, x, deferred=deferred)
        pu, lfr, tr = _pu_lfr(mesh, win, w, gw, x, basis.support_width)
        assert pu and lfr and not tr
    if deg in (2, 3):
        xb = x.copy()
        xb[0] = 0.02
        win, w, gw = oracle.basis_weight(mesh, basis, xb)
        pu, lfr, tr = _pu_lfr(mesh, win, w, gw, xb, basis.support_width)
        assert tr and not pu and not lfr


@pytest.mark.parametrize("dim", [2, 3])
def test_ugimp_unity_reproduction_and_windows(oracle, dim):
    # test/basis.jl:309-346
    mesh = CartesianMesh(0.1, *[(0, 1)] * dim)
    g = uGIMP()
    x = np.array([0.45 + 0.01 * (i + 1) for i in range(dim)])
    win, w, gw = oracle.basis_weight(mesh, g, x, l=0.5 * mesh.h)
    pu, lfr, _ = _pu_lfr(mesh, win, w, gw, x, 3)
    assert pu and lfr
    xb = x.copy()
    xb[0] = 0.02
    win, w, gw = oracle.basis_weight(mesh, g, xb, l=0.5 * mesh.h)
    pu, lfr, tr = _pu_lfr(mesh, win, w, gw, xb, 3)
    assert tr and not pu and not lfr
    x44 = np.full(dim, 0.44)
    win, w, gw = oracle.basis_weight(mesh, g, x44, l=mesh.h)
    assert all(hi - lo + 1 <= 3 for lo, hi in win)
    pu, lfr, _ = _pu_lfr(mesh, win, w, gw, x44, 3)
    assert pu and lfr
    # window sizes along an axis: 3 at x=0.40, 2 at x=0.44 with l=h/2
    w40, _, _ = oracle.basis_weight(mesh, g, np.array([0.40] + [0.5] * (dim - 1)), l=0.5 * mesh.h)
    w44, _, _ = oracle.basis_weight(mesh, g, np.array([0.44] + [0.5] * (dim - 1)), l=0.5 * mesh.h)
    assert w40[0][1] - w40[0][0] + 1 == 3
    assert w44[0][1] - w44[0][0] + 1 == 2


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_fast_path_matches_nodewise(oracle, deg):
    # test/basis.jl:541-561: the generated fast path equals the autodiff jet
    mesh = CartesianMesh(0.1, (0, 1), (0, 1), (0, 1))
    basis = BSpline(deg)
    rng = np.random.default_rng(1234)
    for x in rng.uniform(0.2, 0.8, size=(40, 3)):
        _, w0, g0 = oracle.basis_weight(mesh, basis, x, deferred=False)
        _, w1, g1 = oracle.basis_weight(mesh, basis, x, deferred=True)
        np.testing.assert_allclose(w0, w1, atol=np.sqrt(np.finfo(float).eps))
        np.testing.assert_allclose(g0, g1, atol=np.sqrt(np.finfo(float).eps) / mesh.h)
