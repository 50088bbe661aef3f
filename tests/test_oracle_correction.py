"""The oracle's corrected bases, WLS and KernelCorrection (src/Basis/wls.jl,
src/Basis/kernelcorrection.jl, src/Basis/polynomials.jl).

What pins them:
- the reference's branch tests (test/basis.jl:415-466, 525-540): partition of
  unity and linear field reproduction (Σ w = 1, Σ ∇w = 0, Σ xᵢ w = x,
  Σ xᵢ ⊗ ∇w = I) at an interior and a boundary point, with the truncation flag;
  the plain B-spline at the boundary point fails both (test/basis.jl:264-267),
  which is the mistake the correction exists to remove;
- the reference's positivity maps of KernelCorrection (test/basis.jl:543-630):
  which nodes' corrected weights stay non-negative over their whole support for
  Polynomial(Linear()) and Polynomial(MultiLinear()), quadratic and cubic;
- a numpy brute force of the weighted least-squares fit (np.linalg.inv of the
  moment matrix built from exponent tuples), which every branch equals in exact
  arithmetic: the closed-form moments and the per-axis fits are the general fit
  evaluated analytically;
- the filter semantics of test/basis.jl:481-520: a masked node gets w = ∇w = 0,
  reproduction still holds, a plain kernel refuses a filter;
- stored and deferred weights give the same transfers (test/transfer.jl:960-1040).
"""
import itertools

import numpy as np
import pytest

from oracle_bind import HostState, HostWeights, host_filter
from test_oracle_partition_transfer import _kernel
from tesserae_hip import (BSpline, CartesianMesh, Cubic, KernelCorrection, Linear, MultiLinear, Polynomial,
                          Quadratic, WLS, uGIMP)
from tesserae_hip.scenes import random_box

LIN, ML = Polynomial(Linear()), Polynomial(MultiLinear())


def interior_point(dim):
    return np.array([0.45 + 0.01 * (i + 1) for i in range(dim)])


def boundary_point(dim):
    return np.array([0.02 if i == 0 else 0.45 + 0.01 * (i + 1) for i in range(dim)])


def _nodes(mesh, win):
    """Node multi-indices (0-based) of a 1-based window, column-major, with their storage slots."""
    d = mesh.dim
    sw = None
    ranges = [range(win[k][0] - 1, win[k][1]) for k in range(d)]
    out = []
    for I in itertools.product(*reversed(ranges)):
        out.append(I[::-1])
    return out


def _slot(mesh, win, I, sw):
    s, mul = 0, 1
    for k in range(mesh.dim):
        s += (I[k] - (win[k][0] - 1)) * mul
        mul *= sw
    return s


def _sums(oracle, mesh, basis, x, l=0.0, deferred=False, filter=None):
    """(Σw, Σ∇w, Σ xᵢ w, Σ xᵢ⊗∇w, truncated, weights by node) of one particle."""
    win, w, gw = oracle.basis_weight(mesh, basis, x, l=l, deferred=deferred, filter=filter)
    sw, d = basis.support_width, mesh.dim
    axes = [mesh.axis(k) for k in range(d)]
    S0, S1 = 0.0, np.zeros(d)
    X1, XG = np.zeros(d), np.zeros((d, d))
    by_node = {}
    nodes = _nodes(mesh, win)
    for I in nodes:
        s = _slot(mesh, win, I, sw)
        Xi = np.array([axes[k][I[k]] for k in range(d)])
        S0 += w[s]
        S1 += gw[s]
        X1 += Xi * w[s]
        XG += np.outer(Xi, gw[s])
        by_node[I] = (w[s], gw[s].copy())
    truncated = len(nodes) != sw ** d
    return S0, S1, X1, XG, truncated, by_node


def _pu(S0, S1):
    return np.isclose(S0, 1.0, rtol=np.sqrt(np.finfo(float).eps)) and np.allclose(S1 + 1, 1, rtol=1e-8, atol=0)


def _lfr(X1, XG, x):
    return np.allclose(X1, x, rtol=1e-8, atol=1e-12) and np.allclose(XG, np.eye(len(x)), rtol=1e-8, atol=1e-10)


BRANCHES = [  # (basis, point, truncated) of test/basis.jl:415-466
    (WLS(BSpline(Quadratic())), interior_point, False),
    (WLS(BSpline(Linear())), interior_point, False),
    (WLS(BSpline(Quadratic()), ML), interior_point, False),
    (WLS(BSpline(Quadratic())), boundary_point, True),
    (WLS(uGIMP()), boundary_point, True),
    (KernelCorrection(BSpline(Quadratic())), interior_point, False),
    (KernelCorrection(BSpline(Quadratic())), boundary_point, True),
    (KernelCorrection(uGIMP()), boundary_point, True),
    # coverage of the cubic kernel (test/basis.jl:525-540)
    (WLS(BSpline(Cubic())), interior_point, False),
    (WLS(BSpline(Cubic())), boundary_point, True),
    (KernelCorrection(BSpline(Cubic())), interior_point, False),
    (KernelCorrection(BSpline(Cubic())), boundary_point, True),
    # the other polynomial of each wrapper
    (KernelCorrection(BSpline(Quadratic()), LIN), boundary_point, True),
    (WLS(BSpline(Cubic()), ML), boundary_point, True),
    (WLS(uGIMP(), ML), boundary_point, True),
]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("deferred", [False, True], ids=["stored", "deferred"])
@pytest.mark.parametrize("basis,point,truncated", BRANCHES, ids=[f"{b!r}@{p.__name__}" for b, p, _ in BRANCHES])
def test_branches_reproduce_linear_fields(oracle, dim, deferred, basis, point, truncated):
    mesh = CartesianMesh(0.1, *[(0, 1)] * dim)
    x = point(dim)
    l = 0.5 * mesh.h / 2
    S0, S1, X1, XG, tr, _ = _sums(oracle, mesh, basis, x, l=l, deferred=deferred)
    assert tr is truncated
    assert _pu(S0, S1)
    assert _lfr(X1, XG, x)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", [BSpline(Quadratic()), BSpline(Cubic())], ids=["quadratic", "cubic"])
def test_plain_kernel_at_the_boundary_fails_both(oracle, dim, kernel):
    """The planted mistake: the uncorrected truncated support (test/basis.jl:264-267)."""
    mesh = CartesianMesh(0.1, *[(0, 1)] * dim)
    x = boundary_point(dim)
    S0, S1, X1, XG, tr, _ = _sums(oracle, mesh, kernel, x)
    assert tr
    assert not _pu(S0, S1)
    assert not _lfr(X1, XG, x)


def _exponents(dim, ml):
    if ml:
        return list(itertools.product((0, 1), repeat=dim))
    return [tuple(0 for _ in range(dim))] + [tuple(int(k == j) for k in range(dim)) for j in range(dim)]


def brute_corrected(mesh, basis, x, l=0.0, filter=None, nodes=None):
    """Weights of the weighted least-squares fit by brute force, or the plain kernel where KernelCorrection keeps
    it (whole support that the filter passes).  nodes: the window's node multi-indices."""
    d, h = mesh.dim, mesh.h
    axes = [mesh.axis(k) for k in range(d)]
    kern = BSpline(Linear()) if basis.kernel == 1 else BSpline(Quadratic()) if basis.kernel == 2 else \
        BSpline(Cubic()) if basis.kernel == 3 else uGIMP()
    ml = bool(basis.kind & 0x400)
    lh = l / h
    N, G = {}, {}
    for I in nodes:
        jets = [_kernel(kern, (x[k] - axes[k][I[k]]) / h, lh) for k in range(d)]
        N[I] = np.prod([j[0] for j in jets])
        G[I] = np.array([np.prod([jets[k][1] if k == c else jets[k][0] for k in range(d)]) / h for c in range(d)])
    passes = {I: (filter is None or bool(filter[I])) for I in nodes}
    full = len(nodes) == basis.support_width ** d
    if (basis.kind & 0x200) and full and all(passes.values()):
        return {I: (N[I], G[I]) for I in nodes}
    E = _exponents(d, ml)
    c0 = E.index(tuple(0 for _ in range(d)))
    cl = [E.index(tuple(int(k == j) for k in range(d))) for j in range(d)]

    def P(I):
        dx = [axes[k][I[k]] - x[k] for k in range(d)]
        return np.array([np.prod([dx[k] ** e[k] for k in range(d)]) for e in E])

    M = sum((N[I] * passes[I]) * np.outer(P(I), P(I)) for I in nodes)
    Minv = np.linalg.inv(M)
    out = {}
    for I in nodes:
        q = (N[I] * passes[I]) * (Minv @ P(I))
        out[I] = (q[c0], np.array([q[j] for j in cl]))
    return out


def _random_points(rng, dim, n, h):
    """Points near the low faces and corners (truncated windows) and inside."""
    pts = rng.uniform(0.0, 1.0, (n, dim))
    pts[: n // 2] = rng.uniform(0.0, 2.5 * h, (n // 2, dim))
    pts[n // 2: 3 * n // 4, 0] = rng.uniform(0.0, 1.5 * h, n // 4)
    return pts


CORRECTED = [WLS(BSpline(Linear())), WLS(BSpline(Quadratic())), WLS(BSpline(Cubic())), WLS(uGIMP()),
             WLS(BSpline(Quadratic()), ML), WLS(BSpline(Cubic()), ML), WLS(uGIMP(), ML),
             KernelCorrection(BSpline(Quadratic())), KernelCorrection(BSpline(Cubic())),
             KernelCorrection(uGIMP()), KernelCorrection(BSpline(Quadratic()), LIN),
             KernelCorrection(BSpline(Linear()))]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("deferred", [False, True], ids=["stored", "deferred"])
@pytest.mark.parametrize("basis", CORRECTED, ids=repr)
def test_matches_brute_force_fit(oracle, dim, deferred, basis):
    rng = np.random.default_rng(7 + dim + 3 * deferred)
    mesh = CartesianMesh(0.1, *[(0, 1)] * dim)
    l = 0.4 * mesh.h
    for x in _random_points(rng, dim, 12, mesh.h):
        win, w, gw = oracle.basis_weight(mesh, basis, x, l=l, deferred=deferred)
        nodes = _nodes(mesh, win)
        ref = brute_corrected(mesh, basis, x, l=l, nodes=nodes)
        scale = max(abs(v[0]) for v in ref.values())
        gscale = max(np.abs(v[1]).max() for v in ref.values())
        for I in nodes:
            s = _slot(mesh, win, I, basis.support_width)
            assert abs(w[s] - ref[I][0]) <= 1e-11 * scale, (x, I)
            assert np.abs(gw[s] - ref[I][1]).max() <= 1e-11 * gscale, (x, I)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", [WLS(BSpline(Quadratic())), KernelCorrection(BSpline(Quadratic())),
                                   KernelCorrection(BSpline(Cubic())), WLS(BSpline(Quadratic()), ML)], ids=repr)
def test_filtered_fit(oracle, dim, basis):
    """test/basis.jl:497-520: a masked support node gets zero weight and gradient, reproduction holds, stored and
    deferred agree, and the fit equals the brute force with the node dropped.  (uGIMP's two-node windows leave a
    MultiLinear fit without enough nodes once one is masked: its moment matrix is singular.)"""
    mesh = CartesianMesh(0.1, *[(0, 1)] * dim)
    x = interior_point(dim)
    l = 0.4 * mesh.h
    win0, _, _ = oracle.basis_weight(mesh, basis, x, l=l)
    masked = tuple(w[0] - 1 for w in win0)  # first support node
    filt = np.ones(mesh.size(), bool)
    filt[masked] = False
    S0, S1, X1, XG, _, by = _sums(oracle, mesh, basis, x, l=l, filter=filt)
    assert _pu(S0, S1) and _lfr(X1, XG, x)
    assert by[masked][0] == 0 and np.all(by[masked][1] == 0)
    _, _, _, _, _, byd = _sums(oracle, mesh, basis, x, l=l, deferred=True, filter=filt)
    ref = brute_corrected(mesh, basis, x, l=l, filter=filt, nodes=list(by))
    for I in by:
        assert abs(byd[I][0] - by[I][0]) <= 1e-12 and np.abs(byd[I][1] - by[I][1]).max() <= 1e-9
        assert abs(by[I][0] - ref[I][0]) <= 1e-12 and np.abs(by[I][1] - ref[I][1]).max() <= 1e-9


def test_plain_kernel_refuses_a_filter(oracle):
    mesh = CartesianMesh(0.1, (0, 1), (0, 1))
    filt = np.ones(mesh.size(), bool)
    with pytest.raises(RuntimeError, match="filtered updates"):
        oracle.basis_weight(mesh, BSpline(Quadratic()), interior_point(2), filter=filt)


# ---- positivity (test/basis.jl:543-630) -----------------------------------------

def _kernel_values(oracle, mesh, basis, index, L):
    """The corrected weight of node `index` (1-based) over a 1/11 lattice of its support."""
    d = mesh.dim
    lo = [max(0.0, index[k] - L - 1) for k in range(d)]
    hi = [min(float(mesh.size()[k] - 1), index[k] + L - 1) - np.sqrt(np.finfo(float).eps) for k in range(d)]
    axes = [np.arange(lo[k], hi[k] + 1e-12, 1 / 11) for k in range(d)]
    node = tuple(index[k] - 1 for k in range(d))
    sw = basis.support_width
    out = []
    for x in itertools.product(*axes):
        win, w, _ = oracle.basis_weight(mesh, basis, np.array(x))
        if all(win[k][0] - 1 <= node[k] <= win[k][1] - 1 for k in range(d)):
            out.append(w[_slot(mesh, win, node, sw)])
    return np.array(out)


def _positive(oracle, mesh, basis, index, L):
    return bool(np.all(_kernel_values(oracle, mesh, basis, index, L) > -np.sqrt(np.finfo(float).eps)))


POSITIVITY_2D = [  # kernel, L, (index, poly, expected) on CartesianMesh(1, (0,10), (0,10))
    (BSpline(Quadratic()), 1.5, [((1, 1), LIN, True), ((2, 2), LIN, False), ((2, 2), ML, True)]),
    (BSpline(Cubic()), 2.0, [((1, 1), LIN, True), ((2, 2), LIN, False), ((2, 2), ML, True), ((3, 1), LIN, False),
                             ((3, 1), ML, False)]),
]


@pytest.mark.parametrize("kernel,L,cases", POSITIVITY_2D, ids=["quadratic", "cubic"])
def test_positivity_2d(oracle, kernel, L, cases):
    mesh = CartesianMesh(1.0, (0, 10), (0, 10))
    for index, poly, expected in cases:
        assert _positive(oracle, mesh, KernelCorrection(kernel, poly), index, L) is expected, (index, poly)


@pytest.mark.parametrize("kernel,L,domains", [
    (BSpline(Quadratic()), 1.5, [((1, 1), True), ((3, 2), True)]),
    (BSpline(Cubic()), 2.0, [((1, 1), True), ((2, 1), False)]),
], ids=["quadratic", "cubic"])
def test_positivity_small_domains_2d(oracle, kernel, L, domains):
    """Over CartesianMesh(1, (0,I1), (0,I2)): Linear is not positive at every node, MultiLinear is where the
    reference says so."""
    for I, ml_positive in domains:
        mesh = CartesianMesh(1.0, (0, I[0]), (0, I[1]))
        nodes = list(itertools.product(*[range(1, I[k] + 1) for k in range(2)]))
        assert not all(_positive(oracle, mesh, KernelCorrection(kernel, LIN), i, L) for i in nodes)
        assert all(_positive(oracle, mesh, KernelCorrection(kernel, ML), i, L) for i in nodes) is ml_positive


@pytest.mark.slow
@pytest.mark.parametrize("kernel,L,cases", [
    (BSpline(Quadratic()), 1.5, [((1, 1, 1), LIN, True), ((2, 2, 1), LIN, False), ((2, 2, 1), ML, True)]),
    (BSpline(Cubic()), 2.0, [((1, 1, 1), LIN, True), ((2, 2, 1), LIN, False), ((2, 2, 1), ML, True),
                             ((3, 1, 1), LIN, False), ((3, 1, 1), ML, False)]),
], ids=["quadratic", "cubic"])
def test_positivity_3d(oracle, kernel, L, cases):
    mesh = CartesianMesh(1.0, (0, 10), (0, 10), (0, 10))
    for index, poly, expected in cases:
        assert _positive(oracle, mesh, KernelCorrection(kernel, poly), index, L) is expected, (index, poly)


# ---- transfers (test/transfer.jl:960-1040) ----------------------------------------

def _transfer(oracle, sc, basis, weights):
    st = HostState(sc)
    oracle.p2g(st, basis, sc.params, weights)
    oracle.g2p(st, basis, sc.params, weights)
    return st


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("basis", [WLS(BSpline(Quadratic())), KernelCorrection(BSpline(Quadratic())),
                                   KernelCorrection(uGIMP()), KernelCorrection(BSpline(Cubic()))], ids=repr)
def test_stored_and_deferred_transfers_agree(oracle, dim, basis):
    sc = random_box(dim, 300 if dim == 2 else 400, 3 + dim, basis, h=0.1, lo=-0.02, hi=1.02, stress=1.0,
                    with_l=basis.kernel == 16)
    stored = oracle.update_weights(HostState(sc), basis)
    a = _transfer(oracle, sc, basis, stored)
    b = _transfer(oracle, sc, basis, None)
    for name in ("m", "mv", "f", "v", "x", "grad_v", "F"):
        ra, rb = getattr(a, name), getattr(b, name)
        assert np.abs(ra - rb).max() <= 1e-10 * max(np.abs(ra).max(), 1e-300), name
    if basis.kernel == 16:
        return  # a masked node in a two-node uGIMP window leaves the MultiLinear fit singular
    # the same with a node filter: support nodes of the particles, a hole inside
    mask = np.zeros(sc.mesh.size(), bool)
    inside = np.all((sc.x >= 0) & (sc.x <= 1), axis=1)
    idx = np.floor(sc.x[inside] / sc.mesh.h).astype(int)
    for off in itertools.product(*[range(-1, 3)] * dim):
        J = np.clip(idx + np.array(off), 0, np.array(sc.mesh.size()) - 1)
        mask[tuple(J.T)] = True
    mask[tuple(s // 2 for s in sc.mesh.size())] = False
    st = HostState(sc)
    stored = oracle.update_weights(st, basis, filter=mask)
    a = _transfer(oracle, sc, basis, stored)
    W = HostWeights(basis, dim, sc.n, st.npdt, filter=host_filter(sc.mesh, mask), deferred=True)
    b = _transfer(oracle, sc, basis, W)
    for name in ("m", "mv", "f", "v", "x", "grad_v"):
        ra, rb = getattr(a, name), getattr(b, name)
        assert np.abs(ra - rb).max() <= 1e-9 * max(np.abs(ra).max(), 1e-300), name
