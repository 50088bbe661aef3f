"""Writes transfer_fixture.json: the transfer fixture of the reference's tests
(test/transfer.jl:2-29 geometry, :61-82 particle and grid state) run through
the README program (README.md:71-86) by a plain numpy restatement that is
independent of the oracle's C++ code: linear B-spline weights from the closed
form, dense loops over every node, Float64.

Differences from the reference fixture, forced by the README program: the
particle mass is the program's constant mₚ (1.0 here) instead of the per-particle
m[p]; there is no gravity.  The G2P part starts from the fixture's grid state
(vⁿ_i = (0.08 I₁, -0.04 I₂), v_i = vⁿ_i + (0.03 I₂, -0.02 I₁)), as the
reference's own G2P test does.

Run: python tests/golden/make_transfer_fixture.py
"""
import json
import os

import numpy as np

H = 1.0
NODES = 3                       # CartesianMesh(1.0, (0,2), (0,2)) -> 3x3 nodes
DT = 0.01
MP = 1.0
LAM, MU, V0 = 2.0, 1.5, 0.25


def particles():
    # GridSampling with spacing h/2 (src/particles.jl:9-18), x fastest, 1-based p
    ax = 0.25 + 0.5 * np.arange(4)
    xs = np.array([(x, y) for y in ax for x in ax])
    n = len(xs)
    p = np.arange(1, n + 1, dtype=float)
    V = 0.2 + 0.01 * p
    v = np.stack([0.15 + 0.02 * p, -0.25 + 0.015 * p], axis=1)
    F = np.zeros((n, 2, 2))
    F[:, 0, 0] = 1.0 + 0.002 * p
    F[:, 1, 1] = 1.0 - 0.001 * p
    a = np.stack([1.0 + 0.03 * p, np.full(n, 0.2)], axis=1)
    b = np.stack([np.full(n, 0.4), 0.8 + 0.01 * p], axis=1)
    ab = a[:, :, None] * b[:, None, :]
    S = 0.5 * (ab + ab.transpose(0, 2, 1))   # symmetric(a ⊗ b)
    return xs, V, v, F, S


def w_linear(x, i):
    """Tensor-product linear B-spline weight and gradient of node i (0-based) at x."""
    r = (x - np.asarray(i) * H) / H
    N = np.where(np.abs(r) < 1, 1 - np.abs(r), 0.0)
    dN = np.where(np.abs(r) < 1, -np.sign(r), 0.0) / H
    w = N[0] * N[1]
    g = np.array([dN[0] * N[1], N[0] * dN[1]])
    return w, g


def nodes():
    return [(i, j) for j in range(NODES) for i in range(NODES)]  # x fastest


def p2g(xs, V, v, S):
    m = np.zeros(NODES * NODES)
    mv = np.zeros((NODES * NODES, 2))
    f = np.zeros((NODES * NODES, 2))
    for p in range(len(xs)):
        for k, I in enumerate(nodes()):
            w, g = w_linear(xs[p], I)
            m[k] += w * MP
            mv[k] += w * (MP * v[p])
            f[k] += (-V[p] * S[p]) @ g
    vn = np.zeros_like(mv)
    vg = np.zeros_like(mv)
    for k in range(len(m)):
        minv = 1.0 / m[k] if m[k] > 0 else 0.0
        vn[k] = mv[k] * minv
        vg[k] = vn[k] + DT * f[k] * minv
    return m, mv, f, vn, vg


def g2p(xs, v, F, gvn, gv):
    out = dict(x=[], v=[], grad_v=[], F=[], sigma=[], V=[])
    for p in range(len(xs)):
        dv = np.zeros(2)
        L = np.zeros((2, 2))
        dx = np.zeros(2)
        for k, I in enumerate(nodes()):
            w, g = w_linear(xs[p], I)
            dv += w * (gv[k] - gvn[k])
            L += np.outer(gv[k], g)
            dx += w * gv[k] * DT
        Fn = (np.eye(2) + DT * L) @ F[p]
        J = np.linalg.det(Fn)
        sig = (MU * (Fn @ Fn.T - np.eye(2)) + LAM * np.log(J) * np.eye(2)) / J
        out["x"].append(xs[p] + dx)
        out["v"].append(v[p] + dv)
        out["grad_v"].append(L.reshape(-1, order="F"))
        out["F"].append(Fn.reshape(-1, order="F"))
        out["sigma"].append(sig.reshape(-1, order="F"))
        out["V"].append(V0 * J)
    return {k: np.array(val).tolist() for k, val in out.items()}


def main():
    xs, V, v, F, S = particles()
    m, mv, f, vn, vg = p2g(xs, V, v, S)
    I = np.array(nodes(), dtype=float) + 1.0   # 1-based node indices of the fixture
    gvn = np.stack([0.08 * I[:, 0], -0.04 * I[:, 1]], axis=1)
    gv = gvn + np.stack([0.03 * I[:, 1], -0.02 * I[:, 0]], axis=1)
    data = {
        "source": "test/transfer.jl:2-29,61-82 fixture through README.md:71-86 (numpy restatement)",
        "mesh": {"h": H, "xmin": [0.0, 0.0], "dims": [NODES, NODES]},
        "params": {"dt": DT, "mp": MP, "lambda": LAM, "mu": MU, "V0": V0},
        "particles": {"x": xs.tolist(), "V": V.tolist(), "v": v.tolist(),
                      "F": [a.reshape(-1, order="F").tolist() for a in F],
                      "sigma": [a.reshape(-1, order="F").tolist() for a in S]},
        "p2g": {"m": m.tolist(), "mv": mv.tolist(), "f": f.tolist(), "vn": vn.tolist(), "v": vg.tolist()},
        "g2p_grid": {"vn": gvn.tolist(), "v": gv.tolist()},
        "g2p": g2p(xs, v, F, gvn, gv),
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "transfer_fixture.json")
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
