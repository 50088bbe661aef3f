"""Seeded synthetic scenes for the README step (host numpy state).

Each builder returns a ``Scene``: the mesh, the basis, the program constants
and the initial particle state as float64 numpy arrays in the Julia-native
layout (x/v: (n, d); V: (n,); F/sigma/grad_v: (n, d*d) column-major).  The
same arrays feed the GPU engine and the CPU oracle, so both start from
identical inputs.  Random numbers come from splitmix64, a fixed published
generator (Steele, Lea & Flood 2014), so every machine draws the same values.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .core import BSpline, Basis, CartesianMesh, MPMParams, Quadratic, grid_sampling_points

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """The n outputs of splitmix64 seeded with ``seed``, starting at draw ``offset``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(offset + 1, offset + n + 1, dtype=np.uint64) * _G)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller over two uniform streams."""
    u1 = uniform01(seed, n, offset)
    u2 = uniform01(seed ^ 0x5DEECE66D, n, offset)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Scene:
    name: str
    mesh: CartesianMesh
    basis: Basis
    params: MPMParams
    x: np.ndarray
    v: np.ndarray
    V: np.ndarray
    F: np.ndarray
    sigma: np.ndarray
    grad_v: np.ndarray
    l: np.ndarray | None = None
    logE: np.ndarray | None = None
    steps: int = 1

    @property
    def n(self) -> int:
        return len(self.x)

    def copy(self) -> "Scene":
        c = Scene(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()})
        return c


def _state(x, v, V0, dim):
    n = len(x)
    eye = np.eye(dim).reshape(-1, order="F")
    return dict(x=np.ascontiguousarray(x, dtype=np.float64), v=np.ascontiguousarray(v, dtype=np.float64),
                V=np.full(n, V0, dtype=np.float64), F=np.tile(eye, (n, 1)), sigma=np.zeros((n, dim * dim)),
                grad_v=np.zeros((n, dim * dim)))


def two_disks(dtype=torch.float64, h: float = 0.01, steps: int = 1000) -> Scene:
    """README example (README.md:44-62): two elastic disks colliding on [0,1]^2."""
    mesh = CartesianMesh(h, (0, 1), (0, 1), dtype=dtype)
    pts = grid_sampling_points(mesh)
    V0 = mesh.volume() / len(pts)  # volume / count before the filter
    r = 0.2
    keep = (np.linalg.norm(pts - r, axis=1) < r) | (np.linalg.norm(pts - (1 - r), axis=1) < r)
    x = pts[keep]
    v = np.where(x[:, :1] > 0.5, -0.1, 0.1) * np.ones((1, 2))
    params = MPMParams(dt=1e-3, E=500.0, nu=0.3, rho0=1e3, V0=V0)
    return Scene("two_disks", mesh, BSpline(Quadratic()), params, steps=steps, **_state(x, v, V0, 2))


def block_drop(dtype=torch.float32, dx: float = 1 / 256, seed: int = 1, steps: int = 100,
               box=((0.25, 0.75), (0.25, 0.75), (0.5, 0.9))) -> Scene:
    """3D elastic block falling under gravity with slip walls (2x2x2 particles per cell, jittered)."""
    mesh = CartesianMesh(dx, (0, 1), (0, 1), (0, 1), dtype=dtype)
    s = dx / 2
    axes = []
    for lo, hi in box:
        k = int(math.floor((hi - lo) / s - 0.5 + 1e-9)) + 1  # count of lo + (k+1/2) s < hi
        axes.append(lo + (np.arange(k) + 0.5) * s)
    g = np.meshgrid(*axes, indexing="ij")
    x = np.stack([a.reshape(-1, order="F") for a in g], axis=1)
    n = len(x)
    jit = (uniform01(seed, 3 * n).reshape(n, 3) * 2.0 - 1.0) * (0.1 * dx)
    x = x + jit
    V0 = s ** 3
    params = MPMParams(dt=1e-4, E=1e6, nu=0.3, rho0=1e3, V0=V0, gravity=(0.0, 0.0, -9.81), slip_walls=0b111111)
    return Scene("block_drop", mesh, BSpline(Quadratic()), params, steps=steps,
                 **_state(x, np.zeros_like(x), V0, 3))


def taylor_rod(dtype=torch.float32, radius: float = 0.0032, length: float = 0.0324, per_cell: int = 3,
               seed: int = 1, speed: float = 227.0, basis: Basis | None = None) -> Scene:
    """Geometry of the reference's Taylor-impact GPU example (docs/src/gpu.md:202-393): a rod of radius R and
    length L standing on the floor of a mesh of spacing R/12 over (-3R,3R)^2 x (0,1.1L), moving down at 227 m/s,
    with the floor a slip wall; copper constants E=117e9, nu=0.35, rho0=8.93e3 and dt = 0.8 h / (c + |v|).
    Particles sit on a lattice of spacing h/per_cell (per_cell=3 gives the example's ~1.48M particles) jittered
    by 0.1 of the spacing.  The material is the README's Neo-Hookean solid (the example's von Mises model is
    outside this engine); the basis is KernelCorrection(BSpline(Quadratic())) unless given."""
    from .core import KernelCorrection
    h = radius / 12
    mesh = CartesianMesh(h, (-3 * radius, 3 * radius), (-3 * radius, 3 * radius), (0, 1.1 * length), dtype=dtype)
    s = h / per_cell
    kr = int(math.floor(radius / s))
    ax = (np.arange(-kr, kr) + 0.5) * s
    az = (np.arange(int(math.floor(length / s))) + 0.5) * s
    X, Y = np.meshgrid(ax, ax, indexing="ij")
    keep = X ** 2 + Y ** 2 < radius ** 2
    xy = np.stack([X[keep], Y[keep]], axis=1)
    x = np.concatenate([np.column_stack([xy, np.full(len(xy), z)]) for z in az])
    n = len(x)
    x = x + (uniform01(seed, 3 * n).reshape(n, 3) * 2.0 - 1.0) * (0.1 * s)
    E, nu, rho0 = 117e9, 0.35, 8.93e3
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    dt = 0.8 * h / (math.sqrt((lam + 2 * mu) / rho0) + speed)
    V0 = math.pi * radius ** 2 * length / n
    params = MPMParams(dt=dt, E=E, nu=nu, rho0=rho0, V0=V0, slip_walls=1 << 4)
    v = np.zeros_like(x)
    v[:, 2] = -speed
    return Scene("taylor_rod", mesh, basis or KernelCorrection(BSpline(Quadratic())), params,
                 **_state(x, v, V0, 3))


def sphere_spin(basis: Basis, dtype=torch.float32, dx: float = 1 / 512, seed: int = 7, steps: int = 50,
                block_size_log2: int = 3, radius: float = 0.3) -> Scene:
    """3D sparse-grid sphere: 2x2x2 particles per cell of [0,1]^3 (spacing dx/2) inside the
    sphere of the given radius at the centre, v = 0.5*(rigid rotation about z at ω=2) +
    0.05*N(0,1) per component, Neo-Hookean E=5e5, ν=0.35, ρ⁰=1e3, Δt=5e-5; SpArray blocks
    of 2^block_size_log2 nodes; l = dx/2 for uGIMP.  Positions are generated one z-plane
    of samples at a time."""
    mesh = CartesianMesh(dx, (0, 1), (0, 1), (0, 1), dtype=dtype, block_size_log2=block_size_log2)
    s = dx / 2
    c = 0.5
    k0 = int(math.floor((c - radius) / s - 0.5))
    k1 = int(math.ceil((c + radius) / s - 0.5))
    ax = (np.arange(k0, k1 + 1) + 0.5) * s
    X, Y = np.meshgrid(ax, ax, indexing="ij")
    X, Y = X.reshape(-1, order="F"), Y.reshape(-1, order="F")
    r2xy = (X - c) ** 2 + (Y - c) ** 2
    parts = []
    for z in ax:
        keep = r2xy + (z - c) ** 2 < radius * radius
        if keep.any():
            parts.append(np.stack([X[keep], Y[keep], np.full(int(keep.sum()), z)], axis=1))
    x = np.concatenate(parts) if parts else np.zeros((0, 3))
    n = len(x)
    omega = 2.0
    v = np.stack([-(x[:, 1] - c) * omega, (x[:, 0] - c) * omega, np.zeros(n)], axis=1) * 0.5
    v = v + 0.05 * normal(seed, 3 * n).reshape(n, 3)
    V0 = s ** 3
    params = MPMParams(dt=5e-5, E=5e5, nu=0.35, rho0=1e3, V0=V0)
    l = np.full(n, dx / 2) if basis.kernel == 16 else None
    return Scene("sphere_spin", mesh, basis, params, steps=steps, l=l, **_state(x, v, V0, 3))


def inverse_cube(dtype=torch.float32, dx: float = 1 / 64, seed: int = 5, steps: int = 200,
                 lo: float = 0.3, side: float = 0.4, drift: float = 0.25, noise: float = 0.05) -> Scene:
    """Differentiable 3D inverse problem: a cube of the given side at [lo, lo+side]^3 with 2x2x2
    particles per cell of [0,1]^3 (dx), quadratic B-spline, Δt=2e-4, Neo-Hookean with base
    E=1e5, ν=0.3, ρ⁰=1e3 and per-particle logE = 0 (E_p = E exp(logE_p)).  The initial
    velocity is a drift along x plus seeded noise, v = (drift, 0, 0) + noise*N(0,1).  The
    loss target is x(0) + (0.1, 0, 0)."""
    mesh = CartesianMesh(dx, (0, 1), (0, 1), (0, 1), dtype=dtype)
    s = dx / 2
    k = int(math.floor(side / s - 0.5 + 1e-9)) + 1
    ax = lo + (np.arange(k) + 0.5) * s
    g = np.meshgrid(ax, ax, ax, indexing="ij")
    x = np.stack([a.reshape(-1, order="F") for a in g], axis=1)
    n = len(x)
    v = np.zeros((n, 3))
    v[:, 0] = drift
    v = v + noise * normal(seed, 3 * n).reshape(n, 3)
    V0 = s ** 3
    params = MPMParams(dt=2e-4, E=1e5, nu=0.3, rho0=1e3, V0=V0)
    return Scene("inverse_cube", mesh, BSpline(Quadratic()), params, steps=steps, logE=np.zeros(n),
                 **_state(x, v, V0, 3))


def batched_scenes(nscenes: int = 256, seed: int = 11, dtype=torch.float32, cells=(64, 512),
                   particles=(2e3, 2e5), area=(0.2, 0.6), stiffness=(1e3, 1e6), steps: int = 200) -> list:
    """Uneven 2D scenes for batching: per scene a grid of N cells per side on [0,1]^2 with N
    log-uniform in ``cells``, n_p particles (log-uniform in ``particles``) uniform in a random
    axis-aligned box covering a fraction of the unit square drawn from ``area``, velocities
    U(-0.2, 0.2), E log-uniform in ``stiffness``, ν=0.3, ρ⁰=1e3, V⁰ = box area / n_p, quadratic
    B-spline, Δt = 0.3 h / (c + max|v|) with c = sqrt((λ+2μ)/ρ⁰)."""
    out = []
    for s in range(nscenes):
        u = uniform01(seed * 1000003 + s, 8)
        ncell = int(round(math.exp(math.log(cells[0]) + u[0] * (math.log(cells[1]) - math.log(cells[0])))))
        npart = int(round(math.exp(math.log(particles[0]) + u[1] * (math.log(particles[1]) - math.log(particles[0])))))
        frac = area[0] + u[2] * (area[1] - area[0])
        aspect = math.exp((u[3] - 0.5) * math.log(2.0))  # width/height within [1/sqrt2, sqrt2]
        w = min(math.sqrt(frac * aspect), 0.98)
        hgt = min(frac / w, 0.98)
        x0, y0 = 0.01 + u[4] * (0.98 - w), 0.01 + u[5] * (0.98 - hgt)
        E = math.exp(math.log(stiffness[0]) + u[6] * (math.log(stiffness[1]) - math.log(stiffness[0])))
        h = 1.0 / ncell
        mesh = CartesianMesh(h, (0, 1), (0, 1), dtype=dtype)
        r = uniform01(seed * 7919 + s, 4 * npart).reshape(npart, 4)
        x = np.stack([x0 + w * r[:, 0], y0 + hgt * r[:, 1]], axis=1)
        v = (r[:, 2:4] - 0.5) * 0.4
        V0 = w * hgt / npart
        probe = MPMParams(dt=1.0, E=E, nu=0.3, rho0=1e3, V0=V0)
        c = math.sqrt((probe.lam + 2 * probe.mu) / 1e3)
        dt = 0.3 * h / (c + float(np.linalg.norm(v, axis=1).max()))
        params = MPMParams(dt=dt, E=E, nu=0.3, rho0=1e3, V0=V0)
        out.append(Scene(f"batch_{s}", mesh, BSpline(Quadratic()), params, steps=steps, **_state(x, v, V0, 2)))
    return out


def random_box(dim: int, n: int, seed: int, basis: Basis, dtype=torch.float64, h: float = 0.1,
               extent: float = 1.0, vscale: float = 0.3, lo: float = 0.0, hi: float = 1.0,
               stress: float = 0.0, with_l: bool = False) -> Scene:
    """n particles uniform in [lo, hi]^dim of a mesh on [0, extent]^dim, random v and a small random F."""
    mesh = CartesianMesh(h, *[(0.0, extent)] * dim, dtype=dtype)
    u = uniform01(seed, n * dim).reshape(n, dim)
    x = lo + (hi - lo) * u
    v = (uniform01(seed + 1, n * dim).reshape(n, dim) - 0.5) * 2 * vscale
    V0 = (extent ** dim) / max(n, 1)
    st = _state(x, v, V0, dim)
    if stress:
        d2 = dim * dim
        pert = (uniform01(seed + 2, n * d2).reshape(n, d2) - 0.5) * 2 * stress
        st["F"] = st["F"] + 0.05 * pert
        sym = pert.reshape(n, dim, dim)
        st["sigma"] = (sym + sym.transpose(0, 2, 1)).reshape(n, d2)
    params = MPMParams(dt=1e-3, E=1e3, nu=0.3, rho0=1e3, V0=V0)
    l = np.full(n, 0.5 * h) if with_l else None
    return Scene(f"random_box_{dim}d", mesh, basis, params, l=l, **st)
