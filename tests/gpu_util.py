"""Move scene state between the host (numpy, Julia layout) and tesserae_hip's device containers."""
from __future__ import annotations

import numpy as np
import torch

import tesserae_hip as th


def to_gpu(scene, layout="julia", dtype=None):
    dtype = dtype or scene.mesh.dtype
    mesh = scene.mesh
    if dtype != mesh.dtype:
        mesh = th.CartesianMesh(mesh.h, *[(a, b) for a, b in zip(mesh.xmin, mesh.xmax())], dtype=dtype,
                                block_size_log2=mesh.block_size_log2)
    P = th.Particles(scene.n, mesh.dim, dtype, device="cuda", layout=layout, with_l=scene.l is not None,
                     with_logE=scene.logE is not None)
    for name in ("x", "v", "V", "F", "sigma", "grad_v", "l", "logE"):
        a = getattr(scene, name)
        if a is not None:
            P.set(name, a)
    G = th.generate_grid(mesh, device="cuda", layout=layout)
    return mesh, P, G


def particles_np(P):
    out = {}
    for name in P.FIELDS:
        t = P.tensors[name]
        if t is not None:
            out[name] = P.get(name).detach().double().cpu().numpy().copy()
    return out


def grid_np(G):
    return {name: G.get(name).detach().double().cpu().numpy().copy() for name in G.FIELDS}


def rel_err(a, b):
    """max |a-b| / max |b| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max() if b.size else 0.0
    diff = np.abs(a - b).max() if b.size else 0.0
    return 0.0 if scale == 0 and diff == 0 else diff / max(scale, 1e-300)
