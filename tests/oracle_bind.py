"""ctypes binding of the CPU oracle (oracle/libtess_oracle.so) plus host-side state.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, always as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from tesserae_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libtess_oracle.so")

_P = ctypes.POINTER
_c64 = ctypes.c_int64


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _field(a, ncomp):
    f = _abi.tess_field()
    if a is None:
        f.ptr = None
        return f
    assert a.flags["C_CONTIGUOUS"]
    f.ptr = a.ctypes.data
    f.elem_stride, f.comp_stride = ncomp, 1
    return f


class HostState:
    """Particle and grid arrays on the host in the Julia-native layout."""

    PFIELDS = ("x", "v", "V", "grad_v", "F", "sigma", "l", "logE")

    def __init__(self, scene, dtype=None):
        self.mesh = scene.mesh
        self.dim = scene.mesh.dim
        self.npdt = dtype or (np.float32 if str(scene.mesh.dtype).endswith("float32") else np.float64)
        self.n = scene.n
        for name in self.PFIELDS:
            a = getattr(scene, name)
            setattr(self, name, None if a is None else np.ascontiguousarray(a, dtype=self.npdt).copy())
        N, d = scene.mesh.nnodes(), self.dim
        self.m = np.zeros(N, self.npdt)
        self.mv = np.zeros((N, d), self.npdt)
        self.f = np.zeros((N, d), self.npdt)
        self.vn = np.zeros((N, d), self.npdt)
        self.gv = np.zeros((N, d), self.npdt)

    def ncomp(self, name):
        d = self.dim
        return {"x": d, "v": d, "V": 1, "grad_v": d * d, "F": d * d, "sigma": d * d, "l": 1, "logE": 1}[name]

    def particles_abi(self):
        P = _abi.tess_particles()
        P.n = self.n
        for name in self.PFIELDS:
            setattr(P, name, _field(getattr(self, name), self.ncomp(name)))
        return P

    def grid_abi(self):
        G = _abi.tess_grid()
        G.m = _field(self.m, 1)
        G.mv = _field(self.mv, self.dim)
        G.f = _field(self.f, self.dim)
        G.vn = _field(self.vn, self.dim)
        G.v = _field(self.gv, self.dim)
        return G

    def mesh_abi(self):
        m = self.mesh.abi()
        m.dtype = _abi.TESS_FLOAT32 if self.npdt == np.float32 else _abi.TESS_FLOAT64
        return m


def host_filter(mesh, filter):
    """A Bool array of size(mesh) as one byte per node, column-major (None stays None)."""
    if filter is None:
        return None
    f = np.asarray(filter, bool)
    assert f.shape == tuple(mesh.size())
    return np.ascontiguousarray(f.astype(np.uint8).ravel(order="F"))


class HostWeights:
    """Stored weights on the host; deferred=True keeps no storage and carries only the filter."""

    def __init__(self, basis, dim, n, npdt, filter=None, deferred=False):
        sw = basis.support_width
        self.basis = basis
        self.deferred = deferred
        self.filter = filter  # bytes per node (host_filter) or None
        if not deferred:
            self.w = np.zeros((n, sw ** dim), npdt)
            self.grad_w = np.zeros((n, sw ** dim, dim), npdt)
            self.window = np.zeros((n, 2 * dim), np.int64)
        self.dim = dim

    def abi(self):
        W = _abi.tess_weights()
        W.filter = self.filter.ctypes.data if self.filter is not None else None
        if self.deferred:
            return W
        W.w = _field(self.w, 1)
        W.grad_w = _field(self.grad_w.reshape(len(self.w), -1), self.dim)
        W.w.elem_stride, W.w.comp_stride = 1, 1
        W.grad_w.elem_stride, W.grad_w.comp_stride = self.dim, 1
        W.window = self.window.ctypes.data
        return W


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            build_oracle()
        lib = ctypes.CDLL(path)
        M, PP, G, W, PR = _P(_abi.tess_mesh), _P(_abi.tess_particles), _P(_abi.tess_grid), _P(_abi.tess_weights), \
            _P(_abi.tess_params)
        sig = {
            "oref_last_error": (ctypes.c_char_p, []),
            "oref_basis_weight": (ctypes.c_int, [M, ctypes.c_int, _P(ctypes.c_double), ctypes.c_double, ctypes.c_int,
                                                 _P(_c64), _P(ctypes.c_double), _P(ctypes.c_double)]),
            "oref_basis_weight_filtered": (ctypes.c_int, [M, ctypes.c_int, _P(ctypes.c_double), ctypes.c_double,
                                                          ctypes.c_int, ctypes.c_void_p, _P(_c64),
                                                          _P(ctypes.c_double), _P(ctypes.c_double)]),
            "oref_jet1d": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _P(ctypes.c_double),
                                          _P(ctypes.c_double)]),
            "oref_findblock": (ctypes.c_int, [M, _P(ctypes.c_double), _P(_c64)]),
            "oref_supportnodes_r": (ctypes.c_int, [M, _P(ctypes.c_double), ctypes.c_double, _P(_c64)]),
            "oref_update_weights": (ctypes.c_int, [ctypes.c_int, M, PP, W]),
            "oref_partition_update": (ctypes.c_int, [M, _abi.tess_field, _c64, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p, _P(_c64)]),
            "oref_sparsity_update": (ctypes.c_int, [M, _abi.tess_field, _c64, ctypes.c_void_p, _P(_c64),
                                                    _P(ctypes.c_int32)]),
            "oref_step_adjoint": (ctypes.c_int, [ctypes.c_int, M, G, PP, PR, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p]),
            "oref_p2g": (ctypes.c_int, [ctypes.c_int, M, G, PP, W, PR]),
            "oref_g2p": (ctypes.c_int, [ctypes.c_int, M, G, PP, W, PR]),
            "oref_g2p2g": (ctypes.c_int, [ctypes.c_int, M, G, PP, W, PR]),
            "oref_step": (ctypes.c_int, [ctypes.c_int, M, G, PP, PR, ctypes.c_int]),
            "oref_step_threaded": (ctypes.c_int, [ctypes.c_int, M, G, PP, PR, ctypes.c_int, ctypes.c_int, W]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        self.lib = lib

    def _ok(self, status):
        if status != 0:
            raise RuntimeError(f"oracle error {status}: {self.lib.oref_last_error().decode()}")

    # ---- single-particle helpers
    def basis_weight(self, mesh, basis, x, l=0.0, deferred=False, filter=None):
        """(window (d, 2) 1-based, w (sw^d,), ∇w (sw^d, d)) of one particle; filter: Bool array of size(mesh)."""
        d = mesh.dim
        sw = basis.support_width
        xx = (ctypes.c_double * d)(*x)
        win = (_c64 * (2 * d))()
        w = (ctypes.c_double * (sw ** d))()
        gw = (ctypes.c_double * (sw ** d * d))()
        fb = host_filter(mesh, filter)
        if fb is None:
            self._ok(self.lib.oref_basis_weight(ctypes.byref(mesh.abi()), basis.kind, xx, float(l), int(deferred),
                                                win, w, gw))
        else:
            self._ok(self.lib.oref_basis_weight_filtered(ctypes.byref(mesh.abi()), basis.kind, xx, float(l),
                                                         int(deferred), fb.ctypes.data, win, w, gw))
        return (np.array(win[:], dtype=np.int64).reshape(d, 2), np.array(w[:]),
                np.array(gw[:]).reshape(sw ** d, d))

    def jet1d(self, basis, xi, l=0.0):
        v, dv = ctypes.c_double(), ctypes.c_double()
        self._ok(self.lib.oref_jet1d(basis.kind, float(xi), float(l), ctypes.byref(v), ctypes.byref(dv)))
        return v.value, dv.value

    def findblock(self, mesh, x):
        out = (_c64 * mesh.dim)()
        self._ok(self.lib.oref_findblock(ctypes.byref(mesh.abi()), (ctypes.c_double * mesh.dim)(*x), out))
        b = tuple(out[:])
        return None if all(v == 0 for v in b) else b

    def supportnodes_r(self, mesh, x, r):
        out = (_c64 * (2 * mesh.dim))()
        self._ok(self.lib.oref_supportnodes_r(ctypes.byref(mesh.abi()), (ctypes.c_double * mesh.dim)(*x), float(r),
                                              out))
        return [(out[2 * d], out[2 * d + 1]) for d in range(mesh.dim)]

    # ---- whole-array operations on HostState
    def update_weights(self, st: HostState, basis, filter=None):
        W = HostWeights(basis, st.dim, st.n, st.npdt, filter=host_filter(st.mesh, filter))
        P = st.particles_abi()
        self._ok(self.lib.oref_update_weights(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(P),
                                              ctypes.byref(W.abi())))
        return W

    def partition(self, mesh, x, dtype=np.float64):
        x = np.ascontiguousarray(x, dtype=dtype)
        n = len(x)
        nb = int(np.prod(mesh.nblocks()))
        pidx = np.zeros(max(n, 1), np.int32)
        starts = np.zeros(nb, np.int32)
        stops = np.zeros(nb, np.int32)
        na = _c64(0)
        m = mesh.abi()
        m.dtype = _abi.TESS_FLOAT32 if dtype == np.float32 else _abi.TESS_FLOAT64
        self._ok(self.lib.oref_partition_update(ctypes.byref(m), _field(x, mesh.dim), n, _np_ptr(pidx),
                                                _np_ptr(starts), _np_ptr(stops), ctypes.byref(na)))
        return pidx[:na.value].copy(), starts, stops, na.value

    def sparsity(self, mesh, x, numbering=None, dtype=np.float64):
        """update_sparsity!(spinds, x, mesh): (numbering int32 (nblocks, column-major), storage_len, changed)."""
        x = np.ascontiguousarray(x, dtype=dtype).reshape(-1, mesh.dim)
        nb = int(np.prod(mesh.nblocks()))
        num = np.zeros(nb, np.int32) if numbering is None else np.ascontiguousarray(numbering, np.int32).copy()
        sl, ch = _c64(0), ctypes.c_int32(0)
        m = mesh.abi()
        m.dtype = _abi.TESS_FLOAT32 if dtype == np.float32 else _abi.TESS_FLOAT64
        self._ok(self.lib.oref_sparsity_update(ctypes.byref(m), _field(x, mesh.dim), len(x), _np_ptr(num),
                                               ctypes.byref(sl), ctypes.byref(ch)))
        return num, sl.value, bool(ch.value)

    def p2g(self, st: HostState, basis, params, weights: HostWeights | None = None):
        P, G = st.particles_abi(), st.grid_abi()
        Wp = ctypes.byref(weights.abi()) if weights is not None else None
        self._ok(self.lib.oref_p2g(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(G), ctypes.byref(P), Wp,
                                   ctypes.byref(params.abi())))

    def g2p(self, st: HostState, basis, params, weights: HostWeights | None = None):
        P, G = st.particles_abi(), st.grid_abi()
        Wp = ctypes.byref(weights.abi()) if weights is not None else None
        self._ok(self.lib.oref_g2p(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(G), ctypes.byref(P), Wp,
                                   ctypes.byref(params.abi())))

    def g2p2g(self, st: HostState, basis, params, weights: HostWeights | None = None):
        P, G = st.particles_abi(), st.grid_abi()
        Wp = ctypes.byref(weights.abi()) if weights is not None else None
        self._ok(self.lib.oref_g2p2g(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(G), ctypes.byref(P), Wp,
                                     ctypes.byref(params.abi())))

    def step(self, st: HostState, basis, params, nsteps=1):
        P, G = st.particles_abi(), st.grid_abi()
        self._ok(self.lib.oref_step(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(G), ctypes.byref(P),
                                    ctypes.byref(params.abi()), int(nsteps)))

    def step_adjoint(self, st: HostState, basis, params, ax, av, aF, ath=None):
        """Reverse of one step from the start state ``st`` (unchanged; its grid receives the
        step's forward P2G): ax/av/aF (float64, updated in place) go from the adjoint after
        the step to the adjoint before it; ath += ∂/∂logE."""
        for a in (ax, av, aF) + ((ath,) if ath is not None else ()):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        P, G = st.particles_abi(), st.grid_abi()
        self._ok(self.lib.oref_step_adjoint(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(G),
                                            ctypes.byref(P), ctypes.byref(params.abi()), _np_ptr(ax), _np_ptr(av),
                                            _np_ptr(aF), _np_ptr(ath)))

    def step_threaded(self, st: HostState, basis, params, nthreads, nsteps=1, weights: HostWeights | None = None):
        W = weights or HostWeights(basis, st.dim, st.n, st.npdt)
        P, G = st.particles_abi(), st.grid_abi()
        self._ok(self.lib.oref_step_threaded(basis.kind, ctypes.byref(st.mesh_abi()), ctypes.byref(G),
                                             ctypes.byref(P), ctypes.byref(params.abi()), int(nthreads),
                                             int(nsteps), ctypes.byref(W.abi())))
        return W
