"""Host-side mirror of Tesserae.jl's API for the explicit MPM step, over the C ABI.

Names follow the reference (src/Tesserae.jl:35-129): ``CartesianMesh``,
``generate_grid``, ``generate_particles``, ``generate_basis_weights``,
``update`` (Julia ``update!``), ``Partition``, ``reorder_particles``, and the
README program's ``P2G`` / ``G2P`` transfers.  Device memory is held in torch
tensors (plumbing only); every computation runs in libtesserae_hip.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np
import torch

from . import _abi
from ._abi import (TESS_BSPLINE_CUBIC, TESS_BSPLINE_LINEAR, TESS_BSPLINE_QUADRATIC, TESS_FLOAT32, TESS_FLOAT64,
                   TESS_KERNEL_CORRECTION, TESS_POLY_MULTILINEAR, TESS_UGIMP, TESS_WLS)


# ---- exceptions mirroring the reference's host-side errors -----------------

class TesseraeError(Exception):
    """Base class of errors raised by libtesserae_hip."""


class ArgumentError(TesseraeError, ValueError):
    """Julia ArgumentError."""


class DimensionMismatch(TesseraeError, ValueError):
    """Julia DimensionMismatch."""


class TesseraeAssertionError(TesseraeError, AssertionError):
    """Julia AssertionError raised by the reference's @assert checks."""


class ErrorException(TesseraeError, RuntimeError):
    """Julia ErrorException (the reference's error(...))."""


class DeviceError(TesseraeError, RuntimeError):
    """HIP runtime failure."""


class UnsupportedError(TesseraeError, NotImplementedError):
    """Program, basis or dimension outside the library's fixed set."""


_STATUS = {
    _abi.TESS_ERR_ARGUMENT: ArgumentError,
    _abi.TESS_ERR_DIMENSION: DimensionMismatch,
    _abi.TESS_ERR_ASSERTION: TesseraeAssertionError,
    _abi.TESS_ERR_STATE: ErrorException,
    _abi.TESS_ERR_DEVICE: DeviceError,
    _abi.TESS_ERR_UNSUPPORTED: UnsupportedError,
}


def _check(lib, status: int) -> None:
    if status != _abi.TESS_OK:
        msg = lib.tess_last_error().decode("utf-8", "replace")
        raise _STATUS.get(status, TesseraeError)(msg)


# ---- bases (src/Basis/bspline.jl:15-25, src/Basis/gimp.jl:16-18) ----------

@dataclass(frozen=True)
class Basis:
    kind: int
    name: str

    @property
    def kernel(self) -> int:
        """Code of the kernel (the basis itself, or the one a WLS / KernelCorrection wraps)."""
        return self.kind & 0xFF

    @property
    def corrected(self) -> bool:
        return bool(self.kind & (TESS_WLS | TESS_KERNEL_CORRECTION))

    @property
    def support_width(self) -> int:
        return {TESS_BSPLINE_LINEAR: 2, TESS_BSPLINE_QUADRATIC: 3, TESS_BSPLINE_CUBIC: 4, TESS_UGIMP: 3}[self.kernel]

    def __repr__(self) -> str:
        return self.name


class Linear:
    degree = 1


class Quadratic:
    degree = 2


class Cubic:
    degree = 3


def BSpline(degree) -> Basis:
    """BSpline(Linear()/Quadratic()/Cubic()) (src/Basis/bspline.jl:15-17)."""
    d = degree.degree if hasattr(degree, "degree") else int(degree)
    if d not in (1, 2, 3):
        raise UnsupportedError(f"BSpline degree {d} is outside the library's set (1, 2, 3)")
    return Basis(d, f"BSpline({['', 'Linear', 'Quadratic', 'Cubic'][d]}())")


def uGIMP() -> Basis:
    """uGIMP() (src/Basis/gimp.jl:16); particles need the length field ``l``."""
    return Basis(TESS_UGIMP, "uGIMP()")


# ---- corrected bases (src/Basis/wls.jl, kernelcorrection.jl, polynomials.jl) ----

class MultiLinear:
    """Degree tag of the multilinear polynomial (polynomials.jl:1)."""


@dataclass(frozen=True)
class Polynomial:
    """Polynomial(degree) (polynomials.jl:4-7): Linear() or MultiLinear() for a correction."""
    degree: object

    def __repr__(self) -> str:
        return f"Polynomial({type(self.degree).__name__}())"


def _correction_polynomial(name: str, poly) -> bool:
    """True for MultiLinear; the reference's CorrectionPolynomial check (polynomials.jl:9-21)."""
    if isinstance(poly, Polynomial) and isinstance(poly.degree, Linear):
        return False
    if isinstance(poly, Polynomial) and isinstance(poly.degree, MultiLinear):
        return True
    raise ArgumentError(f"{name} supports Polynomial(Linear()) and Polynomial(MultiLinear()), got {poly}. "
                        "A higher-degree basis leaves the weighted least-squares moment matrix singular on a kernel "
                        "support, producing negative and unbounded weights.")


def _wrap(name: str, flag: int, kernel: Basis, poly) -> Basis:
    if not isinstance(kernel, Basis) or kernel.corrected:
        raise ArgumentError(f"{name} wraps a kernel (BSpline or uGIMP), got {kernel!r}")
    ml = _correction_polynomial(name, poly)
    return Basis(kernel.kind | flag | (TESS_POLY_MULTILINEAR if ml else 0), f"{name}({kernel!r}, {poly!r})")


def WLS(kernel: Basis, poly: Polynomial | None = None) -> Basis:
    """WLS(kernel[, poly]) (wls.jl:1-17): weighted least-squares fit, Polynomial(Linear()) by default."""
    return _wrap("WLS", TESS_WLS, kernel, Polynomial(Linear()) if poly is None else poly)


def KernelCorrection(kernel: Basis, poly: Polynomial | None = None) -> Basis:
    """KernelCorrection(kernel[, poly]) (kernelcorrection.jl:1-24): the kernel on whole supports, WLS where the
    mesh or the filter cuts them; Polynomial(MultiLinear()) by default."""
    return _wrap("KernelCorrection", TESS_KERNEL_CORRECTION, kernel,
                 Polynomial(MultiLinear()) if poly is None else poly)


# ---- mesh (src/Mesh/cartesian.jl:38-103) ----------------------------------

def _covered_axis(h: float, lims: tuple[float, float]) -> tuple[float, int]:
    """Node count of range(xmin, xmax; step=h), expanded to cover xmax (:74-86)."""
    xmin, xmax = float(lims[0]), float(lims[1])
    n = int(math.floor((xmax - xmin) / h + 1e-9)) + 1
    if not math.isclose(xmin + (n - 1) * h, xmax, rel_tol=math.sqrt(2.220446049250313e-16), abs_tol=0.0):
        n += 1
    return xmin, n


class CartesianMesh:
    """CartesianMesh([T,] h, (xmin, xmax), ...; block_size_log2=2)."""

    def __init__(self, h: float, *minmax: Sequence[float], dtype: torch.dtype = torch.float64,
                 block_size_log2: int = 2):
        if len(minmax) not in (2, 3):
            raise UnsupportedError("CartesianMesh: only 2D and 3D meshes are supported")
        if not all(lo < hi for lo, hi in minmax):
            raise TesseraeAssertionError("CartesianMesh: each axis needs xmin < xmax")
        if not isinstance(block_size_log2, int) or block_size_log2 < 0:
            raise ArgumentError(f"block_size_log2 must be non-negative, got {block_size_log2}")
        if dtype not in (torch.float32, torch.float64):
            raise ArgumentError("CartesianMesh dtype must be torch.float32 or torch.float64")
        self.dim = len(minmax)
        self.h = float(h)
        self.h_inv = 1.0 / self.h
        axes = [_covered_axis(self.h, lims) for lims in minmax]
        self.xmin = tuple(a[0] for a in axes)
        self.dims = tuple(a[1] for a in axes)
        self.dtype = dtype
        self.block_size_log2 = block_size_log2

    # Julia accessors
    def size(self) -> tuple[int, ...]:
        return self.dims

    def spacing(self) -> float:
        return self.h

    def nnodes(self) -> int:
        return int(np.prod(self.dims))

    def nblocks(self) -> tuple[int, ...]:
        """nblocks(mesh) (cartesian.jl:241-243)."""
        return tuple(((n - 1) >> self.block_size_log2) + 1 for n in self.dims)

    def xmax(self) -> tuple[float, ...]:
        return tuple(x0 + (n - 1) * self.h for x0, n in zip(self.xmin, self.dims))

    def volume(self) -> float:
        """volume(mesh) = prod(xmax - xmin) (cartesian.jl:66-71)."""
        return float(np.prod([b - a for a, b in zip(self.xmin, self.xmax())]))

    def axis(self, d: int) -> np.ndarray:
        """Node coordinates along axis d, as the library computes them: T(xmin + i*h)."""
        npdt = np.float32 if self.dtype == torch.float32 else np.float64
        return (self.xmin[d] + np.arange(self.dims[d], dtype=np.float64) * self.h).astype(npdt)

    @property
    def tess_dtype(self) -> int:
        return TESS_FLOAT32 if self.dtype == torch.float32 else TESS_FLOAT64

    def abi(self) -> _abi.tess_mesh:
        m = _abi.tess_mesh()
        m.dim = self.dim
        m.dtype = self.tess_dtype
        m.block_size_log2 = self.block_size_log2
        for d in range(3):
            m.dims[d] = self.dims[d] if d < self.dim else 1
            m.xmin[d] = self.xmin[d] if d < self.dim else 0.0
        m.h = self.h
        m.h_inv = self.h_inv
        return m

    def __repr__(self) -> str:
        return f"CartesianMesh{self.dims}(h={self.h}, xmin={self.xmin}, T={self.dtype})"


# ---- StructArray-like containers -------------------------------------------

def _field(t: torch.Tensor | None, ncomp: int, layout: str) -> _abi.tess_field:
    f = _abi.tess_field()
    if t is None:
        f.ptr = None
        return f
    if not t.is_contiguous():
        raise ArgumentError("fields must be contiguous tensors")
    f.ptr = t.data_ptr()
    if layout == "soa" and ncomp > 1:
        f.elem_stride, f.comp_stride = 1, t.shape[-1]
    else:
        f.elem_stride, f.comp_stride = ncomp, 1
    return f


class Particles:
    """Particle StructArray of the README program (README.md:53-54).

    Fields: x (n,d), v (n,d), V (n,), grad_v / F / sigma (n, d*d) with
    column-major components (Tensorial Mat layout), optional l (n,) for uGIMP
    and logE (n,) for per-particle stiffness.  ``layout="soa"`` stores each
    field component-major ((ncomp, n)) instead.
    """

    FIELDS = ("x", "v", "V", "grad_v", "F", "sigma", "l", "logE")

    def __init__(self, n: int, dim: int, dtype: torch.dtype, device="cuda", layout: str = "julia",
                 with_l: bool = False, with_logE: bool = False):
        self.n, self.dim, self.dtype, self.device, self.layout = int(n), dim, dtype, device, layout
        self.tensors: dict[str, torch.Tensor | None] = {}
        for name in self.FIELDS:
            nc = self.ncomp(name)
            if (name == "l" and not with_l) or (name == "logE" and not with_logE):
                self.tensors[name] = None
                continue
            shape = (self.n,) if nc == 1 else ((nc, self.n) if layout == "soa" else (self.n, nc))
            self.tensors[name] = torch.zeros(shape, dtype=dtype, device=device)

    def ncomp(self, name: str) -> int:
        return {"x": self.dim, "v": self.dim, "V": 1, "grad_v": self.dim ** 2, "F": self.dim ** 2,
                "sigma": self.dim ** 2, "l": 1, "logE": 1}[name]

    def __getattr__(self, name):
        t = self.__dict__.get("tensors")
        if t is not None and name in t:
            return t[name]
        raise AttributeError(name)

    def __len__(self) -> int:
        return self.n

    def get(self, name: str) -> torch.Tensor:
        """Field as (n, ncomp) (or (n,)) regardless of layout."""
        t = self.tensors[name]
        return t.t() if (self.layout == "soa" and t is not None and t.dim() == 2) else t

    def set(self, name: str, value) -> None:
        t = self.tensors[name]
        v = torch.as_tensor(value, dtype=self.dtype).to(t.device)
        if self.layout == "soa" and t.dim() == 2:
            t.copy_(v.reshape(self.n, -1).t())
        else:
            t.copy_(v.reshape(t.shape))

    def abi(self) -> _abi.tess_particles:
        P = _abi.tess_particles()
        P.n = self.n
        for name in self.FIELDS:
            setattr(P, name, _field(self.tensors[name], self.ncomp(name), self.layout))
        return P


class Grid:
    """Grid StructArray of the README program: m (N,), mv / f / vn / v (N, d); x-fastest nodes."""

    FIELDS = ("m", "mv", "f", "vn", "v")

    def __init__(self, mesh: CartesianMesh, device="cuda", layout: str = "julia"):
        self.mesh, self.layout = mesh, layout
        N, d = mesh.nnodes(), mesh.dim
        self.tensors = {}
        for name in self.FIELDS:
            if name == "m":
                self.tensors[name] = torch.zeros((N,), dtype=mesh.dtype, device=device)
            else:
                shape = (d, N) if layout == "soa" else (N, d)
                self.tensors[name] = torch.zeros(shape, dtype=mesh.dtype, device=device)

    def __getattr__(self, name):
        t = self.__dict__.get("tensors")
        if t is not None and name in t:
            return t[name]
        raise AttributeError(name)

    def get(self, name: str) -> torch.Tensor:
        t = self.tensors[name]
        return t.t() if (self.layout == "soa" and t.dim() == 2) else t

    def abi(self) -> _abi.tess_grid:
        G = _abi.tess_grid()
        for name in self.FIELDS:
            setattr(G, name, _field(self.tensors[name], 1 if name == "m" else self.mesh.dim, self.layout))
        return G


class SpIndices:
    """SpIndices(mesh) (src/Grids/spindices.jl:57-70): one int32 block number per
    block of 2^L nodes per axis, column-major; 0 = inactive, k = k-th active block."""

    def __init__(self, mesh: CartesianMesh, device="cuda"):
        self.mesh = mesh
        self.blocknumbering = torch.zeros((int(np.prod(mesh.nblocks())),), dtype=torch.int32, device=device)
        self.storage_len = 0

    def blocklength(self) -> int:
        return 1 << (self.mesh.block_size_log2 * self.mesh.dim)

    def numbering(self) -> np.ndarray:
        """blocknumbering as a host array of shape nblocks (Julia column-major)."""
        return self.blocknumbering.cpu().numpy().reshape(self.mesh.nblocks(), order="F")


class SpGrid(Grid):
    """generate_grid(SpArray, GridProp, mesh) (src/Grids/grid.jl:79-101): every field
    is an SpArray data vector over the active blocks of ``spinds``."""

    def __init__(self, mesh: CartesianMesh, device="cuda", layout: str = "julia"):
        if layout != "julia":
            raise ArgumentError("SpGrid storage uses the Julia layout")
        self.mesh, self.layout, self.device = mesh, layout, device
        self.spinds = SpIndices(mesh, device)
        self.tensors = {}
        self._allocate(0)

    def _allocate(self, nodes: int) -> None:
        d = self.mesh.dim
        for name in self.FIELDS:
            shape = (nodes,) if name == "m" else (nodes, d)
            self.tensors[name] = torch.zeros(shape, dtype=self.mesh.dtype, device=self.device)
        self.capacity = nodes

    def resize_fillzero(self, nodes: int) -> None:
        """resize_fillzero_data! of every field (src/Grids/sparray.jl:147-150): grow when
        needed, then zero."""
        if nodes > self.capacity:
            self._allocate(nodes)
        else:
            for t in self.tensors.values():
                t.zero_()

    def sparsity_abi(self) -> _abi.tess_sparsity:
        sp = _abi.tess_sparsity()
        sp.blocknumbering = self.spinds.blocknumbering.data_ptr()
        sp.storage_len = self.spinds.storage_len
        sp.capacity = self.capacity
        return sp

    def to_dense(self, name: str) -> np.ndarray:
        """Field over all mesh nodes (inactive nodes read as zero, as SpArray getindex does)."""
        nb = self.mesh.nblocks()
        num = self.spinds.numbering()
        L, d = self.mesh.block_size_log2, self.mesh.dim
        data = self.tensors[name].detach().double().cpu().numpy()
        ncomp = 1 if name == "m" else d
        out = np.zeros((self.mesh.nnodes(), ncomp))
        idx = sparse_storage_index(self.mesh, num)
        ok = idx >= 0
        out[ok] = data.reshape(-1, ncomp)[idx[ok]]
        return out[:, 0] if ncomp == 1 else out


def sparse_storage_index(mesh: CartesianMesh, numbering: np.ndarray) -> np.ndarray:
    """0-based SpArray storage index of every mesh node (x fastest), -1 if inactive:
    (k-1)*2^(L*dim) + column-major local index (src/Grids/spindices.jl:89,96-102)."""
    L, d = mesh.block_size_log2, mesh.dim
    grids = np.meshgrid(*[np.arange(n) for n in mesh.dims], indexing="ij")
    I = [g.reshape(-1, order="F") for g in grids]
    blk = [i >> L for i in I]
    loc = [i & ((1 << L) - 1) for i in I]
    k = numbering[tuple(blk)]
    lin_local = np.zeros_like(loc[0])
    for a in reversed(range(d)):
        lin_local = (lin_local << L) | loc[a]
    return np.where(k > 0, (k.astype(np.int64) - 1) * (1 << (L * d)) + lin_local, -1)


def generate_grid(mesh_or_kind, mesh: CartesianMesh | None = None, device="cuda", layout: str = "julia"):
    """generate_grid(GridProp, mesh) / generate_grid(SpArray, GridProp, mesh) for the
    README grid schema (src/Grids/grid.jl:70-86); fields zeroed."""
    if isinstance(mesh_or_kind, CartesianMesh):
        return Grid(mesh_or_kind, device=device, layout=layout)
    if mesh_or_kind is SpArray:
        return SpGrid(mesh, device=device, layout=layout)
    raise ArgumentError("generate_grid: expected a mesh or (SpArray, mesh)")


class SpArray:
    """Marker for generate_grid(SpArray, mesh)."""


def grid_sampling_points(mesh: CartesianMesh, spacing: float = 0.5) -> np.ndarray:
    """GridSampling positions (src/particles.jl:9-18): axis (xmin+l/2):l:xmax, l = spacing*h, x fastest."""
    l = spacing * mesh.h
    axes = []
    for x0, x1 in zip(mesh.xmin, mesh.xmax()):
        start = x0 + l / 2
        n = int(math.floor((x1 - start) / l + 1e-9)) + 1
        axes.append(start + np.arange(n, dtype=np.float64) * l)
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.reshape(-1, order="F") for g in grids], axis=1)


def generate_particles(mesh: CartesianMesh, n: int | None = None, positions=None, device="cuda",
                       layout: str = "julia", with_l: bool = False, with_logE: bool = False) -> Particles:
    """generate_particles(ParticleProp, mesh) for the README particle schema.

    With no ``positions`` the GridSampling points of the mesh are used; F starts at I.
    """
    pts = grid_sampling_points(mesh) if positions is None else np.asarray(positions, dtype=np.float64)
    n = len(pts) if n is None else n
    P = Particles(n, mesh.dim, mesh.dtype, device=device, layout=layout, with_l=with_l, with_logE=with_logE)
    P.set("x", pts[:n])
    eye = np.eye(mesh.dim).reshape(-1, order="F")
    P.set("F", np.broadcast_to(eye, (n, mesh.dim ** 2)))
    return P


class BasisWeights:
    """BasisWeightArray (src/Basis/basis_weight.jl:345-358) in stored or deferred mode.

    ``deferred`` weights store nothing (``generate_basis_weights(...; deferred=true)``);
    stored weights switched over by ``update(..., deferred=True)`` keep their
    storage but are evaluated inside the transfers until the next plain update
    (basis_weight.jl:700-721).  ``filter`` is the node mask the last update
    recorded for a corrected basis (None = Trues).
    """

    def __init__(self, basis: Basis, mesh: CartesianMesh, n: int, deferred: bool = False, device="cuda"):
        self.basis, self.mesh, self.n, self.storageless = basis, mesh, int(n), deferred
        self.deferring = False
        self.filter = None
        sw, d = basis.support_width, mesh.dim
        if deferred:
            self.w = self.grad_w = self.window = None
        else:
            self.w = torch.zeros((self.n, sw ** d), dtype=mesh.dtype, device=device)
            self.grad_w = torch.zeros((self.n, sw ** d, d), dtype=mesh.dtype, device=device)
            self.window = torch.zeros((self.n, 2 * d), dtype=torch.int64, device=device)

    @property
    def deferred(self) -> bool:
        """Tesserae.isdeferred: transfers evaluate the basis instead of reading stored values."""
        return self.storageless or self.deferring

    def __len__(self) -> int:
        return self.n

    def abi(self) -> _abi.tess_weights | None:
        """The C-ABI view: None for deferred weights without a filter, a storage-less struct carrying the
        filter for deferred ones with a filter, else the stored arrays (and the filter update! reads)."""
        W = _abi.tess_weights()
        W.filter = self.filter.data_ptr() if self.filter is not None else None
        if self.deferred:
            return W if self.filter is not None else None
        W.w = _field(self.w, 1, "julia")
        W.grad_w = _field(self.grad_w, self.mesh.dim, "julia")
        W.window = self.window.data_ptr()
        return W


def generate_basis_weights(basis: Basis, mesh: CartesianMesh, n: int, deferred: bool = False,
                           device="cuda") -> BasisWeights:
    """generate_basis_weights([T,] basis, mesh, n; deferred) (basis_weight.jl:514-520)."""
    return BasisWeights(basis, mesh, n, deferred=deferred, device=device)


class Partition:
    """Partition(mesh) with the CPU block strategy's ordering (src/Parallel/partition.jl:37-56)."""

    def __init__(self, mesh: CartesianMesh, device="cuda"):
        self.mesh = mesh
        nb = int(np.prod(mesh.nblocks()))
        self.starts = torch.zeros((nb,), dtype=torch.int32, device=device)
        self.stops = torch.zeros((nb,), dtype=torch.int32, device=device)
        self.particleindices = torch.zeros((0,), dtype=torch.int32, device=device)
        self.nassigned = 0
        self.device = device


@dataclass
class MPMParams:
    """Constants of the README program (README.md:46-49,57) plus gravity and slip walls."""

    dt: float
    E: float
    nu: float
    rho0: float
    V0: float
    gravity: tuple = (0.0, 0.0, 0.0)
    slip_walls: int = 0
    lam: float = field(init=False)
    mu: float = field(init=False)
    mp: float = field(init=False)

    def __post_init__(self):
        E, nu = self.E, self.nu
        self.lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
        self.mu = E / (2 * (1 + nu))
        self.mp = self.rho0 * self.V0

    def abi(self) -> _abi.tess_params:
        p = _abi.tess_params()
        p.dt, p.lambda_, p.mu, p.mp, p.V0 = self.dt, self.lam, self.mu, self.mp, self.V0
        g = list(self.gravity) + [0.0] * (3 - len(self.gravity))
        for d in range(3):
            p.gravity[d] = g[d]
        p.slip_walls = int(self.slip_walls)
        return p


# ---- context ---------------------------------------------------------------

class Context:
    """Owns one tess_ctx bound to a device and a HIP stream."""

    def __init__(self, device: int | None = None, stream=None, library: str | None = None):
        # library: the path of another build of libtesserae_hip (e.g. the debug build) for this context
        self.lib = _abi.load_library(library)
        self.device = torch.cuda.current_device() if device is None else int(device)
        s = torch.cuda.current_stream(self.device) if stream is None else stream
        self.stream = s
        handle = ctypes.c_void_p()
        _check(self.lib, self.lib.tess_ctx_create(self.device, ctypes.c_void_p(s.cuda_stream), ctypes.byref(handle)))
        self.handle = handle

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _check(self.lib, self.lib.tess_ctx_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        _check(self.lib, self.lib.tess_sync(self.handle))

    def profile(self, on: bool = True) -> None:
        """Start (or stop) per-kernel HIP-event timing on this context's stream."""
        _check(self.lib, self.lib.tess_profile_enable(self.handle, int(on)))

    def profile_read(self) -> dict:
        """{kernel: (total_ms, calls, algorithmic_bytes)} accumulated since profile(True)."""
        out, i = {}, 0
        name = ctypes.create_string_buffer(64)
        ms, calls, nbytes = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        while True:
            st = self.lib.tess_profile_read(self.handle, i, name, 64, ctypes.byref(ms), ctypes.byref(calls),
                                            ctypes.byref(nbytes))
            if st == _abi.TESS_ERR_ARGUMENT:
                return out
            _check(self.lib, st)
            out[name.value.decode()] = (ms.value, calls.value, nbytes.value)
            i += 1

    def device_copy(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        """The library's plain copy kernel from src to dst (same byte size, a multiple of 16)."""
        nb = src.numel() * src.element_size()
        if dst.numel() * dst.element_size() != nb:
            raise DimensionMismatch("device_copy: src and dst differ in size")
        _check(self.lib, self.lib.tess_device_copy(self.handle, ctypes.c_void_p(src.data_ptr()),
                                                   ctypes.c_void_p(dst.data_ptr()), nb))

    def stats(self) -> dict:
        vals = (ctypes.c_int64 * 8)()
        _check(self.lib, self.lib.tess_stats(self.handle, vals, 8))
        # misplaced: layered neighbour pairs not consecutive in storage (contiguity 1 - misplaced / pairs)
        keys = ("n_inside", "active_tiles", "node_blocks", "active_nodes", "reorders", "misplaced", "movers",
                "kept_binnings")
        return dict(zip(keys, list(vals)))


_contexts: dict = {}


def default_context() -> Context:
    dev = torch.cuda.current_device()
    key = (dev, torch.cuda.current_stream(dev).cuda_stream)
    if key not in _contexts:
        _contexts[key] = Context(dev)
    return _contexts[key]


# ---- operations --------------------------------------------------------------

def _require_device(*tensors) -> None:
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ArgumentError("libtesserae_hip runs on the GPU: move grid, particles and weights to 'cuda'")


def update_sparsity(grid: SpGrid, x: torch.Tensor, ctx: Context | None = None):
    """update_sparsity!(grid, xₚ) (src/Grids/grid.jl:97-101, sparsity.jl:23-40): activates the
    blocks within ±1 block of every occupied block, renumbers them, resizes and zeroes
    the SpArray storage.  Returns the storage length when the active set changed, else
    None (the reference's ``nothing``)."""
    if not isinstance(grid, SpGrid):
        raise ArgumentError("update_sparsity: expected an SpGrid")
    _require_device(x)
    ctx = ctx or default_context()
    sp = grid.sparsity_abi()
    changed = ctypes.c_int32(0)
    f = _field(x, grid.mesh.dim, "julia")
    _check(ctx.lib, ctx.lib.tess_sparsity_update(ctx.handle, ctypes.byref(grid.mesh.abi()), f, x.shape[0],
                                                 ctypes.byref(sp), ctypes.byref(changed)))
    grid.spinds.storage_len = int(sp.storage_len)
    grid.resize_fillzero(grid.spinds.storage_len)
    return grid.spinds.storage_len if changed.value else None


def update(obj, *args, ctx: Context | None = None, **kwargs):
    """Julia's ``update!``: ``update(weights, particles, mesh)`` or ``update(partition, x)``."""
    if isinstance(obj, BasisWeights):
        return _update_weights(obj, *args, ctx=ctx, **kwargs)
    if isinstance(obj, Partition):
        return _update_partition(obj, *args, ctx=ctx)
    raise ArgumentError(f"update: no method for {type(obj).__name__}")


def _node_filter(mesh: CartesianMesh, filter, device) -> torch.Tensor | None:
    """A Bool array of size(mesh) as one byte per node in column-major node order; None for Trues."""
    if filter is None:
        return None
    f = filter.detach().cpu().numpy() if isinstance(filter, torch.Tensor) else np.asarray(filter)
    if tuple(f.shape) != tuple(mesh.size()):
        raise TesseraeAssertionError(f"size(mesh) == size(filter): {tuple(mesh.size())} vs {tuple(f.shape)}")
    return torch.from_numpy(np.ascontiguousarray(f.astype(np.uint8).ravel(order="F"))).to(device)


def _update_weights(weights: BasisWeights, particles: Particles, mesh: CartesianMesh, filter=None, ctx=None,
                    deferred=None):
    """update!(weights, particles, mesh[, filter]; deferred) (basis_weight.jl:700-752)."""
    deferred = weights.storageless if deferred is None else bool(deferred)
    device = particles.x.device
    if deferred or weights.storageless:
        # recorded before the early return: deferred corrected bases read it in the transfers
        weights.filter = _node_filter(mesh, filter, device)
    if weights.storageless:
        if not deferred:
            raise ErrorException("update!: these weights were built with `deferred=true` and have no storage for "
                                 "basis values, so they cannot be materialized. Build them without `deferred=true` "
                                 "to store the values.")
        return weights
    if deferred:
        weights.deferring = True
        return weights
    weights.deferring = False
    if len(weights) < len(particles):
        raise TesseraeAssertionError("length(weights) ≥ n")
    weights.filter = _node_filter(mesh, filter, device)
    if weights.filter is not None and len(particles) != 0 and not weights.basis.corrected:
        raise ArgumentError(f"{weights.basis!r} does not support filtered updates")
    _require_device(particles.x, weights.w)
    ctx = ctx or default_context()
    P = particles.abi()
    _check(ctx.lib, ctx.lib.tess_update_weights(ctx.handle, weights.basis.kind, ctypes.byref(mesh.abi()),
                                                ctypes.byref(P), ctypes.byref(weights.abi())))
    return weights


def _update_partition(partition: Partition, x: torch.Tensor, ctx=None, layout: str = "julia"):
    """update!(partition, x) with CPUBlockStrategy ordering (cpu_block_strategy.jl:62-75)."""
    _require_device(x)
    ctx = ctx or default_context()
    n = x.shape[0] if layout == "julia" else x.shape[-1]
    if partition.particleindices.numel() != n:
        partition.particleindices = torch.zeros((n,), dtype=torch.int32, device=x.device)
    na = ctypes.c_int64(0)
    f = _field(x, partition.mesh.dim, layout)
    _check(ctx.lib, ctx.lib.tess_partition_update(ctx.handle, ctypes.byref(partition.mesh.abi()), f, n,
                                                  ctypes.c_void_p(partition.particleindices.data_ptr()),
                                                  ctypes.c_void_p(partition.starts.data_ptr()),
                                                  ctypes.c_void_p(partition.stops.data_ptr()), ctypes.byref(na)))
    partition.nassigned = int(na.value)
    return partition


def reorder_particles(particles: Particles, partition: Partition, threshold: float = 1.0, ctx=None) -> bool:
    """reorder_particles!(particles, partition; threshold) (cpu_block_strategy.jl:248-269)."""
    ctx = ctx or default_context()
    out = ctypes.c_int32(0)
    P = particles.abi()
    _check(ctx.lib, ctx.lib.tess_reorder_particles(ctx.handle, ctypes.byref(partition.mesh.abi()), ctypes.byref(P),
                                                   ctypes.c_void_p(partition.particleindices.data_ptr()),
                                                   partition.nassigned, float(threshold), ctypes.byref(out)))
    return bool(out.value)


def P2G(grid: Grid, particles: Particles, weights, params: MPMParams, ctx=None) -> None:
    """@P2G of the README program (README.md:71-77), deterministic."""
    basis, Wp = _resolve_weights(weights)
    _require_device(particles.x, grid.m)
    ctx = ctx or default_context()
    P, G = particles.abi(), grid.abi()
    if isinstance(grid, SpGrid):
        _check(ctx.lib, ctx.lib.tess_p2g_sparse(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                                ctypes.byref(grid.sparsity_abi()), ctypes.byref(P), Wp,
                                                ctypes.byref(params.abi())))
        return
    _check(ctx.lib, ctx.lib.tess_p2g(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                     ctypes.byref(P), Wp, ctypes.byref(params.abi())))


def G2P(grid: Grid, particles: Particles, weights, params: MPMParams, ctx=None) -> None:
    """@G2P of the README program (README.md:80-87)."""
    basis, Wp = _resolve_weights(weights)
    _require_device(particles.x, grid.m)
    ctx = ctx or default_context()
    P, G = particles.abi(), grid.abi()
    if isinstance(grid, SpGrid):
        _check(ctx.lib, ctx.lib.tess_g2p_sparse(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                                ctypes.byref(grid.sparsity_abi()), ctypes.byref(P), Wp,
                                                ctypes.byref(params.abi())))
        return
    _check(ctx.lib, ctx.lib.tess_g2p(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                     ctypes.byref(P), Wp, ctypes.byref(params.abi())))


def G2P2G(grid: Grid, particles: Particles, weights, params: MPMParams, ctx=None) -> None:
    """@G2P2G of the README program (src/Transfer/g2p2g.jl:100-159): the G2P half, then
    the P2G and grid-node halves of the next step, both halves with the weights at the
    positions the particles have before the call.  Dense grids."""
    basis, Wp = _resolve_weights(weights)
    _require_device(particles.x, grid.m)
    if isinstance(grid, SpGrid):
        raise UnsupportedError("G2P2G on an SpGrid is not supported")
    ctx = ctx or default_context()
    P, G = particles.abi(), grid.abi()
    _check(ctx.lib, ctx.lib.tess_g2p2g(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                       ctypes.byref(P), Wp, ctypes.byref(params.abi())))


def _resolve_weights(weights):
    """(basis, pointer-to-tess_weights or None for deferred evaluation)."""
    if isinstance(weights, Basis):
        return weights, None
    if not isinstance(weights, BasisWeights):
        raise ArgumentError(f"invalid `BasisWeight`s, got type {type(weights).__name__}")
    W = weights.abi()
    if W is None:
        return weights.basis, None
    weights._abi_keepalive = W
    return weights.basis, ctypes.pointer(W)


def g2p2g_steps(grid: Grid, particles: Particles, basis: Basis, params: MPMParams, nsteps: int = 1,
                reorder_threshold: float = 0.0, ctx=None) -> None:
    """nsteps @G2P2G passes with deferred weights: the time loop of a program written as one @G2P2G block
    (g2p2g.jl:100-159), with the particles reordered into the engine's tile order at the threshold as in
    :func:`step`.  Dense grids."""
    _require_device(particles.x, grid.m)
    if isinstance(grid, SpGrid):
        raise UnsupportedError("G2P2G on an SpGrid is not supported")
    ctx = ctx or default_context()
    P, G = particles.abi(), grid.abi()
    _check(ctx.lib, ctx.lib.tess_g2p2g_steps(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                             ctypes.byref(P), ctypes.byref(params.abi()), int(nsteps),
                                             float(reorder_threshold)))


def step(grid: Grid, particles: Particles, basis: Basis, params: MPMParams, nsteps: int = 1,
         reorder_threshold: float = 0.0, ctx=None) -> None:
    """nsteps of the README loop body with deferred weights (README.md:66-87).

    ``reorder_threshold`` acts like the reference loop's
    ``reorder_particles!(particles, partition; threshold)``: particles are
    permuted into the engine's tile order when the in-place fraction drops
    below it (0 never, 1 every step)."""
    _require_device(particles.x, grid.m)
    ctx = ctx or default_context()
    P = particles.abi()
    if isinstance(grid, SpGrid):
        # update_sparsity!(grid, x) runs inside every step; when the active blocks
        # outgrow the storage the library stops before that step's work and reports
        # how many steps ran; the storage grows and the remaining steps follow
        if grid.capacity == 0:
            update_sparsity(grid, particles.x, ctx=ctx)
        s = 0
        while s < nsteps:
            sp, G = grid.sparsity_abi(), grid.abi()
            done = ctypes.c_int32(0)
            st = ctx.lib.tess_step_sparse(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                          ctypes.byref(sp), ctypes.byref(P), ctypes.byref(params.abi()),
                                          int(nsteps - s), float(reorder_threshold), ctypes.byref(done))
            grid.spinds.storage_len = int(sp.storage_len)
            s += int(done.value)
            if st == _abi.TESS_ERR_STATE and sp.storage_len > grid.capacity:
                grid.resize_fillzero(int(sp.storage_len * 1.25) + 1)
                continue
            _check(ctx.lib, st)
        return
    G = grid.abi()
    _check(ctx.lib, ctx.lib.tess_step(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                      ctypes.byref(P), ctypes.byref(params.abi()), int(nsteps),
                                      float(reorder_threshold)))


# ---- differentiation --------------------------------------------------------------

class Adjoint:
    """Float64 adjoint of the step state: x̄, v̄ (n, d), F̄ (n, d*d) column-major, and the
    accumulated ∂/∂logE (n,)."""

    def __init__(self, n: int, dim: int, device="cuda"):
        self.n, self.dim = int(n), dim
        self.x = torch.zeros((self.n, dim), dtype=torch.float64, device=device)
        self.v = torch.zeros((self.n, dim), dtype=torch.float64, device=device)
        self.F = torch.zeros((self.n, dim * dim), dtype=torch.float64, device=device)
        self.logE = torch.zeros((self.n,), dtype=torch.float64, device=device)

    def abi(self) -> _abi.tess_adjoint:
        A = _abi.tess_adjoint()
        A.x = _field(self.x, self.dim, "julia")
        A.v = _field(self.v, self.dim, "julia")
        A.F = _field(self.F, self.dim * self.dim, "julia")
        A.logE = _field(self.logE, 1, "julia")
        return A


def step_adjoint(grid: Grid, particles: Particles, basis: Basis, params: MPMParams, adj: Adjoint, ctx=None) -> None:
    """Reverse of one step from the start state ``particles`` (unchanged): ``adj`` goes from
    the adjoint after the step to the adjoint before it; ``adj.logE`` accumulates."""
    _require_device(particles.x, grid.m, adj.x)
    ctx = ctx or default_context()
    P, G, A = particles.abi(), grid.abi(), adj.abi()
    _check(ctx.lib, ctx.lib.tess_step_adjoint(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                              ctypes.byref(P), ctypes.byref(params.abi()), ctypes.byref(A)))


def rollout(grid: Grid, particles: Particles, basis: Basis, params: MPMParams, nsteps: int,
            checkpoint_every: int = 20, ctx=None) -> None:
    """nsteps forward steps (no reordering) with a checkpoint every ``checkpoint_every`` steps,
    recorded for :func:`rollout_adjoint`."""
    _require_device(particles.x, grid.m)
    ctx = ctx or default_context()
    P, G = particles.abi(), grid.abi()
    ctx._rollout_keepalive = particles.tensors.get("logE")
    _check(ctx.lib, ctx.lib.tess_rollout(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()), ctypes.byref(G),
                                         ctypes.byref(P), ctypes.byref(params.abi()), int(nsteps),
                                         int(checkpoint_every)))


def rollout_adjoint(grid: Grid, basis: Basis, params: MPMParams, adj: Adjoint, ctx=None) -> None:
    """Pullback of the context's last rollout: ``adj`` holds the adjoint of the final state
    on entry and of the initial state on return; ``adj.logE`` accumulates ∂/∂logE."""
    _require_device(grid.m, adj.x)
    ctx = ctx or default_context()
    G, A = grid.abi(), adj.abi()
    _check(ctx.lib, ctx.lib.tess_rollout_adjoint(ctx.handle, basis.kind, ctypes.byref(grid.mesh.abi()),
                                                 ctypes.byref(G), ctypes.byref(params.abi()), ctypes.byref(A)))


# ---- batches of scenes ---------------------------------------------------------

class Batch:
    """Independent scenes (one dimension, dtype and basis) stepped together: every
    kernel of a step is one launch for the whole batch.  ``scenes`` is a list of
    (grid, particles, params) with Julia-layout fields; the arrays are referenced."""

    def __init__(self, scenes, basis: Basis, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.basis = basis
        self.members = list(scenes)
        arr = (_abi.tess_scene * len(self.members))()
        for k, (grid, particles, params) in enumerate(self.members):
            _require_device(particles.x, grid.m)
            arr[k].mesh = grid.mesh.abi()
            arr[k].grid = grid.abi()
            arr[k].particles = particles.abi()
            arr[k].params = params.abi()
        self._scenes = arr
        handle = ctypes.c_void_p()
        _check(self.ctx.lib, self.ctx.lib.tess_batch_create(self.ctx.handle, basis.kind, len(self.members), arr,
                                                            ctypes.byref(handle)))
        self.handle = handle

    def step(self, nsteps: int = 1) -> None:
        _check(self.ctx.lib, self.ctx.lib.tess_batch_step(self.ctx.handle, self.handle, int(nsteps)))

    def set_reorder(self, threshold: float) -> None:
        """Reorder each scene's particles into the tile order when the contiguity score (the fraction of
        neighbours in that order whose storage indices are consecutive) is below ``threshold`` (0 never, 1 every
        step); permutes the scenes' particle arrays like reorder_particles!."""
        _check(self.ctx.lib, self.ctx.lib.tess_batch_set_reorder(self.handle, float(threshold)))

    def set_variant(self, variant: int) -> None:
        """Launch variant to use (a number autotune returned): in 2D 0..5, in 3D 0 or 2, each
        optionally with bit 3 (+8): graph mode, every step one replay of a captured hipGraph with no
        host readback inside the step (one workgroup per possible tile; +24: persistent grids)."""
        _check(self.ctx.lib, self.ctx.lib.tess_batch_set_variant(self.handle, int(variant)))

    def stats(self) -> dict:
        vals = (ctypes.c_int64 * 6)()
        _check(self.ctx.lib, self.ctx.lib.tess_batch_stats(self.handle, vals, 6))
        return dict(zip(("reorders", "reorders_by_count", "kept_binnings", "variant", "graph_replays", "captures"),
                        list(vals)))

    def autotune(self, steps_per_variant: int = 2) -> int:
        v = ctypes.c_int32(0)
        _check(self.ctx.lib, self.ctx.lib.tess_batch_autotune(self.ctx.handle, self.handle, int(steps_per_variant),
                                                              ctypes.byref(v)))
        return int(v.value)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _check(self.ctx.lib, self.ctx.lib.tess_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
