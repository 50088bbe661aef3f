"""Output of particles and grids in the formats of src/IO/export.jl.

``openvtk(path, x)`` / ``openvtk(path, mesh)`` with ``vtk[name] = field`` and
``close`` write a VTK XML UnstructuredGrid of vertices (.vtu) or a
StructuredGrid of the mesh nodes (.vts) (export.jl:47-55, 89-99);
``openvtm`` collects block files into a multiblock file (.vtm, :101);
``openpvd`` collects (time, file) entries into a ParaView collection (.pvd,
:102); ``write_ply`` writes a binary PLY point cloud (:115-132).  Device
fields are packed into the Float32 VTK layout on the GPU (tess_export_field)
and copied once; the files are written by the library's host writers.
"""
from __future__ import annotations

import ctypes
import os
import xml.etree.ElementTree as ET

import numpy as np
import torch

from . import _abi
from .core import ArgumentError, CartesianMesh, _check, default_context, _field


class Symmetric:
    """Marks a field of full dim×dim tensors as symmetric: written as (xx, yy, zz, xy, yz, zx)
    (export.jl:79-85), NaN where 2D has no component."""

    def __init__(self, data):
        self.data = data


def _host_layout(a: np.ndarray, dim: int, kind: int) -> np.ndarray:
    a = np.asarray(a)
    n = a.shape[0]
    a = a.reshape(n, int(np.prod(a.shape[1:])))
    if kind == _abi.TESS_VTK_POINTS:
        out = np.zeros((n, 3), np.float32)
        out[:, :dim] = a[:, :dim]
        return out
    if kind == _abi.TESS_VTK_SYMMETRIC:
        out = np.full((n, 6), np.nan, np.float32)
        for j, (r, c) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0))):
            if r < dim and c < dim:
                out[:, j] = a[:, r + dim * c]
        return out
    return np.ascontiguousarray(a, dtype=np.float32)


def export_field(data, dim: int, kind: int, ctx=None) -> np.ndarray:
    """f32(vtk_format(data)) as a host (n, components) Float32 array.  A CUDA tensor (Julia layout,
    components contiguous) is packed on the device with one copy back; a host array is converted here."""
    if isinstance(data, torch.Tensor) and data.is_cuda:
        t = data.contiguous()
        n = t.shape[0]
        nc = {_abi.TESS_VTK_SCALAR: 1, _abi.TESS_VTK_VECTOR: dim, _abi.TESS_VTK_POINTS: 3,
              _abi.TESS_VTK_TENSOR: dim * dim, _abi.TESS_VTK_SYMMETRIC: 6}[kind]
        out = np.empty((n, nc), np.float32)
        if n == 0:
            return out
        if t.dtype not in (torch.float32, torch.float64):
            raise ArgumentError("fields are Float32 or Float64")
        ctx = ctx or default_context()
        width = t.reshape(n, -1).shape[1]
        dtype = _abi.TESS_FLOAT32 if t.dtype == torch.float32 else _abi.TESS_FLOAT64
        _check(ctx.lib, ctx.lib.tess_export_field(ctx.handle, dtype, dim, _field(t, width, "julia"), n, kind,
                                                  out.ctypes.data))
        return out
    if isinstance(data, torch.Tensor):
        data = data.numpy()
    return _host_layout(data, dim, kind)


def _kind_of(value, n: int, dim: int):
    if isinstance(value, Symmetric):
        return value.data, _abi.TESS_VTK_SYMMETRIC
    shape = tuple(value.shape)
    if len(shape) == 1 or (len(shape) == 2 and shape[1] == 1):
        return value, _abi.TESS_VTK_SCALAR
    width = int(np.prod(shape[1:]))
    if width == dim:
        return value, _abi.TESS_VTK_VECTOR
    if width == dim * dim:
        return value, _abi.TESS_VTK_TENSOR
    raise ArgumentError(f"cannot write a field of shape {shape} for {n} points in {dim}D")


class VTKFile:
    """One dataset being written: ``vtk[name] = field`` adds point data, ``close()`` writes it."""

    def __init__(self, path: str, geometry, compress: bool = True, ctx=None):
        self.path, self.compress, self.ctx = str(path), bool(compress), ctx
        self.arrays: list[tuple[bytes, np.ndarray]] = []
        if isinstance(geometry, CartesianMesh):
            self.mesh, self.points = geometry, None
            self.dim, self.n = geometry.dim, geometry.nnodes()
            self.filename = self.path + ".vts"
        else:
            self.mesh = None
            self.dim = int(geometry.shape[1])
            self.points = export_field(geometry, self.dim, _abi.TESS_VTK_POINTS, ctx)
            self.n = self.points.shape[0]
            self.filename = self.path + ".vtu"

    def __setitem__(self, name: str, value) -> None:
        data, kind = _kind_of(value, self.n, self.dim)
        if data.shape[0] != self.n:
            raise ArgumentError(f"field {name!r} has {data.shape[0]} entries, the dataset {self.n} points")
        self.arrays.append((name.encode("utf-8"), export_field(data, self.dim, kind, self.ctx)))

    def close(self) -> list[str]:
        lib = _abi.load_library()
        arr = (_abi.tess_vtk_array * max(len(self.arrays), 1))()
        for k, (name, a) in enumerate(self.arrays):
            arr[k].name, arr[k].data, arr[k].ncomp = name, a.ctypes.data, a.shape[1]
        path = self.path.encode("utf-8")
        if self.mesh is not None:
            st = lib.tess_vtk_write_mesh(path, ctypes.byref(self.mesh.abi()), len(self.arrays), arr,
                                         int(self.compress))
        else:
            st = lib.tess_vtk_write_points(path, self.n, self.points.ctypes.data, len(self.arrays), arr,
                                           int(self.compress))
        _check(lib, st)
        return [self.filename]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class VTMFile:
    """A multiblock file (.vtm): datasets opened on it become its blocks (``{path}_{k}``)."""

    def __init__(self, path: str):
        self.path = str(path)
        self.filename = self.path + ".vtm"
        self.blocks: list[str] = []

    def _next_block(self) -> str:
        return f"{self.path}_{len(self.blocks) + 1}"

    def close(self) -> list[str]:
        lib = _abi.load_library()
        base = os.path.dirname(os.path.abspath(self.filename))
        rel = [os.path.relpath(os.path.abspath(b), base).encode("utf-8") for b in self.blocks]
        files = (ctypes.c_char_p * max(len(rel), 1))(*rel)
        _check(lib, lib.tess_vtm_write(self.path.encode("utf-8"), len(rel), files))
        return [self.filename] + list(self.blocks)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class PVDFile:
    """A ParaView collection (.pvd): ``pvd[t] = dataset`` adds a time step; ``append`` keeps the entries of
    an existing file (WriteVTK's paraview_collection(...; append=true))."""

    def __init__(self, path: str, append: bool = False):
        self.path = str(path)
        self.filename = self.path + ".pvd"
        self.entries: list[tuple[float, str]] = []
        if append and os.path.exists(self.filename):
            for ds in ET.parse(self.filename).getroot().iter("DataSet"):
                self.entries.append((float(ds.get("timestep")), ds.get("file")))

    def __setitem__(self, t: float, dataset) -> None:
        fn = dataset.filename if hasattr(dataset, "filename") else str(dataset)
        base = os.path.dirname(os.path.abspath(self.filename))
        self.entries.append((float(t), os.path.relpath(os.path.abspath(fn), base)))

    def close(self) -> list[str]:
        lib = _abi.load_library()
        n = len(self.entries)
        times = (ctypes.c_double * max(n, 1))(*[t for t, _ in self.entries])
        files = (ctypes.c_char_p * max(n, 1))(*[f.encode("utf-8") for _, f in self.entries])
        _check(lib, lib.tess_pvd_write(self.path.encode("utf-8"), n, times, files))
        return [self.filename]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def openvtk(target, geometry, compress: bool = True, ctx=None) -> VTKFile:
    """openvtk(vtk, x) / openvtk(vtk, mesh) (export.jl:89-99); target is a path (no extension) or a VTMFile."""
    if isinstance(target, VTMFile):
        path = target._next_block()
        vtk = VTKFile(path, geometry, compress, ctx)
        target.blocks.append(vtk.filename)
        return vtk
    return VTKFile(target, geometry, compress, ctx)


def openvtm(path: str) -> VTMFile:
    """openvtm(path) (export.jl:101)."""
    return VTMFile(path)


def openpvd(path: str, append: bool = False) -> PVDFile:
    """openpvd(path; append) (export.jl:102)."""
    return PVDFile(path, append=append)


def write_ply(name: str, xs, ctx=None) -> str:
    """write_ply(name, xs) (export.jl:115-132): name.ply with float x, y, z per vertex."""
    dim = int(xs.shape[1])
    pts = export_field(xs, dim, _abi.TESS_VTK_POINTS, ctx)
    lib = _abi.load_library()
    _check(lib, lib.tess_ply_write(str(name).encode("utf-8"), pts.shape[0], pts.ctypes.data))
    return str(name) + ".ply"
