"""GPU packing of fields for output (tess_export_field) against the host conversion of the same data, in
every component layout of export.jl:73-85, FP32 and FP64, Julia and SoA strides; then a whole openvtk of
device particles read back."""
import ctypes

import numpy as np
import pytest
import torch

import tesserae_hip as th
from tesserae_hip import _abi
from tesserae_hip import io as tio
from tesserae_hip.core import _field
from test_export import read_vtk

pytestmark = pytest.mark.gpu

KINDS = [(_abi.TESS_VTK_SCALAR, 1), (_abi.TESS_VTK_VECTOR, None), (_abi.TESS_VTK_POINTS, None),
         (_abi.TESS_VTK_TENSOR, None), (_abi.TESS_VTK_SYMMETRIC, None)]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("layout", ["julia", "soa"])
def test_pack_matches_host(dtype, dim, layout):
    print(f"[export pack] dim={dim} dtype={dtype} layout={layout}", flush=True)
    n = 1001
    g = torch.Generator().manual_seed(dim)
    for kind, width in KINDS:
        w = width or (dim if kind in (_abi.TESS_VTK_VECTOR, _abi.TESS_VTK_POINTS) else dim * dim)
        host = torch.randn(n, w, generator=g, dtype=torch.float64).to(dtype)
        want = tio.export_field(host.numpy(), dim, kind)
        dev = host.cuda() if layout == "julia" else host.t().contiguous().cuda()
        f = _field(dev, w, layout)
        if layout == "soa" and w == 1:
            f.elem_stride, f.comp_stride = 1, n
        ctx = th.default_context()
        out = np.empty_like(want)
        st = ctx.lib.tess_export_field(ctx.handle, _abi.TESS_FLOAT32 if dtype == torch.float32 else _abi.TESS_FLOAT64,
                                       dim, f, n, kind, out.ctypes.data)
        assert st == 0, ctx.lib.tess_last_error()
        np.testing.assert_array_equal(out, want)


def test_openvtk_from_device(tmp_path):
    n, dim = 2000, 3
    P = th.Particles(n, dim, torch.float64, device="cuda")
    P.set("x", np.random.default_rng(0).uniform(size=(n, dim)))
    P.set("v", np.random.default_rng(1).standard_normal((n, dim)))
    P.set("F", np.tile(np.eye(dim).ravel(order="F"), (n, 1)))
    base = str(tmp_path / "pcls")
    with tio.openvtk(base, P.x) as vtk:
        vtk["velocity"] = P.v
        vtk["F"] = P.F
        vtk["stress"] = tio.Symmetric(P.F)
    _, arr = read_vtk(base + ".vtu")
    np.testing.assert_array_equal(arr["Points"], P.get("x").cpu().numpy().astype(np.float32))
    np.testing.assert_array_equal(arr["velocity"], P.get("v").cpu().numpy().astype(np.float32))
    np.testing.assert_array_equal(arr["F"], P.get("F").cpu().numpy().astype(np.float32))
    np.testing.assert_array_equal(arr["stress"], np.tile(np.float32([1, 1, 1, 0, 0, 0]), (n, 1)))
