"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/tesserae_hip.h declares; the ctypes mirror matches the header
layout; host-side argument checks raise the reference's exception classes."""
import ctypes
import os
import re

import pytest

from tesserae_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tesserae_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:tess_status|const char\*|int32_t)\s+(tess_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert set(names) == set(_abi.SIGNATURES), names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_abi.LIB_PATH), "build libtesserae_hip.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(_abi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_library_loads_and_reports_abi_version():
    lib = _abi.load_library()
    assert lib.tess_abi_version() == _abi.TESS_ABI_VERSION == 4
    assert lib.tess_last_error() == b""


def test_struct_layouts_match_header():
    # offsets follow the C declarations (natural alignment, 64-bit host)
    assert ctypes.sizeof(_abi.tess_mesh) == 4 * 4 + 3 * 8 + 3 * 8 + 2 * 8
    assert ctypes.sizeof(_abi.tess_field) == 24
    assert ctypes.sizeof(_abi.tess_particles) == 8 + 8 * 24
    assert ctypes.sizeof(_abi.tess_grid) == 5 * 24
    assert ctypes.sizeof(_abi.tess_weights) == 2 * 24 + 8 + 8
    assert _abi.tess_weights.filter.offset == 56
    assert ctypes.sizeof(_abi.tess_params) == 5 * 8 + 3 * 8 + 8
    assert ctypes.sizeof(_abi.tess_sparsity) == 3 * 8
    assert _abi.tess_mesh.dims.offset == 16 and _abi.tess_mesh.h.offset == 64


def test_host_checks_raise_reference_exception_classes():
    import tesserae_hip as th
    with pytest.raises(th.ArgumentError):
        th.CartesianMesh(0.1, (0, 1), (0, 1), block_size_log2=-1)
    with pytest.raises(th.TesseraeAssertionError):
        th.CartesianMesh(0.1, (1, 0), (0, 1))
    with pytest.raises(th.UnsupportedError):
        th.BSpline(5)


def test_mesh_covered_axis_expansion():
    import tesserae_hip as th
    m = th.CartesianMesh(0.3, (0, 1), (0, 1))  # 1/0.3 is not an integer: upper bound expands
    assert m.dims == (5, 5) and m.xmax()[0] == pytest.approx(1.2)
    assert th.CartesianMesh(1 / 256, (0, 1), (0, 1), (0, 1)).dims == (257, 257, 257)


def test_corrected_basis_codes():
    """WLS / KernelCorrection codes (include/tesserae_hip.h) and the reference's constructor checks
    (src/Basis/polynomials.jl:9-21, wls.jl:14-17, kernelcorrection.jl:20-22)."""
    import tesserae_hip as th
    Q = th.BSpline(th.Quadratic())
    kc = th.KernelCorrection(Q)
    assert kc.kind == _abi.TESS_BSPLINE_QUADRATIC | _abi.TESS_KERNEL_CORRECTION | _abi.TESS_POLY_MULTILINEAR
    assert kc.support_width == 3 and kc.kernel == _abi.TESS_BSPLINE_QUADRATIC and kc.corrected
    assert th.WLS(Q).kind == _abi.TESS_BSPLINE_QUADRATIC | _abi.TESS_WLS
    assert th.WLS(th.uGIMP(), th.Polynomial(th.MultiLinear())).kind == \
        _abi.TESS_UGIMP | _abi.TESS_WLS | _abi.TESS_POLY_MULTILINEAR
    assert th.KernelCorrection(th.BSpline(th.Cubic()), th.Polynomial(th.Linear())).kind == \
        _abi.TESS_BSPLINE_CUBIC | _abi.TESS_KERNEL_CORRECTION
    for poly in (th.Polynomial(th.Quadratic()), th.Polynomial(th.Cubic())):
        with pytest.raises(th.ArgumentError, match="Polynomial"):
            th.WLS(Q, poly)
        with pytest.raises(th.ArgumentError, match="Polynomial"):
            th.KernelCorrection(Q, poly)
    with pytest.raises(th.ArgumentError):
        th.WLS(kc)
