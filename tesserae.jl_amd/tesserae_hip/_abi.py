"""ctypes mirror of include/tesserae_hip.h and the loader of libtesserae_hip.so.

The library is the product: there is no fallback.  If the shared object is
missing, importing the binding raises with the build command to run.
"""
from __future__ import annotations

import ctypes
import os

TESS_BUILD_DEBUG_CHECKS = 1  # tess_build_flags bit of the debug build
TESS_ABI_VERSION = 4  # tess_debug_fail_after (v3: tess_step_sparse reports the steps it ran; v2: weight filter)

TESS_OK = 0
TESS_ERR_ARGUMENT = 1
TESS_ERR_DIMENSION = 2
TESS_ERR_ASSERTION = 3
TESS_ERR_STATE = 4
TESS_ERR_DEVICE = 5
TESS_ERR_UNSUPPORTED = 6

TESS_FLOAT32 = 0
TESS_FLOAT64 = 1

TESS_BSPLINE_LINEAR = 1
TESS_BSPLINE_QUADRATIC = 2
TESS_BSPLINE_CUBIC = 3
TESS_UGIMP = 16

TESS_VTK_SCALAR = 1
TESS_VTK_VECTOR = 2
TESS_VTK_POINTS = 3
TESS_VTK_TENSOR = 4
TESS_VTK_SYMMETRIC = 5

TESS_WLS = 0x100
TESS_KERNEL_CORRECTION = 0x200
TESS_POLY_MULTILINEAR = 0x400

TESS_PARTITION_NOALLOC_MAX = 1048576  # particles of tess_partition_update_noalloc

LIB_NAME = "libtesserae_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)
# the debug build (`make debug`): every binning read back and checked on the host
DEBUG_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtesserae_hip_debug.so")


class tess_mesh(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("block_size_log2", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 3),
        ("xmin", ctypes.c_double * 3),
        ("h", ctypes.c_double),
        ("h_inv", ctypes.c_double),
    ]


class tess_field(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("elem_stride", ctypes.c_int64),
        ("comp_stride", ctypes.c_int64),
    ]


class tess_particles(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("x", tess_field),
        ("v", tess_field),
        ("V", tess_field),
        ("grad_v", tess_field),
        ("F", tess_field),
        ("sigma", tess_field),
        ("l", tess_field),
        ("logE", tess_field),
    ]


class tess_grid(ctypes.Structure):
    _fields_ = [
        ("m", tess_field),
        ("mv", tess_field),
        ("f", tess_field),
        ("vn", tess_field),
        ("v", tess_field),
    ]


class tess_weights(ctypes.Structure):
    _fields_ = [
        ("w", tess_field),
        ("grad_w", tess_field),
        ("window", ctypes.c_void_p),
        ("filter", ctypes.c_void_p),
    ]


class tess_sparsity(ctypes.Structure):
    _fields_ = [
        ("blocknumbering", ctypes.c_void_p),
        ("storage_len", ctypes.c_int64),
        ("capacity", ctypes.c_int64),
    ]


class tess_adjoint(ctypes.Structure):
    _fields_ = [
        ("x", tess_field),
        ("v", tess_field),
        ("F", tess_field),
        ("logE", tess_field),
    ]


class tess_params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("mp", ctypes.c_double),
        ("V0", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("slip_walls", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class tess_scene(ctypes.Structure):
    _fields_ = [
        ("mesh", tess_mesh),
        ("grid", tess_grid),
        ("particles", tess_particles),
        ("params", tess_params),
    ]


# Every symbol include/tesserae_hip.h declares, with its ctypes signature.
class tess_vtk_array(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char_p),
        ("data", ctypes.c_void_p),
        ("ncomp", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_P = ctypes.POINTER
SIGNATURES = {
    "tess_export_field": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, tess_field, ctypes.c_int64,
                                           ctypes.c_int32, ctypes.c_void_p]),
    "tess_vtk_write_points": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32,
                                               _P(tess_vtk_array), ctypes.c_int32]),
    "tess_vtk_write_mesh": (ctypes.c_int32, [ctypes.c_char_p, _P(tess_mesh), ctypes.c_int32, _P(tess_vtk_array),
                                             ctypes.c_int32]),
    "tess_vtm_write": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32, _P(ctypes.c_char_p)]),
    "tess_pvd_write": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32, _P(ctypes.c_double), _P(ctypes.c_char_p)]),
    "tess_ply_write": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p]),
    "tess_abi_version": (ctypes.c_int32, []),
    "tess_last_error": (ctypes.c_char_p, []),
    "tess_debug_fail_after": (ctypes.c_int32, [ctypes.c_int64]),
    "tess_build_flags": (ctypes.c_int32, []),
    "tess_ctx_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_void_p, _P(ctypes.c_void_p)]),
    "tess_ctx_destroy": (ctypes.c_int32, [ctypes.c_void_p]),
    "tess_sync": (ctypes.c_int32, [ctypes.c_void_p]),
    "tess_update_weights": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_particles), _P(tess_weights)],
    ),
    "tess_partition_update": (
        ctypes.c_int32,
        [ctypes.c_void_p, _P(tess_mesh), tess_field, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p, _P(ctypes.c_int64)],
    ),
    "tess_partition_update_noalloc": (
        ctypes.c_int32,
        [_P(tess_mesh), tess_field, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p, ctypes.c_void_p],
    ),
    "tess_reorder_particles": (
        ctypes.c_int32,
        [ctypes.c_void_p, _P(tess_mesh), _P(tess_particles), ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
         _P(ctypes.c_int32)],
    ),
    "tess_p2g": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_weights),
         _P(tess_params)],
    ),
    "tess_g2p2g": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_weights),
         _P(tess_params)],
    ),
    "tess_g2p": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_weights),
         _P(tess_params)],
    ),
    "tess_g2p2g_steps": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_params),
         ctypes.c_int32, ctypes.c_double],
    ),
    "tess_step": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_params),
         ctypes.c_int32, ctypes.c_double],
    ),
    "tess_sparsity_update": (
        ctypes.c_int32,
        [ctypes.c_void_p, _P(tess_mesh), tess_field, ctypes.c_int64, _P(tess_sparsity), _P(ctypes.c_int32)],
    ),
    "tess_p2g_sparse": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_sparsity), _P(tess_particles),
         _P(tess_weights), _P(tess_params)],
    ),
    "tess_g2p_sparse": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_sparsity), _P(tess_particles),
         _P(tess_weights), _P(tess_params)],
    ),
    "tess_step_sparse": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_sparsity), _P(tess_particles),
         _P(tess_params), ctypes.c_int32, ctypes.c_double, _P(ctypes.c_int32)],
    ),
    "tess_step_adjoint": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_params),
         _P(tess_adjoint)],
    ),
    "tess_rollout": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_particles), _P(tess_params),
         ctypes.c_int32, ctypes.c_int32],
    ),
    "tess_rollout_adjoint": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, _P(tess_mesh), _P(tess_grid), _P(tess_params), _P(tess_adjoint)],
    ),
    "tess_batch_create": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _P(tess_scene), _P(ctypes.c_void_p)],
    ),
    "tess_batch_step": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "tess_batch_autotune": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, _P(ctypes.c_int32)]),
    "tess_batch_set_reorder": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double]),
    "tess_batch_set_variant": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "tess_batch_stats": (ctypes.c_int32, [ctypes.c_void_p, _P(ctypes.c_int64), ctypes.c_int32]),
    "tess_batch_destroy": (ctypes.c_int32, [ctypes.c_void_p]),
    "tess_profile_enable": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "tess_profile_read": (
        ctypes.c_int32,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, _P(ctypes.c_double), _P(ctypes.c_int64),
         _P(ctypes.c_double)],
    ),
    "tess_device_copy": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "tess_stats": (ctypes.c_int32, [ctypes.c_void_p, _P(ctypes.c_int64), ctypes.c_int32]),
}

_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libtesserae_hip.so (built in-tree by __graft_entry__.build()); the environment
    variable TESS_HIP_LIBRARY names another build of the same library to load instead."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("TESS_HIP_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{LIB_NAME} not found at {p}; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `make -C tesserae.jl_amd/csrc`)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.tess_abi_version() != TESS_ABI_VERSION:
        raise ImportError(f"{LIB_NAME} ABI version mismatch")
    if path is None:
        _lib = lib
    return lib
