"""tesserae_hip — MI355X (gfx950) engine for Tesserae.jl's explicit MPM step.

Host-side mirror of the reference API over the C ABI of libtesserae_hip
(include/tesserae_hip.h).  Importing this package does not touch the GPU; the
shared library is loaded on first use and its absence is an error.
"""
from .core import (Adjoint, ArgumentError, Batch, Basis, BasisWeights, BSpline, CartesianMesh, Context, Cubic, DeviceError,
                   DimensionMismatch, ErrorException, G2P, G2P2G, Grid, KernelCorrection, Linear, MPMParams, MultiLinear,
                   P2G, Particles, Partition, Polynomial, WLS,
                   Quadratic, SpArray, SpGrid, SpIndices, TesseraeAssertionError, TesseraeError,
                   UnsupportedError, default_context, g2p2g_steps, generate_basis_weights, generate_grid, generate_particles,
                   grid_sampling_points, reorder_particles, sparse_storage_index, step, uGIMP, update,
                   update_sparsity, step_adjoint, rollout, rollout_adjoint)
from ._abi import LIB_PATH, load_library

__all__ = [
    "KernelCorrection", "MultiLinear", "Polynomial", "WLS",
    "Adjoint", "Batch", "step_adjoint", "rollout", "rollout_adjoint", "ArgumentError", "Basis", "BasisWeights", "BSpline", "CartesianMesh", "Context", "Cubic", "DeviceError",
    "DimensionMismatch", "ErrorException", "G2P", "G2P2G", "Grid", "Linear", "MPMParams", "P2G", "Particles", "Partition",
    "Quadratic", "SpArray", "SpGrid", "SpIndices", "sparse_storage_index", "update_sparsity",
    "TesseraeAssertionError", "TesseraeError", "UnsupportedError", "default_context",
    "g2p2g_steps", "generate_basis_weights", "generate_grid", "generate_particles", "grid_sampling_points", "reorder_particles",
    "step", "uGIMP", "update", "LIB_PATH", "load_library",
]
