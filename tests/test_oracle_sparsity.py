"""SpGrid block activation and storage indices (src/Grids/sparsity.jl, spindices.jl).

The oracle's update_sparsity! is checked against a brute force written the way
test/sparray.jl:70-113 states it (occupied blocks from findblock, the ±1-block
neighbourhood, numbering = count of active blocks up to b), against the
reference's reuse cases (:114-136), and the host storage-index map against the
closed form of test/sparray.jl:17-46.
"""
import numpy as np
import pytest

import tesserae_hip as th
from tesserae_hip.core import grid_sampling_points


def brute_force(oracle, mesh, x):
    nb = mesh.nblocks()
    occupied = np.zeros(nb, bool)
    for p in x:
        I = oracle.findblock(mesh, p)
        if I is not None:
            occupied[tuple(i - 1 for i in I)] = True
    active = np.zeros(nb, bool)
    for I in zip(*np.nonzero(occupied)):
        sl = tuple(slice(max(i - 1, 0), min(i + 2, n)) for i, n in zip(I, nb))
        active[sl] = True
    flat = active.reshape(-1, order="F")
    numbering = np.where(flat, np.cumsum(flat), 0).astype(np.int32)
    return numbering, int(flat.sum()) << (mesh.block_size_log2 * mesh.dim)


def test_circle_example_matches_brute_force(oracle):
    # test/sparray.jl:70-113: mesh 0.25 on (0,4)^2, GridSampling particles in the unit circle at (2,2)
    mesh = th.CartesianMesh(0.25, (0, 4), (0, 4))
    x = grid_sampling_points(mesh)
    x = x[((x[:, 0] - 2) ** 2 + (x[:, 1] - 2) ** 2) < 1]
    num, n, changed = oracle.sparsity(mesh, x)
    ref_num, ref_n = brute_force(oracle, mesh, x)
    np.testing.assert_array_equal(num, ref_num)
    assert n == ref_n and changed


@pytest.mark.parametrize("dim,L", [(2, 2), (2, 3), (3, 2), (3, 3)])
def test_random_particles_match_brute_force(oracle, dim, L):
    mesh = th.CartesianMesh(0.05, *[(0, 1.3)] * dim, block_size_log2=L)
    rng = np.random.default_rng(dim * 10 + L)
    x = np.concatenate([rng.uniform(0.1, 0.4, size=(60, dim)), rng.uniform(-0.2, 1.5, size=(40, dim))])
    num, n, _ = oracle.sparsity(mesh, x)
    ref_num, ref_n = brute_force(oracle, mesh, x)
    np.testing.assert_array_equal(num, ref_num)
    assert n == ref_n


def test_reuse_when_active_set_is_unchanged(oracle):
    # test/sparray.jl:114-136
    mesh = th.CartesianMesh(1.0, (0, 16), (0, 4))
    x = np.array([[1.0, 1.0], [2.0, 1.0], [9.0, 1.0]])
    num1, n1, ch1 = oracle.sparsity(mesh, x)
    assert ch1 and n1 > 0
    x[0] = [10.0, 1.0]  # moves into an already occupied block
    num2, n2, ch2 = oracle.sparsity(mesh, x, numbering=num1)
    assert not ch2 and n2 == n1
    np.testing.assert_array_equal(num2, num1)
    num3, n3, ch3 = oracle.sparsity(mesh, np.zeros((0, 2)), numbering=num1)
    assert ch3 and n3 == 0 and not num3.any()


def test_storage_index_closed_form():
    # test/sparray.jl:17-46: index = (count(blkspy[1:b]) - 1) * blocklength + localindex
    mesh = th.CartesianMesh(1.0, (0, 8), (0, 16))  # 9x17 nodes, L=2 -> 3x5 blocks
    rng = np.random.default_rng(5)
    spy = rng.random(mesh.nblocks()) < 0.5
    flat = spy.reshape(-1, order="F")
    numbering = np.where(flat, np.cumsum(flat), 0).reshape(mesh.nblocks(), order="F")
    idx = th.sparse_storage_index(mesh, numbering)
    L, B = mesh.block_size_log2, 1 << mesh.block_size_log2
    for lin in range(mesh.nnodes()):
        i, j = lin % mesh.dims[0], lin // mesh.dims[0]
        b = (i >> L, j >> L)
        if not spy[b]:
            assert idx[lin] == -1
            continue
        nblks = int(flat[: b[0] + b[1] * mesh.nblocks()[0] + 1].sum())
        local = (i & (B - 1)) + (j & (B - 1)) * B + 1  # 1-based column-major
        assert idx[lin] + 1 == (nblks - 1) * B * B + local
