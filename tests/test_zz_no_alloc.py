"""tess_partition_update_noalloc: the partition over the library's fixed device work
area, called through the C ABI, equals tess_partition_update bit for bit (and the
oracle's partition), on two calls in a row with the second one smaller; sizes
beyond its maximum are refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

import tesserae_hip as th
from tesserae_hip import _abi
from tesserae_hip.core import _field


def _noalloc(lib, mesh, x, n, stream):
    nb = int(np.prod(mesh.nblocks()))
    dev = x.device
    # outputs hold stale values: the call must write every word it reports
    pidx = torch.full((max(n, 1),), -7, dtype=torch.int32, device=dev)
    starts = torch.full((nb,), -7, dtype=torch.int32, device=dev)
    stops = torch.full((nb,), -7, dtype=torch.int32, device=dev)
    na = torch.full((1,), -7, dtype=torch.int64, device=dev)
    st = lib.tess_partition_update_noalloc(ctypes.byref(mesh.abi()), _field(x, mesh.dim, "julia"), n,
                                           ctypes.c_void_p(pidx.data_ptr()), ctypes.c_void_p(starts.data_ptr()),
                                           ctypes.c_void_p(stops.data_ptr()), ctypes.c_void_p(na.data_ptr()),
                                           ctypes.c_void_p(stream))
    return st, pidx[:n], starts, stops, na


@pytest.mark.gpu
@pytest.mark.parametrize("dim,L,dtype", [(3, 2, torch.float32), (2, 3, torch.float64)])
def test_noalloc_partition_equals_partition_update(oracle, dim, L, dtype):
    lib = _abi.load_library()
    stream = torch.cuda.current_stream().cuda_stream
    npdt = np.float32 if dtype == torch.float32 else np.float64
    mesh = th.CartesianMesh(0.01 if dim == 3 else 0.004, *[(0.0, 1.0)] * dim, dtype=dtype, block_size_log2=L)
    rng = np.random.default_rng(7 + dim)
    for n in (900_000, 70_001):  # two calls in a row, the second smaller
        print(f"[noalloc partition] dim={dim} L={L} dtype={dtype} n={n}", flush=True)
        x = np.concatenate([rng.uniform(0.1, 0.9, size=(n - 1000, dim)), rng.uniform(-0.2, 1.2, size=(1000, dim))])
        xt = torch.from_numpy(x.astype(npdt)).cuda()
        st, pidx, starts, stops, na = _noalloc(lib, mesh, xt, n, stream)
        assert st == _abi.TESS_OK, lib.tess_last_error()
        torch.cuda.synchronize()
        part = th.update(th.Partition(mesh), xt)
        nassigned = int(na.item())
        assert nassigned == part.nassigned
        np.testing.assert_array_equal(pidx[:nassigned].cpu().numpy(), part.particleindices[:nassigned].cpu().numpy())
        np.testing.assert_array_equal(starts.cpu().numpy(), part.starts.cpu().numpy())
        np.testing.assert_array_equal(stops.cpu().numpy(), part.stops.cpu().numpy())
        if n < 100_000:
            ref_pidx, ref_starts, ref_stops, ref_na = oracle.partition(mesh, x.astype(npdt), dtype=npdt)
            assert nassigned == ref_na
            np.testing.assert_array_equal(pidx[:nassigned].cpu().numpy(), ref_pidx[:nassigned])
            np.testing.assert_array_equal(starts.cpu().numpy(), ref_starts)
            np.testing.assert_array_equal(stops.cpu().numpy(), ref_stops)


@pytest.mark.gpu
def test_noalloc_partition_empty_input():
    lib = _abi.load_library()
    mesh = th.CartesianMesh(0.1, (0.0, 1.0), (0.0, 1.0), dtype=torch.float32)
    x = torch.zeros((0, 2), dtype=torch.float32, device="cuda")
    st, pidx, starts, stops, na = _noalloc(lib, mesh, x, 0, torch.cuda.current_stream().cuda_stream)
    assert st == _abi.TESS_OK
    torch.cuda.synchronize()
    assert int(na.item()) == 0 and not starts.any() and not stops.any()


def test_noalloc_partition_rejects_beyond_maximum():
    """Refused on the host before any device work (runs without a GPU)."""
    lib = _abi.load_library()
    mesh = th.CartesianMesh(0.1, (0.0, 1.0), (0.0, 1.0), dtype=torch.float32)
    f = _abi.tess_field()
    f.ptr, f.elem_stride, f.comp_stride = 1, 2, 1  # never dereferenced
    dummy = ctypes.c_void_p(1)
    st = lib.tess_partition_update_noalloc(ctypes.byref(mesh.abi()), f, _abi.TESS_PARTITION_NOALLOC_MAX + 1,
                                           dummy, dummy, dummy, dummy, None)
    assert st == _abi.TESS_ERR_ARGUMENT
    assert "at most" in lib.tess_last_error().decode()
