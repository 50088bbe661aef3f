"""Output formats of src/IO/export.jl through the library's host writers.

What pins them: every file is read back by an independent reader of the VTK
XML format written here (appended raw data with UInt64 headers, optionally
vtkZLibDataCompressor blocks) and the decoded arrays must equal what was
written; the component layouts follow export.jl:73-85 (a Vec's components;
points padded to 3 with resize(x, 3); a full tensor column-major; a symmetric
tensor as xx, yy, zz, xy, yz, zx with NaN where 2D has none); the mesh points
are the node coordinates in column-major order (export.jl:53-55); the PLY file
is the byte string export.jl:115-132 writes (header lines, then Float32 x, y, z
per vertex, little-endian).  WriteVTK.jl 1.18 itself is not vendored, so its
byte-for-byte output is parity unpinned: the files are pinned by their decoded
content and structure.
"""
import os
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import pytest

import tesserae_hip as th
from tesserae_hip import io as tio

DT = {"Float32": np.float32, "Int64": np.int64, "UInt8": np.uint8}


def read_vtk(path):
    """(root element, {array name: ndarray (n, ncomp)}) of a VTK XML file with appended raw data."""
    raw = open(path, "rb").read()
    marker = b'<AppendedData encoding="raw">\n_'
    k = raw.index(marker)
    head = raw[:k].decode("utf-8")
    data = raw[k + len(marker):]
    root = ET.fromstring(head + "<AppendedData/></VTKFile>")
    compressed = root.get("compressor") == "vtkZLibDataCompressor"
    assert root.get("header_type") == "UInt64" and root.get("byte_order") == "LittleEndian"
    arrays = {}
    for el in root.iter("DataArray"):
        off = int(el.get("offset"))
        if compressed:
            nb, bs, last = struct.unpack_from("<3Q", data, off)
            sizes = struct.unpack_from(f"<{nb}Q", data, off + 24)
            pos = off + 24 + 8 * nb
            chunks = []
            for s in sizes:
                chunks.append(zlib.decompress(data[pos:pos + s]))
                pos += s
            blob = b"".join(chunks)
            assert nb == 0 or (len(blob) == bs * (nb - 1) + last)
        else:
            (nbytes,) = struct.unpack_from("<Q", data, off)
            blob = data[off + 8: off + 8 + nbytes]
        a = np.frombuffer(blob, dtype=DT[el.get("type")])
        arrays[el.get("Name")] = a.reshape(-1, int(el.get("NumberOfComponents", "1")))
    return root, arrays


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("compress", [True, False])
def test_points_round_trip(tmp_path, dim, compress):
    rng = np.random.default_rng(dim)
    n = 37
    x = rng.standard_normal((n, dim))
    v = rng.standard_normal((n, dim))
    V = rng.uniform(size=n)
    F = rng.standard_normal((n, dim * dim))
    S = F + F[:, [0, 2, 1, 3] if dim == 2 else [0, 3, 6, 1, 4, 7, 2, 5, 8]]  # symmetric, full storage
    base = str(tmp_path / "pcls")
    with tio.openvtk(base, x, compress=compress) as vtk:
        vtk["velocity"] = v
        vtk["volume"] = V
        vtk["F"] = F
        vtk["stress"] = tio.Symmetric(S)
    root, arr = read_vtk(base + ".vtu")
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == n and int(piece.get("NumberOfCells")) == n
    pts = np.zeros((n, 3), np.float32)
    pts[:, :dim] = x
    np.testing.assert_array_equal(arr["Points"], pts)
    np.testing.assert_array_equal(arr["connectivity"].ravel(), np.arange(n))
    np.testing.assert_array_equal(arr["offsets"].ravel(), np.arange(1, n + 1))
    assert np.all(arr["types"] == 1)  # VTK_VERTEX
    np.testing.assert_array_equal(arr["velocity"], v.astype(np.float32))
    np.testing.assert_array_equal(arr["volume"].ravel(), V.astype(np.float32))
    np.testing.assert_array_equal(arr["F"], F.astype(np.float32))  # column-major components
    sym = arr["stress"]
    assert sym.shape == (n, 6)
    Sm = S.reshape(n, dim, dim).transpose(0, 2, 1)  # Sm[e, r, c] = component r + dim*c
    np.testing.assert_array_equal(sym[:, 0], Sm[:, 0, 0].astype(np.float32))
    np.testing.assert_array_equal(sym[:, 1], Sm[:, 1, 1].astype(np.float32))
    np.testing.assert_array_equal(sym[:, 3], Sm[:, 0, 1].astype(np.float32))
    if dim == 3:
        np.testing.assert_array_equal(sym[:, 2], Sm[:, 2, 2].astype(np.float32))
        np.testing.assert_array_equal(sym[:, 4], Sm[:, 1, 2].astype(np.float32))
        np.testing.assert_array_equal(sym[:, 5], Sm[:, 2, 0].astype(np.float32))
    else:
        assert np.all(np.isnan(sym[:, [2, 4, 5]]))  # export.jl:79-81


@pytest.mark.parametrize("dim", [2, 3])
def test_mesh_round_trip(tmp_path, dim):
    mesh = th.CartesianMesh(0.25, *[(-0.5, 1.0)] * dim)
    nn = mesh.nnodes()
    m = np.arange(nn, dtype=np.float64)
    v = np.arange(nn * dim, dtype=np.float64).reshape(nn, dim)
    base = str(tmp_path / "grid")
    with tio.openvtk(base, mesh) as vtk:
        vtk["m"] = m
        vtk["velocity"] = v
    root, arr = read_vtk(base + ".vts")
    dims = list(mesh.size()) + [1] * (3 - dim)
    ext = " ".join(f"0 {d - 1}" for d in dims)
    assert root.find("StructuredGrid").get("WholeExtent") == ext
    # node coordinates x-fastest (column-major)
    axes = [mesh.axis(k) for k in range(dim)] + [np.zeros(1)] * (3 - dim)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([a.ravel(order="F") for a in (X, Y, Z)], axis=1).astype(np.float32)
    np.testing.assert_array_equal(arr["Points"], pts)
    np.testing.assert_array_equal(arr["m"].ravel(), m.astype(np.float32))
    np.testing.assert_array_equal(arr["velocity"], v.astype(np.float32))


def test_multiblock_and_collection(tmp_path):
    x = np.random.default_rng(1).standard_normal((5, 3))
    mesh = th.CartesianMesh(0.5, (0, 1), (0, 1), (0, 1))
    pvd_base = str(tmp_path / "paraview")
    tio.openpvd(pvd_base).close()  # create, as the reference's closepvd(openpvd(file))
    for step, t in ((1, 0.5), (2, 1.25)):
        with tio.openpvd(pvd_base, append=True) as pvd:
            with tio.openvtm(f"{pvd_base}{step}") as vtm:
                with tio.openvtk(vtm, x) as vtk:
                    vtk["velocity"] = x
                with tio.openvtk(vtm, mesh) as vtk:
                    vtk["m"] = np.ones(mesh.nnodes())
            pvd[t] = vtm
    root = ET.parse(pvd_base + ".pvd").getroot()
    assert root.get("type") == "Collection"
    ds = root.findall("Collection/DataSet")
    assert [(float(d.get("timestep")), d.get("file")) for d in ds] == [(0.5, "paraview1.vtm"), (1.25, "paraview2.vtm")]
    vtm = ET.parse(str(tmp_path / "paraview2.vtm")).getroot()
    assert vtm.get("type") == "vtkMultiBlockDataSet"
    files = [d.get("file") for d in vtm.findall("vtkMultiBlockDataSet/DataSet")]
    assert files == ["paraview2_1.vtu", "paraview2_2.vts"]
    for f in files:
        assert os.path.exists(tmp_path / f)


@pytest.mark.parametrize("dim", [2, 3])
def test_ply_bytes(tmp_path, dim):
    x = np.random.default_rng(4).standard_normal((9, dim))
    path = tio.write_ply(str(tmp_path / "cloud"), x)
    pts = np.zeros((9, 3), np.float32)
    pts[:, :dim] = x
    want = (b"ply\nformat binary_little_endian 1.0\nelement vertex 9\nproperty float x\nproperty float y\n"
            b"property float z\nend_header\n" + pts.astype("<f4").tobytes())
    assert open(path, "rb").read() == want


def test_empty_and_errors(tmp_path):
    base = str(tmp_path / "empty")
    tio.openvtk(base, np.zeros((0, 3))).close()
    _, arr = read_vtk(base + ".vtu")
    assert arr["Points"].shape == (0, 3)
    vtk = tio.openvtk(str(tmp_path / "bad"), np.zeros((4, 3)))
    with pytest.raises(th.ArgumentError):
        vtk["odd"] = np.zeros((4, 5))
    with pytest.raises(th.ArgumentError):
        vtk["short"] = np.zeros(3)
    lib = th.load_library()
    assert lib.tess_vtm_write(None, 0, None) == th._abi.TESS_ERR_ARGUMENT
    with pytest.raises(th.ErrorException):
        tio.write_ply(str(tmp_path / "no" / "such" / "dir"), np.zeros((1, 3)))
