// Output of particle and grid fields in the formats src/IO/export.jl writes:
// VTK XML datasets through WriteVTK.jl 1.18 (openvtk / openvtm / openpvd,
// export.jl:47-113) and binary PLY point clouds (write_ply, export.jl:115-132).
// The device half packs a field into the Float32 component layout of
// export.jl's vtk_format (:73-85) in one pass and one device-to-host copy; the
// host half writes the VTK XML files (UnstructuredGrid of vertices,
// StructuredGrid of the mesh nodes, multiblock and collection files) with
// appended raw data, zlib-compressed as vtkZLibDataCompressor blocks with
// UInt64 headers, the layout of the published VTK XML file format.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tess_device.hpp"
#include "tess_internal.hpp"

namespace tess {
namespace {

int vtk_components(int kind, int dim) {
  switch (kind) {
    case TESS_VTK_SCALAR: return 1;
    case TESS_VTK_VECTOR: return dim;
    case TESS_VTK_POINTS: return 3;
    case TESS_VTK_TENSOR: return dim * dim;
    case TESS_VTK_SYMMETRIC: return 6;
    default: return 0;
  }
}

// Component j of element e in the layout `kind` (export.jl:73-85): a Vec's
// components; points padded with zeros to 3; a full tensor column-major
// (Tensorial's storage order); a symmetric tensor as (xx, yy, zz, xy, yz, zx),
// in 2D (xx, yy, NaN, xy, NaN, NaN).  Tensors are read from full dim×dim
// storage, element (a, b) at component a + dim*b.
template <typename T>
__global__ __launch_bounds__(256) void k_pack(DField<T> f, int64_t n, int dim, int kind, int nc,
                                              float* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * nc) return;
  const int64_t e = i / nc;
  const int j = int(i - e * nc);
  float v;
  if (kind == TESS_VTK_POINTS) {
    v = j < dim ? float(f(e, j)) : 0.0f;
  } else if (kind == TESS_VTK_SYMMETRIC) {
    // (row, col) of xx, yy, zz, xy, yz, zx
    const int r[6] = {0, 1, 2, 0, 1, 2}, c[6] = {0, 1, 2, 1, 2, 0};
    v = (r[j] < dim && c[j] < dim) ? float(f(e, r[j] + dim * c[j])) : __builtin_nanf("");
  } else {
    v = float(f(e, j));
  }
  out[i] = v;
}

// ---- VTK XML ------------------------------------------------------------------

// One array of the appended section: raw bytes [UInt64 size][data], or a single
// zlib block [UInt64 nblocks=1, block size, last block size, compressed size][data].
std::string encode_block(const void* data, size_t bytes, bool compress) {
  std::string out;
  auto put64 = [&](uint64_t v) { out.append(reinterpret_cast<const char*>(&v), 8); };
  if (!compress) {
    put64(bytes);
    out.append(static_cast<const char*>(data), bytes);
    return out;
  }
  if (bytes == 0) {
    put64(0);
    put64(0);
    put64(0);
    return out;
  }
  uLongf clen = compressBound(uLong(bytes));
  std::vector<Bytef> buf(clen);
  if (compress2(buf.data(), &clen, static_cast<const Bytef*>(data), uLong(bytes), Z_DEFAULT_COMPRESSION) != Z_OK)
    raise(TESS_ERR_STATE, "zlib compression failed");
  put64(1);
  put64(bytes);
  put64(bytes);
  put64(clen);
  out.append(reinterpret_cast<const char*>(buf.data()), clen);
  return out;
}

struct Appended {
  bool compress;
  std::string bytes;
  // the DataArray element of one appended array
  std::string array(const char* type, const std::string& name, int ncomp, const void* data, size_t nbytes) {
    std::string el = "<DataArray type=\"" + std::string(type) + "\" Name=\"" + name + "\"";
    if (ncomp > 1) el += " NumberOfComponents=\"" + std::to_string(ncomp) + "\"";
    el += " format=\"appended\" offset=\"" + std::to_string(bytes.size()) + "\"/>\n";
    bytes += encode_block(data, nbytes, compress);
    return el;
  }
};

std::string xml_escape(const std::string& s) {
  std::string o;
  for (char ch : s) {
    if (ch == '&') o += "&amp;";
    else if (ch == '<') o += "&lt;";
    else if (ch == '>') o += "&gt;";
    else if (ch == '"') o += "&quot;";
    else o += ch;
  }
  return o;
}

void write_file(const std::string& path, const std::string& head, const std::string& appended,
                const std::string& tail) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) raise(TESS_ERR_STATE, "cannot open " + path + " for writing");
  bool ok = std::fwrite(head.data(), 1, head.size(), fp) == head.size();
  ok = ok && std::fwrite(appended.data(), 1, appended.size(), fp) == appended.size();
  ok = ok && std::fwrite(tail.data(), 1, tail.size(), fp) == tail.size();
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) raise(TESS_ERR_STATE, "write to " + path + " failed");
}

std::string file_header(const char* type, bool compress) {
  std::string h = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<VTKFile type=\"" + std::string(type) +
                  "\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\"";
  if (compress) h += " compressor=\"vtkZLibDataCompressor\"";
  return h + ">\n";
}

std::string point_data(Appended& ap, int64_t n, int narrays, const tess_vtk_array* arrays) {
  if (narrays == 0) return "";
  std::string s = "<PointData>\n";
  for (int a = 0; a < narrays; ++a) {
    const tess_vtk_array& A = arrays[a];
    s += ap.array("Float32", xml_escape(A.name), A.ncomp, A.data, size_t(n) * size_t(A.ncomp) * sizeof(float));
  }
  return s + "</PointData>\n";
}

const char* kAppendedOpen = "<AppendedData encoding=\"raw\">\n_";
const char* kAppendedClose = "\n</AppendedData>\n</VTKFile>\n";

void check_arrays(int32_t narrays, const tess_vtk_array* arrays) {
  if (narrays < 0 || (narrays > 0 && !arrays)) raise(TESS_ERR_ARGUMENT, "point-data arrays are NULL");
  for (int a = 0; a < narrays; ++a) {
    if (!arrays[a].name || !arrays[a].name[0]) raise(TESS_ERR_ARGUMENT, "a point-data array has no name");
    if (arrays[a].ncomp < 1) raise(TESS_ERR_ARGUMENT, "a point-data array has no components");
  }
}

}  // namespace

void run_export_field(tess_ctx* ctx, int32_t dtype, int32_t dim, tess_field field, int64_t n, int32_t kind,
                      float* dst) {
  const int nc = vtk_components(kind, dim);
  if (nc == 0) raise(TESS_ERR_ARGUMENT, "unknown VTK component layout");
  if (n == 0) return;
  float* stage = ctx->export_buf.as<float>(size_t(n) * size_t(nc));
  const int64_t total = n * nc;
  const unsigned blocks = unsigned((total + 255) / 256);
  if (dtype == TESS_FLOAT32)
    hipLaunchKernelGGL((k_pack<float>), dim3(blocks), dim3(256), 0, ctx->stream, dfield<float>(field), n, dim, kind,
                       nc, stage);
  else
    hipLaunchKernelGGL((k_pack<double>), dim3(blocks), dim3(256), 0, ctx->stream, dfield<double>(field), n, dim,
                       kind, nc, stage);
  TESS_HIP(hipGetLastError());
  TESS_HIP(hipMemcpyAsync(dst, stage, size_t(total) * sizeof(float), hipMemcpyDefault, ctx->stream));
  TESS_HIP(hipStreamSynchronize(ctx->stream));
}

void run_vtk_write_points(const char* path, int64_t n, const float* xyz, int32_t narrays,
                          const tess_vtk_array* arrays, int32_t compress) {
  check_arrays(narrays, arrays);
  Appended ap{compress != 0, {}};
  const size_t np = static_cast<size_t>(n);
  std::vector<int64_t> conn(np), offs(np);
  std::vector<uint8_t> types(np, uint8_t(1));  // VTK_VERTEX
  for (int64_t i = 0; i < n; ++i) {
    conn[size_t(i)] = i;
    offs[size_t(i)] = i + 1;
  }
  std::string body = "<UnstructuredGrid>\n<Piece NumberOfPoints=\"" + std::to_string(n) + "\" NumberOfCells=\"" +
                     std::to_string(n) + "\">\n<Points>\n";
  body += ap.array("Float32", "Points", 3, xyz, size_t(n) * 3 * sizeof(float));
  body += "</Points>\n<Cells>\n";
  body += ap.array("Int64", "connectivity", 1, conn.data(), conn.size() * sizeof(int64_t));
  body += ap.array("Int64", "offsets", 1, offs.data(), offs.size() * sizeof(int64_t));
  body += ap.array("UInt8", "types", 1, types.data(), types.size());
  body += "</Cells>\n" + point_data(ap, n, narrays, arrays) + "</Piece>\n</UnstructuredGrid>\n";
  write_file(std::string(path) + ".vtu", file_header("UnstructuredGrid", ap.compress) + body + kAppendedOpen,
             ap.bytes, kAppendedClose);
}

void run_vtk_write_mesh(const char* path, const tess_mesh& mesh, int32_t narrays, const tess_vtk_array* arrays,
                        int32_t compress) {
  check_arrays(narrays, arrays);
  Appended ap{compress != 0, {}};
  const int D = mesh.dim;
  int64_t dims[3] = {1, 1, 1};
  for (int d = 0; d < D; ++d) dims[d] = mesh.dims[d];
  const int64_t n = dims[0] * dims[1] * dims[2];
  // node coordinates as the mesh holds them (T(xmin + i*h)), then Float32 (export.jl:53-55)
  std::vector<float> pts(size_t(n) * 3);
  for (int64_t k = 0; k < n; ++k) {
    int64_t r = k;
    for (int d = 0; d < 3; ++d) {
      const int64_t i = r % dims[d];
      r /= dims[d];
      double c = 0.0;
      if (d < D) {
        c = mesh.xmin[d] + double(i) * mesh.h;
        if (mesh.dtype == TESS_FLOAT32) c = double(float(c));
      }
      pts[size_t(k) * 3 + d] = float(c);
    }
  }
  std::string ext = "0 " + std::to_string(dims[0] - 1) + " 0 " + std::to_string(dims[1] - 1) + " 0 " +
                    std::to_string(dims[2] - 1);
  std::string body = "<StructuredGrid WholeExtent=\"" + ext + "\">\n<Piece Extent=\"" + ext + "\">\n<Points>\n";
  body += ap.array("Float32", "Points", 3, pts.data(), pts.size() * sizeof(float));
  body += "</Points>\n" + point_data(ap, n, narrays, arrays) + "</Piece>\n</StructuredGrid>\n";
  write_file(std::string(path) + ".vts", file_header("StructuredGrid", ap.compress) + body + kAppendedOpen,
             ap.bytes, kAppendedClose);
}

void run_vtm_write(const char* path, int32_t nblocks, const char* const* files) {
  if (nblocks < 0 || (nblocks > 0 && !files)) raise(TESS_ERR_ARGUMENT, "block files are NULL");
  std::string s = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<VTKFile type=\"vtkMultiBlockDataSet\" "
                  "version=\"1.0\" byte_order=\"LittleEndian\">\n<vtkMultiBlockDataSet>\n";
  for (int b = 0; b < nblocks; ++b) {
    if (!files[b]) raise(TESS_ERR_ARGUMENT, "a block file is NULL");
    s += "<DataSet index=\"" + std::to_string(b) + "\" file=\"" + xml_escape(files[b]) + "\"/>\n";
  }
  s += "</vtkMultiBlockDataSet>\n</VTKFile>\n";
  write_file(std::string(path) + ".vtm", s, "", "");
}

void run_pvd_write(const char* path, int32_t n, const double* times, const char* const* files) {
  if (n < 0 || (n > 0 && (!files || !times))) raise(TESS_ERR_ARGUMENT, "collection entries are NULL");
  std::string s = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<VTKFile type=\"Collection\" version=\"1.0\" "
                  "byte_order=\"LittleEndian\">\n<Collection>\n";
  for (int k = 0; k < n; ++k) {
    if (!files[k]) raise(TESS_ERR_ARGUMENT, "a collection file is NULL");
    char t[40];
    std::snprintf(t, sizeof(t), "%.17g", times[k]);
    s += "<DataSet timestep=\"" + std::string(t) + "\" part=\"0\" file=\"" + xml_escape(files[k]) + "\"/>\n";
  }
  s += "</Collection>\n</VTKFile>\n";
  write_file(std::string(path) + ".pvd", s, "", "");
}

// write_ply (export.jl:115-132): the header lines, then float x, y, z per vertex.
void run_ply_write(const char* name, int64_t n, const float* xyz) {
  if (n < 0 || (n > 0 && !xyz)) raise(TESS_ERR_ARGUMENT, "vertex coordinates are NULL");
  std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) +
                  "\nproperty float x\nproperty float y\nproperty float z\nend_header\n";
  const std::string body = n > 0 ? std::string(reinterpret_cast<const char*>(xyz), size_t(n) * 3 * sizeof(float))
                                 : std::string();
  write_file(std::string(name) + ".ply", h, body, "");
}

}  // namespace tess
