// VTK legacy ASCII writers (host code) for the outputs of the isosurface and
// topology paths: the format of the reference benchmarks' PrintContentsToStream
// (Benchmarks/MarchingCubes/Pipeline.h:51-98, POLYDATA / POLYGONS;
// Benchmarks/Threshold/Pipeline.h:99-150, UNSTRUCTURED_GRID / CELLS /
// CELL_TYPES).  Numbers are written as an ostream does by default: floats
// with 6 significant digits in the shortest of fixed and scientific form
// (printf's %g), ids in decimal.  Points go three to a line, each followed by
// a space.  The text is formatted in parallel over fixed-size slices of the
// points and cells (std::to_chars into per-slice buffers) and written in
// order.
#include <charconv>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace daxhip {
namespace vtk {

constexpr int64_t kSlice = 1 << 16;  // points or cells formatted per task

inline void put_float(std::string& out, float v) {
  char b[32];
  const auto r = std::to_chars(b, b + sizeof(b), v, std::chars_format::general, 6);
  out.append(b, r.ptr);
}

inline void put_int(std::string& out, int64_t v) {
  char b[24];
  const auto r = std::to_chars(b, b + sizeof(b), v);
  out.append(b, r.ptr);
}

// Formats n items in slices with `fmt(begin, end, out)` on up to
// hardware_concurrency threads; returns the slices' texts in order.
template <class F>
std::vector<std::string> format_slices(int64_t n, F fmt) {
  const int64_t slices = (n + kSlice - 1) / kSlice;
  std::vector<std::string> parts((size_t)slices);
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if ((int64_t)nt > slices) nt = (unsigned)(slices > 0 ? slices : 1);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (int64_t s = t; s < slices; s += nt)
        fmt(s * kSlice, std::min(n, (s + 1) * kSlice), parts[(size_t)s]);
    });
  for (auto& th : pool) th.join();
  return parts;
}

int write_points(FILE* fp, const float* pts, int64_t n) {
  auto parts = format_slices(n, [&](int64_t b, int64_t e, std::string& out) {
    out.reserve((size_t)(e - b) * 30);
    for (int64_t i = b; i < e; ++i) {
      for (int d = 0; d < 3; ++d) {
        put_float(out, pts[3 * i + d]);
        out.push_back(' ');
      }
      if (i % 3 == 2) out.push_back('\n');
    }
  });
  for (auto& p : parts)
    if (fwrite(p.data(), 1, p.size(), fp) != p.size()) return 1;
  if (n % 3 != 0 && fputc('\n', fp) == EOF) return 1;
  return 0;
}

int write_cells(FILE* fp, const int32_t* conn, int64_t n, int nv) {
  auto parts = format_slices(n, [&](int64_t b, int64_t e, std::string& out) {
    out.reserve((size_t)(e - b) * (size_t)(nv + 1) * 8);
    for (int64_t c = b; c < e; ++c) {
      put_int(out, nv);
      for (int v = 0; v < nv; ++v) {
        out.push_back(' ');
        put_int(out, conn[c * nv + v]);
      }
      out.push_back('\n');
    }
  });
  for (auto& p : parts)
    if (fwrite(p.data(), 1, p.size(), fp) != p.size()) return 1;
  return 0;
}

}  // namespace vtk
}  // namespace daxhip

using namespace daxhip;

namespace {

int check_args(const char* fn, const char* path, const float* pts, int64_t n_points,
               const int32_t* conn, int64_t n_cells, int nv) {
  if (!path || n_points < 0 || n_cells < 0 || nv < 1 || nv > 8 ||
      (n_points > 0 && !pts) || (n_cells > 0 && !conn)) {
    set_error("%s: bad arguments", fn);
    return DAX_HIP_BAD_VALUE;
  }
  for (int64_t i = 0; i < n_cells * nv; ++i)
    if (conn[i] < 0 || conn[i] >= n_points) {
      set_error("%s: cell %lld references point %d outside [0, %lld)", fn, (long long)(i / nv),
                conn[i], (long long)n_points);
      return DAX_HIP_BAD_VALUE;
    }
  return DAX_HIP_OK;
}

struct File {
  FILE* fp = nullptr;
  ~File() {
    if (fp) fclose(fp);
  }
};

}  // namespace

extern "C" {

int dax_hip_write_vtk_polydata(const char* path, const float* h_points, int64_t n_points,
                               const int32_t* h_conn, int64_t n_cells, int verts_per_cell) {
  int rc = check_args("dax_hip_write_vtk_polydata", path, h_points, n_points, h_conn, n_cells,
                      verts_per_cell);
  if (rc) return rc;
  File f;
  f.fp = fopen(path, "wb");
  if (!f.fp) {
    set_error("dax_hip_write_vtk_polydata: cannot open %s", path);
    return DAX_HIP_BAD_VALUE;
  }
  const std::string head = "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET POLYDATA\n"
                           "POINTS " + std::to_string(n_points) + " float\n";
  const std::string cells = "POLYGONS " + std::to_string(n_cells) + " " +
                            std::to_string(n_cells * (verts_per_cell + 1)) + "\n";
  if (fputs(head.c_str(), f.fp) == EOF || vtk::write_points(f.fp, h_points, n_points) ||
      fputs(cells.c_str(), f.fp) == EOF ||
      vtk::write_cells(f.fp, h_conn, n_cells, verts_per_cell) || fputc('\n', f.fp) == EOF) {
    set_error("dax_hip_write_vtk_polydata: write to %s failed", path);
    return DAX_HIP_EXEC_ERROR;
  }
  return DAX_HIP_OK;
}

int dax_hip_write_vtk_unstructured(const char* path, const float* h_points, int64_t n_points,
                                   const int32_t* h_conn, int64_t n_cells, int verts_per_cell,
                                   int vtk_cell_type) {
  int rc = check_args("dax_hip_write_vtk_unstructured", path, h_points, n_points, h_conn, n_cells,
                      verts_per_cell);
  if (rc) return rc;
  File f;
  f.fp = fopen(path, "wb");
  if (!f.fp) {
    set_error("dax_hip_write_vtk_unstructured: cannot open %s", path);
    return DAX_HIP_BAD_VALUE;
  }
  const std::string head = "# vtk DataFile Version 3.0\nvtk output\nASCII\n"
                           "DATASET UNSTRUCTURED_GRID\nPOINTS " + std::to_string(n_points) +
                           " float\n";
  const std::string cells = "CELLS " + std::to_string(n_cells) + " " +
                            std::to_string(n_cells * (verts_per_cell + 1)) + "\n";
  std::string types = "\nCELL_TYPES " + std::to_string(n_cells) + "\n";
  const std::string one = std::to_string(vtk_cell_type) + "\n";
  types.reserve(types.size() + (size_t)n_cells * one.size());
  for (int64_t c = 0; c < n_cells; ++c) types += one;
  if (fputs(head.c_str(), f.fp) == EOF || vtk::write_points(f.fp, h_points, n_points) ||
      fputs(cells.c_str(), f.fp) == EOF ||
      vtk::write_cells(f.fp, h_conn, n_cells, verts_per_cell) ||
      fwrite(types.data(), 1, types.size(), f.fp) != types.size()) {
    set_error("dax_hip_write_vtk_unstructured: write to %s failed", path);
    return DAX_HIP_EXEC_ERROR;
  }
  return DAX_HIP_OK;
}

}  // extern "C"
