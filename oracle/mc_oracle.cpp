// ORACLE -- test infrastructure only.
//
// CPU restatement of Dax's Serial device adapter on the marching-cubes
// isosurface path (classify -> inclusive scan -> upper bounds -> visit index
// -> generate -> sort/unique/lower-bounds dedup -> coordinate interpolation),
// plus Dax's CellGradient worklet, the central-difference point normals and an
// FP64 adjoint / finite-difference checker for the backward pass.  Only the
// test-suite, the smoke check and the CPU-baseline leg of the benchmark load
// this library; the product never does.
//
// Every function names the reference lines it follows.  The FP32 operation
// order of the reference is kept (weight, Lerp, CellGradient sum), and the file
// is compiled with -ffp-contract=off so no multiply-add is fused.
//
// Pinning (what checks each function; see tests/test_oracle.py):
//   tables            SHA-256 of the expanded TriTable/NumFaces bytes
//   oracle_mc_count   brute force over all 256 corner sign patterns; known
//                     counts for the 128^3 sphere
//   scans / bounds    the reference conformance suite's golden vectors
//                     (TestingDeviceAdapter.h:621-717, 940-1001)
//   oracle_mc_generate  UnitTestWorkletMarchingCubes.cxx (26^3, x+y+z, iso 70:
//                     3T points without merge, 45 with merge, secondary field
//                     interpolated exactly); closed-form plane intersections;
//                     dedup == 3T-point set with duplicates removed
//   oracle_point_gradient / normals   exact on linear and quadratic fields,
//                     sphere normals vs analytic (p-c)/|p-c|
//   oracle_cell_gradient  linear fields give the exact gradient
//                     (UnitTestWorkletCellGradient.cxx)
//   oracle_mc_backward_f64  oracle_mc_fd_f64 central differences, and the
//                     torch float64 autograd restatement in the tests
//   oracle_mc_pipeline_omp  equality with the serial oracle
#include <algorithm>
#include <fstream>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

#include "mc_tables_oracle.h"

namespace {

typedef int32_t Id;      // dax::Id (dax/internal/Configure.h.in:67-75, 32-bit default)
typedef float Scalar;    // dax::Scalar (dax/Types.h:258-268)

struct Tables {
  uint8_t num_faces[256];
  uint8_t tri[256][16];
  Tables() {
    for (int c = 0; c < 256; ++c) {
      const char* s = kOracleCaseEdges[c];
      int n = (int)std::strlen(s);
      num_faces[c] = (uint8_t)(n / 3);
      for (int k = 0; k < 16; ++k) {
        if (k < n) {
          char ch = s[k];
          tri[c][k] = (uint8_t)(ch <= '9' ? ch - '0' : ch - 'a' + 10);
        } else {
          tri[c][k] = 255;
        }
      }
    }
  }
};
const Tables& tables() {
  static Tables t;
  return t;
}

// dax/worklet/MarchingCubes.h:131-135 -- cell vertex pairs of the 12 edges.
const unsigned char kCellVertEdges[12][2] = {
    {0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
    {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct Grid {
  int64_t nx, ny, nz;
  float origin[3];
  float spacing[3];
  int64_t points() const { return nx * ny * nz; }
  int64_t cx() const { return nx > 1 ? nx - 1 : 0; }
  int64_t cy() const { return ny > 1 ? ny - 1 : 0; }
  int64_t cz() const { return nz > 1 ? nz - 1 : 0; }
  int64_t cells() const { return cx() * cy() * cz(); }
};

Grid make_grid(const int64_t dims[3], const float* origin, const float* spacing) {
  Grid g;
  g.nx = dims[0];
  g.ny = dims[1];
  g.nz = dims[2];
  for (int d = 0; d < 3; ++d) {
    g.origin[d] = origin ? origin[d] : 0.0f;
    g.spacing[d] = spacing ? spacing[d] : 1.0f;
  }
  return g;
}

// dax/exec/internal/TopologyUniform.h:139-159 with the flat cell index
// decomposed as in dax/Extent.h:127-137 (x fastest).
void cell_vertices(const Grid& g, int64_t cell, Id verts[8]) {
  const int64_t cx = g.cx(), cy = g.cy();
  const int64_t i = cell % cx, j = (cell / cx) % cy, k = cell / (cx * cy);
  const Id first = (Id)(i + g.nx * (j + g.ny * k));
  const Id second = (Id)(first + g.nx * g.ny);
  const Id xdim = (Id)g.nx;
  verts[0] = first;
  verts[1] = first + 1;
  verts[2] = first + xdim + 1;
  verts[3] = first + xdim;
  verts[4] = second;
  verts[5] = second + 1;
  verts[6] = second + xdim + 1;
  verts[7] = second + xdim;
}

// dax/worklet/MarchingCubes.h:36-48 (strict '>').
int classification(Scalar iso, const Scalar v[8]) {
  return ((v[0] > iso) << 0 | (v[1] > iso) << 1 | (v[2] > iso) << 2 |
          (v[3] > iso) << 3 | (v[4] > iso) << 4 | (v[5] > iso) << 5 |
          (v[6] > iso) << 6 | (v[7] > iso) << 7);
}

// dax/cont/UniformGrid.h:82-89 (origin + spacing*ijk, Id converted to Scalar).
void point_coords(const Grid& g, Id p, Scalar out[3]) {
  const int64_t i = p % g.nx, j = (p / g.nx) % g.ny, k = p / (g.nx * g.ny);
  out[0] = g.origin[0] + g.spacing[0] * (Scalar)i;
  out[1] = g.origin[1] + g.spacing[1] * (Scalar)j;
  out[2] = g.origin[2] + g.spacing[2] * (Scalar)k;
}

// dax/cont/internal/EdgeInterpolatedGrid.h:28-74
struct PointAsEdgeInterpolation {
  Id first, second;
  Scalar weight;
  bool operator==(const PointAsEdgeInterpolation& o) const {
    return first == o.first && second == o.second;
  }
  bool operator<(const PointAsEdgeInterpolation& o) const {
    return first < o.first || (first == o.first && second < o.second);
  }
};

// dax/worklet/MarchingCubes.h:122-168 -- one output triangle of one cell.
void build_triangle(const Grid& g, const Scalar* field, Scalar iso, int64_t cell,
                    int visit, PointAsEdgeInterpolation out[3]) {
  const Tables& T = tables();
  Id verts[8];
  cell_vertices(g, cell, verts);
  Scalar values[8];
  for (int v = 0; v < 8; ++v) values[v] = field[verts[v]];
  const int voxel_class = classification(iso, values);
  for (int ov = 0; ov < 3; ++ov) {
    const unsigned char edge = T.tri[voxel_class][visit * 3 + ov];
    int a, b;
    if (verts[kCellVertEdges[edge][0]] < verts[kCellVertEdges[edge][1]]) {
      a = kCellVertEdges[edge][0];
      b = kCellVertEdges[edge][1];
    } else {
      a = kCellVertEdges[edge][1];
      b = kCellVertEdges[edge][0];
    }
    const Scalar weight = (iso - values[a]) / (values[b] - values[a]);
    out[ov].first = verts[a];
    out[ov].second = verts[b];
    out[ov].weight = weight;
  }
}

// dax/math/VectorAnalysis.h:33-49,60-74 -- Lerp = a + w*(b-a), per component.
inline Scalar lerp(Scalar a, Scalar b, Scalar w) { return a + w * (b - a); }

// Central-difference point gradient (no reference counterpart).  Interior:
// (f(p+e)-f(p-e))/(2h); first plane (f(p+e)-f(p))/h; last plane
// (f(p)-f(p-e))/h; an axis with one point has zero derivative.
void point_gradient_at(const Grid& g, const Scalar* f, int64_t p, Scalar out[3]) {
  const int64_t idx[3] = {p % g.nx, (p / g.nx) % g.ny, p / (g.nx * g.ny)};
  const int64_t n[3] = {g.nx, g.ny, g.nz};
  const int64_t stride[3] = {1, g.nx, g.nx * g.ny};
  for (int d = 0; d < 3; ++d) {
    const Scalar h = g.spacing[d];
    if (n[d] < 2) {
      out[d] = 0.0f;
    } else if (idx[d] == 0) {
      out[d] = (f[p + stride[d]] - f[p]) / h;
    } else if (idx[d] == n[d] - 1) {
      out[d] = (f[p] - f[p - stride[d]]) / h;
    } else {
      out[d] = (f[p + stride[d]] - f[p - stride[d]]) / (2.0f * h);
    }
  }
}

// Vertex normal: Lerp of the two endpoint gradients with the vertex weight
// (the CompactPointField semantics of DispatcherGenerateInterpolatedCells.h:
// 107-141 applied to the gradient field), then normalised; a zero vector stays
// zero.
void normal_from(const Scalar gA[3], const Scalar gB[3], Scalar w, Scalar out[3]) {
  Scalar m[3];
  for (int d = 0; d < 3; ++d) m[d] = lerp(gA[d], gB[d], w);
  const Scalar len2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
  if (len2 > 0.0f) {
    const Scalar len = std::sqrt(len2);
    for (int d = 0; d < 3; ++d) out[d] = m[d] / len;
  } else {
    out[0] = out[1] = out[2] = 0.0f;
  }
}

}  // namespace

extern "C" {

// Rebuilds the reference's NumFaces[256] and TriTable[256][16] byte layouts.
void oracle_tables_expand(uint8_t* num_faces, uint8_t* tri) {
  const Tables& T = tables();
  std::memcpy(num_faces, T.num_faces, 256);
  std::memcpy(tri, T.tri, 4096);
}

// MarchingCubesCount over every cell (dax/worklet/MarchingCubes.h:53-86),
// scheduled as the Serial adapter's Id3 triple loop
// (dax/cont/internal/DeviceAdapterAlgorithmSerial.h:146-176).  `cases` (one
// byte per cell) is optional; the reference computes but does not store it.
int oracle_mc_count(const float* field, const int64_t* dims, float iso,
                    uint8_t* cases, int32_t* counts) {
  const Grid g = make_grid(dims, nullptr, nullptr);
  const Tables& T = tables();
  const int64_t cx = g.cx(), cy = g.cy(), cz = g.cz();
  int64_t cell = 0;
  for (int64_t k = 0; k < cz; ++k)
    for (int64_t j = 0; j < cy; ++j)
      for (int64_t i = 0; i < cx; ++i, ++cell) {
        Id verts[8];
        cell_vertices(g, cell, verts);
        Scalar values[8];
        for (int v = 0; v < 8; ++v) values[v] = field[verts[v]];
        const int c = classification(iso, values);
        if (cases) cases[cell] = (uint8_t)c;
        counts[cell] = T.num_faces[c];
      }
  return 0;
}

// ScanInclusive: std::partial_sum, returns the last value
// (DeviceAdapterAlgorithmSerial.h:46-68).
int64_t oracle_scan_inclusive_i32(const int32_t* in, int32_t* out, int64_t n) {
  if (n <= 0) return 0;
  std::partial_sum(in, in + n, out);
  return out[n - 1];
}

// ScanExclusive: partial_sum, shift right, out[0]=0, returns the full sum
// (DeviceAdapterAlgorithmSerial.h:70-99).
int64_t oracle_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n) {
  if (n <= 0) return 0;
  std::partial_sum(in, in + n, out);
  const int32_t full = out[n - 1];
  std::copy_backward(out, out + n - 1, out + n);
  out[0] = 0;
  return full;
}

// LowerBounds / UpperBounds: std::lower_bound / std::upper_bound per value
// (DeviceAdapterAlgorithmGeneral.h:258-308, 1044-1092).
void oracle_lower_bounds_i32(const int32_t* sorted, int64_t n, const int32_t* vals,
                             int64_t m, int32_t* out) {
  for (int64_t q = 0; q < m; ++q)
    out[q] = (int32_t)(std::lower_bound(sorted, sorted + n, vals[q]) - sorted);
}
void oracle_upper_bounds_i32(const int32_t* sorted, int64_t n, const int32_t* vals,
                             int64_t m, int32_t* out) {
  for (int64_t q = 0; q < m; ++q)
    out[q] = (int32_t)(std::upper_bound(sorted, sorted + n, vals[q]) - sorted);
}

// The whole generate step for MarchingCubesGenerate on a uniform grid,
// restating DispatcherGenerateInterpolatedCells::GenerateNewTopology and
// ResolveCoordinates (dax/cont/DispatcherGenerateInterpolatedCells.h:158-301):
//   count (MarchingCubes.h:53-86) -> ScanInclusive (:180-183) -> T
//   UpperBounds(scan, Counting(0,T)) = validCellRange (:199-202)
//   visit index = t - LowerBounds(vcr, vcr)[t]  (AddVisitIndexArg.h:44-64,
//     VisitIndexWorklets.h:48-58)
//   BuildTriangle per t into slots 3t..3t+2 (MarchingCubes.h:122-168,
//     GeometryCell.h:80-91)
//   dedup: Copy, Sort, Unique, LowerBounds -> connectivity (:273-290);
//   otherwise connectivity = 0..3T-1 (:292-300)
//   coordinates = InterpolateFieldToField over the kept interpolation records
//     (GenerateWorklets.h:65-100).
// Sizes are always reported.  Outputs are written only when the capacities
// suffice (cap_pts >= 3T always does); returns 1 when they do not.
int oracle_mc_generate(const float* field, const int64_t* dims, const float* origin,
                       const float* spacing, float iso, int dedup, int64_t cap_tris,
                       int64_t cap_pts, int64_t* n_tris, int64_t* n_pts,
                       int32_t* conn, int32_t* edge_first, int32_t* edge_second,
                       float* weight, float* coords) {
  const Grid g = make_grid(dims, origin, spacing);
  const int64_t C = g.cells();
  std::vector<int32_t> count(C);
  if (C > 0) oracle_mc_count(field, dims, iso, nullptr, count.data());
  std::vector<int32_t> scan(C);
  const int64_t T = oracle_scan_inclusive_i32(count.data(), scan.data(), C);
  *n_tris = T;
  *n_pts = 0;
  if (T == 0) return 0;
  std::vector<int32_t> tids(T);
  std::iota(tids.begin(), tids.end(), 0);
  std::vector<int32_t> vcr(T);
  oracle_upper_bounds_i32(scan.data(), C, tids.data(), T, vcr.data());
  std::vector<int32_t> lb(T);
  oracle_lower_bounds_i32(vcr.data(), T, vcr.data(), T, lb.data());
  std::vector<PointAsEdgeInterpolation> interp(3 * T);
  for (int64_t t = 0; t < T; ++t) {
    const int visit = (int)(t - lb[t]);
    build_triangle(g, field, iso, vcr[t], visit, &interp[3 * t]);
  }
  std::vector<PointAsEdgeInterpolation> weights(interp);
  std::vector<int32_t> connections(3 * T);
  if (dedup) {
    std::sort(weights.begin(), weights.end());
    weights.erase(std::unique(weights.begin(), weights.end()), weights.end());
    for (int64_t q = 0; q < 3 * T; ++q)
      connections[q] = (int32_t)(std::lower_bound(weights.begin(), weights.end(), interp[q]) -
                                 weights.begin());
  } else {
    std::iota(connections.begin(), connections.end(), 0);
  }
  const int64_t U = (int64_t)weights.size();
  *n_pts = U;
  if (cap_tris < T || cap_pts < U) return 1;
  if (conn) std::memcpy(conn, connections.data(), sizeof(int32_t) * 3 * T);
  for (int64_t u = 0; u < U; ++u) {
    const PointAsEdgeInterpolation& pe = weights[u];
    if (edge_first) edge_first[u] = pe.first;
    if (edge_second) edge_second[u] = pe.second;
    if (weight) weight[u] = pe.weight;
    if (coords) {
      Scalar a[3], b[3];
      point_coords(g, pe.first, a);
      point_coords(g, pe.second, b);
      for (int d = 0; d < 3; ++d) coords[3 * u + d] = lerp(a[d], b[d], pe.weight);
    }
  }
  return 0;
}

// CompactPointField for a scalar point field (GenerateWorklets.h:65-100).
void oracle_interpolate_scalar(const float* in, const int32_t* first, const int32_t* second,
                               const float* w, int64_t n, float* out) {
  for (int64_t u = 0; u < n; ++u) out[u] = lerp(in[first[u]], in[second[u]], w[u]);
}

// Central-difference point gradient of the whole field, float32, [P][3].
void oracle_point_gradient(const float* field, const int64_t* dims, const float* spacing,
                           float* grad) {
  const Grid g = make_grid(dims, nullptr, spacing);
  const int64_t P = g.points();
  for (int64_t p = 0; p < P; ++p) point_gradient_at(g, field, p, grad + 3 * p);
}

// Vertex normals for interpolation records (first, second, w).
void oracle_mc_normals(const float* field, const int64_t* dims, const float* spacing,
                       const int32_t* first, const int32_t* second, const float* w,
                       int64_t n, float* normals) {
  const Grid g = make_grid(dims, nullptr, spacing);
  for (int64_t u = 0; u < n; ++u) {
    Scalar gA[3], gB[3];
    point_gradient_at(g, field, first[u], gA);
    point_gradient_at(g, field, second[u], gB);
    normal_from(gA, gB, w[u], normals + 3 * u);
  }
}

// CellGradient on a uniform grid (dax/worklet/CellGradient.h:28-48):
// CellDerivative(Voxel) -> CellDerivativeAxisAligned (dax/exec/Derivative.h:
// 37-75) with DerivativeWeights(Voxel) at the parametric centre
// (dax/exec/internal/DerivativeWeights.h:30-69, ParametricCoordinates.h:77-81);
// widths = coords[6] - coords[0].
void oracle_cell_gradient(const float* field, const int64_t* dims, const float* origin,
                          const float* spacing, float* out) {
  const Grid g = make_grid(dims, origin, spacing);
  const int64_t C = g.cells();
  const Scalar pc[3] = {0.5f, 0.5f, 0.5f};
  const Scalar rc[3] = {1.0f - pc[0], 1.0f - pc[1], 1.0f - pc[2]};
  Scalar w[8][3];
  w[0][0] = -rc[1] * rc[2]; w[0][1] = -rc[0] * rc[2]; w[0][2] = -rc[0] * rc[1];
  w[1][0] = rc[1] * rc[2];  w[1][1] = -pc[0] * rc[2]; w[1][2] = -pc[0] * rc[1];
  w[2][0] = pc[1] * rc[2];  w[2][1] = pc[0] * rc[2];  w[2][2] = -pc[0] * pc[1];
  w[3][0] = -pc[1] * rc[2]; w[3][1] = rc[0] * rc[2];  w[3][2] = -rc[0] * pc[1];
  w[4][0] = -rc[1] * pc[2]; w[4][1] = -rc[0] * pc[2]; w[4][2] = rc[0] * rc[1];
  w[5][0] = rc[1] * pc[2];  w[5][1] = -pc[0] * pc[2]; w[5][2] = pc[0] * rc[1];
  w[6][0] = pc[1] * pc[2];  w[6][1] = pc[0] * pc[2];  w[6][2] = pc[0] * pc[1];
  w[7][0] = -pc[1] * pc[2]; w[7][1] = rc[0] * pc[2];  w[7][2] = rc[0] * pc[1];
  for (int64_t c = 0; c < C; ++c) {
    Id verts[8];
    cell_vertices(g, c, verts);
    Scalar sum[3] = {0.0f, 0.0f, 0.0f};
    for (int v = 0; v < 8; ++v) {
      const Scalar f = field[verts[v]];
      for (int d = 0; d < 3; ++d) sum[d] = sum[d] + f * w[v][d];
    }
    Scalar c0[3], c6[3];
    point_coords(g, verts[0], c0);
    point_coords(g, verts[6], c6);
    for (int d = 0; d < 3; ++d) out[3 * c + d] = sum[d] / (c6[d] - c0[d]);
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FP64 adjoint and finite differences (no reference counterpart).
//
// Forward in FP64 for vertex u on edge (A,B):
//   t = (iso - fA)/(fB - fA),  x = cA + t (cB - cA)
//   m = gA + t (gB - gA),      n = m/|m|
// with g the central-difference gradient above.  L = sum_u dLdx_u.x_u +
// dLdn_u.n_u with topology (the edge list) held fixed.
namespace {

void point_gradient_at_f64(const Grid& g, const float* f, const std::unordered_map<int64_t, double>* over,
                           int64_t p, double out[3]) {
  auto val = [&](int64_t q) -> double {
    if (over) {
      auto it = over->find(q);
      if (it != over->end()) return it->second;
    }
    return (double)f[q];
  };
  const int64_t idx[3] = {p % g.nx, (p / g.nx) % g.ny, p / (g.nx * g.ny)};
  const int64_t n[3] = {g.nx, g.ny, g.nz};
  const int64_t stride[3] = {1, g.nx, g.nx * g.ny};
  for (int d = 0; d < 3; ++d) {
    const double h = (double)g.spacing[d];
    if (n[d] < 2)
      out[d] = 0.0;
    else if (idx[d] == 0)
      out[d] = (val(p + stride[d]) - val(p)) / h;
    else if (idx[d] == n[d] - 1)
      out[d] = (val(p) - val(p - stride[d])) / h;
    else
      out[d] = (val(p + stride[d]) - val(p - stride[d])) / (2.0 * h);
  }
}

double vertex_loss_f64(const Grid& g, const float* f, const std::unordered_map<int64_t, double>* over,
                       float iso, int64_t A, int64_t B, const float* dldx, const float* dldn) {
  auto val = [&](int64_t q) -> double {
    if (over) {
      auto it = over->find(q);
      if (it != over->end()) return it->second;
    }
    return (double)f[q];
  };
  const double fA = val(A), fB = val(B);
  const double t = ((double)iso - fA) / (fB - fA);
  double loss = 0.0;
  const int64_t ia[3] = {A % g.nx, (A / g.nx) % g.ny, A / (g.nx * g.ny)};
  const int64_t ib[3] = {B % g.nx, (B / g.nx) % g.ny, B / (g.nx * g.ny)};
  for (int d = 0; d < 3; ++d) {
    const double ca = (double)g.origin[d] + (double)g.spacing[d] * (double)ia[d];
    const double cb = (double)g.origin[d] + (double)g.spacing[d] * (double)ib[d];
    loss += (double)dldx[d] * (ca + t * (cb - ca));
  }
  double gA[3], gB[3], m[3];
  point_gradient_at_f64(g, f, over, A, gA);
  point_gradient_at_f64(g, f, over, B, gB);
  for (int d = 0; d < 3; ++d) m[d] = gA[d] + t * (gB[d] - gA[d]);
  const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  if (len > 0.0)
    for (int d = 0; d < 3; ++d) loss += (double)dldn[d] * (m[d] / len);
  return loss;
}

}  // namespace

extern "C" {

// Slice (dax/worklet/Slice.h:28-173) on a uniform grid: SliceCount and
// SliceGenerate classify and interpolate the per-point value
// dot(normal, c - origin) of the point coordinates c = grid origin + spacing
// * ijk (dax/cont/UniformGrid.h:82-89) against the isovalue
// dot(normal, origin) (Slice.h:44-57,104-120), dot summed from component 0
// (dax/Types.h:611-621).  With that value as the point field the two
// worklets are MarchingCubesCount / MarchingCubesGenerate (same table, same
// edge pairs; on a uniform grid every edge pair lists the lower point id
// first, so the weights match MarchingCubes.h:148-161).  Returns the
// isovalue; writes the values of all points.
float oracle_slice_field(const int64_t* dims, const float* origin, const float* spacing,
                         const float* plane_origin, const float* plane_normal, float* out) {
  const Grid g = make_grid(dims, origin, spacing);
  for (int64_t p = 0; p < g.points(); ++p) {
    const int64_t ijk[3] = {p % g.nx, (p / g.nx) % g.ny, p / (g.nx * g.ny)};
    float c[3];
    for (int d = 0; d < 3; ++d) c[d] = g.origin[d] + g.spacing[d] * (float)ijk[d];
    float r = plane_normal[0] * (c[0] - plane_origin[0]);
    r += plane_normal[1] * (c[1] - plane_origin[1]);
    r += plane_normal[2] * (c[2] - plane_origin[2]);
    out[p] = r;
  }
  float iso = plane_normal[0] * plane_origin[0];
  iso += plane_normal[1] * plane_origin[1];
  iso += plane_normal[2] * plane_origin[2];
  return iso;
}

// Marching tetrahedra, Serial-adapter sequence (dax/worklet/MarchingTetrahedra.h:29-143
// through DispatcherGenerateInterpolatedCells.h:138-301): per tetrahedron
// NumFaces[case] (case = sum_v (f_v > iso) << v, :37-43); inclusive scan;
// triangle -> cell by upper bounds and visit index by lower bounds (as the
// MC path above); per triangle corner the edge from TriTable, endpoints
// ordered by point id (:117-127), weight (iso - f_A) / (f_B - f_A)
// (:130-131); dedup: sort by (first, second), unique on the pair, each
// corner's point by lower bound (:261-301); coordinates Lerp (:107-141).
// Table rows: the edges of each case's triangles (MarchingTetrahedraTable.h:28-54).
static const char* const kTetTri[16] = {"",       "032", "014",    "143132", "125",    "031135",
                                        "254024", "354", "345",    "204245", "015053", "152",
                                        "341312", "041", "023",    ""};
static const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 1}, {3, 2}};

// The tetrahedron tables in the reference's byte layout: NumFaces[16],
// TriTable[16][8] with 255 for unused entries.
void oracle_tet_tables(uint8_t* num_faces, uint8_t* tri) {
  for (int c = 0; c < 16; ++c) {
    const size_t n = strlen(kTetTri[c]);
    num_faces[c] = (uint8_t)(n / 3);
    for (int k = 0; k < 8; ++k) tri[8 * c + k] = k < (int)n ? (uint8_t)(kTetTri[c][k] - '0') : 255;
  }
}

}  // extern "C"

namespace {

// Cells of one shape with NV vertices: case_edges(c) is the case's TriTable
// row as a string of edge digits (hex digits for the hexahedron), ends the
// edge -> vertex pairs.  The Serial-adapter sequence shared by marching
// tetrahedra and marching cubes on an unstructured grid.
template <int NV, class CaseEdges>
int unstructured_generate(const int32_t* cells, int64_t n_cells, const float* field,
                          const float* points, float iso, int dedup, int64_t cap,
                          int64_t* n_tris, int64_t* n_pts, int32_t* conn, int32_t* first,
                          int32_t* second, float* weight, float* coords, CaseEdges case_edges,
                          const int (*ends)[2]) {
  auto digit = [](char ch) { return ch <= '9' ? ch - '0' : ch - 'a' + 10; };
  std::vector<int> cls(n_cells);
  std::vector<Id> incl(n_cells);
  Id run = 0;
  for (int64_t c = 0; c < n_cells; ++c) {
    int k = 0;
    for (int v = 0; v < NV; ++v) k |= (field[cells[NV * c + v]] > iso) << v;
    cls[c] = k;
    run += (Id)(strlen(case_edges(k)) / 3);
    incl[c] = run;
  }
  const int64_t T = run;
  *n_tris = T;
  std::vector<PointAsEdgeInterpolation> rec(3 * T);
  for (int64_t t = 0; t < T; ++t) {
    const int64_t c = std::upper_bound(incl.begin(), incl.end(), (Id)t) - incl.begin();
    const int64_t first_t = c == 0 ? 0 : incl[c - 1];
    const int s = (int)(t - first_t);
    const int32_t* id = cells + NV * c;
    for (int v = 0; v < 3; ++v) {
      const int e = digit(case_edges(cls[c])[3 * s + v]);
      int va = ends[e][0], vb = ends[e][1];
      if (!(id[va] < id[vb])) std::swap(va, vb);
      const Scalar w = (iso - field[id[va]]) / (field[id[vb]] - field[id[va]]);
      rec[3 * t + v] = PointAsEdgeInterpolation{id[va], id[vb], w};
    }
  }
  std::vector<PointAsEdgeInterpolation> pts;
  if (dedup) {
    pts = rec;
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    for (int64_t i = 0; i < 3 * T; ++i)
      conn[i] = (int32_t)(std::lower_bound(pts.begin(), pts.end(), rec[i]) - pts.begin());
  } else {
    pts = rec;
    for (int64_t i = 0; i < 3 * T; ++i) conn[i] = (int32_t)i;
  }
  *n_pts = (int64_t)pts.size();
  if ((int64_t)pts.size() > cap) return 1;
  for (size_t u = 0; u < pts.size(); ++u) {
    first[u] = pts[u].first;
    second[u] = pts[u].second;
    weight[u] = pts[u].weight;
    for (int d = 0; d < 3; ++d) {
      const Scalar a = points[3 * (int64_t)pts[u].first + d], b = points[3 * (int64_t)pts[u].second + d];
      coords[3 * u + d] = a + pts[u].weight * (b - a);
    }
  }
  return 0;
}

// Hexahedron edges -> vertex pairs (dax/worklet/MarchingCubes.h:131-135).
const int kHexEdge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                             {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

}  // namespace

extern "C" {

int oracle_mt_generate(const int32_t* tets, int64_t n_tets, const float* field,
                       const float* points, float iso, int dedup, int64_t cap,
                       int64_t* n_tris, int64_t* n_pts, int32_t* conn, int32_t* first,
                       int32_t* second, float* weight, float* coords) {
  return unstructured_generate<4>(tets, n_tets, field, points, iso, dedup, cap, n_tris, n_pts,
                                  conn, first, second, weight, coords,
                                  [](int c) { return kTetTri[c]; }, kTetEdge);
}

// Marching cubes on an unstructured hexahedral grid
// (dax/worklet/MarchingCubes.h:53-164 on UnstructuredGrid<Hexahedron>): the
// same sequence with the marching-cubes table (GetHexahedronClassification,
// :38-48) and the hexahedron's edges.
int oracle_mc_hex_generate(const int32_t* hexes, int64_t n_hexes, const float* field,
                           const float* points, float iso, int dedup, int64_t cap,
                           int64_t* n_tris, int64_t* n_pts, int32_t* conn, int32_t* first,
                           int32_t* second, float* weight, float* coords) {
  return unstructured_generate<8>(hexes, n_hexes, field, points, iso, dedup, cap, n_tris, n_pts,
                                  conn, first, second, weight, coords,
                                  [](int c) { return kOracleCaseEdges[c]; }, kHexEdge);
}

// Threshold, Serial-adapter sequence: ThresholdCount per cell (every point
// value v with lo <= v <= hi, dax/worklet/Threshold.h:30-96) through
// DispatcherMapCell; DispatcherGenerateTopology
// (dax/cont/DispatcherGenerateTopology.h:146-256): inclusive scan of the
// counts, UpperBounds of [0, T) in it (the input cell of each output cell),
// ThresholdTopology copies the input cell's vertices (:98-110); with dedup
// the point mask of the used ids (:211-229), StreamCompact of the
// coordinates by it, and the connectivity replaced by LowerBounds in the
// sorted unique used ids (:231-254).  cells == nullptr: the uniform grid
// `dims` (voxels x-fastest, corners in the hexahedron's order, coordinates
// origin + spacing * ijk).  Returns 1 if cap_cells output cells do not fit.
int oracle_threshold(const int32_t* cells, int nv, int64_t n_cells, const int64_t* dims,
                     const float* origin, const float* spacing, const float* points,
                     int64_t n_points, const float* field, float lo, float hi, int dedup,
                     int64_t cap_cells, int32_t* counts, int64_t* n_out_cells,
                     int64_t* n_out_points, int32_t* conn, float* coords, int32_t* mask) {
  Grid g{};
  if (!cells) {
    g = make_grid(dims, origin, spacing);
    nv = 8;
    n_cells = g.cells();
    n_points = g.points();
  }
  auto verts_of = [&](int64_t c, Id* v) {
    if (cells)
      for (int q = 0; q < nv; ++q) v[q] = cells[c * nv + q];
    else
      cell_vertices(g, c, v);  // the uniform grid's voxel topology
  };
  std::vector<Id> scanned(n_cells);
  Id run = 0;
  for (int64_t c = 0; c < n_cells; ++c) {
    Id v[8];
    verts_of(c, v);
    int valid = 1;
    for (int q = 0; q < nv; ++q) valid &= field[v[q]] >= lo && field[v[q]] <= hi;
    counts[c] = valid;
    run += valid;
    scanned[c] = run;
  }
  const int64_t T = run;
  *n_out_cells = T;
  *n_out_points = 0;
  if (T > cap_cells) return 1;
  for (int64_t t = 0; t < T; ++t) {
    const int64_t c = std::upper_bound(scanned.begin(), scanned.end(), (Id)t) - scanned.begin();
    Id v[8];
    verts_of(c, v);
    for (int q = 0; q < nv; ++q) conn[t * nv + q] = v[q];
  }
  if (!dedup) return 0;
  std::fill(mask, mask + n_points, 0);
  for (int64_t i = 0; i < T * nv; ++i) mask[conn[i]] = 1;
  int64_t u = 0;
  for (int64_t p = 0; p < n_points; ++p) {
    if (!mask[p]) continue;
    for (int d = 0; d < 3; ++d) {
      float x;
      if (cells) {
        x = points[3 * p + d];
      } else {
        const int64_t idx = d == 0 ? p % g.nx : d == 1 ? (p / g.nx) % g.ny : p / (g.nx * g.ny);
        x = g.origin[d] + g.spacing[d] * (float)idx;
      }
      coords[3 * u + d] = x;
    }
    ++u;
  }
  *n_out_points = u;
  std::vector<Id> used(conn, conn + T * nv);
  std::sort(used.begin(), used.end());
  used.erase(std::unique(used.begin(), used.end()), used.end());
  for (int64_t i = 0; i < T * nv; ++i)
    conn[i] = (Id)(std::lower_bound(used.begin(), used.end(), conn[i]) - used.begin());
  return 0;
}

// VTK legacy writers, the reference benchmarks' PrintContentsToStream stream
// sequence written with an ofstream: POLYDATA/POLYGONS
// (Benchmarks/MarchingCubes/Pipeline.h:51-98) and UNSTRUCTURED_GRID/CELLS/
// CELL_TYPES (Benchmarks/Threshold/Pipeline.h:99-150).  Returns 0 on success.
int oracle_write_vtk(const char* path, int polydata, const float* pts, int64_t n_points,
                     const int32_t* conn, int64_t n_cells, int nv, int cell_type) {
  std::ofstream stream(path);
  if (!stream) return 1;
  stream << "# vtk DataFile Version 3.0" << std::endl;
  stream << "vtk output" << std::endl;
  stream << "ASCII" << std::endl;
  stream << (polydata ? "DATASET POLYDATA" : "DATASET UNSTRUCTURED_GRID") << std::endl;
  stream << "POINTS " << n_points << " float" << std::endl;
  for (int64_t i = 0; i < n_points; ++i) {
    stream << pts[3 * i] << " " << pts[3 * i + 1] << " " << pts[3 * i + 2] << " ";
    if (i % 3 == 2) stream << std::endl;
  }
  if (n_points % 3 != 0) stream << std::endl;
  stream << (polydata ? "POLYGONS " : "CELLS ") << n_cells << " " << n_cells * (nv + 1)
         << std::endl;
  for (int64_t c = 0; c < n_cells; ++c) {
    stream << nv << " ";
    for (int v = 0; v < nv; ++v) {
      stream << conn[c * nv + v];
      if (v + 1 < nv) stream << " ";
    }
    stream << std::endl;
  }
  stream << std::endl;
  if (!polydata) {
    stream << "CELL_TYPES " << n_cells << std::endl;
    for (int64_t c = 0; c < n_cells; ++c) stream << cell_type << std::endl;
  }
  return stream.good() ? 0 : 1;
}

// Analytic adjoint in FP64, scatter form: dL/df for every point.
void oracle_mc_backward_f64(const float* field, const int64_t* dims, const float* origin,
                            const float* spacing, float iso, const int32_t* first,
                            const int32_t* second, int64_t n, const float* dldx,
                            const float* dldn, double* dldf) {
  const Grid g = make_grid(dims, origin, spacing);
  const int64_t P = g.points();
  std::fill(dldf, dldf + P, 0.0);
  const int64_t n3[3] = {g.nx, g.ny, g.nz};
  const int64_t stride[3] = {1, g.nx, g.nx * g.ny};
  // d g_p[d] / d f : list of (point, coefficient)
  auto grad_adj = [&](int64_t p, const double G[3]) {
    const int64_t idx[3] = {p % g.nx, (p / g.nx) % g.ny, p / (g.nx * g.ny)};
    for (int d = 0; d < 3; ++d) {
      const double h = (double)g.spacing[d];
      if (n3[d] < 2) continue;
      if (idx[d] == 0) {
        dldf[p + stride[d]] += G[d] / h;
        dldf[p] -= G[d] / h;
      } else if (idx[d] == n3[d] - 1) {
        dldf[p] += G[d] / h;
        dldf[p - stride[d]] -= G[d] / h;
      } else {
        dldf[p + stride[d]] += G[d] / (2.0 * h);
        dldf[p - stride[d]] -= G[d] / (2.0 * h);
      }
    }
  };
  for (int64_t u = 0; u < n; ++u) {
    const int64_t A = first[u], B = second[u];
    const double fA = field[A], fB = field[B];
    const double t = ((double)iso - fA) / (fB - fA);
    const int64_t ia[3] = {A % g.nx, (A / g.nx) % g.ny, A / (g.nx * g.ny)};
    const int64_t ib[3] = {B % g.nx, (B / g.nx) % g.ny, B / (g.nx * g.ny)};
    double dLdt = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double ca = (double)g.origin[d] + (double)g.spacing[d] * (double)ia[d];
      const double cb = (double)g.origin[d] + (double)g.spacing[d] * (double)ib[d];
      dLdt += (double)dldx[3 * u + d] * (cb - ca);
    }
    double gA[3], gB[3], m[3];
    point_gradient_at_f64(g, field, nullptr, A, gA);
    point_gradient_at_f64(g, field, nullptr, B, gB);
    for (int d = 0; d < 3; ++d) m[d] = gA[d] + t * (gB[d] - gA[d]);
    const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    double dLdm[3] = {0.0, 0.0, 0.0};
    if (len > 0.0) {
      double nv[3], dot = 0.0;
      for (int d = 0; d < 3; ++d) {
        nv[d] = m[d] / len;
        dot += nv[d] * (double)dldn[3 * u + d];
      }
      for (int d = 0; d < 3; ++d) dLdm[d] = ((double)dldn[3 * u + d] - nv[d] * dot) / len;
    }
    for (int d = 0; d < 3; ++d) dLdt += dLdm[d] * (gB[d] - gA[d]);
    const double den = fB - fA;
    dldf[A] += dLdt * (-(1.0 - t) / den);
    dldf[B] += dLdt * (-t / den);
    double GA[3], GB[3];
    for (int d = 0; d < 3; ++d) {
      GA[d] = (1.0 - t) * dLdm[d];
      GB[d] = t * dLdm[d];
    }
    grad_adj(A, GA);
    grad_adj(B, GB);
  }
}

// FP64 central differences of L with respect to selected points (step h),
// evaluating only the vertices whose value depends on the perturbed point.
void oracle_mc_fd_f64(const float* field, const int64_t* dims, const float* origin,
                      const float* spacing, float iso, const int32_t* first,
                      const int32_t* second, int64_t n, const float* dldx, const float* dldn,
                      const int64_t* points, int64_t npoints, double h, double* out) {
  const Grid g = make_grid(dims, origin, spacing);
  // point -> incident vertices
  std::unordered_map<int64_t, std::vector<int64_t>> incident;
  incident.reserve((size_t)(2 * n));
  for (int64_t u = 0; u < n; ++u) {
    incident[first[u]].push_back(u);
    incident[second[u]].push_back(u);
  }
  const int64_t stride[3] = {1, g.nx, g.nx * g.ny};
  const int64_t n3[3] = {g.nx, g.ny, g.nz};
  for (int64_t s = 0; s < npoints; ++s) {
    const int64_t q = points[s];
    const int64_t idx[3] = {q % g.nx, (q / g.nx) % g.ny, q / (g.nx * g.ny)};
    std::vector<int64_t> nb{q};
    for (int d = 0; d < 3; ++d) {
      if (idx[d] > 0) nb.push_back(q - stride[d]);
      if (idx[d] < n3[d] - 1) nb.push_back(q + stride[d]);
    }
    std::vector<int64_t> verts;
    for (int64_t p : nb) {
      auto it = incident.find(p);
      if (it != incident.end()) verts.insert(verts.end(), it->second.begin(), it->second.end());
    }
    std::sort(verts.begin(), verts.end());
    verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
    double lp = 0.0, lm = 0.0;
    std::unordered_map<int64_t, double> over;
    over[q] = (double)field[q] + h;
    for (int64_t u : verts)
      lp += vertex_loss_f64(g, field, &over, iso, first[u], second[u], dldx + 3 * u, dldn + 3 * u);
    over[q] = (double)field[q] - h;
    for (int64_t u : verts)
      lm += vertex_loss_f64(g, field, &over, iso, first[u], second[u], dldx + 3 * u, dldn + 3 * u);
    out[s] = (lp - lm) / (2.0 * h);
  }
}

// The same adjoint in FP32, in the gather form and operation order of the
// device adjoint, so that its output can be compared bit for bit:
//   per vertex (A, B = A + e_axis): t, gA, gB, m, |m|, dL/dm, dL/dt and the
//     record (sA, GA, sB, GB) with sA = dL/dt * (-(1-t)/den), sB = dL/dt *
//     (-t/den), GA = (1-t) dL/dm, GB = t dL/dm;
//   per point p: S_p and G_p summed from 0 over the edges leaving p (axis
//     0, 1, 2: lower halves) and then entering p (axis 0, 1, 2: upper
//     halves);
//   per point q: S_q, then the one-sided terms G_q[d] * (-1/h) on the first
//     and G_q[d] * (1/h) on the last plane of axis d, then for d = 0, 1, 2
//     the terms of q - e_d (+G/h next to the first plane, else +G/(2h)) and
//     of q + e_d (-(G/h) next to the last plane, else -(G/(2h))).
void oracle_mc_backward_f32(const float* field, const int64_t* dims, const float* origin,
                            const float* spacing, float iso, const int32_t* first,
                            const int32_t* second, int64_t n, const float* dldx,
                            const float* dldn, float* dldf) {
  const Grid g = make_grid(dims, origin, spacing);
  const int64_t P = g.points();
  const int64_t n3[3] = {g.nx, g.ny, g.nz};
  const int64_t stride[3] = {1, g.nx, g.nx * g.ny};
  auto grad = [&](int64_t p, float out[3]) {
    const int64_t idx[3] = {p % g.nx, (p / g.nx) % g.ny, p / (g.nx * g.ny)};
    for (int d = 0; d < 3; ++d) {
      const float h = g.spacing[d];
      if (n3[d] < 2)
        out[d] = 0.0f;
      else if (idx[d] == 0)
        out[d] = (field[p + stride[d]] - field[p]) / h;
      else if (idx[d] == n3[d] - 1)
        out[d] = (field[p] - field[p - stride[d]]) / h;
      else
        out[d] = (field[p + stride[d]] - field[p - stride[d]]) / (2.0f * h);
    }
  };
  // records and the vertex of each (point, axis) as lower / upper endpoint
  std::vector<float> rec((size_t)n * 8);
  std::vector<int64_t> lower((size_t)P * 3, -1), upper((size_t)P * 3, -1);
  for (int64_t u = 0; u < n; ++u) {
    const int64_t A = first[u], B = second[u];
    const int axis = B - A == 1 ? 0 : (B - A == g.nx ? 1 : 2);
    lower[3 * A + axis] = u;
    upper[3 * B + axis] = u;
    const float fA = field[A], fB = field[B];
    const float den = fB - fA;
    const float t = (iso - fA) / den;
    float gA[3], gB[3], m[3];
    grad(A, gA);
    grad(B, gB);
    for (int d = 0; d < 3; ++d) m[d] = gA[d] + t * (gB[d] - gA[d]);
    const float len2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    const float dn[3] = {dldn[3 * u], dldn[3 * u + 1], dldn[3 * u + 2]};
    float dm[3] = {0.0f, 0.0f, 0.0f};
    if (len2 > 0.0f) {
      const float len = std::sqrt(len2);
      float nv[3];
      for (int d = 0; d < 3; ++d) nv[d] = m[d] / len;
      const float dot = nv[0] * dn[0] + nv[1] * dn[1] + nv[2] * dn[2];
      for (int d = 0; d < 3; ++d) dm[d] = (dn[d] - nv[d] * dot) / len;
    }
    const int64_t ia[3] = {A % g.nx, (A / g.nx) % g.ny, A / (g.nx * g.ny)};
    float dldt = 0.0f;
    for (int d = 0; d < 3; ++d) {
      const float a = g.origin[d] + g.spacing[d] * (float)ia[d];
      const float b = g.origin[d] + g.spacing[d] * (float)(ia[d] + (d == axis ? 1 : 0));
      dldt += dldx[3 * u + d] * (b - a);
    }
    for (int d = 0; d < 3; ++d) dldt += dm[d] * (gB[d] - gA[d]);
    float* r = &rec[(size_t)u * 8];
    r[0] = dldt * (-(1.0f - t) / den);
    r[4] = dldt * (-t / den);
    for (int d = 0; d < 3; ++d) {
      r[1 + d] = (1.0f - t) * dm[d];
      r[5 + d] = t * dm[d];
    }
  }
  // (S, G) per point
  std::vector<float> sg((size_t)P * 4, 0.0f);
  for (int64_t p = 0; p < P; ++p) {
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; ++a) {
      const int64_t u = lower[3 * p + a];
      if (u >= 0)
        for (int c = 0; c < 4; ++c) acc[c] += rec[(size_t)u * 8 + c];
    }
    for (int a = 0; a < 3; ++a) {
      const int64_t u = upper[3 * p + a];
      if (u >= 0)
        for (int c = 0; c < 4; ++c) acc[c] += rec[(size_t)u * 8 + 4 + c];
    }
    for (int c = 0; c < 4; ++c) sg[(size_t)p * 4 + c] = acc[c];
  }
  for (int64_t q = 0; q < P; ++q) {
    const int64_t co[3] = {q % g.nx, (q / g.nx) % g.ny, q / (g.nx * g.ny)};
    float acc = sg[(size_t)q * 4];
    for (int d = 0; d < 3; ++d) {
      const float h = g.spacing[d];
      if (n3[d] < 2) continue;
      const float G = sg[(size_t)q * 4 + 1 + d];
      if (co[d] == 0)
        acc += G * (-1.0f / h);
      else if (co[d] == n3[d] - 1)
        acc += G * (1.0f / h);
    }
    for (int d = 0; d < 3; ++d) {
      const float h = g.spacing[d];
      if (n3[d] < 2) continue;
      if (co[d] >= 1) {
        const float G = sg[(size_t)(q - stride[d]) * 4 + 1 + d];
        acc += (co[d] - 1 == 0) ? G / h : G / (2.0f * h);
      }
      if (co[d] <= n3[d] - 2) {
        const float G = sg[(size_t)(q + stride[d]) * 4 + 1 + d];
        acc += (co[d] + 1 == n3[d] - 1) ? -(G / h) : -(G / (2.0f * h));
      }
    }
    dldf[q] = acc;
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CPU baseline: the same step sequence as Dax's OpenMP adapter
// (dax/openmp/cont/internal/DeviceAdapterAlgorithmOpenMP.h:34-144 over the
// Thrust algorithms of DeviceAdapterAlgorithmThrust.h): parallel classify,
// parallel inclusive scan, parallel upper/lower bounds, parallel generate,
// parallel sort, unique, parallel lower bounds and coordinate interpolation.
// Timed like Benchmarks/MarchingCubes/Pipeline.h:110-132 (classify + generate
// with dedup, field already resident).  Returns seconds.
extern "C" double oracle_mc_pipeline_omp(const float* field, const int64_t* dims,
                                         const float* origin, const float* spacing,
                                         float iso, int dedup, int threads,
                                         int64_t* n_tris, int64_t* n_pts) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  const double t0 = omp_get_wtime();
#endif
  const Grid g = make_grid(dims, origin, spacing);
  const Tables& TB = tables();
  const int64_t C = g.cells();
  std::vector<int32_t> count(C), scan(C);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < C; ++c) {
    Id verts[8];
    cell_vertices(g, c, verts);
    Scalar values[8];
    for (int v = 0; v < 8; ++v) values[v] = field[verts[v]];
    count[c] = TB.num_faces[classification(iso, values)];
  }
  // blocked two-pass inclusive scan
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  std::vector<int64_t> part(nth + 1, 0);
#pragma omp parallel num_threads(nth)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    const int64_t lo = C * tid / nth, hi = C * (tid + 1) / nth;
    int64_t s = 0;
    for (int64_t c = lo; c < hi; ++c) s += count[c];
    part[tid + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < nth; ++i) part[i + 1] += part[i];
    int64_t run = part[tid];
    for (int64_t c = lo; c < hi; ++c) {
      run += count[c];
      scan[c] = (int32_t)run;
    }
  }
  const int64_t T = C > 0 ? scan[C - 1] : 0;
  *n_tris = T;
  *n_pts = 0;
  if (T > 0) {
    std::vector<int32_t> vcr(T), visit(T);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; ++t)
      vcr[t] = (int32_t)(std::upper_bound(scan.begin(), scan.end(), (int32_t)t) - scan.begin());
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; ++t)
      visit[t] = (int32_t)(t - (std::lower_bound(vcr.begin(), vcr.end(), vcr[t]) - vcr.begin()));
    std::vector<PointAsEdgeInterpolation> interp(3 * T);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; ++t) build_triangle(g, field, iso, vcr[t], visit[t], &interp[3 * t]);
    std::vector<PointAsEdgeInterpolation> weights(interp);
    std::vector<int32_t> connections(3 * T);
    if (dedup) {
#ifdef _OPENMP
      __gnu_parallel::sort(weights.begin(), weights.end());
#else
      std::sort(weights.begin(), weights.end());
#endif
      weights.erase(std::unique(weights.begin(), weights.end()), weights.end());
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < 3 * T; ++q)
        connections[q] = (int32_t)(std::lower_bound(weights.begin(), weights.end(), interp[q]) -
                                   weights.begin());
    } else {
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < 3 * T; ++q) connections[q] = (int32_t)q;
    }
    const int64_t U = (int64_t)weights.size();
    std::vector<float> coords(3 * U);
#pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < U; ++u) {
      Scalar a[3], b[3];
      point_coords(g, weights[u].first, a);
      point_coords(g, weights[u].second, b);
      for (int d = 0; d < 3; ++d) coords[3 * u + d] = lerp(a[d], b[d], weights[u].weight);
    }
    *n_pts = U;
  }
#ifdef _OPENMP
  return omp_get_wtime() - t0;
#else
  return -1.0;
#endif
}

// ---------------------------------------------------------------------------
// CPU baselines of the unstructured-grid and threshold paths: the
// Serial-adapter sequences of unstructured_generate and oracle_threshold
// with every step on the OpenMP threads, as Dax's OpenMP adapter runs them
// (dax/openmp/cont/internal/DeviceAdapterAlgorithmOpenMP.h:34-144 over the
// Thrust algorithms of DeviceAdapterAlgorithmThrust.h).  Inputs already in
// memory; each returns seconds and the output counts.
namespace {

int omp_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double wall_seconds() {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return -1.0;
#endif
}

// Blocked two-pass inclusive scan of n counts (Thrust inclusive_scan on
// the OpenMP backend); returns the total.
int64_t omp_inclusive_scan(const int32_t* count, Id* scan, int64_t n) {
  const int nth = omp_threads();
  std::vector<int64_t> part(nth + 1, 0);
#pragma omp parallel num_threads(nth)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    const int64_t lo = n * tid / nth, hi = n * (tid + 1) / nth;
    int64_t s = 0;
    for (int64_t c = lo; c < hi; ++c) s += count[c];
    part[tid + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < nth; ++i) part[i + 1] += part[i];
    int64_t run = part[tid];
    for (int64_t c = lo; c < hi; ++c) {
      run += count[c];
      scan[c] = (Id)run;
    }
  }
  return part[nth];
}

template <class It>
void omp_sort(It b, It e) {
#ifdef _OPENMP
  __gnu_parallel::sort(b, e);
#else
  std::sort(b, e);
#endif
}

// unstructured_generate's sequence (MarchingTetrahedra.h:29-143 /
// MarchingCubes.h:53-164 through DispatcherGenerateInterpolatedCells.h:
// 138-301): classify, scan, upper bounds (triangle -> cell), lower bounds
// (visit index), the corners' interpolation records, sort + unique + lower
// bounds (dedup), coordinates.
template <int NV, class CaseEdges>
double unstructured_pipeline_omp(const int32_t* cells, int64_t n_cells, const float* field,
                                 const float* points, float iso, int dedup,
                                 CaseEdges case_edges, const int (*ends)[2], int64_t* n_tris,
                                 int64_t* n_pts) {
  const double t0 = wall_seconds();
  auto digit = [](char ch) { return ch <= '9' ? ch - '0' : ch - 'a' + 10; };
  int32_t num_faces[1 << NV];
  for (int k = 0; k < (1 << NV); ++k) num_faces[k] = (int32_t)(strlen(case_edges(k)) / 3);
  std::vector<uint8_t> cls(n_cells);
  std::vector<int32_t> count(n_cells);
  std::vector<Id> incl(n_cells);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n_cells; ++c) {
    int k = 0;
    for (int v = 0; v < NV; ++v) k |= (field[cells[NV * c + v]] > iso) << v;
    cls[c] = (uint8_t)k;
    count[c] = num_faces[k];
  }
  const int64_t T = omp_inclusive_scan(count.data(), incl.data(), n_cells);
  *n_tris = T;
  *n_pts = 0;
  if (T == 0) return wall_seconds() - t0;
  std::vector<Id> vcr(T), visit(T);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < T; ++t)
    vcr[t] = (Id)(std::upper_bound(incl.begin(), incl.end(), (Id)t) - incl.begin());
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < T; ++t)
    visit[t] = (Id)(t - (std::lower_bound(vcr.begin(), vcr.end(), vcr[t]) - vcr.begin()));
  std::vector<PointAsEdgeInterpolation> rec(3 * T);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < T; ++t) {
    const int64_t c = vcr[t];
    const int32_t* id = cells + NV * c;
    for (int v = 0; v < 3; ++v) {
      const int e = digit(case_edges(cls[c])[3 * visit[t] + v]);
      int va = ends[e][0], vb = ends[e][1];
      if (!(id[va] < id[vb])) std::swap(va, vb);
      const Scalar w = (iso - field[id[va]]) / (field[id[vb]] - field[id[va]]);
      rec[3 * t + v] = PointAsEdgeInterpolation{id[va], id[vb], w};
    }
  }
  std::vector<PointAsEdgeInterpolation> pts(rec);
  std::vector<int32_t> conn(3 * T);
  if (dedup) {
    omp_sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 3 * T; ++i)
      conn[i] = (int32_t)(std::lower_bound(pts.begin(), pts.end(), rec[i]) - pts.begin());
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 3 * T; ++i) conn[i] = (int32_t)i;
  }
  const int64_t U = (int64_t)pts.size();
  std::vector<float> coords(3 * U);
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < U; ++u)
    for (int d = 0; d < 3; ++d) {
      const Scalar a = points[3 * (int64_t)pts[u].first + d];
      const Scalar b = points[3 * (int64_t)pts[u].second + d];
      coords[3 * u + d] = a + pts[u].weight * (b - a);
    }
  *n_pts = U;
  return wall_seconds() - t0;
}

}  // namespace

extern "C" {

// Marching tetrahedra (nv 4) or marching cubes on hexahedra (nv 8).
double oracle_unstructured_pipeline_omp(const int32_t* cells, int nv, int64_t n_cells,
                                        const float* field, const float* points, float iso,
                                        int dedup, int threads, int64_t* n_tris,
                                        int64_t* n_pts) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  if (nv == 4)
    return unstructured_pipeline_omp<4>(cells, n_cells, field, points, iso, dedup,
                                        [](int c) { return kTetTri[c]; }, kTetEdge, n_tris,
                                        n_pts);
  return unstructured_pipeline_omp<8>(cells, n_cells, field, points, iso, dedup,
                                      [](int c) { return kOracleCaseEdges[c]; }, kHexEdge,
                                      n_tris, n_pts);
}

// oracle_threshold's sequence on the uniform grid `dims` (ThresholdCount
// through DispatcherMapCell, then DispatcherGenerateTopology.h:146-256):
// per-cell count, scan, upper bounds and the cells' vertex ids, and with
// dedup the point mask, StreamCompact of the coordinates by it, sort +
// unique of the used ids and lower bounds of the connectivity in them, and
// the field compacted by the mask (CompactPointField).
double oracle_threshold_pipeline_omp(const int64_t* dims, const float* origin,
                                     const float* spacing, const float* field, float lo,
                                     float hi, int dedup, int threads, int64_t* n_out_cells,
                                     int64_t* n_out_points) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const double t0 = wall_seconds();
  const Grid g = make_grid(dims, origin, spacing);
  const int64_t n_cells = g.cells(), n_points = g.points();
  std::vector<int32_t> count(n_cells);
  std::vector<Id> scanned(n_cells);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n_cells; ++c) {
    Id v[8];
    cell_vertices(g, c, v);
    int valid = 1;
    for (int q = 0; q < 8; ++q) valid &= field[v[q]] >= lo && field[v[q]] <= hi;
    count[c] = valid;
  }
  const int64_t T = omp_inclusive_scan(count.data(), scanned.data(), n_cells);
  *n_out_cells = T;
  *n_out_points = 0;
  std::vector<Id> conn(8 * T);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < T; ++t) {
    const int64_t c = std::upper_bound(scanned.begin(), scanned.end(), (Id)t) - scanned.begin();
    cell_vertices(g, c, &conn[8 * t]);
  }
  if (!dedup) return wall_seconds() - t0;
  std::vector<int32_t> mask(n_points);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n_points; ++p) mask[p] = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < 8 * T; ++i) __atomic_store_n(&mask[conn[i]], 1, __ATOMIC_RELAXED);
  std::vector<Id> pos(n_points);
  const int64_t U = omp_inclusive_scan(mask.data(), pos.data(), n_points);
  std::vector<float> coords(3 * U), values(U);  // CompactPointField of the field too
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n_points; ++p) {
    if (!mask[p]) continue;
    const int64_t u = pos[p] - 1;
    const int64_t ijk[3] = {p % g.nx, (p / g.nx) % g.ny, p / (g.nx * g.ny)};
    for (int d = 0; d < 3; ++d) coords[3 * u + d] = g.origin[d] + g.spacing[d] * (float)ijk[d];
    values[u] = field[p];
  }
  std::vector<Id> used(conn);
  omp_sort(used.begin(), used.end());
  used.erase(std::unique(used.begin(), used.end()), used.end());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < 8 * T; ++i)
    conn[i] = (Id)(std::lower_bound(used.begin(), used.end(), conn[i]) - used.begin());
  *n_out_points = U;
  return wall_seconds() - t0;
}

}  // extern "C"
