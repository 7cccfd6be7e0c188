// Sanitizer run of the oracle (test infrastructure only): every exported
// function of mc_oracle.cpp on small grids, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by `make -C oracle sanitize`.  A report from
// either sanitizer, a failed return code or a wrong known answer exits
// non-zero.  The known answers are those of the reference's unit tests:
// 26^3 points, f = x + y + z, iso 70 gives 45 merged points
// (dax/worklet/testing/UnitTestWorkletMarchingCubes.cxx:39-40,161-200), and
// the same field thresholded to [70, 82] keeps cells with all corners in
// range (dax/worklet/testing/UnitTestWorkletThreshold.cxx:38-243).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
void oracle_tables_expand(uint8_t* num_faces, uint8_t* tri);
int oracle_mc_count(const float* field, const int64_t* dims, float iso, uint8_t* cases,
                    int32_t* counts);
int64_t oracle_scan_inclusive_i32(const int32_t* in, int32_t* out, int64_t n);
int64_t oracle_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n);
void oracle_lower_bounds_i32(const int32_t* sorted, int64_t n, const int32_t* vals, int64_t m,
                             int32_t* out);
void oracle_upper_bounds_i32(const int32_t* sorted, int64_t n, const int32_t* vals, int64_t m,
                             int32_t* out);
int oracle_mc_generate(const float* field, const int64_t* dims, const float* origin,
                       const float* spacing, float iso, int dedup, int64_t cap_tris,
                       int64_t cap_pts, int64_t* n_tris, int64_t* n_pts, int32_t* conn,
                       int32_t* edge_first, int32_t* edge_second, float* weight, float* coords);
void oracle_interpolate_scalar(const float* in, const int32_t* first, const int32_t* second,
                               const float* w, int64_t n, float* out);
void oracle_point_gradient(const float* field, const int64_t* dims, const float* spacing,
                           float* grad);
void oracle_mc_normals(const float* field, const int64_t* dims, const float* spacing,
                       const int32_t* first, const int32_t* second, const float* w, int64_t n,
                       float* normals);
void oracle_cell_gradient(const float* field, const int64_t* dims, const float* origin,
                          const float* spacing, float* out);
float oracle_slice_field(const int64_t* dims, const float* origin, const float* spacing,
                         const float* plane_origin, const float* plane_normal, float* out);
void oracle_tet_tables(uint8_t* num_faces, uint8_t* tri);
int oracle_mt_generate(const int32_t* tets, int64_t n_tets, const float* field,
                       const float* points, float iso, int dedup, int64_t cap, int64_t* n_tris,
                       int64_t* n_pts, int32_t* conn, int32_t* first, int32_t* second,
                       float* weight, float* coords);
int oracle_mc_hex_generate(const int32_t* hexes, int64_t n_hexes, const float* field,
                           const float* points, float iso, int dedup, int64_t cap,
                           int64_t* n_tris, int64_t* n_pts, int32_t* conn, int32_t* first,
                           int32_t* second, float* weight, float* coords);
int oracle_threshold(const int32_t* cells, int nv, int64_t n_cells, const int64_t* dims,
                     const float* origin, const float* spacing, const float* points,
                     int64_t n_points, const float* field, float lo, float hi, int dedup,
                     int64_t cap_cells, int32_t* counts, int64_t* n_out_cells,
                     int64_t* n_out_points, int32_t* conn, float* coords, int32_t* mask);
int oracle_write_vtk(const char* path, int polydata, const float* pts, int64_t n_points,
                     const int32_t* conn, int64_t n_cells, int nv, int cell_type);
void oracle_mc_backward_f64(const float* field, const int64_t* dims, const float* origin,
                            const float* spacing, float iso, const int32_t* first,
                            const int32_t* second, int64_t n, const float* dldx,
                            const float* dldn, double* dldf);
void oracle_mc_fd_f64(const float* field, const int64_t* dims, const float* origin,
                      const float* spacing, float iso, const int32_t* first,
                      const int32_t* second, int64_t n, const float* dldx, const float* dldn,
                      const int64_t* points, int64_t npoints, double h, double* out);
void oracle_mc_backward_f32(const float* field, const int64_t* dims, const float* origin,
                            const float* spacing, float iso, const int32_t* first,
                            const int32_t* second, int64_t n, const float* dldx,
                            const float* dldn, float* dldf);
double oracle_mc_pipeline_omp(const float* field, const int64_t* dims, const float* origin,
                              const float* spacing, float iso, int dedup, int threads,
                              int64_t* n_tris, int64_t* n_pts);
double oracle_unstructured_pipeline_omp(const int32_t* cells, int nv, int64_t n_cells,
                                        const float* field, const float* points, float iso,
                                        int dedup, int threads, int64_t* n_tris, int64_t* n_pts);
double oracle_threshold_pipeline_omp(const int64_t* dims, const float* origin,
                                     const float* spacing, const float* field, float lo,
                                     float hi, int dedup, int threads, int64_t* n_out_cells,
                                     int64_t* n_out_points);
}

namespace {

int g_fail = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "sanitize: check failed: %s\n", what);
    g_fail = 1;
  }
}

struct Grid {
  int64_t dims[3];
  std::vector<float> f;
  int64_t points() const { return dims[0] * dims[1] * dims[2]; }
  int64_t cells() const { return (dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1); }
};

Grid linear_grid(int64_t n) {  // f = x + y + z
  Grid g{{n, n, n}, {}};
  g.f.resize((size_t)g.points());
  for (int64_t k = 0; k < n; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) g.f[(size_t)(i + n * (j + n * k))] = (float)(i + j + k);
  return g;
}

Grid random_grid(int64_t nx, int64_t ny, int64_t nz, unsigned seed) {
  Grid g{{nx, ny, nz}, {}};
  g.f.resize((size_t)g.points());
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  for (auto& v : g.f) v = u(rng);
  return g;
}

// Voxels of a uniform grid as cells of `split` (rows of corner indices in
// the voxel's corner order) and the points' coordinates.
void voxel_cells(const Grid& g, const std::vector<std::vector<int>>& split,
                 std::vector<int32_t>& cells, std::vector<float>& pts) {
  const int64_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                            {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int64_t k = 0; k + 1 < nz; ++k)
    for (int64_t j = 0; j + 1 < ny; ++j)
      for (int64_t i = 0; i + 1 < nx; ++i)
        for (const auto& s : split)
          for (int v : s)
            cells.push_back((int32_t)((i + corner[v][0]) +
                                      nx * ((j + corner[v][1]) + ny * (k + corner[v][2]))));
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        pts.push_back((float)i);
        pts.push_back((float)j);
        pts.push_back((float)k);
      }
}

struct Contour {
  int64_t T = 0, U = 0;
  std::vector<int32_t> conn, first, second;
  std::vector<float> weight, coords;
};

Contour generate(const Grid& g, float iso, int dedup, const float* origin, const float* spacing) {
  std::vector<uint8_t> cases((size_t)std::max<int64_t>(g.cells(), 1));
  std::vector<int32_t> counts(cases.size());
  check(oracle_mc_count(g.f.data(), g.dims, iso, cases.data(), counts.data()) == 0, "count");
  int64_t T = 0;
  for (int64_t c = 0; c < g.cells(); ++c) T += counts[(size_t)c];
  Contour r;
  const size_t cap = (size_t)std::max<int64_t>(3 * T, 1);
  r.conn.resize(cap);
  r.first.resize(cap);
  r.second.resize(cap);
  r.weight.resize(cap);
  r.coords.resize(3 * cap);
  check(oracle_mc_generate(g.f.data(), g.dims, origin, spacing, iso, dedup, T, (int64_t)cap, &r.T,
                           &r.U, r.conn.data(), r.first.data(), r.second.data(), r.weight.data(),
                           r.coords.data()) == 0,
        "generate");
  check(r.T == T, "generate triangle count");
  return r;
}

}  // namespace

int main() {
  std::vector<uint8_t> nf(256), tri(256 * 16);
  oracle_tables_expand(nf.data(), tri.data());
  int sum = 0;
  for (int c = 0; c < 256; ++c) sum += nf[(size_t)c];
  check(sum == 820, "table face sum");
  std::vector<uint8_t> tnf(16), ttri(16 * 8);
  oracle_tet_tables(tnf.data(), ttri.data());

  // scans and bounds: the golden vectors of TestingDeviceAdapter.h:621-717
  {
    const int32_t in[8] = {500, 955, 955, 120, 320, 955, 986, 281};
    std::vector<int32_t> uniq(in, in + 8), out(8), lo(8), hi(8);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    const int64_t nu = (int64_t)uniq.size();
    oracle_lower_bounds_i32(uniq.data(), nu, in, 8, lo.data());
    oracle_upper_bounds_i32(uniq.data(), nu, in, 8, hi.data());
    const int32_t want_lo[8] = {3, 4, 4, 0, 2, 4, 5, 1}, want_hi[8] = {4, 5, 5, 1, 3, 5, 6, 2};
    for (int i = 0; i < 8; ++i) check(lo[i] == want_lo[i] && hi[i] == want_hi[i], "bounds");
    check(oracle_scan_inclusive_i32(in, out.data(), 8) == 5072, "scan inclusive");
    check(oracle_scan_exclusive_i32(in, out.data(), 8) == 5072 && out[0] == 0, "scan exclusive");
    check(oracle_scan_inclusive_i32(in, out.data(), 0) == 0, "empty scan");
  }

  const float org0[3] = {0, 0, 0}, sp1[3] = {1, 1, 1};
  const float org[3] = {-1.5f, 0.25f, 2.0f}, sp[3] = {0.5f, 1.25f, 2.0f};

  // the reference unit test: 26^3, x + y + z, iso 70
  Grid lin = linear_grid(26);
  Contour merged = generate(lin, 70.0f, 1, org0, sp1);
  Contour plain = generate(lin, 70.0f, 0, org0, sp1);
  check(merged.U == 45, "45 merged points");
  check(plain.U == 3 * plain.T, "3T points without merging");
  {
    std::vector<float> vals((size_t)merged.U);
    oracle_interpolate_scalar(lin.f.data(), merged.first.data(), merged.second.data(),
                              merged.weight.data(), merged.U, vals.data());
    for (float v : vals) check(std::fabs(v - 70.0f) < 1e-4f, "interpolated field equals iso");
  }

  // a random ragged field with origin and spacing, both dedup modes
  Grid rnd = random_grid(37, 11, 9, 7);
  Contour rc = generate(rnd, 0.1f, 1, org, sp);
  Contour rp = generate(rnd, 0.1f, 0, org, sp);
  check(rc.T == rp.T && rc.T > 0, "random field triangles");
  {
    std::vector<float> grad((size_t)(3 * rnd.points())), nrm((size_t)(3 * rc.U));
    oracle_point_gradient(rnd.f.data(), rnd.dims, sp, grad.data());
    oracle_mc_normals(rnd.f.data(), rnd.dims, sp, rc.first.data(), rc.second.data(),
                      rc.weight.data(), rc.U, nrm.data());
    std::vector<float> cg((size_t)(3 * rnd.cells()));
    oracle_cell_gradient(rnd.f.data(), rnd.dims, org, sp, cg.data());
    for (float v : nrm) check(std::isfinite(v), "finite normals");
  }
  // degenerate grids: a single plane, a single point
  {
    Grid flat = random_grid(9, 7, 1, 3);
    Contour c = generate(flat, 0.0f, 1, org0, sp1);
    check(c.T == 0 && c.U == 0, "no cells, no output");
    Grid one = random_grid(1, 1, 1, 4);
    Contour d = generate(one, 0.0f, 0, org0, sp1);
    check(d.T == 0, "one point");
  }

  // slice field
  {
    Grid s{{12, 9, 7}, {}};
    s.f.resize((size_t)s.points());
    const float po[3] = {4.0f, 3.0f, 2.5f}, pn[3] = {0.36f, 0.48f, 0.8f};
    const float iso = oracle_slice_field(s.dims, org0, sp1, po, pn, s.f.data());
    Contour c = generate(s, iso, 1, org0, sp1);
    check(c.T > 0, "slice output");
  }

  // marching tetrahedra and hexahedra on the voxels of the random field
  {
    std::vector<int32_t> tets, hexes;
    std::vector<float> pts_t, pts_h;
    voxel_cells(rnd, {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 5, 6, 4}, {1, 5, 6, 0},
                      {1, 6, 2, 0}},
                tets, pts_t);
    voxel_cells(rnd, {{0, 1, 2, 3, 4, 5, 6, 7}}, hexes, pts_h);
    const int64_t nt = (int64_t)tets.size() / 4, nh = (int64_t)hexes.size() / 8;
    for (int dedup = 0; dedup < 2; ++dedup) {
      const size_t cap_t = (size_t)(2 * nt * 3), cap_h = (size_t)(5 * nh * 3);
      std::vector<int32_t> conn(cap_t), first(cap_t), second(cap_t);
      std::vector<float> w(cap_t), xyz(3 * cap_t);
      int64_t T = 0, U = 0;
      check(oracle_mt_generate(tets.data(), nt, rnd.f.data(), pts_t.data(), 0.1f, dedup,
                               (int64_t)cap_t, &T, &U, conn.data(), first.data(), second.data(),
                               w.data(), xyz.data()) == 0,
            "tetrahedra");
      int64_t To = 0, Uo = 0;
      oracle_unstructured_pipeline_omp(tets.data(), 4, nt, rnd.f.data(), pts_t.data(), 0.1f,
                                       dedup, 2, &To, &Uo);
      check(To == T && Uo == U, "OpenMP tetrahedra equal the serial ones");
      std::vector<int32_t> hconn(cap_h), hfirst(cap_h), hsecond(cap_h);
      std::vector<float> hw(cap_h), hxyz(3 * cap_h);
      check(oracle_mc_hex_generate(hexes.data(), nh, rnd.f.data(), pts_h.data(), 0.1f, dedup,
                                   (int64_t)cap_h, &T, &U, hconn.data(), hfirst.data(),
                                   hsecond.data(), hw.data(), hxyz.data()) == 0,
            "hexahedra");
      oracle_unstructured_pipeline_omp(hexes.data(), 8, nh, rnd.f.data(), pts_h.data(), 0.1f,
                                       dedup, 2, &To, &Uo);
      check(To == T && Uo == U, "OpenMP hexahedra equal the serial ones");
      const Contour& ref = dedup ? rc : rp;
      Contour unit = generate(rnd, 0.1f, dedup, org0, sp1);
      check(T == ref.T && U == unit.U, "hexahedra equal the uniform grid");
    }
    // threshold: uniform and unstructured (hexahedra)
    std::vector<int32_t> counts((size_t)nh), tconn((size_t)(8 * nh)), mask((size_t)rnd.points());
    std::vector<float> tcoords((size_t)(3 * rnd.points()));
    int64_t oc = 0, op = 0, oc2 = 0, op2 = 0;
    check(oracle_threshold(nullptr, 8, nh, rnd.dims, org0, sp1, nullptr, rnd.points(),
                           rnd.f.data(), -0.5f, 0.5f, 1, nh, counts.data(), &oc, &op,
                           tconn.data(), tcoords.data(), mask.data()) == 0,
          "threshold uniform");
    const int64_t zdims[3] = {0, 0, 0};
    check(oracle_threshold(hexes.data(), 8, nh, zdims, org0, sp1, pts_h.data(), rnd.points(),
                           rnd.f.data(), -0.5f, 0.5f, 1, nh, counts.data(), &oc2, &op2,
                           tconn.data(), tcoords.data(), mask.data()) == 0,
          "threshold hexahedra");
    check(oc == oc2 && op == op2, "threshold uniform equals hexahedra");
    // the reference's threshold test field: x + y + z in [70, 82] on 26^3
    std::vector<int32_t> lc((size_t)lin.cells()), lconn((size_t)(8 * lin.cells())),
        lmask((size_t)lin.points());
    std::vector<float> lcoords((size_t)(3 * lin.points()));
    check(oracle_threshold(nullptr, 8, lin.cells(), lin.dims, org0, sp1, nullptr, lin.points(),
                           lin.f.data(), 70.0f, 82.0f, 1, lin.cells(), lc.data(), &oc, &op,
                           lconn.data(), lcoords.data(), lmask.data()) == 0,
          "threshold 26^3");
    check(oc > 0, "threshold 26^3 keeps cells");
    int64_t oc3 = 0, op3 = 0;
    oracle_threshold_pipeline_omp(lin.dims, org0, sp1, lin.f.data(), 70.0f, 82.0f, 1, 2, &oc3,
                                  &op3);
    check(oc3 == oc && op3 == op, "OpenMP threshold equals the serial one");
    // VTK writer
    check(oracle_write_vtk("/tmp/dax_oracle_sanitize.vtk", 1, rc.coords.data(), rc.U,
                           rc.conn.data(), rc.T, 3, 5) == 0,
          "vtk polydata");
    std::remove("/tmp/dax_oracle_sanitize.vtk");
  }

  // adjoint: FP64, FP32 and FP64 central differences at a few points
  {
    std::mt19937 rng(9);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    std::vector<float> dldx((size_t)(3 * rc.U)), dldn((size_t)(3 * rc.U));
    for (auto& v : dldx) v = nd(rng);
    for (auto& v : dldn) v = nd(rng);
    std::vector<double> g64((size_t)rnd.points());
    std::vector<float> g32((size_t)rnd.points());
    oracle_mc_backward_f64(rnd.f.data(), rnd.dims, org, sp, 0.1f, rc.first.data(),
                           rc.second.data(), rc.U, dldx.data(), dldn.data(), g64.data());
    oracle_mc_backward_f32(rnd.f.data(), rnd.dims, org, sp, 0.1f, rc.first.data(),
                           rc.second.data(), rc.U, dldx.data(), dldn.data(), g32.data());
    double scale = 1e-30, worst = 0.0;
    for (size_t i = 0; i < g64.size(); ++i) {
      scale = std::max(scale, std::fabs(g64[i]));
      worst = std::max(worst, std::fabs(g64[i] - (double)g32[i]));
      check(std::isfinite(g64[i]) && std::isfinite(g32[i]), "finite adjoint");
    }
    check(worst <= 1e-3 * scale, "FP32 adjoint near the FP64 one");
    const int64_t pts[3] = {0, rnd.points() / 2, rnd.points() - 1};
    double fd[3];
    oracle_mc_fd_f64(rnd.f.data(), rnd.dims, org, sp, 0.1f, rc.first.data(), rc.second.data(),
                     rc.U, dldx.data(), dldn.data(), pts, 3, 1e-5, fd);
    for (double v : fd) check(std::isfinite(v), "finite differences");
  }

  // the OpenMP pipeline on two threads
  {
    int64_t T = 0, U = 0;
    oracle_mc_pipeline_omp(lin.f.data(), lin.dims, org0, sp1, 70.0f, 1, 2, &T, &U);
    check(T == merged.T && U == merged.U, "OpenMP pipeline equals the serial one");
  }

  if (g_fail) return 1;
  std::printf("sanitize ok\n");
  return 0;
}
