// Marching tetrahedra on an unstructured tetrahedral grid: the reference's
// MarchingTetrahedraCount / MarchingTetrahedraGenerate
// (dax/worklet/MarchingTetrahedra.h:29-143) dispatched through
// DispatcherMapCell and DispatcherGenerateInterpolatedCells
// (dax/cont/DispatcherGenerateInterpolatedCells.h:138-301).
//   count    : a thread per tetrahedron: case = sum_v (f_v > iso) << v,
//              NumFaces[case]; tetrahedra with an id outside the points are
//              counted as 0 and reported
//   scan     : exclusive scan of the counts (first triangle per tetrahedron)
//   generate : a thread per tetrahedron writes its triangles' corners:
//              edge from the table, endpoints ordered by point id,
//              weight (iso - f_A) / (f_B - f_A)
//   dedup    : the general path (no grid structure to number edges by):
//              (A << 32 | B) keys sorted, unique, each corner's id by lower
//              bound, as the reference's Sort / Unique / LowerBounds on
//              (EdgeIdFirst, EdgeIdSecond) (:261-301)
//   points   : Lerp(coords[A], coords[B], w) (CompactPointField, :107-141)
#include "common.h"

namespace daxhip {
namespace mt {

// Per case: number of triangles (bits 0-1) and up to 6 edge numbers (3 bits
// each from bit 2) of MarchingTetrahedraTable.h:28-54, in the table's order.
__constant__ uint32_t c_tet_cases[16] = {0,      609,    1057, 314246, 1349, 706914,
                                         558250, 1197,   1421, 726026, 476450, 677,
                                         285070, 385,    833,  0};
// Endpoints (cell vertices) of the 6 tetrahedron edges (MarchingTetrahedra.h:110-113).
__constant__ uint8_t c_tet_edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 1}, {3, 2}};

__global__ void k_mt_count(const int32_t* __restrict__ tets, int64_t n_tets,
                           const float* __restrict__ field, int64_t n_points, float iso,
                           int32_t* __restrict__ counts, int32_t* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tets) return;
  int id[4];
  bool ok = true;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    id[v] = tets[4 * t + v];
    ok = ok && id[v] >= 0 && id[v] < n_points;
  }
  if (!ok) {
    counts[t] = 0;
    *bad = 1;
    return;
  }
  int c = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) c |= (field[id[v]] > iso) << v;
  counts[t] = (int32_t)(c_tet_cases[c] & 3u);
}

__device__ __forceinline__ void lerp3(const float* __restrict__ pts, int64_t a, int64_t b, float w,
                                      float* out) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float pa = pts[3 * a + d], pb = pts[3 * b + d];
    out[d] = pa + w * (pb - pa);
  }
}

// Corners of every triangle: with keys, the (A << 32 | B) key of each
// corner (dedup); without, the corner's point record, coordinates and
// identity connectivity.
__global__ void k_mt_generate(const int32_t* __restrict__ tets, int64_t n_tets,
                              const float* __restrict__ field, const float* __restrict__ pts,
                              float iso, const int32_t* __restrict__ counts,
                              const int32_t* __restrict__ offs, int64_t* __restrict__ keys,
                              int32_t* __restrict__ first, int32_t* __restrict__ second,
                              float* __restrict__ weight, float* __restrict__ verts,
                              int32_t* __restrict__ conn) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tets) return;
  const int nf = counts[t];
  if (nf == 0) return;
  int id[4];
  float f[4];
  int c = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    id[v] = tets[4 * t + v];
    f[v] = field[id[v]];
    c |= (f[v] > iso) << v;
  }
  const uint32_t code = c_tet_cases[c];
  const int64_t base = 3 * (int64_t)offs[t];
  for (int k = 0; k < 3 * nf; ++k) {
    const int e = (int)((code >> (2 + 3 * k)) & 7u);
    int va = c_tet_edge[e][0], vb = c_tet_edge[e][1];
    if (!(id[va] < id[vb])) {
      const int s = va;
      va = vb;
      vb = s;
    }
    const int64_t i = base + k;
    if (keys) {
      keys[i] = ((int64_t)id[va] << 32) | (int64_t)(uint32_t)id[vb];
    } else {
      const float w = (iso - f[va]) / (f[vb] - f[va]);
      first[i] = id[va];
      second[i] = id[vb];
      weight[i] = w;
      float xyz[3];
      lerp3(pts, id[va], id[vb], w, xyz);
      verts[3 * i + 0] = xyz[0];
      verts[3 * i + 1] = xyz[1];
      verts[3 * i + 2] = xyz[2];
      conn[i] = (int32_t)i;
    }
  }
}

// The unique points (sorted keys): records, weights and coordinates.
__global__ void k_mt_points(const int64_t* __restrict__ ukeys, int64_t n,
                            const float* __restrict__ field, const float* __restrict__ pts,
                            float iso, int32_t* __restrict__ first, int32_t* __restrict__ second,
                            float* __restrict__ weight, float* __restrict__ verts) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int64_t key = ukeys[u];
  const int32_t a = (int32_t)(key >> 32), b = (int32_t)(key & 0xffffffff);
  const float fa = field[a], fb = field[b];
  const float w = (iso - fa) / (fb - fa);
  if (first) first[u] = a;
  if (second) second[u] = b;
  if (weight) weight[u] = w;
  float xyz[3];
  lerp3(pts, a, b, w, xyz);
  verts[3 * u + 0] = xyz[0];
  verts[3 * u + 1] = xyz[1];
  verts[3 * u + 2] = xyz[2];
}

}  // namespace mt
}  // namespace daxhip

using namespace daxhip;
using namespace daxhip::mt;

extern "C" {

int dax_hip_mt_count(const int32_t* d_tets, int64_t n_tets, const float* d_field,
                     int64_t n_points, float iso, int32_t* d_counts, int32_t* d_offsets,
                     int64_t* h_tris, void* stream) {
  if (n_tets < 0 || n_points < 0 || n_points > 0x7fffffffLL || !h_tris ||
      (n_tets > 0 && (!d_tets || !d_field || !d_counts || !d_offsets))) {
    set_error("dax_hip_mt_count: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  *h_tris = 0;
  if (n_tets == 0) return DAX_HIP_OK;
  hipStream_t s = (hipStream_t)stream;
  struct Flag {  // device flag of out-of-range ids, freed on every return
    int32_t* p = nullptr;
    ~Flag() {
      if (p) (void)hipFree(p);
    }
  } bad_flag;
  DAXHIP_CHECK(hipMalloc((void**)&bad_flag.p, sizeof(int32_t)));
  DAXHIP_CHECK(hipMemsetAsync(bad_flag.p, 0, sizeof(int32_t), s));
  k_mt_count<<<dim3((unsigned)ceil_div(n_tets, 256)), dim3(256), 0, s>>>(
      d_tets, n_tets, d_field, n_points, iso, d_counts, bad_flag.p);
  DAXHIP_LAUNCH_CHECK("k_mt_count");
  int32_t total = 0, bad = 0;
  int rc = dax_hip_scan_exclusive_i32(d_counts, d_offsets, n_tets, &total, stream);
  if (rc) return rc;
  DAXHIP_CHECK(hipMemcpyAsync(&bad, bad_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  if (bad) {
    set_error("dax_hip_mt_count: a tetrahedron references a point outside [0, %lld)",
              (long long)n_points);
    return DAX_HIP_BAD_VALUE;
  }
  *h_tris = total;
  return DAX_HIP_OK;
}

size_t dax_hip_mt_scratch_bytes(int64_t n_tris) {
  return n_tris > 0 ? (size_t)n_tris * 3 * 2 * sizeof(int64_t) : 0;
}

int dax_hip_mt_generate(const int32_t* d_tets, int64_t n_tets, const float* d_field,
                        const float* d_points, int64_t n_points, float iso, unsigned flags,
                        const int32_t* d_counts, const int32_t* d_offsets, int64_t n_tris,
                        void* d_scratch, size_t scratch_bytes, float* d_vertices,
                        int32_t* d_conn, int32_t* d_first, int32_t* d_second, float* d_weight,
                        int64_t* h_points, void* stream) {
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  if (n_tets < 0 || n_tris < 0 || 3 * n_tris > 0x7fffffffLL || !h_points ||
      (n_tris > 0 && (!d_tets || !d_field || !d_points || !d_counts || !d_offsets ||
                      !d_vertices || !d_conn)) ||
      (!dedup && n_tris > 0 && (!d_first || !d_second || !d_weight)) ||
      (dedup && n_tris > 0 && (!d_scratch || scratch_bytes < dax_hip_mt_scratch_bytes(n_tris)))) {
    set_error("dax_hip_mt_generate: bad arguments (dedup needs %zu bytes of scratch)",
              dax_hip_mt_scratch_bytes(n_tris));
    return DAX_HIP_BAD_VALUE;
  }
  *h_points = 0;
  if (n_tris == 0) return DAX_HIP_OK;
  hipStream_t s = (hipStream_t)stream;
  const dim3 gt((unsigned)ceil_div(n_tets, 256)), b(256);
  if (!dedup) {
    k_mt_generate<<<gt, b, 0, s>>>(d_tets, n_tets, d_field, d_points, iso, d_counts, d_offsets,
                                   nullptr, d_first, d_second, d_weight, d_vertices, d_conn);
    DAXHIP_LAUNCH_CHECK("k_mt_generate");
    *h_points = 3 * n_tris;
    return DAX_HIP_OK;
  }
  const int64_t n = 3 * n_tris;
  int64_t* keys = (int64_t*)d_scratch;
  int64_t* sorted = keys + n;
  k_mt_generate<<<gt, b, 0, s>>>(d_tets, n_tets, d_field, d_points, iso, d_counts, d_offsets,
                                 keys, nullptr, nullptr, nullptr, nullptr, nullptr);
  DAXHIP_LAUNCH_CHECK("k_mt_generate");
  DAXHIP_CHECK(hipMemcpyAsync(sorted, keys, (size_t)n * sizeof(int64_t),
                              hipMemcpyDeviceToDevice, s));
  int rc = dax_hip_sort_i64(sorted, n, 0, stream);
  if (rc) return rc;
  int64_t u = 0;
  if ((rc = dax_hip_unique_i64(sorted, n, &u, stream))) return rc;
  if ((rc = dax_hip_lower_bounds_i64(sorted, u, keys, n, 0, d_conn, nullptr, stream))) return rc;
  k_mt_points<<<dim3((unsigned)ceil_div(u, 256)), b, 0, s>>>(sorted, u, d_field, d_points, iso,
                                                              d_first, d_second, d_weight,
                                                              d_vertices);
  DAXHIP_LAUNCH_CHECK("k_mt_points");
  *h_points = u;
  return DAX_HIP_OK;
}

}  // extern "C"
