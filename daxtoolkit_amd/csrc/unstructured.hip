// Isosurfaces of unstructured grids of one cell shape: the reference's
// MarchingTetrahedraCount / MarchingTetrahedraGenerate
// (dax/worklet/MarchingTetrahedra.h:29-143) on UnstructuredGrid<Tetrahedron>
// and MarchingCubesCount / MarchingCubesGenerate (dax/worklet/MarchingCubes.h:
// 53-164) on UnstructuredGrid<Hexahedron>, dispatched through
// DispatcherMapCell and DispatcherGenerateInterpolatedCells
// (dax/cont/DispatcherGenerateInterpolatedCells.h:138-301).
//   count    : a thread per cell: case = sum_v (f_v > iso) << v,
//              NumFaces[case]; cells with an id outside the points are
//              counted as 0 and reported
//   scan     : exclusive scan of the counts (first triangle per cell)
//   generate : a thread per cell writes its triangles' corners: edge from
//              the table, endpoints ordered by point id, weight
//              (iso - f_A) / (f_B - f_A)
//   dedup    : no grid structure to number edges by, so the general path:
//              each corner's edge as the key A * P + B (P points; the order
//              of the keys is the order of the (A, B) pairs), keys sorted
//              over their significant bits, the unique keys numbered by a
//              scan of the run starts, each corner's point the number of
//              its key's first position -- the reference's Sort / Unique /
//              LowerBounds on (EdgeIdFirst, EdgeIdSecond) (:261-301)
//   points   : Lerp(coords[A], coords[B], w) (CompactPointField, :107-141)
#include "algorithms.h"
#include "mc_case_table.h"

namespace daxhip {
namespace ucell {

// Tetrahedra: per case the number of triangles (bits 0-1) and up to 6 edge
// numbers (3 bits each from bit 2) of MarchingTetrahedraTable.h:28-54, in
// the table's order.
__constant__ uint32_t c_tet_cases[16] = {0,      609,    1057, 314246, 1349, 706914,
                                         558250, 1197,   1421, 726026, 476450, 677,
                                         285070, 385,    833,  0};
// Endpoints (cell vertices) of the 6 tetrahedron edges (MarchingTetrahedra.h:110-113).
__constant__ uint8_t c_tet_edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 1}, {3, 2}};

// Hexahedra: the marching-cubes case table (packed as in mc_case_table.h)
// and the endpoints of the 12 hexahedron edges (MarchingCubes.h:131-135).
__constant__ uint64_t c_hex_cases[256] = DAX_HIP_MC_CASE_TABLE_INIT;
__constant__ uint8_t c_hex_edge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                          {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

template <int NV>
struct Shape;

template <>
struct Shape<4> {
  __device__ static int faces(int c) { return (int)(c_tet_cases[c] & 3u); }
  __device__ static int edge(int c, int k) { return (int)((c_tet_cases[c] >> (2 + 3 * k)) & 7u); }
  __device__ static int end0(int e) { return c_tet_edge[e][0]; }
  __device__ static int end1(int e) { return c_tet_edge[e][1]; }
};

template <>
struct Shape<8> {
  __device__ static int faces(int c) { return (int)(c_hex_cases[c] >> 60); }
  __device__ static int edge(int c, int k) { return (int)((c_hex_cases[c] >> (4 * k)) & 15u); }
  __device__ static int end0(int e) { return c_hex_edge[e][0]; }
  __device__ static int end1(int e) { return c_hex_edge[e][1]; }
};

// The NV point ids of cell t; with kVec the rows are read as 16-byte loads
// (the connectivity is 16-byte aligned).
template <int NV, bool kVec>
__device__ __forceinline__ void load_ids(const int32_t* __restrict__ cells, int64_t t, int* id) {
  if (kVec) {
    const int4* p = (const int4*)(cells + NV * t);
#pragma unroll
    for (int q = 0; q < NV / 4; ++q) {
      const int4 v = p[q];
      id[4 * q + 0] = v.x;
      id[4 * q + 1] = v.y;
      id[4 * q + 2] = v.z;
      id[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) id[v] = cells[NV * t + v];
  }
}

template <int NV, bool kVec>
__global__ __launch_bounds__(256) void k_count(const int32_t* __restrict__ cells, int64_t n_cells,
                                               const float* __restrict__ field, int64_t n_points,
                                               float iso, int32_t* __restrict__ counts,
                                               int32_t* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_cells) return;
  int id[NV];
  load_ids<NV, kVec>(cells, t, id);
  bool ok = true;
#pragma unroll
  for (int v = 0; v < NV; ++v) ok = ok && id[v] >= 0 && id[v] < n_points;
  if (!ok) {
    counts[t] = 0;
    *bad = 1;
    return;
  }
  int c = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) c |= (field[id[v]] > iso) << v;
  counts[t] = Shape<NV>::faces(c);
}

__device__ __forceinline__ void lerp3(const float* __restrict__ pts, int64_t a, int64_t b, float w,
                                      float* out) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float pa = pts[3 * a + d], pb = pts[3 * b + d];
    out[d] = pa + w * (pb - pa);
  }
}

// Corners of every triangle: with keys, the edge key A * P + B of each
// corner and, after the 3T keys, the corner's index (dedup: the index moves
// with its key through the sort); without, the corner's point record,
// coordinates and identity connectivity.
//
// A thread per output triangle (the reference's UpperBounds and visit index,
// DispatcherGenerateInterpolatedCells.h:196-212): its cell is the last one
// whose exclusive offset is <= the triangle, its visit index the difference.
template <int NV, bool kVec>
__global__ __launch_bounds__(256) void k_generate(
    const int32_t* __restrict__ cells, int64_t n_cells, const float* __restrict__ field,
    const float* __restrict__ pts, int64_t n_points, float iso, int64_t n_tris,
    const int32_t* __restrict__ offs, uint64_t* __restrict__ keys, int32_t* __restrict__ first,
    int32_t* __restrict__ second, float* __restrict__ weight, float* __restrict__ verts,
    int32_t* __restrict__ conn) {
  const int64_t tri = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tri >= n_tris) return;
  int64_t lo = 0, hi = n_cells;  // offs[0] = 0 <= tri
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)offs[mid] <= tri)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t t = lo;
  const int visit = (int)(tri - offs[t]);
  int id[NV];
  load_ids<NV, kVec>(cells, t, id);
  float f[NV];
  int c = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    f[v] = field[id[v]];
    c |= (f[v] > iso) << v;
  }
  if (visit >= Shape<NV>::faces(c)) return;  // offsets not from this field's counts
  const int64_t base = 3 * (int64_t)offs[t];
  for (int k = 3 * visit; k < 3 * visit + 3; ++k) {
    const int e = Shape<NV>::edge(c, k);
    int va = Shape<NV>::end0(e), vb = Shape<NV>::end1(e);
    int ia = 0, ib = 0;
    float fa = 0.0f, fb = 0.0f;
    // registers indexed by a run-time vertex: select instead of a local array
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (v == va) {
        ia = id[v];
        fa = f[v];
      }
      if (v == vb) {
        ib = id[v];
        fb = f[v];
      }
    }
    if (!(ia < ib)) {  // endpoints ordered by point id (ties keep the swap)
      const int si = ia;
      ia = ib;
      ib = si;
      const float sf = fa;
      fa = fb;
      fb = sf;
    }
    const int64_t i = base + k;
    if (keys) {
      keys[i] = (uint64_t)ia * (uint64_t)n_points + (uint64_t)ib;
      ((int32_t*)(keys + 3 * n_tris))[i] = (int32_t)i;
    } else {
      const float w = (iso - fa) / (fb - fa);
      first[i] = ia;
      second[i] = ib;
      weight[i] = w;
      float xyz[3];
      lerp3(pts, ia, ib, w, xyz);
      verts[3 * i + 0] = xyz[0];
      verts[3 * i + 1] = xyz[1];
      verts[3 * i + 2] = xyz[2];
      conn[i] = (int32_t)i;
    }
  }
}

// {triangle total, bad-id flag} for one read-back: the last offset plus the
// last count.
__global__ void k_count_result(const int32_t* __restrict__ offs,
                               const int32_t* __restrict__ counts, int64_t n,
                               const int32_t* __restrict__ bad, int32_t* __restrict__ out2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out2[0] = offs[n - 1] + counts[n - 1];
    out2[1] = *bad;
  }
}

// 1 where a sorted key starts a run of equal keys.
__global__ __launch_bounds__(256) void k_run_starts(const uint64_t* __restrict__ sorted,
                                                    int64_t n, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || sorted[i] != sorted[i - 1]) ? 1 : 0;
}

// Each corner's point (the reference's LowerBounds of its key in the unique
// keys): sorted position i holds corner idx[i], whose point is the number of
// its run, i.e. the run starts up to and including i, less one; `rank` is
// the exclusive scan of the run starts, so a position inside a run (not a
// start) subtracts one.
__global__ __launch_bounds__(256) void k_corner_points(const uint64_t* __restrict__ sorted,
                                                       const int32_t* __restrict__ idx,
                                                       const int32_t* __restrict__ rank, int64_t n,
                                                       int32_t* __restrict__ conn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool start = i == 0 || sorted[i] != sorted[i - 1];
  const int32_t c = idx[i];
  if (c >= 0 && c < n) conn[c] = rank[i] - (start ? 0 : 1);  // always in range: a permutation
}

// The unique points (sorted keys): records, weights and coordinates.
__global__ __launch_bounds__(256) void k_points(const uint64_t* __restrict__ sorted,
                                                const int32_t* __restrict__ rank, int64_t n,
                                                uint64_t n_points, const float* __restrict__ field,
                                                const float* __restrict__ pts, float iso,
                                                int32_t* __restrict__ first,
                                                int32_t* __restrict__ second,
                                                float* __restrict__ weight,
                                                float* __restrict__ verts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (i > 0 && sorted[i] == sorted[i - 1])) return;  // a run's first key only
  const int64_t u = rank[i];
  const uint64_t key = sorted[i];
  const int32_t a = (int32_t)(key / n_points), b = (int32_t)(key - (uint64_t)a * n_points);
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

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <int NV>
int count_cells(const char* name, const int32_t* d_cells, int64_t n_cells, const float* d_field,
                int64_t n_points, float iso, int32_t* d_counts, int32_t* d_offsets,
                int64_t* h_tris, void* stream) {
  if (n_cells < 0 || n_points < 0 || n_points > 0x7fffffffLL || !h_tris ||
      n_cells > 0x7fffffffLL / NV ||
      (n_cells > 0 && (!d_cells || !d_field || !d_counts || !d_offsets))) {
    set_error("%s: bad arguments", name);
    return DAX_HIP_BAD_VALUE;
  }
  *h_tris = 0;
  if (n_cells == 0) return DAX_HIP_OK;
  hipStream_t s = (hipStream_t)stream;
  TempBuf flag(s);  // set when a cell references a point outside [0, n_points)
  int rc = flag.alloc(sizeof(int32_t));
  if (rc) return rc;
  DAXHIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int32_t), s));
  const dim3 g((unsigned)ceil_div(n_cells, 256)), b(256);
  if (aligned16(d_cells))
    k_count<NV, true><<<g, b, 0, s>>>(d_cells, n_cells, d_field, n_points, iso, d_counts,
                                      (int32_t*)flag.p);
  else
    k_count<NV, false><<<g, b, 0, s>>>(d_cells, n_cells, d_field, n_points, iso, d_counts,
                                       (int32_t*)flag.p);
  DAXHIP_LAUNCH_CHECK("k_count");
  if ((rc = alg::scan_exclusive_i32_async(d_counts, d_offsets, n_cells, s))) return rc;
  TempBuf res(s);  // {total, bad} read back together
  if ((rc = res.alloc(2 * sizeof(int32_t)))) return rc;
  k_count_result<<<1, 64, 0, s>>>(d_offsets, d_counts, n_cells, (const int32_t*)flag.p,
                                  (int32_t*)res.p);
  DAXHIP_LAUNCH_CHECK("k_count_result");
  int32_t h_res[2] = {0, 0};
  DAXHIP_CHECK(hipMemcpyAsync(h_res, res.p, sizeof(h_res), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  const int32_t total = h_res[0], bad = h_res[1];
  if (bad) {
    set_error("%s: a cell references a point outside [0, %lld)", name, (long long)n_points);
    return DAX_HIP_BAD_VALUE;
  }
  *h_tris = total;
  return DAX_HIP_OK;
}

inline size_t scratch_bytes(int64_t n_tris) {
  return n_tris > 0 ? (size_t)n_tris * 3 * 2 * sizeof(uint64_t) : 0;
}

template <int NV>
int generate_cells(const char* name, const int32_t* d_cells, int64_t n_cells, const float* d_field,
                   const float* d_points, int64_t n_points, float iso, unsigned flags,
                   const int32_t* d_counts, const int32_t* d_offsets, int64_t n_tris,
                   void* d_scratch, size_t scratch_bytes_, float* d_vertices, int32_t* d_conn,
                   int32_t* d_first, int32_t* d_second, float* d_weight, int64_t* h_points,
                   void* stream) {
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  if (n_cells < 0 || n_tris < 0 || 3 * n_tris > 0x7fffffffLL || !h_points || n_points < 0 ||
      n_points > 0x7fffffffLL || (flags & ~(unsigned)DAX_HIP_MC_DEDUP) ||
      (n_tris > 0 && (!d_cells || !d_field || !d_points || !d_counts || !d_offsets ||
                      !d_vertices || !d_conn)) ||
      (!dedup && n_tris > 0 && (!d_first || !d_second || !d_weight)) ||
      (dedup && n_tris > 0 && (!d_scratch || scratch_bytes_ < scratch_bytes(n_tris)))) {
    set_error("%s: bad arguments (only DAX_HIP_MC_DEDUP; dedup needs %zu bytes of scratch)", name,
              scratch_bytes(n_tris));
    return DAX_HIP_BAD_VALUE;
  }
  *h_points = 0;
  if (n_tris == 0) return DAX_HIP_OK;
  hipStream_t s = (hipStream_t)stream;
  const dim3 gt((unsigned)ceil_div(n_tris, 256)), b(256);
  const bool vec = aligned16(d_cells);
  uint64_t* keys = dedup ? (uint64_t*)d_scratch : nullptr;
#define DAXHIP_UGEN(VEC)                                                                        \
  k_generate<NV, VEC><<<gt, b, 0, s>>>(d_cells, n_cells, d_field, d_points, n_points, iso,      \
                                       n_tris, d_offsets, keys, d_first, d_second, d_weight,    \
                                       d_vertices, d_conn)
  if (vec)
    DAXHIP_UGEN(true);
  else
    DAXHIP_UGEN(false);
#undef DAXHIP_UGEN
  DAXHIP_LAUNCH_CHECK("k_generate");
  if (!dedup) {
    *h_points = 3 * n_tris;
    return DAX_HIP_OK;
  }
  const int64_t n = 3 * n_tris;
  // the keys are sorted in place, carrying their corner indices (generate
  // wrote them to the second half of the scratch)
  uint64_t* sorted = keys;
  int32_t* idx = (int32_t*)(keys + n);
  const dim3 gn((unsigned)ceil_div(n, 256));
  // keys < P^2: sort over the bits that can be set
  const uint64_t kmax = (uint64_t)n_points * (uint64_t)n_points - 1;
  const int bits = kmax ? 64 - __builtin_clzll(kmax) : 1;
  int rc = alg::sort_u64_bits(sorted, n, bits, s, idx);
  if (rc) return rc;
  // unique points: the runs of equal sorted keys, numbered by a scan of the
  // run starts (the reference's Unique); each corner's point is its run's
  // number (the reference's LowerBounds)
  TempBuf rk(s);
  if ((rc = rk.alloc((size_t)n * sizeof(int32_t)))) return rc;
  int32_t* rank = (int32_t*)rk.p;
  k_run_starts<<<gn, b, 0, s>>>(sorted, n, rank);
  DAXHIP_LAUNCH_CHECK("k_run_starts");
  int32_t u = 0;
  if ((rc = dax_hip_scan_exclusive_i32(rank, rank, n, &u, stream))) return rc;
  k_points<<<gn, b, 0, s>>>(sorted, rank, n, (uint64_t)n_points, d_field, d_points, iso, d_first,
                            d_second, d_weight, d_vertices);
  DAXHIP_LAUNCH_CHECK("k_points");
  k_corner_points<<<gn, b, 0, s>>>(sorted, idx, rank, n, d_conn);
  DAXHIP_LAUNCH_CHECK("k_corner_points");
  *h_points = u;
  return DAX_HIP_OK;
}

}  // namespace ucell
}  // namespace daxhip

using namespace daxhip;
using namespace daxhip::ucell;

extern "C" {

int dax_hip_mt_count(const int32_t* d_tets, int64_t n_tets, const float* d_field,
                     int64_t n_points, float iso, int32_t* d_counts, int32_t* d_offsets,
                     int64_t* h_tris, void* stream) {
  return count_cells<4>("dax_hip_mt_count", d_tets, n_tets, d_field, n_points, iso, d_counts,
                        d_offsets, h_tris, stream);
}

size_t dax_hip_mt_scratch_bytes(int64_t n_tris) { return scratch_bytes(n_tris); }

int dax_hip_mt_generate(const int32_t* d_tets, int64_t n_tets, const float* d_field,
                        const float* d_points, int64_t n_points, float iso, unsigned flags,
                        const int32_t* d_counts, const int32_t* d_offsets, int64_t n_tris,
                        void* d_scratch, size_t scratch_bytes_, float* d_vertices,
                        int32_t* d_conn, int32_t* d_first, int32_t* d_second, float* d_weight,
                        int64_t* h_points, void* stream) {
  return generate_cells<4>("dax_hip_mt_generate", d_tets, n_tets, d_field, d_points, n_points, iso,
                           flags, d_counts, d_offsets, n_tris, d_scratch, scratch_bytes_,
                           d_vertices, d_conn, d_first, d_second, d_weight, h_points, stream);
}

int dax_hip_mc_hex_count(const int32_t* d_hexes, int64_t n_hexes, const float* d_field,
                         int64_t n_points, float iso, int32_t* d_counts, int32_t* d_offsets,
                         int64_t* h_tris, void* stream) {
  return count_cells<8>("dax_hip_mc_hex_count", d_hexes, n_hexes, d_field, n_points, iso,
                        d_counts, d_offsets, h_tris, stream);
}

size_t dax_hip_mc_hex_scratch_bytes(int64_t n_tris) { return scratch_bytes(n_tris); }

int dax_hip_mc_hex_generate(const int32_t* d_hexes, int64_t n_hexes, const float* d_field,
                            const float* d_points, int64_t n_points, float iso, unsigned flags,
                            const int32_t* d_counts, const int32_t* d_offsets, int64_t n_tris,
                            void* d_scratch, size_t scratch_bytes_, float* d_vertices,
                            int32_t* d_conn, int32_t* d_first, int32_t* d_second,
                            float* d_weight, int64_t* h_points, void* stream) {
  return generate_cells<8>("dax_hip_mc_hex_generate", d_hexes, n_hexes, d_field, d_points,
                           n_points, iso, flags, d_counts, d_offsets, n_tris, d_scratch,
                           scratch_bytes_, d_vertices, d_conn, d_first, d_second, d_weight,
                           h_points, stream);
}

}  // extern "C"
