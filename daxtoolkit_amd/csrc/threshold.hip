// Threshold: the reference's ThresholdCount (dax/worklet/Threshold.h:30-96,
// a cell passes when every point value v has min <= v <= max) dispatched by
// DispatcherMapCell, and ThresholdTopology (:98-110, the output cell's
// vertices are the input cell's) dispatched by DispatcherGenerateTopology
// (dax/cont/DispatcherGenerateTopology.h:146-256), on a uniform grid (voxels
// become hexahedra with the voxel's corner order) or an unstructured grid of
// one cell shape.
//   count    : per cell 1 or 0 (int32, the reference's dax::Id count)
//   prepare  : exclusive scan of the counts (first output cell of each input
//              cell, the reference's inclusive scan + upper bounds); with
//              duplicate removal the point mask (1 where a kept cell uses the
//              point, FillPointMask :211-229) and its exclusive scan, which
//              is each used point's rank among the used points -- the value
//              the reference gets by Sort/Unique/LowerBounds of the used ids
//              (ResolveDuplicatePoints :231-254)
//   generate : every kept cell writes count x V vertex ids (remapped to the
//              ranks with duplicate removal), every used point its
//              coordinates at its rank (StreamCompact of the coordinates)
// Counts may be any non-negative values (DispatcherGenerateTopology repeats
// a cell count times); Threshold's are 0 or 1.
#include <algorithm>

#include "common.h"

namespace daxhip {
namespace topo {

constexpr int kTX = 64, kTY = 4;  // uniform-grid blocks: 64 x-cells by 4 rows
constexpr int kZRun = 16;         // planes one thread marches through

// How prepare laid out the workspace (its first int): per-cell offsets and
// per-point ranks (any counts, any grid), or bit words (counts 0/1 on a
// uniform grid: a bit per cell and per point, 64 to a word along x, with
// per-word exclusive popcount prefixes).  The kernels of both layouts are
// launched; each returns at once unless the workspace holds its layout.
constexpr int kModeCells = 0, kModeBits = 1;

__device__ __forceinline__ bool in_range(float v, float lo, float hi) {
  return v >= lo && v <= hi;
}

// Uniform grid: thread (i, j) marches over kZRun cell planes keeping the
// lower plane's four in-range flags in registers.
__global__ __launch_bounds__(kTX* kTY) void k_count_uniform(const float* __restrict__ f, int nx,
                                                             int ny, int nz, float lo, float hi,
                                                             int32_t* __restrict__ counts) {
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int i = blockIdx.x * kTX + threadIdx.x, j = blockIdx.y * kTY + threadIdx.y;
  const int k0 = blockIdx.z * kZRun;
  if (i >= cx || j >= cy || k0 >= cz) return;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const float* p = f + i + sy * j + sz * k0;
  bool a00 = in_range(p[0], lo, hi), a10 = in_range(p[1], lo, hi);
  bool a01 = in_range(p[sy], lo, hi), a11 = in_range(p[sy + 1], lo, hi);
  const int k1 = min(k0 + kZRun, cz);
  int32_t* out = counts + i + (int64_t)cx * (j + (int64_t)cy * k0);
  const int64_t cplane = (int64_t)cx * cy;
#pragma unroll 4
  for (int k = k0; k < k1; ++k) {
    p += sz;
    const bool b00 = in_range(p[0], lo, hi), b10 = in_range(p[1], lo, hi);
    const bool b01 = in_range(p[sy], lo, hi), b11 = in_range(p[sy + 1], lo, hi);
    *out = (a00 && a10 && a01 && a11 && b00 && b10 && b01 && b11) ? 1 : 0;
    out += cplane;
    a00 = b00;
    a10 = b10;
    a01 = b01;
    a11 = b11;
  }
}

// The counts of a wave's 256 cells (lane l holds cells 4l .. 4l+3 as bits
// of `kept`) written through the wave's LDS slice so that each store
// instruction covers 64 consecutive cells; n = cells of the 256 in the row.
__device__ __forceinline__ void store_counts256(int32_t* s_w, unsigned kept, int32_t* out, int n) {
  const int lane = threadIdx.x & 63;
  int4 v;
  v.x = (int)(kept & 1u);
  v.y = (int)((kept >> 1) & 1u);
  v.z = (int)((kept >> 2) & 1u);
  v.w = (int)((kept >> 3) & 1u);
  *(int4*)(s_w + 4 * lane) = v;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int x = lane + 64 * q;
    const int32_t c = s_w[x];
    if (x < n) out[x] = c;
  }
  __builtin_amdgcn_wave_barrier();
}

// The same with rows of 16-byte aligned float4s (nx % 4 == 0): lane l of a
// wave covers cells x0 .. x0+3 (x0 = 4 * (64 * blockIdx.x + l)) from one
// float4 per point row and the next lane's first value (lane 63 loads it).
// The loads (load_row5) and the range bits (row5_bits) are separate so that
// a plane's rows can be loaded while the previous plane is processed.
struct Row5 {
  float4 v;
  float tail;  // lane 63: the next lane's first value
};

__device__ __forceinline__ Row5 load_row5(const float* __restrict__ row, int x0, int nx) {
  Row5 r;
  r.v = make_float4(0.f, 0.f, 0.f, 0.f);
  r.tail = 0.f;
  if (x0 < nx) r.v = *(const float4*)(row + x0);
  if ((threadIdx.x & 63) == 63 && x0 + 4 < nx) r.tail = row[x0 + 4];
  return r;
}

__device__ __forceinline__ unsigned row5_bits(const Row5& r, float lo, float hi) {
  float nxt = __shfl_down(r.v.x, 1, 64);
  if ((threadIdx.x & 63) == 63) nxt = r.tail;
  return (unsigned)in_range(r.v.x, lo, hi) | (unsigned)in_range(r.v.y, lo, hi) << 1 |
         (unsigned)in_range(r.v.z, lo, hi) << 2 | (unsigned)in_range(r.v.w, lo, hi) << 3 |
         (unsigned)in_range(nxt, lo, hi) << 4;
}

__device__ __forceinline__ unsigned range_bits5(const float* __restrict__ row, int x0, int nx,
                                                float lo, float hi) {
  return row5_bits(load_row5(row, x0, nx), lo, hi);
}

__global__ __launch_bounds__(kTX* kTY) void k_count_uniform4(const float* __restrict__ f, int nx,
                                                              int ny, int nz, float lo, float hi,
                                                              int32_t* __restrict__ counts) {
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int x0 = 4 * (blockIdx.x * kTX + threadIdx.x), j = blockIdx.y * kTY + threadIdx.y;
  const int k0 = blockIdx.z * kZRun;
  if (j >= cy || k0 >= cz) return;  // whole waves (one row per wave)
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const float* p = f + sy * j + sz * k0;
  // bit q: point x0+q in range in both rows j and j+1 of the plane
  __shared__ __align__(16) int32_t s_cnt[kTY][256];
  unsigned a = range_bits5(p, x0, nx, lo, hi) & range_bits5(p + sy, x0, nx, lo, hi);
  const int k1 = min(k0 + kZRun, cz);
  const int xw = 256 * (int)blockIdx.x;  // the wave's first cell
  int32_t* out = counts + xw + (int64_t)cx * (j + (int64_t)cy * k0);
  const int64_t cplane = (int64_t)cx * cy;
  // point plane k + 1 is loaded while plane k is processed (k + 1 <= k1 <= nz - 1)
  p += sz;
  Row5 n0 = load_row5(p, x0, nx), n1 = load_row5(p + sy, x0, nx);
  for (int k = k0; k < k1; ++k) {
    const unsigned b = row5_bits(n0, lo, hi) & row5_bits(n1, lo, hi);
    if (k + 1 < k1) {
      p += sz;
      n0 = load_row5(p, x0, nx);
      n1 = load_row5(p + sy, x0, nx);
    }
    const unsigned both = a & b, kept = both & (both >> 1);  // cell q: points q and q+1
    store_counts256(s_cnt[threadIdx.y], kept, out, min(256, cx - xw));
    out += cplane;
    a = b;
  }
}

// Bits 0..15 of x moved to bits 0, 4, 8, ..., 60.
__device__ __forceinline__ uint64_t spread16(uint64_t x) {
  x &= 0xFFFFull;
  x = (x | (x << 24)) & 0x000000FF000000FFull;
  x = (x | (x << 12)) & 0x000F000F000F000Full;
  x = (x | (x << 6)) & 0x0303030303030303ull;
  x = (x | (x << 3)) & 0x1111111111111111ull;
  return x;
}

// k_count_uniform4 that also writes the cell words of the bit-word layout: a
// wave covers cells 256 * blockIdx.x .. +255 of its row, i.e. words
// 4 * blockIdx.x .. +3; lane l holds cells 4l .. 4l+3, so word q is the four
// ballots of the lanes' bits, each spread to every fourth bit.
__global__ __launch_bounds__(kTX* kTY) void k_threshold_words4(
    const float* __restrict__ f, int nx, int ny, int nz, float lo, float hi, int wcx,
    int32_t* __restrict__ counts, uint64_t* __restrict__ cbits, int32_t* __restrict__ cpop) {
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int lane = threadIdx.x;
  const int x0 = 4 * (blockIdx.x * kTX + lane), j = blockIdx.y * kTY + threadIdx.y;
  const int k0 = blockIdx.z * kZRun;
  if (j >= cy || k0 >= cz) return;  // whole waves (one row per wave)
  __shared__ __align__(16) int32_t s_cnt[kTY][256];
  const unsigned valid = x0 >= cx ? 0u : (x0 + 4 <= cx ? 0xFu : (1u << (cx - x0)) - 1u);
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const float* p = f + sy * j + sz * k0;
  unsigned a = range_bits5(p, x0, nx, lo, hi) & range_bits5(p + sy, x0, nx, lo, hi);
  const int k1 = min(k0 + kZRun, cz);
  const int xw = 256 * (int)blockIdx.x;  // the wave's first cell
  int32_t* out = counts + xw + (int64_t)cx * (j + (int64_t)cy * k0);
  const int64_t cplane = (int64_t)cx * cy;
  // point plane k + 1 is loaded while plane k is processed (k + 1 <= k1 <= nz - 1)
  p += sz;
  Row5 n0 = load_row5(p, x0, nx), n1 = load_row5(p + sy, x0, nx);
  for (int k = k0; k < k1; ++k) {
    const unsigned b = row5_bits(n0, lo, hi) & row5_bits(n1, lo, hi);
    if (k + 1 < k1) {
      p += sz;
      n0 = load_row5(p, x0, nx);
      n1 = load_row5(p + sy, x0, nx);
    }
    const unsigned both = a & b, kept = both & (both >> 1) & valid;
    store_counts256(s_cnt[threadIdx.y], kept, out, min(256, cx - xw));
    const uint64_t b0 = __ballot(kept & 1u), b1 = __ballot((kept >> 1) & 1u);
    const uint64_t b2 = __ballot((kept >> 2) & 1u), b3 = __ballot((kept >> 3) & 1u);
    const int wx = 4 * (int)blockIdx.x + lane;
    if (lane < 4 && wx < wcx) {
      const int sh = 16 * lane;
      const uint64_t w = spread16(b0 >> sh) | spread16(b1 >> sh) << 1 | spread16(b2 >> sh) << 2 |
                         spread16(b3 >> sh) << 3;
      const int64_t wi = ((int64_t)j + (int64_t)cy * k) * wcx + wx;
      cbits[wi] = w;
      cpop[wi] = __popcll(w);
    }
    out += cplane;
    a = b;
  }
}

// Unstructured grid: a thread per cell of nv vertices (nv <= 8); an id
// outside [0, n_points) sets *bad and counts 0.
__global__ __launch_bounds__(256) void k_count_cells(const int32_t* __restrict__ cells, int nv,
                                                     int64_t n_cells,
                                                     const float* __restrict__ f,
                                                     int64_t n_points, float lo, float hi,
                                                     int32_t* __restrict__ counts,
                                                     int32_t* __restrict__ bad) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  bool ok = true, keep = true;
  for (int v = 0; v < nv; ++v) {
    const int32_t id = cells[c * nv + v];
    if (id < 0 || id >= n_points) {
      ok = false;
      break;
    }
    keep = keep && in_range(f[id], lo, hi);
  }
  if (!ok) *bad = 1;
  counts[c] = ok && keep ? 1 : 0;
}

// Uniform grid point mask: point (i, j, k) is used when one of its (up to 8)
// cells has a non-zero count.  A wave per 64 points of a row (blockDim.x is
// the wave), 4 rows per block, marching over z with the lower cell plane's
// result in a register; per plane each lane reads its cell column x = i of
// rows j-1 and j and takes column i-1 from the lane below (lane 0 reads it).
__global__ __launch_bounds__(kTX* kTY) void k_mask_uniform(const int32_t* __restrict__ counts,
                                                            int nx, int ny, int nz,
                                                            int32_t* __restrict__ mask) {
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int i = blockIdx.x * kTX + threadIdx.x, j = blockIdx.y * kTY + threadIdx.y;
  const int k0 = blockIdx.z * kZRun;
  if (j >= ny || k0 >= nz) return;  // whole waves (one row per wave)
  const bool lane0 = threadIdx.x == 0;
  const int64_t row_lo = (int64_t)cx * (j - 1), row_hi = (int64_t)cx * j;
  const bool has_lo = j - 1 >= 0, has_hi = j < cy;
  // cell columns x of rows j-1 and j of cell plane kc: any non-zero count
  auto column = [&](const int32_t* q, int x) -> int {
    if (x < 0 || x >= cx) return 0;
    int any = 0;
    if (has_lo) any |= q[row_lo + x] != 0;
    if (has_hi) any |= q[row_hi + x] != 0;
    return any;
  };
  auto plane_any = [&](int kc) -> int {
    const bool ok = kc >= 0 && kc < cz;
    const int32_t* q = counts + (int64_t)cx * cy * (ok ? kc : 0);
    const int own = ok ? column(q, i) : 0;
    int left = __shfl_up(own, 1, 64);
    if (lane0) left = ok ? column(q, i - 1) : 0;
    return own | left;
  };
  int below = plane_any(k0 - 1);
  const int k1 = min(k0 + kZRun, nz);
  int32_t* out = mask + i + (int64_t)nx * (j + (int64_t)ny * k0);
  for (int k = k0; k < k1; ++k) {
    const int above = plane_any(k);
    if (i < nx) *out = below | above;
    out += (int64_t)nx * ny;
    below = above;
  }
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Unstructured point mask: the kept cells store 1 at their vertices (every
// store writes the same value); an id outside [0, n_points) sets *bad.
__global__ __launch_bounds__(256) void k_mark_cells(const int32_t* __restrict__ cells, int nv,
                                                    int64_t n_cells, int64_t n_points,
                                                    const int32_t* __restrict__ counts,
                                                    int32_t* __restrict__ mask,
                                                    int32_t* __restrict__ bad) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells || counts[c] <= 0) return;
  for (int v = 0; v < nv; ++v) {
    const int32_t id = cells[c * nv + v];
    if (id >= 0 && id < n_points)
      mask[id] = 1;
    else
      *bad = 1;
  }
}

// Vertex ids of voxel (i, j, k) in the hexahedron's corner order
// (TopologyUniform.h:139-159).
__device__ __forceinline__ void voxel_ids(int64_t c, int nx, int ny, int64_t* ids) {
  const int cx = nx - 1, cy = ny - 1;
  const int64_t i = c % cx, r = c / cx, j = r % cy, k = r / cy;
  const int64_t p = i + (int64_t)nx * (j + (int64_t)ny * k), sy = nx, sz = (int64_t)nx * ny;
  ids[0] = p;
  ids[1] = p + 1;
  ids[2] = p + 1 + sy;
  ids[3] = p + sy;
  ids[4] = p + sz;
  ids[5] = p + 1 + sz;
  ids[6] = p + 1 + sy + sz;
  ids[7] = p + sy + sz;
}

// Every cell with a non-zero count writes its vertex ids count times at its
// first output cell; with `rank` the ids are the used points' ranks.  A
// thread per 4 consecutive cells (one 16-byte load of their counts when
// `vec`: the counts are 16-byte aligned).
__global__ __launch_bounds__(256) void k_emit(const int32_t* __restrict__ mode, int nv,
                                              const int32_t* __restrict__ cells,
                                              int64_t n_cells, int nx, int ny,
                                              const int32_t* __restrict__ counts, bool vec,
                                              const int32_t* __restrict__ offs,
                                              const int32_t* __restrict__ rank,
                                              int32_t* __restrict__ conn) {
  const int64_t c0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (c0 >= n_cells || *mode != kModeCells) return;
  int cnt[4];
  if (vec && c0 + 4 <= n_cells) {
    const int4 v = *(const int4*)(counts + c0);
    cnt[0] = v.x;
    cnt[1] = v.y;
    cnt[2] = v.z;
    cnt[3] = v.w;
  } else {
    for (int q = 0; q < 4; ++q) cnt[q] = c0 + q < n_cells ? counts[c0 + q] : 0;
  }
  for (int q = 0; q < 4; ++q) {
    const int n = cnt[q];
    if (n <= 0) continue;
    const int64_t c = c0 + q;
    int64_t ids[8];
    if (cells) {
      for (int v = 0; v < nv; ++v) ids[v] = cells[c * nv + v];
    } else {
      voxel_ids(c, nx, ny, ids);
    }
    int32_t out[8];
    for (int v = 0; v < nv; ++v) out[v] = rank ? rank[ids[v]] : (int32_t)ids[v];
    int32_t* dst = conn + (int64_t)offs[c] * nv;
    for (int r = 0; r < n; ++r)
      for (int v = 0; v < nv; ++v) dst[(int64_t)r * nv + v] = out[v];
  }
}

// The mask of points p0 .. p0+3 (one 16-byte load when `vec`).
__device__ __forceinline__ void load_mask4(const int32_t* __restrict__ mask, int64_t p0, int64_t n,
                                           bool vec, int* m) {
  if (vec && p0 + 4 <= n) {
    const int4 v = *(const int4*)(mask + p0);
    m[0] = v.x;
    m[1] = v.y;
    m[2] = v.z;
    m[3] = v.w;
  } else {
    for (int q = 0; q < 4; ++q) m[q] = p0 + q < n ? mask[p0 + q] : 0;
  }
}

// Coordinates of the used points at their ranks: from `points`, or on a
// uniform grid origin + spacing * (i, j, k).  A thread per 4 points.
__global__ __launch_bounds__(256) void k_compact_points(const float* __restrict__ points,
                                                        int64_t n_points, int nx, int ny,
                                                        float ox, float oy, float oz, float hx,
                                                        float hy, float hz,
                                                        const int32_t* __restrict__ mask, bool vec,
                                                        const int32_t* __restrict__ rank,
                                                        const int32_t* __restrict__ mode,
                                                        float* __restrict__ out) {
  const int64_t p0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p0 >= n_points || *mode != kModeCells) return;
  int m[4];
  load_mask4(mask, p0, n_points, vec, m);
  for (int q = 0; q < 4; ++q) {
    if (!m[q]) continue;
    const int64_t p = p0 + q, u = rank[p];
    float x, y, z;
    if (points) {
      x = points[3 * p];
      y = points[3 * p + 1];
      z = points[3 * p + 2];
    } else {
      const int64_t i = p % nx, r = p / nx, j = r % ny, k = r / ny;
      x = ox + hx * (float)i;
      y = oy + hy * (float)j;
      z = oz + hz * (float)k;
    }
    out[3 * u] = x;
    out[3 * u + 1] = y;
    out[3 * u + 2] = z;
  }
}

// CompactPointField with the ranks of the prepare step: the used points'
// values, in point order (a StreamCompact by the point mask).  A thread per
// 4 points.
__global__ __launch_bounds__(256) void k_compact_values(const uint32_t* __restrict__ in, int words,
                                                        int64_t n_points,
                                                        const int32_t* __restrict__ mask, bool vec,
                                                        const int32_t* __restrict__ rank,
                                                        const int32_t* __restrict__ mode,
                                                        uint32_t* __restrict__ out) {
  const int64_t p0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p0 >= n_points || *mode != kModeCells) return;
  int m[4];
  load_mask4(mask, p0, n_points, vec, m);
  for (int q = 0; q < 4; ++q) {
    if (!m[q]) continue;
    const int64_t p = p0 + q, u = rank[p];
    for (int w = 0; w < words; ++w) out[u * words + w] = in[p * words + w];
  }
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// ---- bit-word layout (uniform grids, counts 0/1) ----
// Cell words: cell row (j, k) has wcx = ceil(cx / 64) words, word index
// (j + cy * k) * wcx + w.  Point words: point row (j, k) has wpx =
// ceil(nx / 64) words, index (j + ny * k) * wpx + w.
struct BitGeom {
  int nx, ny, nz, cx, cy, cz, wcx, wpx;
  int64_t n_cwords, n_pwords;
};

inline BitGeom bit_geom(int nx, int ny, int nz) {
  BitGeom b;
  b.nx = nx;
  b.ny = ny;
  b.nz = nz;
  b.cx = nx - 1;
  b.cy = ny - 1;
  b.cz = nz - 1;
  b.wcx = (b.cx + 63) / 64;
  b.wpx = (nx + 63) / 64;
  b.n_cwords = (int64_t)b.wcx * b.cy * b.cz;
  b.n_pwords = (int64_t)b.wpx * ny * nz;
  return b;
}

// The row kernels run a wave per grid row (4 waves per block), looping over
// the row's words; lane l stands for bit l of a word.
constexpr int kRowWaves = 4;

__device__ __forceinline__ int wave_row() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kRowWaves + (threadIdx.x >> 6)));
}

// Cell row r = j + cy * k: per word the ballot of count != 0 over its 64
// cells and its popcount; *general = 1 if a count is outside {0, 1}.
__global__ __launch_bounds__(64 * kRowWaves) void k_cell_bits(BitGeom g,
                                                              const int32_t* __restrict__ counts,
                                                              uint64_t* __restrict__ cbits,
                                                              int32_t* __restrict__ cpop,
                                                              int32_t* __restrict__ general) {
  const int r = wave_row(), lane = threadIdx.x & 63;
  if (r >= g.cy * g.cz) return;  // whole waves
  const int32_t* row = counts + (int64_t)r * g.cx;
  bool bad = false;
  for (int w0 = 0; w0 < g.wcx; w0 += 4) {  // four words' loads in flight
    int32_t cnt[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int x = (w0 + q) * 64 + lane;
      cnt[q] = x < g.cx ? row[x] : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (w0 + q >= g.wcx) break;
      const uint64_t bits = __ballot(cnt[q] != 0);
      bad = bad || cnt[q] < 0 || cnt[q] > 1;
      if (lane == 0) {
        cbits[(int64_t)r * g.wcx + w0 + q] = bits;
        cpop[(int64_t)r * g.wcx + w0 + q] = __popcll(bits);
      }
    }
  }
  if (bad) *general = 1;
}

// Point row (j, k): point (i, j, k) is used when a cell of cell rows j-1, j
// in cell planes k-1, k at x = i-1 or i is kept, i.e. the word is the OR over
// those four cell rows of c | c << 1 | (the previous word's bit 63).  Writes
// the words, their popcounts and, if `mask`, the int32 mask of the row.
__global__ __launch_bounds__(64 * kRowWaves) void k_mask_bits(BitGeom g,
                                                              const uint64_t* __restrict__ cbits,
                                                              uint64_t* __restrict__ mbits,
                                                              int32_t* __restrict__ mpop,
                                                              int32_t* __restrict__ mask) {
  const int pr = wave_row(), lane = threadIdx.x & 63;
  if (pr >= g.ny * g.nz) return;  // whole waves
  const int j = pr % g.ny, k = pr / g.ny;
  // the four cell rows (j-1 or j, k-1 or k); a missing one reads row 0 and is
  // masked out
  const uint64_t* rows[4];
  uint64_t keep[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int jc = j - 1 + (q & 1), kc = k - 1 + (q >> 1);
    const bool ok = kc >= 0 && kc < g.cz && jc >= 0 && jc < g.cy;
    rows[q] = cbits + (ok ? ((int64_t)jc + (int64_t)g.cy * kc) * g.wcx : 0);
    keep[q] = ok ? ~0ull : 0ull;
  }
  uint64_t carry[4] = {0, 0, 0, 0};  // bit 63 of each cell row's previous word
  // with 16-byte aligned rows the mask goes out 4 words (256 points) at a
  // time, lane l writing points 4l .. 4l+3 as one 16-byte store
  const bool vec = mask && (g.nx & 3) == 0 && ((uintptr_t)mask & 15u) == 0;
  int32_t* mrow = mask ? mask + (int64_t)pr * g.nx : nullptr;
  for (int w0 = 0; w0 < g.wpx; w0 += 4) {
    uint64_t m4[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int wx = w0 + t;
      uint64_t m = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint64_t c = (wx < g.wcx ? rows[q][wx] : 0) & keep[q];  // wx <= wcx
        m |= c | (c << 1) | carry[q];
        carry[q] = c >> 63;
      }
      m4[t] = wx < g.wpx ? m : 0;
      if (lane == 0 && wx < g.wpx) {
        mbits[(int64_t)pr * g.wpx + wx] = m;
        mpop[(int64_t)pr * g.wpx + wx] = __popcll(m);
      }
    }
    if (!mask) continue;
    if (vec) {
      const int x = w0 * 64 + 4 * lane;  // points x .. x+3, all in word w0 + lane / 16
      if (x < g.nx) {
        const uint64_t m = m4[lane >> 4];
        const int sh = x & 63;
        int4 v;
        v.x = (int)((m >> sh) & 1u);
        v.y = (int)((m >> (sh + 1)) & 1u);
        v.z = (int)((m >> (sh + 2)) & 1u);
        v.w = (int)((m >> (sh + 3)) & 1u);
        *(int4*)(mrow + x) = v;
      }
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int x = (w0 + t) * 64 + lane;
        if (x < g.nx) mrow[x] = (int32_t)((m4[t] >> lane) & 1u);
      }
    }
  }
}

// The point words without the int32 mask: a lane per word (32-bit index
// arithmetic; every point word has its own independent loads).
__global__ __launch_bounds__(256) void k_mask_words(BitGeom g, const uint64_t* __restrict__ cbits,
                                                    uint64_t* __restrict__ mbits,
                                                    int32_t* __restrict__ mpop) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= (uint32_t)g.n_pwords) return;
  const uint32_t prow = w / (uint32_t)g.wpx;
  const int wx = (int)(w - prow * (uint32_t)g.wpx);
  const int j = (int)(prow % (uint32_t)g.ny), k = (int)(prow / (uint32_t)g.ny);
  uint64_t m = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int jc = j - 1 + (q & 1), kc = k - 1 + (q >> 1);
    if (kc < 0 || kc >= g.cz || jc < 0 || jc >= g.cy) continue;
    const uint64_t* r = cbits + ((int64_t)jc + (int64_t)g.cy * kc) * g.wcx;
    const uint64_t c = wx < g.wcx ? r[wx] : 0;     // wx <= wcx
    const uint64_t prev = wx > 0 ? r[wx - 1] : 0;  // wx - 1 < wcx
    m |= c | (c << 1) | (prev >> 63);
  }
  mbits[w] = m;
  mpop[w] = __popcll(m);
}

// Rank of point (i, j, k) among the used points.
__device__ __forceinline__ int32_t bit_rank(const BitGeom& g, const uint64_t* __restrict__ mbits,
                                            const int32_t* __restrict__ moff, int i, int j, int k) {
  const int64_t w = ((int64_t)j + (int64_t)g.ny * k) * g.wpx + (i >> 6);
  const uint64_t below = (1ull << (i & 63)) - 1ull;
  return moff[w] + __popcll(mbits[w] & below);
}

// Last word w with off[w] <= v (off: exclusive prefix of the words'
// popcounts, off[0] = 0 <= v): the word holding the v-th set bit.
__device__ __forceinline__ int64_t word_of(const int32_t* __restrict__ off, int64_t n, int64_t v) {
  int64_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)off[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Position of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ int nth_bit(uint64_t m, int k) {
  int lo = 0, hi = 63;  // largest i with popc(m & ((1 << i) - 1)) <= k
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (__popcll(m & ((1ull << mid) - 1ull)) <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// A thread per output cell t: its cell is the t-th kept cell (word by binary
// search over the cell words' prefixes, then the bit); it writes the 8
// corner ids (ranks with dedup).
__global__ __launch_bounds__(256) void k_emit_rank(BitGeom g, const int32_t* __restrict__ mode,
                                                   const uint64_t* __restrict__ cbits,
                                                   const int32_t* __restrict__ coff,
                                                   const uint64_t* __restrict__ mbits,
                                                   const int32_t* __restrict__ moff, bool dedup,
                                                   int64_t n_out, int32_t* __restrict__ conn) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_out || *mode != kModeBits) return;
  const int64_t w = word_of(coff, g.n_cwords, t);
  const uint64_t bits = cbits[w];
  const int r = (int)(t - coff[w]);
  if (r >= __popcll(bits)) return;  // prefixes not of these words
  const int b = nth_bit(bits, r);
  const int64_t row = w / g.wcx;
  const int wx = (int)(w - row * g.wcx);
  const int j = (int)(row % g.cy), k = (int)(row / g.cy), i = wx * 64 + b;
  static constexpr int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  int32_t* dst = conn + t * 8;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int pi = i + corner[v][0], pj = j + corner[v][1], pk = k + corner[v][2];
    dst[v] = dedup ? bit_rank(g, mbits, moff, pi, pj, pk)
                   : (int32_t)(pi + (int64_t)g.nx * (pj + (int64_t)g.ny * pk));
  }
}

// A thread per used point u: its point is the u-th used one; writes its
// coordinates (origin + spacing * ijk) or values (`words` 4-byte words).
__global__ __launch_bounds__(256) void k_compact_rank(
    BitGeom g, const int32_t* __restrict__ mode, const uint64_t* __restrict__ mbits,
    const int32_t* __restrict__ moff, int64_t n_used, float ox, float oy, float oz, float hx,
    float hy, float hz, float* __restrict__ coords, const uint32_t* __restrict__ in, int words,
    uint32_t* __restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_used || *mode != kModeBits) return;
  const int64_t w = word_of(moff, g.n_pwords, u);
  const uint64_t m = mbits[w];
  const int r = (int)(u - moff[w]);
  if (r >= __popcll(m)) return;  // prefixes not of these words
  const int b = nth_bit(m, r);
  const int64_t prow = w / g.wpx;
  const int wx = (int)(w - prow * g.wpx);
  const int i = wx * 64 + b;
  if (coords) {
    const int j = (int)(prow % g.ny), k = (int)(prow / g.ny);
    coords[3 * u] = ox + hx * (float)i;
    coords[3 * u + 1] = oy + hy * (float)j;
    coords[3 * u + 2] = oz + hz * (float)k;
  } else {
    const int64_t p = prow * g.nx + i;
    for (int q = 0; q < words; ++q) out[u * words + q] = in[p * words + q];
  }
}

// Header of the workspace: [0] layout mode, [1] flag of counts outside
// {0, 1}, [2..4] the uniform grid's nx, ny, nz, [5] 1 when prepare computed
// the point ranks (duplicate removal), [6] output cells, [7] used points.
constexpr int kHeaderInts = 8;

// Sum of the caller's per-cell counts in 64 bits (out[0]) and a flag for
// negative counts (out[1]).  The int32 scan of the counts wraps past 2^31-1
// without notice (three counts of 2^31-1 wrap to a positive total), so the
// general path checks the exact sum before it scans.
__global__ __launch_bounds__(256) void k_count_total_i64(const int32_t* __restrict__ counts,
                                                         int64_t n,
                                                         unsigned long long* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long sum = 0ull;
  bool neg = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t c = counts[i];
    neg |= c < 0;
    sum += c > 0 ? (unsigned long long)c : 0ull;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  const bool any_neg = __ballot(neg) != 0ull;
  if ((threadIdx.x & 63) == 0) {
    if (sum) atomicAdd(&out[0], sum);
    if (any_neg) atomicOr(&out[1], 1ull);
  }
}

__global__ void k_set_sizes(int32_t* __restrict__ hdr, int32_t cells, int32_t points) {
  if (threadIdx.x == 0) {
    hdr[6] = cells;
    hdr[7] = points;
  }
}
__global__ void k_set_header(int32_t* __restrict__ hdr, int mode, int nx, int ny, int nz,
                             int ranked) {
  if (threadIdx.x == 0) {
    hdr[0] = mode;
    hdr[1] = 0;
    hdr[2] = nx;
    hdr[3] = ny;
    hdr[4] = nz;
    hdr[5] = ranked;
  }
}

// Workspace: a 256-byte header (the layout's mode, a flag), region A (per-
// cell offsets, or cell words and their prefixes), region B (per-point
// ranks, or point words and their prefixes).
constexpr size_t kHeader = 256;

struct TopoWs {
  size_t off_a, off_b, bytes;
};

inline TopoWs topo_ws(int64_t n_cells, int64_t n_points) {
  TopoWs w;
  w.off_a = kHeader;
  w.off_b = w.off_a + align_up((size_t)(n_cells > 0 ? n_cells : 1) * 4, 256);
  w.bytes = w.off_b + align_up((size_t)(n_points > 0 ? n_points : 1) * 4, 256);
  return w;
}

// The bit-word layout inside regions A and B, if it fits there.
struct BitWs {
  size_t cbits, coff, mbits, moff;
  bool fits;
};

inline BitWs bit_ws(const TopoWs& w, const BitGeom& g) {
  BitWs b;
  b.cbits = w.off_a;
  b.coff = b.cbits + align_up((size_t)g.n_cwords * 8, 256);
  b.mbits = w.off_b;
  b.moff = b.mbits + align_up((size_t)g.n_pwords * 8, 256);
  b.fits = b.coff + (size_t)g.n_cwords * 4 <= w.off_b && b.moff + (size_t)g.n_pwords * 4 <= w.bytes;
  return b;
}

struct Geometry {  // a uniform grid (cells == nullptr) or an unstructured one
  const int32_t* cells;
  int nv;
  int64_t n_cells, n_points;
  int nx, ny, nz;
};

inline bool uniform_ok(int64_t nx, int64_t ny, int64_t nz) {
  return nx >= 2 && ny >= 2 && nz >= 2 && nx <= 65535 * (int64_t)kTX && ny <= 65535 * kTY &&
         nz <= 65535 * (int64_t)kZRun && nx * ny * nz <= 0x7fffffffLL;
}

// words_ready: the cell words and popcounts of the bit-word layout are
// already in the workspace (k_threshold_words4) and the counts are 0/1.
int prepare(const Geometry& g, const int32_t* d_counts, unsigned flags, void* d_ws,
            size_t ws_bytes, int32_t* d_mask, int64_t* h_cells, int64_t* h_points,
            hipStream_t s, bool words_ready = false) {
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  const bool mask_optional = (flags & DAX_HIP_TOPOLOGY_MASK_OPTIONAL) != 0;
  const TopoWs w = topo_ws(g.n_cells, g.n_points);
  if (!d_counts || !h_cells || !h_points || !d_ws || ws_bytes < w.bytes || (dedup && !d_mask) ||
      (flags & ~(unsigned)(DAX_HIP_MC_DEDUP | DAX_HIP_TOPOLOGY_MASK_OPTIONAL))) {
    set_error("topology prepare: bad arguments (workspace of %zu bytes needed)", w.bytes);
    return DAX_HIP_BAD_VALUE;
  }
  *h_cells = 0;
  *h_points = 0;
  char* ws = (char*)d_ws;
  int32_t* hdr = (int32_t*)ws;
  const dim3 b(256);
  int rc = DAX_HIP_OK;
  if (!words_ready) {
    k_set_header<<<1, 64, 0, s>>>(hdr, kModeCells, g.nx, g.ny, g.nz, dedup ? 1 : 0);
    DAXHIP_LAUNCH_CHECK("k_set_header");
  }
  if (!g.cells) {
    // counts 0/1 on a uniform grid: the bit-word layout
    const BitGeom bg = bit_geom(g.nx, g.ny, g.nz);
    const BitWs bw = bit_ws(w, bg);
    if (bw.fits) {
      uint64_t* cbits = (uint64_t*)(ws + bw.cbits);
      int32_t* coff = (int32_t*)(ws + bw.coff);
      if (!words_ready) {
        k_cell_bits<<<dim3((unsigned)ceil_div((int64_t)bg.cy * bg.cz, kRowWaves)),
                      dim3(64 * kRowWaves), 0, s>>>(bg, d_counts, cbits, coff, hdr + 1);
        DAXHIP_LAUNCH_CHECK("k_cell_bits");
      }
      int32_t total = 0, general = 0;
      if ((rc = dax_hip_scan_exclusive_i32(coff, coff, bg.n_cwords, &total, s))) return rc;
      if (!words_ready) {
        DAXHIP_CHECK(hipMemcpyAsync(&general, hdr + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        DAXHIP_CHECK(hipStreamSynchronize(s));
      }
      if (!general) {
        if ((int64_t)total * 8 > 0x7fffffffLL) {
          set_error("topology prepare: %lld output cells exceed 2^31 ids", (long long)total);
          return DAX_HIP_BAD_VALUE;
        }
        k_set_header<<<1, 64, 0, s>>>(hdr, kModeBits, g.nx, g.ny, g.nz, dedup ? 1 : 0);
        DAXHIP_LAUNCH_CHECK("k_set_header");
        *h_cells = total;
        if (!dedup) {
          k_set_sizes<<<1, 64, 0, s>>>(hdr, total, 0);
          DAXHIP_LAUNCH_CHECK("k_set_sizes");
          return DAX_HIP_OK;
        }
        uint64_t* mbits = (uint64_t*)(ws + bw.mbits);
        int32_t* moff = (int32_t*)(ws + bw.moff);
        if (mask_optional) {
          k_mask_words<<<dim3((unsigned)ceil_div(bg.n_pwords, 256)), b, 0, s>>>(bg, cbits, mbits,
                                                                                moff);
          DAXHIP_LAUNCH_CHECK("k_mask_words");
        } else {
          k_mask_bits<<<dim3((unsigned)ceil_div((int64_t)bg.ny * bg.nz, kRowWaves)),
                        dim3(64 * kRowWaves), 0, s>>>(bg, cbits, mbits, moff, d_mask);
          DAXHIP_LAUNCH_CHECK("k_mask_bits");
        }
        int32_t used = 0;
        if ((rc = dax_hip_scan_exclusive_i32(moff, moff, bg.n_pwords, &used, s))) return rc;
        k_set_sizes<<<1, 64, 0, s>>>(hdr, total, used);
        DAXHIP_LAUNCH_CHECK("k_set_sizes");
        *h_points = used;
        return DAX_HIP_OK;
      }
    }
  }
  // any counts: per-cell offsets and per-point ranks.  The exact sum and the
  // sign of every count are checked first: the scan below is int32.
  {
    TempBuf chk(s);
    if ((rc = chk.alloc(2 * sizeof(unsigned long long)))) return rc;
    unsigned long long* d_chk = (unsigned long long*)chk.p;
    k_fill_i32<<<dim3(1), b, 0, s>>>((int32_t*)d_chk, 4, 0);
    DAXHIP_LAUNCH_CHECK("k_fill_i32");
    if (g.n_cells > 0) {
      const int64_t blocks = std::min<int64_t>(ceil_div(g.n_cells, 256), 2048);
      k_count_total_i64<<<dim3((unsigned)blocks), b, 0, s>>>(d_counts, g.n_cells, d_chk);
      DAXHIP_LAUNCH_CHECK("k_count_total_i64");
    }
    unsigned long long h_chk[2] = {0ull, 0ull};
    DAXHIP_CHECK(hipMemcpyAsync(h_chk, d_chk, sizeof(h_chk), hipMemcpyDeviceToHost, s));
    DAXHIP_CHECK(hipStreamSynchronize(s));
    if (h_chk[1]) {
      set_error("topology prepare: a cell count is negative");
      return DAX_HIP_BAD_VALUE;
    }
    if (h_chk[0] > 0x7fffffffull / (unsigned long long)std::max(g.nv, 1)) {
      set_error("topology prepare: %llu output cells of %d vertices exceed 2^31 ids", h_chk[0],
                g.nv);
      return DAX_HIP_BAD_VALUE;
    }
  }
  int32_t* offs = (int32_t*)(ws + w.off_a);
  int32_t* rank = (int32_t*)(ws + w.off_b);
  int32_t total = 0;
  if ((rc = dax_hip_scan_exclusive_i32(d_counts, offs, g.n_cells, &total, s))) return rc;
  *h_cells = total;
  if (!dedup) return DAX_HIP_OK;
  const dim3 gp((unsigned)ceil_div(g.n_points, 256));
  if (g.cells) {
    TempBuf flag(s);
    if ((rc = flag.alloc(sizeof(int32_t)))) return rc;
    k_fill_i32<<<dim3(1), b, 0, s>>>((int32_t*)flag.p, 1, 0);
    DAXHIP_LAUNCH_CHECK("k_fill_i32");
    k_fill_i32<<<gp, b, 0, s>>>(d_mask, g.n_points, 0);
    DAXHIP_LAUNCH_CHECK("k_fill_i32");
    k_mark_cells<<<dim3((unsigned)ceil_div(g.n_cells, 256)), b, 0, s>>>(
        g.cells, g.nv, g.n_cells, g.n_points, d_counts, d_mask, (int32_t*)flag.p);
    DAXHIP_LAUNCH_CHECK("k_mark_cells");
    int32_t bad = 0;
    DAXHIP_CHECK(hipMemcpyAsync(&bad, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DAXHIP_CHECK(hipStreamSynchronize(s));
    if (bad) {
      set_error("topology prepare: a kept cell references a point outside [0, %lld)",
                (long long)g.n_points);
      return DAX_HIP_BAD_VALUE;
    }
  } else {
    k_mask_uniform<<<dim3((unsigned)ceil_div(g.nx, kTX), (unsigned)ceil_div(g.ny, kTY),
                          (unsigned)ceil_div(g.nz, kZRun)),
                     dim3(kTX, kTY), 0, s>>>(d_counts, g.nx, g.ny, g.nz, d_mask);
    DAXHIP_LAUNCH_CHECK("k_mask_uniform");
  }
  int32_t used = 0;
  if ((rc = dax_hip_scan_exclusive_i32(d_mask, rank, g.n_points, &used, s))) return rc;
  *h_points = used;
  return DAX_HIP_OK;
}

int generate(const Geometry& g, const float* d_points, const float* origin, const float* spacing,
             const int32_t* d_counts, unsigned flags, const void* d_ws, size_t ws_bytes,
             const int32_t* d_mask, int32_t* d_conn, float* d_coords, hipStream_t s) {
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  const TopoWs w = topo_ws(g.n_cells, g.n_points);
  if (!d_counts || !d_ws || ws_bytes < w.bytes || !d_conn || (dedup && (!d_mask || !d_coords)) ||
      (flags & ~(unsigned)DAX_HIP_MC_DEDUP)) {
    set_error("topology generate: bad arguments (workspace of %zu bytes needed)", w.bytes);
    return DAX_HIP_BAD_VALUE;
  }
  const char* ws = (const char*)d_ws;
  const int32_t* hdr = (const int32_t*)ws;
  const int32_t* offs = (const int32_t*)(ws + w.off_a);
  const int32_t* rank = (const int32_t*)(ws + w.off_b);
  float o[3] = {0, 0, 0}, h[3] = {1, 1, 1};
  for (int d = 0; d < 3; ++d) {
    if (origin) o[d] = origin[d];
    if (spacing) h[d] = spacing[d];
  }
  const dim3 b(256);
  // the layout prepare chose (the kernels also check it) and whether it
  // ranked the points
  int32_t head[kHeaderInts];
  DAXHIP_CHECK(hipMemcpyAsync(head, hdr, sizeof(head), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  const int32_t mode = head[0];
  if (dedup && !head[5]) {
    set_error("topology generate: duplicate removal after a prepare without it");
    return DAX_HIP_BAD_VALUE;
  }
  if (mode == kModeCells) {
    k_emit<<<dim3((unsigned)ceil_div(g.n_cells, 1024)), b, 0, s>>>(
        hdr, g.nv, g.cells, g.n_cells, g.nx, g.ny, d_counts, aligned16(d_counts), offs,
        dedup ? rank : nullptr, d_conn);
    DAXHIP_LAUNCH_CHECK("k_emit");
    if (dedup) {
      k_compact_points<<<dim3((unsigned)ceil_div(g.n_points, 1024)), b, 0, s>>>(
          d_points, g.n_points, g.nx, g.ny, o[0], o[1], o[2], h[0], h[1], h[2], d_mask,
          aligned16(d_mask), rank, hdr, d_coords);
      DAXHIP_LAUNCH_CHECK("k_compact_points");
    }
  } else if (!g.cells) {
    const BitGeom bg = bit_geom(g.nx, g.ny, g.nz);
    const BitWs bw = bit_ws(w, bg);
    if (!bw.fits) {
      set_error("topology generate: workspace of another grid");
      return DAX_HIP_BAD_VALUE;
    }
    {
      const uint64_t* cbits = (const uint64_t*)(ws + bw.cbits);
      const int32_t* coff = (const int32_t*)(ws + bw.coff);
      const uint64_t* mbits = (const uint64_t*)(ws + bw.mbits);
      const int32_t* moff = (const int32_t*)(ws + bw.moff);
      const int64_t n_out = head[6], n_used = head[7];
      if (n_out > 0) {
        k_emit_rank<<<dim3((unsigned)ceil_div(n_out, 256)), b, 0, s>>>(
            bg, hdr, cbits, coff, mbits, moff, dedup, n_out, d_conn);
        DAXHIP_LAUNCH_CHECK("k_emit_rank");
      }
      if (dedup && n_used > 0) {
        k_compact_rank<<<dim3((unsigned)ceil_div(n_used, 256)), b, 0, s>>>(
            bg, hdr, mbits, moff, n_used, o[0], o[1], o[2], h[0], h[1], h[2], d_coords, nullptr,
            0, nullptr);
        DAXHIP_LAUNCH_CHECK("k_compact_rank");
      }
    }
  }
  return DAX_HIP_OK;
}

}  // namespace topo
}  // namespace daxhip

using namespace daxhip;
using namespace daxhip::topo;

extern "C" {

int dax_hip_threshold_count_uniform(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                                    float min_value, float max_value, int32_t* d_counts,
                                    void* stream) {
  if (!uniform_ok(nx, ny, nz) || !d_field || !d_counts) {
    set_error("dax_hip_threshold_count_uniform: bad arguments (grid %lld x %lld x %lld)",
              (long long)nx, (long long)ny, (long long)nz);
    return DAX_HIP_BAD_VALUE;
  }
  const int cx = (int)nx - 1, cy = (int)ny - 1, cz = (int)nz - 1;
  if ((nx & 3) == 0 && ((uintptr_t)d_field & 15u) == 0) {
    k_count_uniform4<<<dim3((unsigned)ceil_div(nx, 4 * kTX), (unsigned)ceil_div(cy, kTY),
                            (unsigned)ceil_div(cz, kZRun)),
                       dim3(kTX, kTY), 0, (hipStream_t)stream>>>(
        d_field, (int)nx, (int)ny, (int)nz, min_value, max_value, d_counts);
    DAXHIP_LAUNCH_CHECK("k_count_uniform4");
    return DAX_HIP_OK;
  }
  k_count_uniform<<<dim3((unsigned)ceil_div(cx, kTX), (unsigned)ceil_div(cy, kTY),
                         (unsigned)ceil_div(cz, kZRun)),
                    dim3(kTX, kTY), 0, (hipStream_t)stream>>>(d_field, (int)nx, (int)ny, (int)nz,
                                                              min_value, max_value, d_counts);
  DAXHIP_LAUNCH_CHECK("k_count_uniform");
  return DAX_HIP_OK;
}

int dax_hip_threshold_count_cells(const int32_t* d_cells, int nv, int64_t n_cells,
                                  const float* d_field, int64_t n_points, float min_value,
                                  float max_value, int32_t* d_counts, void* stream) {
  if (nv < 1 || nv > 8 || n_cells < 0 || n_points < 0 || n_points > 0x7fffffffLL ||
      n_cells > 0x7fffffffLL / nv || (n_cells > 0 && (!d_cells || !d_field || !d_counts))) {
    set_error("dax_hip_threshold_count_cells: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  if (n_cells == 0) return DAX_HIP_OK;
  hipStream_t s = (hipStream_t)stream;
  TempBuf flag(s);
  int rc = flag.alloc(sizeof(int32_t));
  if (rc) return rc;
  DAXHIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int32_t), s));
  k_count_cells<<<dim3((unsigned)ceil_div(n_cells, 256)), dim3(256), 0, s>>>(
      d_cells, nv, n_cells, d_field, n_points, min_value, max_value, d_counts, (int32_t*)flag.p);
  DAXHIP_LAUNCH_CHECK("k_count_cells");
  int32_t bad = 0;
  DAXHIP_CHECK(hipMemcpyAsync(&bad, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  if (bad) {
    set_error("dax_hip_threshold_count_cells: a cell references a point outside [0, %lld)",
              (long long)n_points);
    return DAX_HIP_BAD_VALUE;
  }
  return DAX_HIP_OK;
}

int dax_hip_threshold_prepare_uniform(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                                      float min_value, float max_value, unsigned flags,
                                      int32_t* d_counts, void* d_workspace,
                                      size_t workspace_bytes, int32_t* d_point_mask,
                                      int64_t* h_cells, int64_t* h_points, void* stream) {
  if (!uniform_ok(nx, ny, nz) || !d_field || !d_counts) {
    set_error("dax_hip_threshold_prepare_uniform: bad arguments (grid %lld x %lld x %lld)",
              (long long)nx, (long long)ny, (long long)nz);
    return DAX_HIP_BAD_VALUE;
  }
  const Geometry g{nullptr, 8, (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz, (int)nx, (int)ny,
                   (int)nz};
  const TopoWs w = topo_ws(g.n_cells, g.n_points);
  const BitGeom bg = bit_geom(g.nx, g.ny, g.nz);
  const BitWs bw = bit_ws(w, bg);
  hipStream_t s = (hipStream_t)stream;
  const bool fused = (nx & 3) == 0 && ((uintptr_t)d_field & 15u) == 0 && bw.fits &&
                     d_workspace && workspace_bytes >= w.bytes;
  if (!fused) {  // ThresholdCount, then the generic prepare
    int rc = dax_hip_threshold_count_uniform(d_field, nx, ny, nz, min_value, max_value, d_counts,
                                             stream);
    if (rc) return rc;
    return prepare(g, d_counts, flags, d_workspace, workspace_bytes, d_point_mask, h_cells,
                   h_points, s);
  }
  char* ws = (char*)d_workspace;
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  k_set_header<<<1, 64, 0, s>>>((int32_t*)ws, kModeCells, g.nx, g.ny, g.nz, dedup ? 1 : 0);
  DAXHIP_LAUNCH_CHECK("k_set_header");
  k_threshold_words4<<<dim3((unsigned)ceil_div(nx, 4 * kTX), (unsigned)ceil_div(bg.cy, kTY),
                            (unsigned)ceil_div(bg.cz, kZRun)),
                       dim3(kTX, kTY), 0, s>>>(d_field, (int)nx, (int)ny, (int)nz, min_value,
                                               max_value, bg.wcx, d_counts,
                                               (uint64_t*)(ws + bw.cbits),
                                               (int32_t*)(ws + bw.coff));
  DAXHIP_LAUNCH_CHECK("k_threshold_words4");
  return prepare(g, d_counts, flags, d_workspace, workspace_bytes, d_point_mask, h_cells, h_points,
                 s, true);
}

size_t dax_hip_topology_workspace_bytes(int64_t n_cells, int64_t n_points) {
  return topo_ws(n_cells, n_points).bytes;
}

int dax_hip_topology_prepare_uniform(int64_t nx, int64_t ny, int64_t nz, const int32_t* d_counts,
                                     unsigned flags, void* d_workspace, size_t workspace_bytes,
                                     int32_t* d_point_mask, int64_t* h_cells, int64_t* h_points,
                                     void* stream) {
  if (!uniform_ok(nx, ny, nz)) {
    set_error("dax_hip_topology_prepare_uniform: unsupported grid");
    return DAX_HIP_BAD_VALUE;
  }
  const Geometry g{nullptr, 8, (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz, (int)nx, (int)ny,
                   (int)nz};
  return prepare(g, d_counts, flags, d_workspace, workspace_bytes, d_point_mask, h_cells, h_points,
                 (hipStream_t)stream);
}

int dax_hip_topology_generate_uniform(int64_t nx, int64_t ny, int64_t nz, const float* h_origin3,
                                      const float* h_spacing3, const int32_t* d_counts,
                                      unsigned flags, const void* d_workspace,
                                      size_t workspace_bytes, const int32_t* d_point_mask,
                                      int32_t* d_conn, float* d_coords, void* stream) {
  if (!uniform_ok(nx, ny, nz)) {
    set_error("dax_hip_topology_generate_uniform: unsupported grid");
    return DAX_HIP_BAD_VALUE;
  }
  const Geometry g{nullptr, 8, (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz, (int)nx, (int)ny,
                   (int)nz};
  return generate(g, nullptr, h_origin3, h_spacing3, d_counts, flags, d_workspace, workspace_bytes,
                  d_point_mask, d_conn, d_coords, (hipStream_t)stream);
}

int dax_hip_topology_prepare_cells(const int32_t* d_cells, int nv, int64_t n_cells,
                                   int64_t n_points, const int32_t* d_counts, unsigned flags,
                                   void* d_workspace, size_t workspace_bytes,
                                   int32_t* d_point_mask, int64_t* h_cells, int64_t* h_points,
                                   void* stream) {
  if (nv < 1 || nv > 8 || n_cells < 1 || n_points < 1 || n_points > 0x7fffffffLL ||
      n_cells > 0x7fffffffLL / nv || !d_cells) {
    set_error("dax_hip_topology_prepare_cells: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  const Geometry g{d_cells, nv, n_cells, n_points, 0, 0, 0};
  return prepare(g, d_counts, flags, d_workspace, workspace_bytes, d_point_mask, h_cells, h_points,
                 (hipStream_t)stream);
}

int dax_hip_topology_generate_cells(const int32_t* d_cells, int nv, int64_t n_cells,
                                    const float* d_points, int64_t n_points,
                                    const int32_t* d_counts, unsigned flags,
                                    const void* d_workspace, size_t workspace_bytes,
                                    const int32_t* d_point_mask, int32_t* d_conn,
                                    float* d_coords, void* stream) {
  if (nv < 1 || nv > 8 || n_cells < 1 || n_points < 1 || n_points > 0x7fffffffLL ||
      n_cells > 0x7fffffffLL / nv || !d_cells ||
      ((flags & DAX_HIP_MC_DEDUP) && !d_points)) {
    set_error("dax_hip_topology_generate_cells: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  const Geometry g{d_cells, nv, n_cells, n_points, 0, 0, 0};
  return generate(g, d_points, nullptr, nullptr, d_counts, flags, d_workspace, workspace_bytes,
                  d_point_mask, d_conn, d_coords, (hipStream_t)stream);
}

int dax_hip_topology_compact_point_field(const void* d_in, size_t value_bytes, int64_t n_cells,
                                         int64_t n_points, const int32_t* d_point_mask,
                                         const void* d_workspace, size_t workspace_bytes,
                                         void* d_out, void* stream) {
  const TopoWs w = topo_ws(n_cells, n_points);
  if (value_bytes == 0 || value_bytes % 4 || value_bytes > 1024 || n_points < 0 ||
      n_points > 0x7fffffffLL || !d_workspace || workspace_bytes < w.bytes ||
      (n_points > 0 && (!d_in || !d_out))) {
    set_error("dax_hip_topology_compact_point_field: bad arguments (values of 4k bytes, "
              "workspace of %zu bytes)", w.bytes);
    return DAX_HIP_BAD_VALUE;
  }
  if (n_points == 0) return DAX_HIP_OK;
  hipStream_t s = (hipStream_t)stream;
  const char* ws = (const char*)d_workspace;
  int32_t hdr[kHeaderInts];
  DAXHIP_CHECK(hipMemcpyAsync(hdr, ws, sizeof(hdr), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  if (!hdr[5]) {
    set_error("dax_hip_topology_compact_point_field: the prepare step did not remove duplicates");
    return DAX_HIP_BAD_VALUE;
  }
  const int words = (int)(value_bytes / 4);
  if (hdr[0] == kModeBits) {
    const BitGeom bg = bit_geom(hdr[2], hdr[3], hdr[4]);
    const BitWs bw = bit_ws(w, bg);
    if ((int64_t)hdr[2] * hdr[3] * hdr[4] != n_points || !bw.fits) {
      set_error("dax_hip_topology_compact_point_field: workspace of another grid");
      return DAX_HIP_BAD_VALUE;
    }
    const int64_t n_used = hdr[7];
    if (n_used > 0) {
      k_compact_rank<<<dim3((unsigned)ceil_div(n_used, 256)), dim3(256), 0, s>>>(
          bg, (const int32_t*)ws, (const uint64_t*)(ws + bw.mbits), (const int32_t*)(ws + bw.moff),
          n_used, 0, 0, 0, 0, 0, 0, nullptr, (const uint32_t*)d_in, words, (uint32_t*)d_out);
      DAXHIP_LAUNCH_CHECK("k_compact_rank");
    }
    return DAX_HIP_OK;
  }
  if (!d_point_mask) {
    set_error("dax_hip_topology_compact_point_field: the point mask is needed for this layout");
    return DAX_HIP_BAD_VALUE;
  }
  k_compact_values<<<dim3((unsigned)ceil_div(n_points, 1024)), dim3(256), 0, s>>>(
      (const uint32_t*)d_in, words, n_points, d_point_mask, aligned16(d_point_mask),
      (const int32_t*)(ws + w.off_b), (const int32_t*)ws, (uint32_t*)d_out);
  DAXHIP_LAUNCH_CHECK("k_compact_values");
  return DAX_HIP_OK;
}

}  // extern "C"
