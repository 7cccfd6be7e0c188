// Marching-cubes forward on gfx950: classify -> count/scan -> generate.
//
// Reference path (Kitware/DaxToolkit):
//   MarchingCubesCount          dax/worklet/MarchingCubes.h:53-86
//   ScanInclusive + UpperBounds + visit index
//                               dax/cont/DispatcherGenerateInterpolatedCells.h:180-223
//   MarchingCubesGenerate       dax/worklet/MarchingCubes.h:90-169
//   Sort/Unique/LowerBounds dedup + coordinate interpolation
//                               dax/cont/DispatcherGenerateInterpolatedCells.h:261-301
//
// Design: the field is read once, as one coalesced stream, into inside-bits
// (classify).  Everything topological after that works on the bits: 64 cells
// per 64-bit word op.  Vertex numbering uses active-edge flags in
// (point id, axis) order, which is the order the reference's sort/unique
// produces, so no sort is needed.  The count/scan kernel is a single pass
// with decoupled look-back; the generate kernel revisits only tiles that
// have output and touches the field only around the surface.
#include <algorithm>

#include "mc_common.h"

namespace daxhip {
namespace mc {

__constant__ uint64_t c_case_table[256] = DAX_HIP_MC_CASE_TABLE_INIT;

// ---------------------------------------------------------------------------
// K1: classify.  One wavefront builds UNROLL row words; lane l reads point
// 64q+l of the word's row, a ballot of (f > iso) is the word.  The first
// threads also clear the scan state of the following count/scan launch.
template <int UNROLL, bool ALIGNED>
__global__ __launch_bounds__(256) void k_classify(VolGeom g, const float* __restrict__ field,
                                                  uint64_t* __restrict__ bits,
                                                  unsigned long long* __restrict__ zero_buf,
                                                  int64_t zero_words) {
  const int lane = threadIdx.x & 63;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid < zero_words) zero_buf[gtid] = 0ull;
  const int64_t wave = gtid >> 6;
  const int64_t w0 = wave * UNROLL;
  const float* f = field + g.field_off;
  float v[UNROLL];
  bool ok[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    const int64_t w = w0 + u;
    ok[u] = false;
    v[u] = 0.0f;
    if (w < g.nwords) {
      if (ALIGNED) {
        v[u] = __builtin_nontemporal_load(f + w * 64 + lane);
        ok[u] = true;
      } else {
        const int64_t row = w / g.wpr;
        const int x = (int)(w - row * g.wpr) * 64 + lane;
        if (x < g.nx) {
          v[u] = __builtin_nontemporal_load(f + row * g.nx + x);
          ok[u] = true;
        }
      }
    }
  }
  uint64_t mine = 0;
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    const uint64_t b = __ballot(ok[u] && v[u] > g.iso);
    if (lane == u) mine = b;
  }
  if (lane < UNROLL && w0 + lane < g.nwords) bits[g.word_off + w0 + lane] = mine;
}

// ---------------------------------------------------------------------------
// K2: per-word counts and the single-pass exclusive scan.
//
// One block per tile of kTileWords words; tiles are taken in order from an
// atomic counter so every predecessor of a tile has started.  The tile's
// words (plus the halo words of the +x/+y/+z neighbours) are staged in LDS.
// Counts: active edges = popcount of the x/y/z flags; triangles =
// sum of NumFaces[case] over the active cells (mixed-corner cells found with
// 64-bit word ops, so all-in / all-out cells cost nothing).
// Look-back: wave 0 reads 64 predecessor tile states at a time.
__device__ __forceinline__ uint32_t faces_of(const uint8_t* s_faces, int c) { return s_faces[c]; }

__global__ __launch_bounds__(kScanThreads) void k_count_scan(
    VolGeom g, const uint64_t* __restrict__ bits, uint32_t* __restrict__ base_v,
    uint32_t* __restrict__ base_t, unsigned long long* __restrict__ status, Ctrl* __restrict__ ctrl,
    int64_t ntiles) {
  __shared__ uint64_t s_b0[kTileWords + kMaxWpr + 2];
  __shared__ uint64_t s_b1[kTileWords + kMaxWpr + 2];
  __shared__ uint8_t s_faces[256];
  __shared__ uint32_t s_wave_v[kScanThreads / 64], s_wave_t[kScanThreads / 64];
  __shared__ uint32_t s_prefix[2];
  __shared__ int64_t s_tile;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_tile = (int64_t)atomicAdd(&ctrl->tile_counter, 1u);
  s_faces[tid] = (uint8_t)(c_case_table[tid] >> 60);
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile >= ntiles) return;  // scan state not reset by a classify launch
  const int64_t w_begin = tile * kTileWords;
  const int nw = (int)min((int64_t)kTileWords, g.nwords - w_begin);
  const uint64_t* b = bits + g.word_off;

  // stage plane k (rows j, j+1 and one word past) and plane k+1
  const int seg = nw + g.wpr + 1;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  for (int i = tid; i < seg; i += kScanThreads) {
    const int64_t w = w_begin + i;
    s_b0[i] = (w < g.nwords) ? b[w] : 0ull;
    s_b1[i] = (w + plane < g.nwords) ? b[w + plane] : 0ull;
  }
  __syncthreads();

  uint32_t cv[kWordsPerThread], ct[kWordsPerThread];
  uint32_t sum_v = 0, sum_t = 0;
#pragma unroll
  for (int u = 0; u < kWordsPerThread; ++u) {
    const int i = tid * kWordsPerThread + u;
    cv[u] = 0;
    ct[u] = 0;
    if (i < nw) {
      const WordPos p = word_pos(g, w_begin + i);
      const bool nx1 = p.q + 1 < g.wpr;
      const uint64_t A = s_b0[i], A1 = nx1 ? s_b0[i + 1] : 0ull;
      const uint64_t B = s_b0[i + g.wpr], B1 = nx1 ? s_b0[i + g.wpr + 1] : 0ull;
      const uint64_t C = s_b1[i], C1 = nx1 ? s_b1[i + 1] : 0ull;
      const uint64_t D = s_b1[i + g.wpr], D1 = nx1 ? s_b1[i + g.wpr + 1] : 0ull;
      uint64_t fx, fy, fz;
      edge_flags(p, A, A1, B, C, fx, fy, fz);
      cv[u] = (uint32_t)(popc(fx) + popc(fy) + popc(fz));
      CellPlanes cp;
      cell_planes(p, A, A1, B, B1, C, C1, D, D1, cp);
      uint64_t act = cp.active;
      uint32_t t = 0;
      while (act) {
        const int bit = __builtin_ctzll(act);
        act &= act - 1;
        t += faces_of(s_faces, cell_case(cp, bit));
      }
      ct[u] = t;
    }
    sum_v += cv[u];
    sum_t += ct[u];
  }

  // block exclusive scan of the per-thread sums
  const uint32_t inc_v = wave_incl_scan(sum_v), inc_t = wave_incl_scan(sum_t);
  if (lane == 63) {
    s_wave_v[wv] = inc_v;
    s_wave_t[wv] = inc_t;
  }
  __syncthreads();
  uint32_t wave_off_v = 0, wave_off_t = 0, agg_v = 0, agg_t = 0;
#pragma unroll
  for (int q = 0; q < kScanThreads / 64; ++q) {
    if (q < wv) {
      wave_off_v += s_wave_v[q];
      wave_off_t += s_wave_t[q];
    }
    agg_v += s_wave_v[q];
    agg_t += s_wave_t[q];
  }

  // decoupled look-back (wave 0)
  if (wv == 0) {
    uint32_t ex_v = 0, ex_t = 0;
    if (tile == 0) {
      if (lane == 0)
        __hip_atomic_store(&status[0], pack_status(agg_v, agg_t, kStatusPrefix), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(&status[tile], pack_status(agg_v, agg_t, kStatusAgg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      int64_t pos = tile - 1;
      while (true) {
        const int64_t idx = pos - lane;
        unsigned long long s = pack_status(0, 0, kStatusPrefix);
        if (idx >= 0) {
          s = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          while ((s >> 62) == 0ull) {
            __builtin_amdgcn_s_sleep(1);
            s = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        const uint64_t pm = __ballot((s >> 62) == kStatusPrefix);
        uint32_t sv = (uint32_t)(s & 0x7fffffffull), st = (uint32_t)((s >> 31) & 0x7fffffffull);
        if (pm) {
          const int first = __builtin_ctzll(pm);
          if (lane > first) {
            sv = 0;
            st = 0;
          }
        }
        ex_v += wave_sum(sv);
        ex_t += wave_sum(st);
        if (pm) break;
        pos -= 64;
      }
      if (lane == 0)
        __hip_atomic_store(&status[tile], pack_status(ex_v + agg_v, ex_t + agg_t, kStatusPrefix),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) {
      s_prefix[0] = ex_v;
      s_prefix[1] = ex_t;
      atomicAdd(&ctrl->total_v, (unsigned long long)agg_v);
      atomicAdd(&ctrl->total_t, (unsigned long long)agg_t);
    }
  }
  __syncthreads();

  uint32_t run_v = s_prefix[0] + wave_off_v + inc_v - sum_v;
  uint32_t run_t = s_prefix[1] + wave_off_t + inc_t - sum_t;
  const int64_t gw = g.word_off + w_begin;
#pragma unroll
  for (int u = 0; u < kWordsPerThread; ++u) {
    const int i = tid * kWordsPerThread + u;
    if (i < nw) {
      base_v[gw + i] = run_v;
      base_t[gw + i] = run_t;
      run_v += cv[u];
      run_t += ct[u];
      if (w_begin + i == g.nwords - 1) {
        base_v[gw + i + 1] = run_v;
        base_t[gw + i + 1] = run_t;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K3: generate.  One block per tile.  Stages the tile's bits for planes
// k, k+1, k+2 and the vertex bases for planes k, k+1 in LDS, then
//   vertices : thread per output vertex (consecutive ids -> coalesced
//              writes): select its edge in its word, weight, coordinates
//              (Lerp of the corner coordinates, MarchingCubes.h:160-161 and
//              GenerateWorklets.h:65-100), central-difference normal
//   triangles: thread per output triangle: find its cell and visit index in
//              its word, look up the three edges in the case table
//              (MarchingCubes.h:141-167) and turn each into its vertex id.
struct OutPtrs {
  float* vertices;
  float* normals;
  int32_t* conn;
  uint32_t* edges;
  int32_t* interp_first;
  int32_t* interp_second;
  float* interp_weight;
};

__device__ __forceinline__ float load_f(const float* f, int64_t i) { return f[i]; }

// Central-difference gradient at point (x,y,z) of a volume.
__device__ __forceinline__ void point_gradient(const VolGeom& g, const float* f, int x, int y, int z,
                                               float out[3]) {
  const int64_t sx = 1, sy = g.nx, sz = (int64_t)g.nx * g.ny;
  const int64_t p = x + sy * y + sz * z;
  const int c[3] = {x, y, z};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int64_t st[3] = {sx, sy, sz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float h = g.spacing[d];
    if (n[d] < 2) {
      out[d] = 0.0f;
    } else if (c[d] == 0) {
      out[d] = (load_f(f, p + st[d]) - load_f(f, p)) / h;
    } else if (c[d] == n[d] - 1) {
      out[d] = (load_f(f, p) - load_f(f, p - st[d])) / h;
    } else {
      out[d] = (load_f(f, p + st[d]) - load_f(f, p - st[d])) / (2.0f * h);
    }
  }
}

// Output record of one interpolated point on edge (lower point (x,y,z), axis).
__device__ __forceinline__ void emit_point(const VolGeom& g, const float* f, int x, int y, int z,
                                           int axis, bool normals, float xyz[3], float nrm[3],
                                           float& w, int64_t& pa, int64_t& pb) {
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  pa = x + sy * y + sz * z;
  const int64_t st = axis == 0 ? 1 : (axis == 1 ? sy : sz);
  pb = pa + st;
  const float fa = load_f(f, pa), fb = load_f(f, pb);
  w = (g.iso - fa) / (fb - fa);
  const int ca[3] = {x, y, z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float a = g.origin[d] + g.spacing[d] * (float)ca[d];
    const float b = g.origin[d] + g.spacing[d] * (float)(ca[d] + (d == axis ? 1 : 0));
    xyz[d] = a + w * (b - a);
  }
  if (normals) {
    float ga[3], gb[3];
    point_gradient(g, f, x, y, z, ga);
    point_gradient(g, f, x + (axis == 0), y + (axis == 1), z + (axis == 2), gb);
    float m[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = ga[d] + w * (gb[d] - ga[d]);
    const float len2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    if (len2 > 0.0f) {
      const float len = sqrtf(len2);
#pragma unroll
      for (int d = 0; d < 3; ++d) nrm[d] = m[d] / len;
    } else {
      nrm[0] = nrm[1] = nrm[2] = 0.0f;
    }
  }
}

// upper_bound over s[0..n] (inclusive end) minus one: index i with
// s[i] <= v < s[i+1].
__device__ __forceinline__ int find_word(const uint32_t* s, int n, uint32_t v) {
  int lo = 0, hi = n;  // s[lo] <= v < s[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kScanThreads) void k_generate(
    VolGeom g, const float* __restrict__ field, const uint64_t* __restrict__ bits,
    const uint32_t* __restrict__ base_v, const uint32_t* __restrict__ base_t,
    const Ctrl* __restrict__ ctrl, OutPtrs out, unsigned flags, int64_t cap_points,
    int64_t cap_tris, int64_t* __restrict__ counts, int64_t vol_base_v, int64_t vol_base_t) {
  constexpr int kSeg = kTileWords + 2 * kMaxWpr + 2;
  __shared__ uint64_t s_b[3][kSeg];
  __shared__ uint32_t s_v[2][kTileWords + kMaxWpr + 2];
  __shared__ uint32_t s_t[kTileWords + 1];
  __shared__ uint64_t s_case[256];

  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  const bool normals = (flags & DAX_HIP_MC_NORMALS) != 0 && out.normals != nullptr;
  const uint64_t total_v = ctrl->total_v, total_t = ctrl->total_t;
  const uint64_t n_points = dedup ? total_v : 3ull * total_t;
  if (blockIdx.x == 0 && threadIdx.x == 0 && counts) {
    counts[0] = (int64_t)total_t;
    counts[1] = (int64_t)n_points;
  }
  if (total_v > 0x7fffffffull || 3ull * total_t > 0x7fffffffull) return;
  if ((int64_t)total_t > cap_tris || (int64_t)n_points > cap_points) return;

  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int64_t w_begin = tile * kTileWords;
  const int nw = (int)min((int64_t)kTileWords, g.nwords - w_begin);
  const int64_t gw = g.word_off + w_begin;
  const uint32_t tv0 = base_v[gw], tv1 = base_v[gw + nw];
  const uint32_t tt0 = base_t[gw], tt1 = base_t[gw + nw];
  if (tv0 == tv1 && tt0 == tt1) return;  // no output in this tile

  const uint64_t* b = bits + g.word_off;
  const uint32_t* bv = base_v + g.word_off;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int seg_b = nw + 2 * g.wpr + 2;
  const int seg_v = nw + g.wpr + 2;
  for (int i = tid; i < seg_b; i += kScanThreads) {
    const int64_t w = w_begin + i;
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const int64_t wz = w + z * plane;
      s_b[z][i] = (wz < g.nwords) ? b[wz] : 0ull;
    }
  }
  for (int i = tid; i < seg_v; i += kScanThreads) {
    const int64_t w = w_begin + i;
#pragma unroll
    for (int z = 0; z < 2; ++z) {
      const int64_t wz = w + z * plane;
      s_v[z][i] = (wz <= g.nwords) ? bv[wz] : 0u;
    }
  }
  for (int i = tid; i <= nw; i += kScanThreads) s_t[i] = base_t[gw + i];
  s_case[tid] = c_case_table[tid];
  __syncthreads();

  const float* f = field + g.field_off;

  // ---- vertices (dedup: one per active edge)
  if (dedup) {
    for (uint32_t v = tv0 + tid; v < tv1; v += kScanThreads) {
      const int li = find_word(s_v[0], nw, v);
      const WordPos p = word_pos(g, w_begin + li);
      const bool nx1 = p.q + 1 < g.wpr;
      const uint64_t A = s_b[0][li], A1 = nx1 ? s_b[0][li + 1] : 0ull;
      const uint64_t B = s_b[0][li + g.wpr], C = s_b[1][li];
      uint64_t fx, fy, fz;
      edge_flags(p, A, A1, B, C, fx, fy, fz);
      int bit, axis;
      edge_select(fx, fy, fz, (int)(v - s_v[0][li]), bit, axis);
      const int x = p.q * 64 + bit;
      float xyz[3], nrm[3], w;
      int64_t pa, pb;
      emit_point(g, f, x, p.j, p.k, axis, normals, xyz, nrm, w, pa, pb);
      const int64_t ov = (int64_t)v - vol_base_v;
      const int64_t gv = (int64_t)v;
      out.vertices[3 * gv + 0] = xyz[0];
      out.vertices[3 * gv + 1] = xyz[1];
      out.vertices[3 * gv + 2] = xyz[2];
      if (normals) {
        out.normals[3 * gv + 0] = nrm[0];
        out.normals[3 * gv + 1] = nrm[1];
        out.normals[3 * gv + 2] = nrm[2];
      }
      if (out.edges) out.edges[gv] = ((uint32_t)axis << 30) | (uint32_t)pa;
      if (out.interp_first) out.interp_first[gv] = (int32_t)pa;
      if (out.interp_second) out.interp_second[gv] = (int32_t)pb;
      if (out.interp_weight) out.interp_weight[gv] = w;
      (void)ov;
    }
  }

  // ---- triangles
  for (uint32_t t = tt0 + tid; t < tt1; t += kScanThreads) {
    const int li = find_word(s_t, nw, t);
    const WordPos p = word_pos(g, w_begin + li);
    const bool nx1 = p.q + 1 < g.wpr;
    CellPlanes cp;
    cell_planes(p, s_b[0][li], nx1 ? s_b[0][li + 1] : 0ull, s_b[0][li + g.wpr],
                nx1 ? s_b[0][li + g.wpr + 1] : 0ull, s_b[1][li], nx1 ? s_b[1][li + 1] : 0ull,
                s_b[1][li + g.wpr], nx1 ? s_b[1][li + g.wpr + 1] : 0ull, cp);
    int k = (int)(t - s_t[li]);
    uint64_t act = cp.active;
    int bit = 0, cs = 0, nf = 0;
    bool found = false;
    while (act) {
      bit = __builtin_ctzll(act);
      act &= act - 1;
      cs = cell_case(cp, bit);
      nf = (int)(s_case[cs] >> 60);
      if (k < nf) {
        found = true;
        break;
      }
      k -= nf;
    }
    if (!found) continue;  // unreachable when the scan and this kernel agree
    const uint64_t packed = s_case[cs];
    int32_t ids[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const int e = (int)((packed >> (4 * (3 * k + v))) & 0xF);
      const uint8_t code = edge_code(e);
      const int dx = code & 1, dy = (code >> 1) & 1, dz = (code >> 2) & 1, axis = code >> 3;
      const int x = p.q * 64 + bit + dx;
      if (dedup) {
        // the edge's word: same row word or the next one in x, row +dy, plane +dz
        const int qq = x >> 6, bb = x & 63;
        const int lw2 = li + (qq - p.q) + dy * g.wpr;
        WordPos p2;
        p2.q = qq;
        p2.j = p.j + dy;
        p2.k = p.k + dz;
        p2.has_y = p2.j + 1 < g.ny;
        p2.has_z = p2.k + 1 < g.nz;
        p2.pmask = low_mask(g.nx - qq * 64);
        p2.xmask = low_mask(g.nx - 1 - qq * 64);
        const bool nx2 = qq + 1 < g.wpr;
        const uint64_t A = s_b[dz][lw2], A1 = nx2 ? s_b[dz][lw2 + 1] : 0ull;
        const uint64_t B = s_b[dz][lw2 + g.wpr], C = s_b[dz + 1][lw2];
        uint64_t fx, fy, fz;
        edge_flags(p2, A, A1, B, C, fx, fy, fz);
        ids[v] = (int32_t)((int64_t)s_v[dz][lw2] + edge_rank(fx, fy, fz, bb, axis) - vol_base_v);
      } else {
        const int64_t corner = 3 * (int64_t)t + v;
        float xyz[3], nrm[3], w;
        int64_t pa, pb;
        emit_point(g, f, x, p.j + dy, p.k + dz, axis, normals, xyz, nrm, w, pa, pb);
        out.vertices[3 * corner + 0] = xyz[0];
        out.vertices[3 * corner + 1] = xyz[1];
        out.vertices[3 * corner + 2] = xyz[2];
        if (normals) {
          out.normals[3 * corner + 0] = nrm[0];
          out.normals[3 * corner + 1] = nrm[1];
          out.normals[3 * corner + 2] = nrm[2];
        }
        if (out.interp_first) out.interp_first[corner] = (int32_t)pa;
        if (out.interp_second) out.interp_second[corner] = (int32_t)pb;
        if (out.interp_weight) out.interp_weight[corner] = w;
        ids[v] = (int32_t)(corner - 3 * vol_base_t);
      }
    }
    out.conn[3 * (int64_t)t + 0] = ids[0];
    out.conn[3 * (int64_t)t + 1] = ids[1];
    out.conn[3 * (int64_t)t + 2] = ids[2];
  }
}

// ---------------------------------------------------------------------------
// Per-cell outputs of the count step: case byte, NumFaces[case] and the
// exclusive triangle offset.  One wavefront per row word; lane = cell.
__global__ __launch_bounds__(256) void k_cell_outputs(VolGeom g, const uint64_t* __restrict__ bits,
                                                      const uint32_t* __restrict__ base_t,
                                                      uint8_t* __restrict__ cases,
                                                      int32_t* __restrict__ count,
                                                      int32_t* __restrict__ offsets) {
  const int lane = threadIdx.x & 63;
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (w >= g.nwords) return;
  const WordPos p = word_pos(g, w);
  if (!(p.has_y && p.has_z)) return;  // no cells in this row
  const uint64_t* b = bits + g.word_off;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const bool nx1 = p.q + 1 < g.wpr;
  CellPlanes cp;
  cell_planes(p, b[w], nx1 ? b[w + 1] : 0ull, b[w + g.wpr], nx1 ? b[w + g.wpr + 1] : 0ull,
              b[w + plane], nx1 ? b[w + plane + 1] : 0ull, b[w + plane + g.wpr],
              nx1 ? b[w + plane + g.wpr + 1] : 0ull, cp);
  const bool valid = ((p.xmask >> lane) & 1ull) != 0;
  const int cs = cell_case(cp, lane);
  const uint32_t nf = valid ? (uint32_t)(c_case_table[cs] >> 60) : 0u;
  const uint32_t incl = wave_incl_scan(nf);
  if (!valid) return;
  const int64_t cx = g.nx - 1, cy = g.ny - 1;
  const int64_t cell = (int64_t)p.q * 64 + lane + cx * (p.j + cy * (int64_t)p.k);
  if (cases) cases[cell] = (uint8_t)cs;
  if (count) count[cell] = (int32_t)nf;
  if (offsets) offsets[cell] = (int32_t)(base_t[g.word_off + w] + incl - nf);
}

__global__ void k_interpolate_scalar(const float* __restrict__ in, const int32_t* __restrict__ first,
                                     const int32_t* __restrict__ second,
                                     const float* __restrict__ weight, int64_t n,
                                     float* __restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const float a = in[first[u]], b = in[second[u]];
  out[u] = a + weight[u] * (b - a);
}

__global__ void k_point_gradient(VolGeom g, const float* __restrict__ f, float* __restrict__ grad) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = (int64_t)g.nx * g.ny * g.nz;
  if (p >= P) return;
  const int x = (int)(p % g.nx), y = (int)((p / g.nx) % g.ny), z = (int)(p / ((int64_t)g.nx * g.ny));
  float gr[3];
  point_gradient(g, f, x, y, z, gr);
  grad[3 * p + 0] = gr[0];
  grad[3 * p + 1] = gr[1];
  grad[3 * p + 2] = gr[2];
}

// CellGradient (dax/worklet/CellGradient.h:28-48): weights of
// DerivativeWeights(Voxel) at the centre are +-0.25 in the corner order of
// TopologyUniform.h:139-159; sum in vertex order, divided by the voxel width
// coords[6]-coords[0] (Derivative.h:37-75).
__global__ void k_cell_gradient(VolGeom g, const float* __restrict__ f, float* __restrict__ out) {
  const int64_t cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cx * cy * cz) return;
  const int i = (int)(c % cx), j = (int)((c / cx) % cy), k = (int)(c / (cx * cy));
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  const int64_t p0 = i + sy * j + sz * k;
  const int64_t pv[8] = {p0, p0 + 1, p0 + sy + 1, p0 + sy, p0 + sz, p0 + sz + 1, p0 + sz + sy + 1,
                         p0 + sz + sy};
  // sign of the x/y/z weight per corner (all magnitudes 0.25)
  const float sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
  const float syy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  const float szz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const float fv = f[pv[v]];
    s0 = s0 + fv * (0.25f * sx[v]);
    s1 = s1 + fv * (0.25f * syy[v]);
    s2 = s2 + fv * (0.25f * szz[v]);
  }
  const int ijk[3] = {i, j, k};
  float wd[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float a = g.origin[d] + g.spacing[d] * (float)ijk[d];
    const float b = g.origin[d] + g.spacing[d] * (float)(ijk[d] + 1);
    wd[d] = b - a;
  }
  out[3 * c + 0] = s0 / wd[0];
  out[3 * c + 1] = s1 / wd[1];
  out[3 * c + 2] = s2 / wd[2];
}

}  // namespace mc
}  // namespace daxhip

// ---------------------------------------------------------------------------
// C-ABI
using namespace daxhip;
using namespace daxhip::mc;

namespace {

int make_geom(int64_t nx, int64_t ny, int64_t nz, const float* origin, const float* spacing,
              float iso, VolGeom& g) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("dax_hip_mc: grid dimensions must be >= 1 (got %lld x %lld x %lld)", (long long)nx,
              (long long)ny, (long long)nz);
    return DAX_HIP_BAD_VALUE;
  }
  if (nx * ny * nz > 0x7fffffffLL) {
    set_error("dax_hip_mc: %lld points exceed the int32 dax::Id range",
              (long long)(nx * ny * nz));
    return DAX_HIP_BAD_VALUE;
  }
  if (ceil_div(nx, 64) > kMaxWpr) {
    set_error("dax_hip_mc: nx=%lld exceeds the supported row length %d", (long long)nx,
              kMaxWpr * 64);
    return DAX_HIP_BAD_VALUE;
  }
  g.nx = (int32_t)nx;
  g.ny = (int32_t)ny;
  g.nz = (int32_t)nz;
  g.wpr = (int32_t)ceil_div(nx, 64);
  g.iso = iso;
  for (int d = 0; d < 3; ++d) {
    g.origin[d] = origin ? origin[d] : 0.0f;
    g.spacing[d] = spacing ? spacing[d] : 1.0f;
  }
  g.field_off = 0;
  g.word_off = 0;
  g.nwords = (int64_t)g.wpr * ny * nz;
  return DAX_HIP_OK;
}

int check_ws(const VolGeom& g, size_t ws_bytes, Workspace& L, void* ws) {
  L = workspace_layout(g.nwords);
  if (!ws || ws_bytes < L.bytes) {
    set_error("dax_hip_mc: workspace of %zu bytes is smaller than the %zu needed", ws_bytes,
              L.bytes);
    return DAX_HIP_BAD_VALUE;
  }
  return DAX_HIP_OK;
}

constexpr int kClassifyUnroll = 16;

int launch_bits(const VolGeom& g, const float* field, char* ws, const Workspace& L, hipStream_t s) {
  uint64_t* bits = (uint64_t*)(ws + L.off_bits);
  unsigned long long* status = (unsigned long long*)(ws + L.off_status);
  const int64_t zero_words = L.ntiles + (int64_t)(sizeof(Ctrl) / 8);
  const int64_t waves = ceil_div(g.nwords, kClassifyUnroll);
  int64_t blocks = ceil_div(waves * 64, 256);
  blocks = std::max<int64_t>(blocks, ceil_div(zero_words, 256));
  const bool aligned = (g.nx % 64) == 0;
  if (aligned)
    k_classify<kClassifyUnroll, true><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        g, field, bits, status, zero_words);
  else
    k_classify<kClassifyUnroll, false><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        g, field, bits, status, zero_words);
  DAXHIP_LAUNCH_CHECK("k_classify");
  return DAX_HIP_OK;
}

int launch_scan(const VolGeom& g, char* ws, const Workspace& L, hipStream_t s) {
  k_count_scan<<<dim3((unsigned)L.ntiles), dim3(kScanThreads), 0, s>>>(
      g, (const uint64_t*)(ws + L.off_bits), (uint32_t*)(ws + L.off_base_v),
      (uint32_t*)(ws + L.off_base_t), (unsigned long long*)(ws + L.off_status),
      (Ctrl*)(ws + L.off_ctrl), L.ntiles);
  DAXHIP_LAUNCH_CHECK("k_count_scan");
  return DAX_HIP_OK;
}

int launch_generate(const VolGeom& g, const float* field, const char* ws, const Workspace& L,
                    unsigned flags, const OutPtrs& o, int64_t cap_points, int64_t cap_tris,
                    int64_t* counts, hipStream_t s) {
  k_generate<<<dim3((unsigned)L.ntiles), dim3(kScanThreads), 0, s>>>(
      g, field, (const uint64_t*)(ws + L.off_bits), (const uint32_t*)(ws + L.off_base_v),
      (const uint32_t*)(ws + L.off_base_t), (const Ctrl*)(ws + L.off_ctrl), o, flags, cap_points,
      cap_tris, counts, 0, 0);
  DAXHIP_LAUNCH_CHECK("k_generate");
  return DAX_HIP_OK;
}

int check_outputs(const VolGeom& g, const float* field, unsigned flags, const float* vertices,
                  const int32_t* conn, const uint32_t* edges, const int64_t* counts) {
  if (!field || !vertices || !conn || !counts) {
    set_error("dax_hip_mc: field, vertices, connectivity and counts are required");
    return DAX_HIP_BAD_VALUE;
  }
  if (edges && (!(flags & DAX_HIP_MC_DEDUP) || (int64_t)g.nx * g.ny * g.nz > (1LL << 30))) {
    set_error("dax_hip_mc: the compact edge list needs dedup and at most 2^30 points");
    return DAX_HIP_BAD_VALUE;
  }
  return DAX_HIP_OK;
}

__global__ void k_write_counts(const Ctrl* ctrl, int64_t* counts, unsigned flags) {
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  counts[0] = (int64_t)ctrl->total_t;
  counts[1] = dedup ? (int64_t)ctrl->total_v : 3 * (int64_t)ctrl->total_t;
}

}  // namespace

extern "C" {

size_t dax_hip_mc_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  return workspace_layout(ceil_div(nx, 64) * ny * nz).bytes;
}

int dax_hip_mc_bits(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                    void* d_workspace, size_t workspace_bytes, void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, nullptr, iso, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  if (!d_field) {
    set_error("dax_hip_mc_bits: field is NULL");
    return DAX_HIP_BAD_VALUE;
  }
  return launch_bits(g, d_field, (char*)d_workspace, L, (hipStream_t)stream);
}

int dax_hip_mc_scan(int64_t nx, int64_t ny, int64_t nz, void* d_workspace, size_t workspace_bytes,
                    void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, nullptr, 0.0f, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  return launch_scan(g, (char*)d_workspace, L, (hipStream_t)stream);
}

int dax_hip_mc_generate(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                        const float* h_origin3, const float* h_spacing3, float iso,
                        unsigned flags, const void* d_workspace, size_t workspace_bytes,
                        float* d_vertices, float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                        int32_t* d_interp_first, int32_t* d_interp_second,
                        float* d_interp_weight, int64_t cap_points, int64_t cap_tris,
                        int64_t* d_counts, void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, h_origin3, h_spacing3, iso, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, (void*)d_workspace))) return rc;
  if ((rc = check_outputs(g, d_field, flags, d_vertices, d_conn, d_edges, d_counts))) return rc;
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, d_interp_first, d_interp_second,
            d_interp_weight};
  return launch_generate(g, d_field, (const char*)d_workspace, L, flags, o, cap_points, cap_tris,
                         d_counts, (hipStream_t)stream);
}

int dax_hip_mc_classify(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                        void* d_workspace, size_t workspace_bytes, int64_t* d_counts,
                        void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, nullptr, iso, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  if (!d_field) {
    set_error("dax_hip_mc_classify: field is NULL");
    return DAX_HIP_BAD_VALUE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  if ((rc = launch_bits(g, d_field, ws, L, s))) return rc;
  if ((rc = launch_scan(g, ws, L, s))) return rc;
  if (d_counts) {
    k_write_counts<<<1, 1, 0, s>>>((const Ctrl*)(ws + L.off_ctrl), d_counts, DAX_HIP_MC_DEDUP);
    DAXHIP_LAUNCH_CHECK("k_write_counts");
  }
  return DAX_HIP_OK;
}

int dax_hip_mc_forward(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                       const float* h_origin3, const float* h_spacing3, float iso,
                       unsigned flags, void* d_workspace, size_t workspace_bytes,
                       float* d_vertices, float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                       int32_t* d_interp_first, int32_t* d_interp_second,
                       float* d_interp_weight, int64_t cap_points, int64_t cap_tris,
                       int64_t* d_counts, void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, h_origin3, h_spacing3, iso, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  if ((rc = check_outputs(g, d_field, flags, d_vertices, d_conn, d_edges, d_counts))) return rc;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  if ((rc = launch_bits(g, d_field, ws, L, s))) return rc;
  if ((rc = launch_scan(g, ws, L, s))) return rc;
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, d_interp_first, d_interp_second,
            d_interp_weight};
  return launch_generate(g, d_field, ws, L, flags, o, cap_points, cap_tris, d_counts, s);
}

int dax_hip_mc_read_counts(const int64_t* d_counts, int64_t cap_points, int64_t cap_tris,
                           int64_t* h_tris, int64_t* h_points, void* stream) {
  int64_t c[2];
  DAXHIP_CHECK(hipMemcpyAsync(c, d_counts, sizeof(c), hipMemcpyDeviceToHost, (hipStream_t)stream));
  DAXHIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  if (h_tris) *h_tris = c[0];
  if (h_points) *h_points = c[1];
  if (c[1] > 0x7fffffffLL || 3 * c[0] > 0x7fffffffLL) {
    set_error("dax_hip_mc: %lld triangles / %lld points exceed the int32 dax::Id range",
              (long long)c[0], (long long)c[1]);
    return DAX_HIP_EXEC_ERROR;
  }
  if (c[0] > cap_tris || c[1] > cap_points) {
    set_error("dax_hip_mc: output needs %lld triangles / %lld points, capacity %lld / %lld",
              (long long)c[0], (long long)c[1], (long long)cap_tris, (long long)cap_points);
    return DAX_HIP_CAPACITY;
  }
  return DAX_HIP_OK;
}

int dax_hip_mc_cell_outputs(int64_t nx, int64_t ny, int64_t nz, const void* d_workspace,
                            size_t workspace_bytes, uint8_t* d_cases, int32_t* d_count,
                            int32_t* d_offsets, void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, nullptr, 0.0f, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, (void*)d_workspace))) return rc;
  const char* ws = (const char*)d_workspace;
  const int64_t blocks = ceil_div(g.nwords * 64, 256);
  if (blocks == 0) return DAX_HIP_OK;
  k_cell_outputs<<<dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream>>>(
      g, (const uint64_t*)(ws + L.off_bits), (const uint32_t*)(ws + L.off_base_t), d_cases,
      d_count, d_offsets);
  DAXHIP_LAUNCH_CHECK("k_cell_outputs");
  return DAX_HIP_OK;
}

int dax_hip_mc_interpolate_scalar(const float* d_in, const int32_t* d_first,
                                  const int32_t* d_second, const float* d_weight, int64_t n,
                                  float* d_out, void* stream) {
  if (n <= 0) return DAX_HIP_OK;
  k_interpolate_scalar<<<dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream>>>(
      d_in, d_first, d_second, d_weight, n, d_out);
  DAXHIP_LAUNCH_CHECK("k_interpolate_scalar");
  return DAX_HIP_OK;
}

int dax_hip_point_gradient(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                           const float* h_spacing3, float* d_grad, void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, h_spacing3, 0.0f, g);
  if (rc) return rc;
  const int64_t P = nx * ny * nz;
  k_point_gradient<<<dim3((unsigned)ceil_div(P, 256)), dim3(256), 0, (hipStream_t)stream>>>(
      g, d_field, d_grad);
  DAXHIP_LAUNCH_CHECK("k_point_gradient");
  return DAX_HIP_OK;
}

int dax_hip_cell_gradient(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                          const float* h_origin3, const float* h_spacing3, float* d_grad,
                          void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, h_origin3, h_spacing3, 0.0f, g);
  if (rc) return rc;
  const int64_t C = (nx - 1) * (ny - 1) * (nz - 1);
  if (C <= 0) return DAX_HIP_OK;
  k_cell_gradient<<<dim3((unsigned)ceil_div(C, 256)), dim3(256), 0, (hipStream_t)stream>>>(
      g, d_field, d_grad);
  DAXHIP_LAUNCH_CHECK("k_cell_gradient");
  return DAX_HIP_OK;
}

}  // extern "C"
