// Device bodies of the marching-cubes kernels, shared by the single-volume
// launches (mc_forward.hip) and the batched persistent launches
// (mc_batched.hip).
//
//   classify   : field -> inside-bits                         (one pass over f)
//   count/scan : per-word edge and triangle counts, exclusive scan with
//                decoupled look-back over scan tiles
//   generate   : per generate tile, maps of the tile's output vertices and
//                triangles built in LDS, then one thread per vertex and per
//                triangle with coalesced writes
#pragma once

#include "mc_common.h"

namespace daxhip {
namespace mc {

// Case table in constant memory (one copy per translation unit).
static __constant__ uint64_t c_case_table[256] = DAX_HIP_MC_CASE_TABLE_INIT;

// ---------------------------------------------------------------------------
// Incremental word positions (32-bit arithmetic, one division per start).
struct WordIter {
  int q, j, k;
  __device__ __forceinline__ void init(const VolGeom& g, int lw) {
    const uint32_t row = fast_div((uint32_t)lw, g.div_wpr);
    const uint32_t kk = fast_div(row, g.div_ny);
    q = lw - (int)row * g.wpr;
    j = (int)(row - kk * (uint32_t)g.ny);
    k = (int)kk;
  }
  __device__ __forceinline__ void next(const VolGeom& g) {
    if (++q == g.wpr) {
      q = 0;
      if (++j == g.ny) {
        j = 0;
        ++k;
      }
    }
  }
  __device__ __forceinline__ WordPos pos(const VolGeom& g) const {
    WordPos p;
    p.q = q;
    p.j = j;
    p.k = k;
    p.has_y = j + 1 < g.ny;
    p.has_z = k + 1 < g.nz;
    const bool has_cells = g.nx > 1 && g.ny > 1 && g.nz > 1;
    const int x0 = q * 64;
    p.pmask = has_cells ? low_mask(g.nx - x0) : 0ull;
    p.xmask = has_cells ? low_mask(g.nx - 1 - x0) : 0ull;
    return p;
  }
};

// ---------------------------------------------------------------------------
// classify, one float per lane: a wave builds UNROLL consecutive words.
template <int UNROLL, bool ALIGNED>
__device__ __forceinline__ void classify_words(const VolGeom& g, const float* __restrict__ f,
                                               uint64_t* __restrict__ bits, int64_t w0,
                                               int64_t w_end) {
  const int lane = __lane_id();
  float v[UNROLL];
  bool ok[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    const int64_t w = w0 + u;
    ok[u] = false;
    v[u] = 0.0f;
    if (w < w_end) {
      if (ALIGNED) {
        v[u] = __builtin_nontemporal_load(f + w * 64 + lane);
        ok[u] = true;
      } else {
        const int64_t row = fast_div((uint32_t)w, g.div_wpr);
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
  if (lane < UNROLL && w0 + lane < w_end) bits[w0 + lane] = mine;
}

// classify, float4 per lane (rows of a multiple of 64 points): a wave reads
// UNROLL chunks of 256 consecutive points; lane l holds points 4l..4l+3 of a
// chunk, and a DPP OR-scan over each 16-lane row assembles the chunk's four
// 64-bit words.
template <int UNROLL>
__device__ __forceinline__ void classify_chunks_v4(const VolGeom& g, const float* __restrict__ f,
                                                   uint64_t* __restrict__ bits, int64_t c0,
                                                   int64_t nwords) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const int lane = __lane_id();
  const f32x4* f4 = reinterpret_cast<const f32x4*>(f);
  f32x4 v[UNROLL];
  bool ok[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    const int64_t c = c0 + u;
    ok[u] = (4 * c + (lane >> 4)) < nwords;
    v[u] = ok[u] ? __builtin_nontemporal_load(f4 + c * 64 + lane) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  }
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    const float iso = g.iso;
    const uint32_t nib = (uint32_t)(v[u].x > iso) | ((uint32_t)(v[u].y > iso) << 1) |
                         ((uint32_t)(v[u].z > iso) << 2) | ((uint32_t)(v[u].w > iso) << 3);
    const int sh = 4 * (lane & 15);
    uint32_t lo = sh < 32 ? nib << sh : 0u;
    uint32_t hi = sh >= 32 ? nib << (sh - 32) : 0u;
    lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x111, 0xf, 0xf, false);
    hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x111, 0xf, 0xf, false);
    lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x112, 0xf, 0xf, false);
    hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x112, 0xf, 0xf, false);
    lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x114, 0xf, 0xf, false);
    hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x114, 0xf, 0xf, false);
    lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x118, 0xf, 0xf, false);
    hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x118, 0xf, 0xf, false);
    if ((lane & 15) == 15 && ok[u]) bits[4 * (c0 + u) + (lane >> 4)] = ((uint64_t)hi << 32) | lo;
  }
}

// ---------------------------------------------------------------------------
// count/scan.  A scan tile is NT threads x WPT consecutive row words.  Each
// thread loads its words and their +x/+y/+z neighbour words straight from
// global memory (L2), counts, and the block scans the per-thread sums; wave 0
// then runs the decoupled look-back for the tile's global prefix.
struct ScanShared {
  uint32_t wave_v[16], wave_t[16];
  uint32_t prefix[2];
};

// Active edges (vertices) and triangles of one row word from its eight
// neighbour words (A = the word, 1 = the next word along x, B = +y row,
// C = +z plane, D = +y+z).
__device__ __forceinline__ void word_counts_regs(const VolGeom& g, const WordPos& p, uint64_t A,
                                                 uint64_t A1, uint64_t B, uint64_t B1, uint64_t C,
                                                 uint64_t C1, uint64_t D, uint64_t D1,
                                                 uint32_t& nv, uint32_t& nt) {
  uint64_t fx, fy, fz;
  edge_flags(p, A, A1, B, C, fx, fy, fz);
  nv = (uint32_t)(popc(fx) + popc(fy) + popc(fz));
  CellPlanes cp;
  cell_planes(p, A, A1, B, B1, C, C1, D, D1, cp);
  nt = cp.active ? word_triangles(cp) : 0u;
}

// Loads the eight neighbour words of row word w (volume-local) and counts it.
__device__ __forceinline__ void word_counts(const VolGeom& g, const uint64_t* __restrict__ b,
                                            int64_t w, uint32_t& nv, uint32_t& nt) {
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int wpr = g.wpr;
  const WordPos p = word_pos(g, w);
  const bool nx1 = p.q + 1 < wpr;
  const bool hy = w + wpr < g.nwords, hz = w + plane < g.nwords, hyz = w + plane + wpr < g.nwords;
  const uint64_t A = b[w], A1 = nx1 ? b[w + 1] : 0ull;
  const uint64_t B = hy ? b[w + wpr] : 0ull, B1 = (hy && nx1) ? b[w + wpr + 1] : 0ull;
  const uint64_t C = hz ? b[w + plane] : 0ull, C1 = (hz && nx1) ? b[w + plane + 1] : 0ull;
  const uint64_t D = hyz ? b[w + plane + wpr] : 0ull;
  const uint64_t D1 = (hyz && nx1) ? b[w + plane + wpr + 1] : 0ull;
  word_counts_regs(g, p, A, A1, B, B1, C, C1, D, D1, nv, nt);
}

// Two count tiles of kCountTile row words, one wave: lane l counts the four
// consecutive words 2*kCountTile*pair + 4l .. +3 (lanes 0-31 the first
// tile, 32-63 the second), reusing each loaded word as its predecessor's +x
// neighbour, and writes each word's exclusive prefix within its tile
// (v | t << 16).  Returns the two tiles' sums (v | t << 32, wave-uniform).
__device__ __forceinline__ void count_tile_pair(const VolGeom& g, const uint64_t* __restrict__ b,
                                                uint32_t* __restrict__ lpre, int64_t pair,
                                                int64_t n_ctiles, unsigned long long& sum0,
                                                unsigned long long& sum1) {
  static_assert(kCountTile == 128, "32 lanes of four words per count tile");
  const int lane = __lane_id();
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int wpr = g.wpr;
  const int64_t w0 = pair * (2 * kCountTile) + 4 * lane;
  uint64_t A[5], B[5], C[5], D[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const int64_t w = w0 + i;
    A[i] = w < g.nwords ? b[w] : 0ull;
    B[i] = w + wpr < g.nwords ? b[w + wpr] : 0ull;
    C[i] = w + plane < g.nwords ? b[w + plane] : 0ull;
    D[i] = w + plane + wpr < g.nwords ? b[w + plane + wpr] : 0ull;
  }
  uint32_t nv[4] = {0u, 0u, 0u, 0u}, nt[4] = {0u, 0u, 0u, 0u};
  if (w0 < g.nwords) {
    WordIter wi;
    wi.init(g, (int)w0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (w0 + i < g.nwords) {
        const WordPos p = wi.pos(g);
        const bool nx1 = p.q + 1 < wpr;
        word_counts_regs(g, p, A[i], nx1 ? A[i + 1] : 0ull, B[i], nx1 ? B[i + 1] : 0ull, C[i],
                         nx1 ? C[i + 1] : 0ull, D[i], nx1 ? D[i + 1] : 0ull, nv[i], nt[i]);
        wi.next(g);
      }
    }
  }
  const uint32_t sv = nv[0] + nv[1] + nv[2] + nv[3], st = nt[0] + nt[1] + nt[2] + nt[3];
  uint32_t iv = wave_incl_scan(sv), it = wave_incl_scan(st);
  const uint32_t hv = (uint32_t)__builtin_amdgcn_readlane((int)iv, 31);
  const uint32_t ht = (uint32_t)__builtin_amdgcn_readlane((int)it, 31);
  const uint32_t ev = (uint32_t)__builtin_amdgcn_readlane((int)iv, 63);
  const uint32_t et = (uint32_t)__builtin_amdgcn_readlane((int)it, 63);
  sum0 = (unsigned long long)hv | ((unsigned long long)ht << 32);
  sum1 = (unsigned long long)(ev - hv) | ((unsigned long long)(et - ht) << 32);
  if (lane >= 32) {  // the second tile's scan starts at lane 32
    iv -= hv;
    it -= ht;
  }
  const int64_t tile = pair * 2 + (lane >> 5);
  if (tile < n_ctiles) {
    uint32_t rv = iv - sv, rt = it - st;
    uint32_t out[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      out[i] = rv | (rt << 16);
      rv += nv[i];
      rt += nt[i];
    }
    *reinterpret_cast<uint4*>(lpre + w0) = make_uint4(out[0], out[1], out[2], out[3]);
  }
}

// Decoupled look-back by wave 0 over the tile states; returns the exclusive
// prefix of `tile` (vertex, triangle) in sh.prefix.
__device__ __forceinline__ void look_back(unsigned long long* __restrict__ status, int64_t tile,
                                          uint32_t agg_v, uint32_t agg_t, Ctrl* __restrict__ ctrl,
                                          uint32_t* prefix) {
  const int lane = __lane_id();
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
    prefix[0] = ex_v;
    prefix[1] = ex_t;
    atomicAdd(&ctrl->total_v, (unsigned long long)agg_v);
    atomicAdd(&ctrl->total_t, (unsigned long long)agg_t);
  }
}

// LDS words of a count/scan tile: one u32 per row word, one pad word per 32
// so that both the coalesced (word u*NT + tid) and the per-thread
// (word tid*WPT + u) access patterns are free of bank conflicts.
template <int NT, int WPT>
struct ScanTileLds {
  static constexpr int kWords = NT * WPT;
  static constexpr int kSize = kWords + kWords / 32;
};
__device__ __forceinline__ int pad32(int i) { return i + (i >> 5); }

// One scan tile of a volume: words [w0, w0 + NT*WPT) (volume-local), status
// slot `tile` in the global tile order.
//   1. per-word vertex / triangle counts, consecutive threads on consecutive
//      words (coalesced row loads), packed into LDS
//   2. each thread sums WPT consecutive words; block scan; decoupled
//      look-back for the tile's prefix
//   3. the bases go back through LDS and out with coalesced stores
template <int NT, int WPT>
__device__ __forceinline__ void count_scan_tile(const VolGeom& g, const uint64_t* __restrict__ b,
                                                uint32_t* __restrict__ base_v,
                                                uint32_t* __restrict__ base_t,
                                                unsigned long long* __restrict__ status,
                                                Ctrl* __restrict__ ctrl, int64_t tile,
                                                int64_t tile_w0, ScanShared& sh,
                                                uint32_t* __restrict__ s_pk) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int wpr = g.wpr;
#pragma unroll
  for (int u = 0; u < WPT; ++u) {
    const int i = u * NT + tid;
    const int64_t w = tile_w0 + i;
    uint32_t pk = 0;
    if (w < g.nwords) {
      const WordPos p = word_pos(g, w);
      const bool nx1 = p.q + 1 < wpr;
      const bool hy = w + wpr < g.nwords, hz = w + plane < g.nwords,
                 hyz = w + plane + wpr < g.nwords;
      const uint64_t A = b[w], A1 = nx1 ? b[w + 1] : 0ull;
      const uint64_t B = hy ? b[w + wpr] : 0ull, B1 = (hy && nx1) ? b[w + wpr + 1] : 0ull;
      const uint64_t C = hz ? b[w + plane] : 0ull, C1 = (hz && nx1) ? b[w + plane + 1] : 0ull;
      const uint64_t D = hyz ? b[w + plane + wpr] : 0ull;
      const uint64_t D1 = (hyz && nx1) ? b[w + plane + wpr + 1] : 0ull;
      uint32_t nv, nt;
      word_counts_regs(g, p, A, A1, B, B1, C, C1, D, D1, nv, nt);
      pk = nv | (nt << 16);
    }
    s_pk[pad32(i)] = pk;
  }
  __syncthreads();
  const int64_t w0 = tile_w0 + (int64_t)tid * WPT;
  uint32_t packed[WPT];
  uint32_t sum_v = 0, sum_t = 0;
#pragma unroll
  for (int u = 0; u < WPT; ++u) {
    packed[u] = s_pk[pad32(tid * WPT + u)];
    sum_v += packed[u] & 0xffffu;
    sum_t += packed[u] >> 16;
  }
  const uint32_t inc_v = wave_incl_scan(sum_v), inc_t = wave_incl_scan(sum_t);
  if (lane == 63) {
    sh.wave_v[wv] = inc_v;
    sh.wave_t[wv] = inc_t;
  }
  __syncthreads();
  uint32_t off_v = 0, off_t = 0, agg_v = 0, agg_t = 0;
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) {
    if (q < wv) {
      off_v += sh.wave_v[q];
      off_t += sh.wave_t[q];
    }
    agg_v += sh.wave_v[q];
    agg_t += sh.wave_t[q];
  }
  if (wv == 0) look_back(status, tile, agg_v, agg_t, ctrl, sh.prefix);
  __syncthreads();
  // vertex bases, then triangle bases, each through LDS
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    uint32_t run = sh.prefix[pass] + (pass == 0 ? off_v + inc_v - sum_v : off_t + inc_t - sum_t);
    uint32_t* base = pass == 0 ? base_v : base_t;
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      s_pk[pad32(tid * WPT + u)] = run;
      run += pass == 0 ? (packed[u] & 0xffffu) : (packed[u] >> 16);
      if (w0 + u == g.nwords - 1) base[g.word_off + g.nwords] = run;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int i = u * NT + tid;
      const int64_t w = tile_w0 + i;
      if (w < g.nwords) base[g.word_off + w] = s_pk[pad32(i)];
    }
    __syncthreads();
  }
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so neighbouring tiles share its L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
  const int64_t per = n / 8, rem = n % 8;
  const int64_t x = b % 8, i = b / 8;
  return x * per + (x < rem ? x : rem) + i;
}

// ---------------------------------------------------------------------------
// generate
struct OutPtrs {
  float* vertices;
  float* normals;
  int32_t* conn;
  uint32_t* edges;
  int32_t* interp_first;
  int32_t* interp_second;
  float* interp_weight;
};

// Point-field sources of the forward: a float32 field in memory, or the
// signed plane distance of a slice computed from the point's coordinates.
// src(x, y, z, p) is the value at point (x, y, z) of flat index p (a volume
// holds fewer than 2^31 points, so 32-bit index arithmetic is exact).
struct FieldArray {
  const float* __restrict__ f;
  __device__ __forceinline__ float operator()(int, int, int, int32_t p) const { return f[p]; }
};

// The field SliceCount / SliceGenerate classify and interpolate
// (dax/worklet/Slice.h:44-57,104-120): dot(normal, c - origin) at the point
// coordinates c = grid origin + spacing * ijk (dax/cont/UniformGrid.h:82-89),
// the dot product summed from component 0 (dax/Types.h:611-621), against
// the isovalue dot(normal, origin).
struct FieldPlane {
  float o[3], s[3];  // grid origin and spacing
  float po[3], pn[3];  // plane origin and normal
  __device__ __forceinline__ float operator()(int x, int y, int z, int32_t) const {
    const int c[3] = {x, y, z};
    float r = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float t = pn[d] * ((o[d] + s[d] * (float)c[d]) - po[d]);
      r = d == 0 ? t : r + t;
    }
    return r;
  }
};

// Central-difference gradient at (x,y,z): interior (f(p+e)-f(p-e))/(2h),
// first plane (f(p+e)-f(p))/h, last plane (f(p)-f(p-e))/h, and 0 along an
// axis with a single point.  Split into the loads (branch-free, so they all
// issue at once) and the arithmetic, so callers can overlap other work with
// the loads' latency.
template <class Src>
__device__ __forceinline__ void gradient_loads(const VolGeom& g, const Src& f, int x, int y, int z,
                                               float lo_v[3], float hi_v[3]) {
  const int32_t sy = g.nx, sz = g.nx * g.ny;
  const int32_t p = x + sy * y + sz * z;
  const int c[3] = {x, y, z};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int32_t st[3] = {1, sy, sz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int dl = c[d] > 0 ? 1 : 0, dh = c[d] < n[d] - 1 ? 1 : 0;
    lo_v[d] = f(x - (d == 0) * dl, y - (d == 1) * dl, z - (d == 2) * dl, p - dl * st[d]);
    hi_v[d] = f(x + (d == 0) * dh, y + (d == 1) * dh, z + (d == 2) * dh, p + dh * st[d]);
  }
}

__device__ __forceinline__ void gradient_finish(const VolGeom& g, int x, int y, int z,
                                                const float lo_v[3], const float hi_v[3],
                                                float out[3]) {
  const int c[3] = {x, y, z};
  const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool wide = c[d] > 0 && c[d] < n[d] - 1;
    const float h = g.spacing[d];
    const float diff = hi_v[d] - lo_v[d];
    if (n[d] < 2)
      out[d] = 0.0f;
    else if ((g.pow2_h >> d) & 1u)  // exact: the same value as the division
      out[d] = diff * (wide ? 0.5f * g.rcp_h[d] : g.rcp_h[d]);
    else
      out[d] = diff / (wide ? 2.0f * h : h);
  }
}

template <class Src>
__device__ __forceinline__ void point_gradient(const VolGeom& g, const Src& f, int x, int y,
                                               int z, float out[3]) {
  float lo_v[3], hi_v[3];
  gradient_loads(g, f, x, y, z, lo_v, hi_v);
  gradient_finish(g, x, y, z, lo_v, hi_v, out);
}

// Field values of one interpolated point on the edge with lower endpoint
// (x,y,z) along axis: the end values and, with normals, the central-
// difference neighbours of both ends.
struct PointLoads {
  float fa, fb;
  float alo[3], ahi[3], blo[3], bhi[3];
};

template <class Src>
__device__ __forceinline__ void point_loads(const VolGeom& g, const Src& f, int x, int y, int z,
                                            int axis, bool normals, PointLoads& L, int32_t& pa,
                                            int32_t& pb) {
  const int32_t sy = g.nx, sz = g.nx * g.ny;
  pa = x + sy * y + sz * z;
  pb = pa + (axis == 0 ? 1 : (axis == 1 ? sy : sz));
  L.fa = f(x, y, z, pa);
  L.fb = f(x + (axis == 0), y + (axis == 1), z + (axis == 2), pb);
  if (normals) {
    gradient_loads(g, f, x, y, z, L.alo, L.ahi);
    gradient_loads(g, f, x + (axis == 0), y + (axis == 1), z + (axis == 2), L.blo, L.bhi);
  }
}

// Coordinates, unit normal (central differences interpolated along the edge)
// and weight of the point whose field values L holds.
__device__ __forceinline__ void point_finish(const VolGeom& g, int x, int y, int z, int axis,
                                             bool normals, const PointLoads& L, float xyz[3],
                                             float nrm[3], float& w) {
  float ga[3], gb[3];
  if (normals) {
    gradient_finish(g, x, y, z, L.alo, L.ahi, ga);
    gradient_finish(g, x + (axis == 0), y + (axis == 1), z + (axis == 2), L.blo, L.bhi, gb);
  }
  w = (g.iso - L.fa) / (L.fb - L.fa);
  const int ca[3] = {x, y, z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float a = g.origin[d] + g.spacing[d] * (float)ca[d];
    const float b = g.origin[d] + g.spacing[d] * (float)(ca[d] + (d == axis ? 1 : 0));
    xyz[d] = a + w * (b - a);
  }
  if (normals) {
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

// One interpolated point on the edge with lower endpoint (x,y,z) along axis.
template <class Src>
__device__ __forceinline__ void emit_point(const VolGeom& g, const Src& f, int x, int y, int z,
                                           int axis, bool normals, float xyz[3], float nrm[3],
                                           float& w, int32_t& pa, int32_t& pb) {
  PointLoads L;
  point_loads(g, f, x, y, z, axis, normals, L, pa, pb);
  point_finish(g, x, y, z, axis, normals, L, xyz, nrm, w);
}

// index i with s[i] <= v < s[i+1], s[0..n]
__device__ __forceinline__ int find_word(const uint32_t* s, int n, uint32_t v) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Position of the k-th set bit of m (k < popcount(m)): the half of m that
// holds it, then halving steps on 32 bits.
__device__ __forceinline__ int select_bit(uint64_t m, int k) {
  const uint32_t lo = (uint32_t)m;
  const int nlo = __popc(lo);
  const bool up = k >= nlo;
  uint32_t x = up ? (uint32_t)(m >> 32) : lo;
  int r = up ? k - nlo : k, pos = up ? 32 : 0;
#pragma unroll
  for (int w = 16; w > 0; w >>= 1) {
    const int c = __popc(x & ((1u << w) - 1u));
    const bool hi = r >= c;
    r = hi ? r - c : r;
    x = hi ? x >> w : x;
    pos += hi ? w : 0;
  }
  return pos;
}

// Edge codes of the 12 cell edges, 5 bits each (edge_code).
__host__ __device__ constexpr uint64_t edge_codes_packed() {
  uint64_t r = 0;
  for (int e = 0; e < 12; ++e) r |= (uint64_t)edge_code(e) << (5 * e);
  return r;
}
constexpr uint64_t kEdgeCodes = edge_codes_packed();

// ---------------------------------------------------------------------------
// Dispatch orders of tile grids (generate tiles, the adjoint's gather tiles):
// tiles of class 0 first, then class 1, then class 2, each in index order,
// from a class byte per tile.  The grid's blocks take order[1 + blockIdx].
constexpr uint32_t kOrderHasOutput = 0x80000000u;  // order entry flag: class 0 or 1

// A thread of the ordering block owns a chunk of K consecutive tiles.  Up to
// kOrderRegTiles of them have their classes loaded before the scan, two bits
// each in a register (the latencies overlap the scan's): byte loads for K
// up to 8, 4-byte loads for K a multiple of 4 up to 32.  Larger chunks (K a
// multiple of 16, volumes of more than 2^28 points) read their class bytes
// in 16-byte vectors while the order is written.  The class arrays end in
// 16 bytes of padding.
constexpr int kOrderRegTiles = 32;
__device__ __forceinline__ uint32_t order_chunk(uint32_t ntl, uint32_t threads) {
  const uint32_t k = (ntl + threads - 1) / threads;
  return k <= 8u ? k : (k <= (uint32_t)kOrderRegTiles ? (k + 3u) / 4u * 4u : (k + 15u) / 16u * 16u);
}
// The loads are unconditional (an address outside the thread's tiles is
// replaced by the first byte, its value dropped), so they issue together
// instead of one conditional load and wait at a time.
__device__ __forceinline__ uint64_t load_classes(const uint8_t* __restrict__ cls, uint32_t t0,
                                                 uint32_t K, uint32_t ntl) {
  uint64_t r = 0ull;
  if (K <= 8u) {
    uint32_t b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool ok = (uint32_t)k < K && t0 + k < ntl;
      b[k] = cls[ok ? t0 + k : 0u];
      b[k] = ok ? b[k] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) r |= (uint64_t)(b[k] & 3u) << (2 * k);
  } else if (K <= (uint32_t)kOrderRegTiles) {
    uint32_t w[kOrderRegTiles / 4];
#pragma unroll
    for (int v = 0; v < kOrderRegTiles / 4; ++v) {
      const uint32_t t = t0 + 4u * v;
      const bool ok = 4u * v < K && t < ntl;  // t0 and K multiples of 4: aligned
      w[v] = *reinterpret_cast<const uint32_t*>(cls + (ok ? t : 0u));
      w[v] = ok ? w[v] : 0u;
    }
#pragma unroll
    for (int v = 0; v < kOrderRegTiles / 4; ++v)
#pragma unroll
      for (int j = 0; j < 4; ++j) r |= (uint64_t)((w[v] >> (8 * j)) & 3u) << (2 * (4 * v + j));
  }
  return r;
}

// Exclusive block scans of two counts per thread (all threads call it; s_a
// and s_b hold a word per wave).
__device__ __forceinline__ void block_scan2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb,
                                            uint32_t& ta, uint32_t& tb, uint32_t* s_a,
                                            uint32_t* s_b) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t ia = wave_incl_scan(a), ib = wave_incl_scan(b);
  if (lane == 63) {
    s_a[wv] = ia;
    s_b[wv] = ib;
  }
  __syncthreads();
  uint32_t oa = 0u, ob = 0u;
  ta = tb = 0u;
  const int nwv = (int)(blockDim.x >> 6);
  for (int q = 0; q < nwv; ++q) {
    if (q < wv) {
      oa += s_a[q];
      ob += s_b[q];
    }
    ta += s_a[q];
    tb += s_b[q];
  }
  __syncthreads();  // s_a / s_b reusable
  ea = oa + ia - a;
  eb = ob + ib - b;
}

// Calls fn(k, class) for the thread's tiles k = 0 .. n-1 in order: from reg
// (K <= kOrderRegTiles) or from the class bytes at cls + t0.
template <bool LARGE, class Fn>
__device__ __forceinline__ void for_classes(uint64_t reg, const uint8_t* __restrict__ cls,
                                            uint32_t t0, uint32_t K, uint32_t n, Fn&& fn) {
  if (!LARGE || K <= (uint32_t)kOrderRegTiles) {
    for (uint32_t k = 0; k < n; ++k) fn(k, (uint32_t)(reg >> (2 * k)) & 3u);
    return;
  }
  for (uint32_t v = 0; v < n; v += 16u) {
    const uint4 q = *reinterpret_cast<const uint4*>(cls + t0 + v);
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (v + j < n) fn(v + j, (w[j >> 2] >> (8 * (j & 3))) & 3u);
  }
}

// The next slot of class c's run (cursors s0, s1, s2), in selects and adds:
// a conditional post-increment of one of three cursors is lowered to a
// dynamically indexed private array, one scratch round trip per tile.
__device__ __forceinline__ uint32_t next_slot(uint32_t c, uint32_t& s0, uint32_t& s1,
                                              uint32_t& s2) {
  const uint32_t slot = c == 0u ? s0 : (c == 1u ? s1 : s2);
  s0 += c == 0u ? 1u : 0u;
  s1 += c == 1u ? 1u : 0u;
  s2 += c >= 2u ? 1u : 0u;
  return slot;
}

// Stores order[1 + slot] = v for the (slot, v) pairs that emit_all(put)
// passes to put (every slot in [0, n) once).  Scattered 4-byte stores of
// one block all go through one CU's address unit, a cache line at a time
// (a 16384-entry order took ~10 us); with s_stage (LDS of `stage` entries)
// the pairs are placed in LDS one window of slots at a time, emit_all
// running once per window, and each window leaves in coalesced stores.
// Without s_stage the pairs go straight to global memory.
template <class EmitAll>
__device__ __forceinline__ void store_order(uint32_t* order, uint32_t n, uint32_t* s_stage,
                                            uint32_t stage, EmitAll&& emit_all) {
  if (s_stage == nullptr) {
    emit_all([&](uint32_t slot, uint32_t v) { order[1u + slot] = v; });
    return;
  }
  for (uint32_t lo = 0; lo < n; lo += stage) {
    emit_all([&](uint32_t slot, uint32_t v) {
      if (slot - lo < stage) s_stage[slot - lo] = v;
    });
    __syncthreads();
    const uint32_t m = min(stage, n - lo);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) order[1u + lo + i] = s_stage[i];
    __syncthreads();
  }
}

// order[1 ..] = tiles [0, ntl) of class 0, then 1, then 2, each by index,
// flagged kOrderHasOutput in classes 0 and 1; order[0] = tag.  The thread
// owns tiles [tid * K, tid * K + K).  s_stage may alias s_a and s_b.
template <bool LARGE>
__device__ __forceinline__ void write_order(uint64_t reg, const uint8_t* __restrict__ cls,
                                            uint32_t K, uint32_t ntl, uint32_t* order,
                                            uint32_t tag, uint32_t* s_a, uint32_t* s_b,
                                            uint32_t* s_stage = nullptr, uint32_t stage = 0u) {
  const uint32_t t0 = (uint32_t)threadIdx.x * K;
  const uint32_t n = t0 < ntl ? min(K, ntl - t0) : 0u;
  uint32_t c0 = 0u, c1 = 0u;
  for_classes<LARGE>(reg, cls, t0, K, n, [&](uint32_t, uint32_t c) {
    c0 += c == 0u;
    c1 += c == 1u;
  });
  uint32_t e0, e1, tot0, tot1;
  block_scan2(c0, c1, e0, e1, tot0, tot1, s_a, s_b);
  const uint32_t b0 = e0, b1 = tot0 + e1, b2 = tot0 + tot1 + (t0 - e0 - e1);
  store_order(order, ntl, s_stage, stage, [&](auto&& put) {
    uint32_t s0 = b0, s1 = b1, s2 = b2;
    for_classes<LARGE>(reg, cls, t0, K, n, [&](uint32_t k, uint32_t c) {
      put(next_slot(c, s0, s1, s2), (t0 + k) | (c < 2u ? kOrderHasOutput : 0u));
    });
  });
  if (threadIdx.x == 0) order[0] = tag;
}

// The same order over units of G tiles that lie a plane apart: unit u =
// zg * sp + r (sp tiles per plane) holds tiles (G zg + m) sp + r, m < G,
// and takes the lowest class of its tiles.  The unit of rank R puts its
// tiles at positions 8 G (R / 8) + R % 8 + 8 m: blocks go to the XCDs
// round robin, so a unit's tiles run on one XCD, in consecutive rounds, and
// the field planes they share (three of four for neighbouring planes) are
// fetched into that XCD's L2 once.  Needs every unit whole and nunits a
// multiple of 8; K G <= kOrderRegTiles (the tile classes are loaded before
// the scan, two bits each, by load_unit_classes).
constexpr uint32_t kOrderUnitTiles = 2;  // G
__device__ __forceinline__ uint64_t load_unit_classes(const uint8_t* __restrict__ cls, uint32_t K,
                                                      uint32_t nunits, uint32_t sp) {
  constexpr uint32_t G = kOrderUnitTiles;
  const uint32_t u0 = (uint32_t)threadIdx.x * K;
  const uint32_t n = u0 < nunits ? min(K, nunits - u0) : 0u;
  uint32_t zg = n ? u0 / sp : 0u, r = n ? u0 - zg * sp : 0u;
  uint32_t c[kOrderRegTiles];  // unconditional loads, as in load_classes
#pragma unroll
  for (int k = 0; k < kOrderRegTiles / (int)G; ++k) {
    const bool ok = (uint32_t)k < n;
#pragma unroll
    for (uint32_t m = 0; m < G; ++m) {
      c[k * G + m] = cls[ok ? (G * zg + m) * sp + r : 0u];
      c[k * G + m] = ok ? c[k * G + m] : 0u;
    }
    if (++r == sp) {
      r = 0u;
      ++zg;
    }
  }
  uint64_t reg = 0ull;  // class of tile m of unit k at bits 2 (k G + m)
#pragma unroll
  for (int k = 0; k < kOrderRegTiles; ++k) reg |= (uint64_t)(c[k] & 3u) << (2 * k);
  return reg;
}
__device__ __forceinline__ void write_order_units(uint64_t reg, uint32_t K, uint32_t nunits,
                                                  uint32_t sp, uint32_t* order, uint32_t tag,
                                                  uint32_t* s_a, uint32_t* s_b,
                                                  uint32_t* s_stage = nullptr,
                                                  uint32_t stage = 0u) {
  constexpr uint32_t G = kOrderUnitTiles;
  const uint32_t u0 = (uint32_t)threadIdx.x * K;
  const uint32_t n = u0 < nunits ? min(K, nunits - u0) : 0u;
  auto unit_class = [&](uint32_t k) {
    uint32_t c = 3u;
#pragma unroll
    for (uint32_t m = 0; m < G; ++m) c = min(c, (uint32_t)(reg >> (2u * (k * G + m))) & 3u);
    return c;
  };
  uint32_t c0 = 0u, c1 = 0u;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t c = unit_class(k);
    c0 += c == 0u;
    c1 += c == 1u;
  }
  uint32_t e0, e1, tot0, tot1;
  block_scan2(c0, c1, e0, e1, tot0, tot1, s_a, s_b);
  const uint32_t b0 = e0, b1 = tot0 + e1, b2 = tot0 + tot1 + (u0 - e0 - e1);
  store_order(order, nunits * G, s_stage, stage, [&](auto&& put) {
    uint32_t s0 = b0, s1 = b1, s2 = b2;
    uint32_t zg = n ? u0 / sp : 0u, r = n ? u0 - zg * sp : 0u;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t R = next_slot(unit_class(k), s0, s1, s2);
      const uint32_t base = 8u * G * (R >> 3) + (R & 7u);
#pragma unroll
      for (uint32_t m = 0; m < G; ++m) {
        const uint32_t tc = (uint32_t)(reg >> (2u * (k * G + m))) & 3u;
        put(base + 8u * m, ((G * zg + m) * sp + r) | (tc < 2u ? kOrderHasOutput : 0u));
      }
      if (++r == sp) {
        r = 0u;
        ++zg;
      }
    }
  });
  if (threadIdx.x == 0) order[0] = tag;
}

// LDS layout of a generate tile of TW words (wpr known at run time).
constexpr int kMaxGenThreads = 256;
template <int TW>
struct GenLayout {
  static constexpr int kRoundTris = 5 * kMaxGenThreads;  // triangles of one round of cells
  int segb, segv;
  __host__ __device__ GenLayout(int wpr) : segb(TW + 2 * wpr + 2), segv(TW + wpr + 2) {}
  // u64: bits (3 planes), case table, active cells; u32: vertex bases
  // (2 planes), triangle bases, active-cell slots, scan scratch, round
  // offsets and cell info; u8: the round's triangle -> cell map
  __host__ __device__ size_t bytes() const {
    return (size_t)(3 * segb + 256 + TW) * 8 +
           (size_t)(2 * segv + 2 * (TW + 1) + 32 + 2 * kMaxGenThreads + 1) * 4 + kRoundTris;
  }
};

// Edge flags of the word at LDS index lw of plane sA (its +z plane sC) with
// position p; branch-free (the +x word is read and masked).
__device__ __forceinline__ void lds_flags(const VolGeom& g, const uint64_t* sA, const uint64_t* sC,
                                          int lw, const WordPos& p, uint64_t& fx, uint64_t& fy,
                                          uint64_t& fz) {
  const uint64_t A = sA[lw], A1n = sA[lw + 1], B = sA[lw + g.wpr], C = sC[lw];
  edge_flags(p, A, p.q + 1 < g.wpr ? A1n : 0ull, B, C, fx, fy, fz);
}

__device__ __forceinline__ WordPos make_pos(const VolGeom& g, int q, int j, int k) {
  WordPos p;
  p.q = q;
  p.j = j;
  p.k = k;
  p.has_y = j + 1 < g.ny;
  p.has_z = k + 1 < g.nz;
  p.pmask = low_mask(g.nx - q * 64);
  p.xmask = low_mask(g.nx - 1 - q * 64);
  return p;
}

// Exclusive scan of one value per thread over the block (thread order);
// returns the thread's offset and the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total, uint32_t* s_w) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int q = 0; q < nwv; ++q) {
    if (q < wv) off += s_w[q];
    tot += s_w[q];
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

// Where the generate kernel of the single-volume forward takes its word
// bases from: the count kernel's per-word prefixes within their count tiles
// (lpre, v | t << 16), every count tile's prefix within its count block of
// tpb tiles (tloc, v | t << 32) and the count blocks' exclusive prefixes
// (bpre, the totals after the last); the tiles per block are a power of
// two.  It writes the bases of its words to
// base_v / base_t, which the adjoint and the per-cell outputs read.
struct CountSrc {
  const uint32_t* __restrict__ lpre;
  const unsigned long long* __restrict__ tloc;
  const unsigned long long* __restrict__ bpre;
  int64_t n_ctiles;
  int tpb_shift;  // log2 of the count tiles per count block
  uint32_t* __restrict__ base_v;
  uint32_t* __restrict__ base_t;
};

__device__ __forceinline__ unsigned long long add_vt(unsigned long long a, unsigned long long b) {
  return (unsigned long long)((uint32_t)a + (uint32_t)b) |
         ((unsigned long long)((uint32_t)(a >> 32) + (uint32_t)(b >> 32)) << 32);
}

// Exclusive vertex / triangle prefix (v | t << 32) of count tile c; for c at
// or past the last tile, the totals.
__device__ __forceinline__ unsigned long long ctile_prefix(const CountSrc& cs, int64_t c) {
  if (c >= cs.n_ctiles)
    return cs.bpre[(cs.n_ctiles + (1 << cs.tpb_shift) - 1) >> cs.tpb_shift];
  return add_vt(cs.bpre[c >> cs.tpb_shift], cs.tloc[c]);
}

// Vertex and triangle base of word w (v | t << 32); past the last count
// tile the totals.
__device__ __forceinline__ unsigned long long word_bases(const CountSrc& cs, int64_t w) {
  const int64_t ct = w / kCountTile;
  const unsigned long long p = ctile_prefix(cs, ct);
  if (ct >= cs.n_ctiles) return p;
  const uint32_t l = cs.lpre[w];
  return add_vt(p, (unsigned long long)(l & 0xffffu) | ((unsigned long long)(l >> 16) << 32));
}

// The same with every load issued (clamped indices, results selected), so
// that the loads of several words can be in flight together.
__device__ __forceinline__ unsigned long long word_bases_nb(const CountSrc& cs, int64_t w) {
  const int64_t ct = w / kCountTile;
  const bool past = ct >= cs.n_ctiles;
  const int64_t bi = past ? (cs.n_ctiles + (1 << cs.tpb_shift) - 1) >> cs.tpb_shift
                          : ct >> cs.tpb_shift;
  const unsigned long long bp = cs.bpre[bi];
  const unsigned long long tl = cs.tloc[past ? 0 : ct];
  const uint32_t l = cs.lpre[past ? 0 : w];
  const unsigned long long full =
      add_vt(add_vt(bp, tl), (unsigned long long)(l & 0xffffu) | ((unsigned long long)(l >> 16) << 32));
  return past ? bp : full;
}

// One generate tile: words [w_begin, w_begin + TW) of a volume.
//   1. stage bits (planes k, k+1, k+2), vertex bases (k, k+1), triangle bases
//   2. per word: active-cell mask; block scan of the active-cell counts
//      gives every active cell a slot
//   3. one thread per output vertex: word by binary search over the vertex
//      bases, edge by popcount select; coordinates, normal, records
//   4. rounds of up to blockDim active cells: one thread per cell (word by
//      binary search over the slots, cell by popcount select) forms its
//      case; a block scan of NumFaces gives its first triangle and the cell
//      marks its triangles' entries of the round's triangle -> cell map; one
//      thread per triangle turns the triangle's edges into vertex ids through
//      the edge flags, formed where they are read from the staged bits
//
// FROM_COUNTS: the bases come from the count kernel's per-word prefixes and
// count-tile prefixes (cs) instead of bv / bt; the tile writes the bases of
// its words (and, for the last tile, entry nwords) to cs.base_v / base_t,
// and with emit false only that.  has_output skips the check for a tile
// without output (one round trip).
template <int TW, class Src, bool FROM_COUNTS = false>
__device__ __forceinline__ void generate_tile(const VolGeom& g, const Src& f,
                                              const uint64_t* __restrict__ b,
                                              const uint32_t* __restrict__ bv,
                                              const uint32_t* __restrict__ bt, int64_t w_begin,
                                              const OutPtrs& out, bool dedup, bool normals,
                                              int64_t vol_base_v, int64_t vol_base_t,
                                              uint64_t* s_dyn, const CountSrc& cs = CountSrc{},
                                              bool emit = true, bool has_output = false) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nw = (int)min((int64_t)TW, g.nwords - w_begin);
  if constexpr (FROM_COUNTS) {
    // count tiles without output (has_output: known not to be the case):
    // constant bases, nothing else to do
    const int64_t c0 = w_begin / kCountTile;
    const int64_t c1 = min((w_begin + nw + kCountTile - 1) / kCountTile, cs.n_ctiles);
    const unsigned long long p0 = has_output ? 0ull : ctile_prefix(cs, c0);
    const unsigned long long p1 = has_output ? 1ull : ctile_prefix(cs, c1);
    if (p0 == p1) {
      const uint32_t v0 = (uint32_t)p0, t0 = (uint32_t)(p0 >> 32);
      for (int i = tid; i < nw; i += nt) {
        cs.base_v[w_begin + i] = v0;
        cs.base_t[w_begin + i] = t0;
      }
      if (tid == 0 && w_begin + nw == g.nwords) {
        cs.base_v[g.nwords] = v0;
        cs.base_t[g.nwords] = t0;
      }
      return;
    }
  } else {
    if (bv[w_begin] == bv[w_begin + nw] && bt[w_begin] == bt[w_begin + nw])
      return;  // uniform: no output in this tile
  }

  const GenLayout<TW> L(g.wpr);
  uint64_t* s_b0 = s_dyn;
  uint64_t* s_b1 = s_b0 + L.segb;
  uint64_t* s_b2 = s_b1 + L.segb;
  uint64_t* s_case = s_b2 + L.segb;
  uint64_t* s_act = s_case + 256;
  uint32_t* s_v0 = reinterpret_cast<uint32_t*>(s_act + TW);
  uint32_t* s_v1 = s_v0 + L.segv;
  uint32_t* s_t = s_v1 + L.segv;
  uint32_t* s_cs = s_t + TW + 1;
  uint32_t* s_w = s_cs + TW + 1;                     // scan scratch (<= 16 waves)
  uint32_t* s_rfirst = s_w + 32;                     // [nt + 1]
  uint32_t* s_rinfo = s_rfirst + kMaxGenThreads + 1;  // [nt]: li << 16 | bit << 8 | case
  uint8_t* s_tcell = reinterpret_cast<uint8_t*>(s_rinfo + kMaxGenThreads);  // [kRoundTris]

  // staging: each thread issues the loads of up to kStageU of its words
  // (clamped indices, every load unconditional) before storing any of them,
  // so they are in flight together rather than one conditional load and
  // wait at a time
  constexpr int kStageU = 4;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  for (int i0 = tid; i0 < L.segb; i0 += kStageU * nt) {
    uint64_t v[kStageU][3];
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
      const int64_t w = w_begin + i0 + u * nt;
#pragma unroll
      for (int z = 0; z < 3; ++z) v[u][z] = b[w + z * plane < g.nwords ? w + z * plane : 0];
    }
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
      const int i = i0 + u * nt;
      if (i < L.segb) {
        const int64_t w = w_begin + i;
        s_b0[i] = (w < g.nwords) ? v[u][0] : 0ull;
        s_b1[i] = (w + plane < g.nwords) ? v[u][1] : 0ull;
        s_b2[i] = (w + 2 * plane < g.nwords) ? v[u][2] : 0ull;
      }
    }
  }
  if constexpr (FROM_COUNTS) {
    for (int i0 = tid; i0 < L.segv; i0 += kStageU * nt) {
      unsigned long long p0[kStageU], p1[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const int i = i0 + u * nt;
        p0[u] = word_bases_nb(cs, w_begin + i);
        p1[u] = word_bases_nb(cs, w_begin + plane + i);
      }
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const int i = i0 + u * nt;
        if (i < L.segv) {
          s_v0[i] = (uint32_t)p0[u];
          s_v1[i] = (uint32_t)p1[u];
          if (i <= nw) s_t[i] = (uint32_t)(p0[u] >> 32);
        }
      }
    }
  } else {
    for (int i0 = tid; i0 < L.segv; i0 += kStageU * nt) {
      uint32_t a0[kStageU], a1[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const int64_t w = w_begin + i0 + u * nt;
        a0[u] = bv[w <= g.nwords ? w : 0];
        a1[u] = bv[w + plane <= g.nwords ? w + plane : 0];
      }
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const int i = i0 + u * nt;
        if (i < L.segv) {
          const int64_t w = w_begin + i;
          s_v0[i] = (w <= g.nwords) ? a0[u] : 0u;
          s_v1[i] = (w + plane <= g.nwords) ? a1[u] : 0u;
        }
      }
    }
    for (int i = tid; i <= nw; i += nt) s_t[i] = bt[w_begin + i];
  }
  for (int i = tid; i < 256; i += nt) s_case[i] = c_case_table[i];
  __syncthreads();
  const uint32_t tv0 = s_v0[0], tv1 = s_v0[nw];
  const uint32_t tt0 = s_t[0], tt1 = s_t[nw];
  if constexpr (FROM_COUNTS) {
    for (int i = tid; i < nw; i += nt) {
      cs.base_v[w_begin + i] = s_v0[i];
      cs.base_t[w_begin + i] = s_t[i];
    }
    if (tid == 0 && w_begin + nw == g.nwords) {
      cs.base_v[g.nwords] = tv1;
      cs.base_t[g.nwords] = tt1;
    }
    if (!emit || (tv0 == tv1 && tt0 == tt1)) return;  // bases only, or no output here
  }

  // 2. per word (each thread a run of consecutive words): active cells
  const int wpt = (TW + nt - 1) / nt;
  uint32_t ncell = 0;
  {
    const int i0 = tid * wpt;
    if (i0 < nw) {
      WordIter it;
      it.init(g, (int)(w_begin + i0));
      for (int u = 0; u < wpt && i0 + u < nw; ++u) {
        const int li = i0 + u;
        const WordPos p = it.pos(g);
        const bool nx1 = p.q + 1 < g.wpr;
        CellPlanes cp;
        cell_planes(p, s_b0[li], nx1 ? s_b0[li + 1] : 0ull, s_b0[li + g.wpr],
                    nx1 ? s_b0[li + g.wpr + 1] : 0ull, s_b1[li], nx1 ? s_b1[li + 1] : 0ull,
                    s_b1[li + g.wpr], nx1 ? s_b1[li + g.wpr + 1] : 0ull, cp);
        s_act[li] = cp.active;
        ncell += (uint32_t)popc(cp.active);
        it.next(g);
      }
    }
  }
  uint32_t n_cells;
  {
    uint32_t run = block_excl_scan(ncell, n_cells, s_w);
    const int i0 = tid * wpt;
    for (int u = 0; u < wpt && i0 + u < nw; ++u) {
      s_cs[i0 + u] = run;
      run += (uint32_t)popc(s_act[i0 + u]);
    }
    if (tid == 0) s_cs[nw] = n_cells;
  }
  __syncthreads();

  // 3. vertices (dedup: one per active edge, in id order).  A thread's
  //    first vertex is located and its field values loaded here, and
  //    finished after the triangle rounds: the loads' latency overlaps the
  //    rounds' LDS work (their barriers wait for LDS operations only).
  const int nV = (int)(tv1 - tv0);
  auto locate = [&](int vl, int& x, int& y, int& z, int& axis) {
    const int li = find_word(s_v0, nw, tv0 + (uint32_t)vl);
    const WordPos p = word_pos(g, w_begin + li);
    uint64_t fx, fy, fz;
    lds_flags(g, s_b0, s_b1, li, p, fx, fy, fz);
    int bit;
    edge_select(fx, fy, fz, (int)(tv0 + (uint32_t)vl - s_v0[li]), bit, axis);
    x = p.q * 64 + bit;
    y = p.j;
    z = p.k;
  };
  auto store_vertex = [&](int vl, int axis, const float xyz[3], const float nrm[3], float w,
                          int32_t pa, int32_t pb) {
    const int64_t gv = (int64_t)tv0 + vl;
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
  };
  const bool first = dedup && tid < nV;
  int fx0 = 0, fy0 = 0, fz0 = 0, fax = 0;
  int32_t fpa = 0, fpb = 0;
  PointLoads fl;
  if (first) {
    locate(tid, fx0, fy0, fz0, fax);
    point_loads(g, f, fx0, fy0, fz0, fax, normals, fl, fpa, fpb);
  }

  // 4. triangles
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < n_cells; c0 += (uint32_t)nt) {  // uniform
    const uint32_t slot = c0 + (uint32_t)tid;
    uint32_t nf = 0;
    if (slot < n_cells) {
      const int li = find_word(s_cs, nw, slot);
      const int bit = select_bit(s_act[li], (int)(slot - s_cs[li]));
      const WordPos p = word_pos(g, w_begin + li);
      const bool nx1 = p.q + 1 < g.wpr;
      CellPlanes cp;
      cell_planes(p, s_b0[li], nx1 ? s_b0[li + 1] : 0ull, s_b0[li + g.wpr],
                  nx1 ? s_b0[li + g.wpr + 1] : 0ull, s_b1[li], nx1 ? s_b1[li + 1] : 0ull,
                  s_b1[li + g.wpr], nx1 ? s_b1[li + g.wpr + 1] : 0ull, cp);
      const int cs = cell_case(cp, bit);
      nf = (uint32_t)(s_case[cs] >> 60);
      s_rinfo[tid] = ((uint32_t)li << 16) | ((uint32_t)bit << 8) | (uint32_t)cs;
    }
    uint32_t round_tris;
    const uint32_t first = block_excl_scan(nf, round_tris, s_w);
    for (uint32_t t = 0; t < nf; ++t) s_tcell[first + t] = (uint8_t)tid;  // <= 5 per cell
    s_rfirst[tid] = first;
    __syncthreads();
    for (uint32_t tl = (uint32_t)tid; tl < round_tris; tl += (uint32_t)nt) {
      const int c = (int)s_tcell[tl];
      const uint32_t info = s_rinfo[c];
      const int li = (int)(info >> 16), bit = (int)((info >> 8) & 0xff), cs = (int)(info & 0xff);
      const int visit = (int)(tl - s_rfirst[c]);
      const WordPos p = word_pos(g, w_begin + li);
      const uint64_t packed = s_case[cs];
      const int64_t t = (int64_t)tt0 + carry + tl;
      int32_t ids[3];
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        const int e = (int)((packed >> (4 * (3 * visit + v))) & 0xF);
        const int code = (int)((kEdgeCodes >> (5 * e)) & 31u);
        const int dx = code & 1, dy = (code >> 1) & 1, dz = (code >> 2) & 1, axis = code >> 3;
        const int x = p.q * 64 + bit + dx;
        if (dedup) {
          const int qq = x >> 6, bb = x & 63;
          const int lw2 = li + (qq - p.q) + dy * g.wpr;  // <= nw + wpr
          const WordPos p2 = make_pos(g, qq, p.j + dy, p.k + dz);
          uint64_t fx, fy, fz;
          lds_flags(g, dz ? s_b1 : s_b0, dz ? s_b2 : s_b1, lw2, p2, fx, fy, fz);
          const uint32_t base = (dz ? s_v1 : s_v0)[lw2];
          ids[v] = (int32_t)((int64_t)base + edge_rank(fx, fy, fz, bb, axis) - vol_base_v);
        } else {
          const int64_t corner = 3 * t + v;
          float xyz[3], nrm[3], w;
          int32_t pa, pb;
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
      out.conn[3 * t + 0] = ids[0];
      out.conn[3 * t + 1] = ids[1];
      out.conn[3 * t + 2] = ids[2];
    }
    carry += round_tris;
    __syncthreads();  // s_rfirst / s_rinfo / s_tcell reused by the next round
  }

  if (first) {
    float xyz[3], nrm[3], w;
    point_finish(g, fx0, fy0, fz0, fax, normals, fl, xyz, nrm, w);
    store_vertex(tid, fax, xyz, nrm, w, fpa, fpb);
  }
  if (dedup) {
    for (int vl = tid + nt; vl < nV; vl += nt) {
      int x, y, z, axis;
      locate(vl, x, y, z, axis);
      float xyz[3], nrm[3], w;
      int32_t pa, pb;
      emit_point(g, f, x, y, z, axis, normals, xyz, nrm, w, pa, pb);
      store_vertex(vl, axis, xyz, nrm, w, pa, pb);
    }
  }
}

}  // namespace mc
}  // namespace daxhip
