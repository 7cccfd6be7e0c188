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
//
// Tile shapes (classify unroll and load width, scan sub-tiles, generate tile
// words) are chosen per grid shape by the autotuner at the end of this file.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mc_kernels.h"

namespace daxhip {
namespace mc {

// Classify row words [w_lo, w_hi) (w_lo a multiple of 4 for the float4 form)
// and clear zero_words words of count/scan state.
template <int UNROLL, bool ALIGNED>
__global__ __launch_bounds__(256) void k_classify(VolGeom g, const float* __restrict__ field,
                                                  uint64_t* __restrict__ bits,
                                                  unsigned long long* __restrict__ zero_buf,
                                                  int64_t zero_words, int64_t w_lo, int64_t w_hi) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid < zero_words) zero_buf[gtid] = 0ull;
  const int64_t w0 = w_lo + (gtid >> 6) * UNROLL;
  classify_words<UNROLL, ALIGNED>(g, field + g.field_off, bits + g.word_off, w0, w_hi);
}

template <int UNROLL>
__global__ __launch_bounds__(256) void k_classify_v4(VolGeom g, const float* __restrict__ field,
                                                     uint64_t* __restrict__ bits,
                                                     unsigned long long* __restrict__ zero_buf,
                                                     int64_t zero_words, int64_t w_lo,
                                                     int64_t w_hi) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid < zero_words) zero_buf[gtid] = 0ull;
  const int64_t c0 = w_lo / 4 + (gtid >> 6) * UNROLL;
  classify_chunks_v4<UNROLL>(g, field + g.field_off, bits + g.word_off, c0, w_hi);
}

// ---- count and scan --------------------------------------------------------
// 1. k_count: a block of WAVES waves, each wave two count tiles of
//    kCountTile row words (four consecutive words per lane, coalesced loads
//    of the words and their +x/+y/+z neighbours): every word's prefix within
//    its tile, every tile's prefix within the block and output size, and the
//    block's sums.
//    No inter-block dependency.
// 2. k_scan_tsums: one block scans the blocks' sums into exclusive prefixes
//    (the totals after the last) and the 64-bit totals of the control block.
// 3. the bases of every word (block + tile + local prefix): the generate
//    kernel stores them for its words; k_bases stores them alone, for calls
//    that classify without generating.
// Dispatch classes of the generate and gather tiles (a byte per tile):
// generate tiles 0 heavy (at least 5/4 outputs, vertices + triangles, per
// word), 1 other with output, 2 none; gather tiles 0 with output of their
// own, 2 none.  tw = 0: no generate classes (the count blocks do not cover
// whole generate tiles).
struct TileClasses {
  uint8_t* gcls;
  uint8_t* tcls;
  int tw;
};

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_count(VolGeom g, const uint64_t* __restrict__ bits,
                                                      uint32_t* __restrict__ lpre,
                                                      unsigned long long* __restrict__ tloc,
                                                      unsigned long long* __restrict__ bsum,
                                                      TileClasses tc, int64_t n_ctiles) {
  constexpr int kBlockTiles = 2 * WAVES;  // count tiles per block
  __shared__ unsigned long long s_ts[kBlockTiles];
  const int wv = threadIdx.x >> 6;
  // XCD-contiguous block order: the block one plane ahead, whose words this
  // block reads as +z neighbours, runs on the same XCD (same L2)
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t pair = blk * WAVES + wv;
  unsigned long long s0 = 0ull, s1 = 0ull;
  if (2 * pair < n_ctiles) count_tile_pair(g, bits + g.word_off, lpre, pair, n_ctiles, s0, s1);
  if ((threadIdx.x & 63) == 0) {
    s_ts[2 * wv] = s0;
    s_ts[2 * wv + 1] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long run = 0ull;
    for (int i = 0; i < kBlockTiles; ++i) {
      const int64_t tile = blk * kBlockTiles + i;
      if (tile < n_ctiles) tloc[tile] = run;
      run = add_vt(run, s_ts[i]);
    }
    bsum[blk] = run;
  } else if (threadIdx.x < 64) {  // the block's generate and gather tiles
    const int lane = threadIdx.x;
    auto weight = [&](int first, int n) -> uint32_t {  // outputs of count tiles [first, +n)
      uint32_t w = 0u;
      for (int i = 0; i < n; ++i)
        if (blk * kBlockTiles + first + i < n_ctiles)
          w += (uint32_t)s_ts[first + i] + (uint32_t)(s_ts[first + i] >> 32);
      return w;
    };
    constexpr int gct = kGatherTileWords / kCountTile;
    const int l = lane - 1;
    if (l >= 0 && l < (kBlockTiles + gct - 1) / gct) {
      const int64_t c0 = blk * kBlockTiles + l * gct;
      if (c0 < n_ctiles) {
        const uint32_t w = weight(l * gct, gct);
        tc.tcls[c0 / gct] = w == 0u ? 2 : (4u * w >= 8u * (uint32_t)kGatherTileWords ? 0 : 1);
      }
    }
    const int cpt = tc.tw / kCountTile;
    if (tc.tw > 0 && l >= 0 && l < kBlockTiles / cpt) {
      const int64_t c0 = blk * kBlockTiles + l * cpt;
      if (c0 < n_ctiles) {
        const uint32_t w = weight(l * cpt, cpt);
        tc.gcls[c0 / cpt] = w == 0u ? 2 : (4ull * w >= 5ull * (uint32_t)tc.tw ? 0 : 1);
      }
    }
  }
}

// Dispatch order of the generate tiles, written by the scan block from the
// count kernel's tile classes: heavy tiles first, then the other tiles with
// output, then the empty ones; within a class, index order.  A grid
// dispatches its blocks roughly in index order as slots free up, so the
// longest tiles start first and the short ones fill the tail (a coarse
// longest-processing-time-first schedule).  Every tile writes at fixed
// offsets, so an order changes no output.  Each thread of the scan block
// takes a contiguous chunk of tiles; two block scans of its class counts
// give its slots.  The adjoint orders its gather tiles the same way, from
// the count kernel's gather classes.
struct GenOrder {
  const uint8_t* __restrict__ gcls;  // classes of the generate tiles
  int64_t nwords;
  int tw;            // generate tile words: 64, 128, 256, 512 or 1024
  bool gen_classes;  // gcls holds the classes
  uint32_t unit_sp;  // > 0: order units of kOrderUnitTiles tiles a plane (unit_sp tiles) apart
  uint32_t* order;   // [0]: tw (0: no order), [1 + i]: the i-th tile to dispatch
  uint32_t* gorder;  // the gather tiles' order: entry 0 cleared (the adjoint writes it)
};

constexpr int kTsumThreads = 1024;
constexpr int kTsumWaves = kTsumThreads / 64;
constexpr uint32_t kOrderStage = 16384;  // 64 KiB of LDS
constexpr int kTsumPerThread = 4;        // block sums per thread and round

// One block scans the n count-block sums (v | t << 32, each below 2^32)
// 4096 at a time: four consecutive sums per thread, wave scans of the
// thread totals' halves, the block's wave sums in 64 bits, 64-bit carries
// between rounds.  The same block then
// writes the two dispatch orders.
// LARGE: some thread's chunk exceeds kOrderRegTiles tiles (a separate
// instance, so the common one stays small).
template <bool LARGE>
__global__ __launch_bounds__(kTsumThreads) void k_scan_tsums(const unsigned long long* __restrict__ bsum,
                                                             unsigned long long* __restrict__ bpre,
                                                             int64_t n, Ctrl* __restrict__ ctrl,
                                                             GenOrder go) {
  constexpr int kWaves = kTsumWaves;
  // the scan's wave sums, then the dispatch order's LDS window
  __shared__ uint32_t s_stage[kOrderStage];
  uint32_t* s_v = s_stage;
  uint32_t* s_t = s_stage + kWaves;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t ng = (uint32_t)ceil_div(go.nwords, (int64_t)go.tw);
  const uint32_t Kg = order_chunk(ng, kTsumThreads);
  const bool ordered = go.gen_classes;  // uniform
  const bool units = ordered && go.unit_sp > 0u;
  const uint32_t nunits = units ? ng / kOrderUnitTiles : 0u;
  const uint32_t Ku = units ? order_chunk(nunits, kTsumThreads) : 0u;
  const uint64_t gcls =
      units ? load_unit_classes(go.gcls, Ku, nunits, go.unit_sp)
            : (ordered ? load_classes(go.gcls, (uint32_t)tid * Kg, Kg, ng) : 0ull);

  unsigned long long cv = 0, ct = 0;  // 64-bit carries
  constexpr int kPer = kTsumPerThread;
  for (int64_t e0 = 0; e0 < n; e0 += kPer * kTsumThreads) {
    const int64_t e = e0 + (int64_t)kPer * tid;  // the thread's kPer consecutive sums
    uint32_t v[kPer], t[kPer], sv = 0u, st = 0u;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const unsigned long long x = e + i < n ? bsum[e + i] : 0ull;
      v[i] = (uint32_t)x;
      t[i] = (uint32_t)(x >> 32);
      sv += v[i];
      st += t[i];
    }
    // a round's sums stay below 2^32 only for the wave partials (a count
    // block holds at most 3 * 2^17 vertices); the cross-wave totals are
    // formed in 64 bits
    const uint32_t iv = wave_incl_scan(sv), it = wave_incl_scan(st);
    if (lane == 63) {
      s_v[wv] = iv;
      s_t[wv] = it;
    }
    __syncthreads();
    unsigned long long ov = 0, ot = 0, rv = 0, rt = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      if (q < wv) {
        ov += s_v[q];
        ot += s_t[q];
      }
      rv += s_v[q];
      rt += s_t[q];
    }
    __syncthreads();
    unsigned long long pv = cv + ov + iv - sv, pt = ct + ot + it - st;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      if (e + i < n) bpre[e + i] = (pv & 0xffffffffull) | (pt << 32);
      pv += v[i];
      pt += t[i];
    }
    cv += rv;
    ct += rt;
  }
  if (tid == 0) {
    bpre[n] = (cv & 0xffffffffull) | (ct << 32);
    ctrl->total_v = cv;
    ctrl->total_t = ct;
  }

  if (units)
    write_order_units(gcls, Ku, nunits, go.unit_sp, go.order, (uint32_t)go.tw, s_v, s_t, s_stage,
                      kOrderStage);
  else if (ordered)
    write_order<LARGE>(gcls, go.gcls, Kg, ng, go.order, (uint32_t)go.tw, s_v, s_t, s_stage,
                       kOrderStage);
  else if (tid == 0)
    go.order[0] = 0u;
  // the adjoint builds the gather tiles' order itself; until then there is none
  if (tid == 0) go.gorder[0] = 0u;
}

// Bases of every row word (and entry nwords): tile prefix + local prefix.
__global__ __launch_bounds__(256) void k_bases(VolGeom g, CountSrc cs) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w > g.nwords) return;
  const unsigned long long p = word_bases(cs, w);
  cs.base_v[g.word_off + w] = (uint32_t)p;
  cs.base_t[g.word_off + w] = (uint32_t)(p >> 32);
}

template <int TW, class Src>
__global__ __launch_bounds__(kScanThreads, 5) void k_generate(
    VolGeom g, Src field, const uint64_t* __restrict__ bits, CountSrc cs,
    const Ctrl* __restrict__ ctrl, OutPtrs out, unsigned flags, int64_t cap_points,
    int64_t cap_tris, int64_t* __restrict__ counts, int xcd_order, int64_t tile_lo,
    int64_t ntiles, const uint32_t* __restrict__ order) {
  extern __shared__ uint64_t s_dyn[];
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  const bool normals = (flags & DAX_HIP_MC_NORMALS) != 0 && out.normals != nullptr;
  const uint64_t total_v = ctrl->total_v, total_t = ctrl->total_t;
  const uint64_t n_points = dedup ? total_v : 3ull * total_t;
  if (blockIdx.x == 0 && threadIdx.x == 0 && counts) {
    counts[0] = (int64_t)total_t;
    counts[1] = (int64_t)n_points;
  }
  // too many ids for dax::Id or too little room: the bases only (the adjoint
  // and the per-cell outputs read them), no output
  const bool emit = total_v <= 0x7fffffffull && 3ull * total_t <= 0x7fffffffull &&
                    (int64_t)total_t <= cap_tris && (int64_t)n_points <= cap_points;
  // the scan's dispatch order when it was built for this tile size (whole
  // volume launches only), else index order
  uint32_t tag = 0u, ord = 0u;
  if (order != nullptr) {
    tag = order[0];
    ord = order[1 + blockIdx.x];
  }
  const bool ordered = tag == (uint32_t)TW && (int64_t)(ord & ~kOrderHasOutput) < ntiles;
  const int64_t tile =
      ordered ? (int64_t)(ord & ~kOrderHasOutput)
              : tile_lo + (xcd_order ? xcd_tile(blockIdx.x, ntiles) : (int64_t)blockIdx.x);
  generate_tile<TW, Src, true>(g, field, bits + g.word_off, nullptr, nullptr, tile * TW, out,
                               dedup, normals, 0, 0, s_dyn, cs, emit,
                               ordered && (ord & kOrderHasOutput) != 0u);
}

// Slice classify: a wave per kPlaneRowsPerWave grid rows.  Along a row the
// plane value pn0 * ((o0 + s0 * x) - po0) + t1 + t2 (t1, t2: the row's y
// and z terms)
// is a monotonic function of x: int -> float conversion, multiplication by
// a constant, addition of a constant and their round-to-nearest are all
// monotonic.  So "value > iso" changes at most once along the row, and the
// row's bits are a prefix or a suffix.  A 64-way search (each lane
// evaluates one candidate x with exactly the operations FieldPlane uses,
// a ballot finds the first change) locates the change in ceil(log64 nx) + 1
// rounds; the lanes then write the row's words.  The same launch clears the
// count/scan state.
constexpr int kPlaneRowsPerWave = 8;

// One grid row of the slice classify (all lanes of a wave).
__device__ __forceinline__ void classify_plane_row(const VolGeom& g, const FieldPlane& pl,
                                                   uint64_t* __restrict__ bits, uint32_t row,
                                                   int lane) {
  const uint32_t k = fast_div(row, g.div_ny), j = row - k * (uint32_t)g.ny;
  const float t1 = pl.pn[1] * ((pl.o[1] + pl.s[1] * (float)(int)j) - pl.po[1]);
  const float t2 = pl.pn[2] * ((pl.o[2] + pl.s[2] * (float)(int)k) - pl.po[2]);
  auto inside = [&](int x) -> bool {
    const float t0 = pl.pn[0] * ((pl.o[0] + pl.s[0] * (float)x) - pl.po[0]);
    return (t0 + t1) + t2 > g.iso;
  };
  const bool p0 = inside(0);
  // first x with inside(x) != p0, nx if none: invariant lo differs not, hi
  // differs (nx stands for "differs"), the change lies in (lo, hi]
  int lo = 0, hi = g.nx;
  while (hi - lo > 1) {
    const int step = (hi - lo + 63) / 64;
    const int x = lo + (lane + 1) * step;
    const uint64_t d = __ballot(x >= hi || inside(x) != p0);
    const int f = __builtin_ctzll(d);  // lane 63 reaches hi, so d != 0
    const int nlo = lo + f * step;
    hi = min(nlo + step, hi);
    lo = nlo;
  }
  const int change = hi;
  uint64_t* out = bits + (int64_t)row * g.wpr;
  for (int q = lane; q < g.wpr; q += 64) {
    const uint64_t below = low_mask(change - q * 64);  // points x < change
    out[q] = (p0 ? below : ~below) & low_mask(g.nx - q * 64);
  }
}

__global__ __launch_bounds__(256) void k_classify_plane(VolGeom g, FieldPlane pl,
                                                        uint64_t* __restrict__ bits,
                                                        unsigned long long* __restrict__ zero_buf,
                                                        int64_t zero_words) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid < zero_words) zero_buf[gtid] = 0ull;
  const int lane = __lane_id();
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  const uint32_t rows = (uint32_t)g.ny * (uint32_t)g.nz;
  for (uint32_t row = wave * kPlaneRowsPerWave;
       row < min(rows, (wave + 1) * kPlaneRowsPerWave); ++row)
    classify_plane_row(g, pl, bits, row, lane);
}

// ---------------------------------------------------------------------------
// Tuning: per grid shape, the classify variant, scan sub-tiles and generate
// tile words.  Defaults are used until dax_hip_mc_autotune (or
// dax_hip_mc_set_tuning) records a choice for a shape.
// The defaults are the 512^3 workload's autotuned choice (classify falls
// back to one float per lane, 16 words per wave, on rows that are not
// 64-point aligned).
struct Tuning {
  int cls_unroll = 4;   // words (vec 1) or 256-point chunks (vec 4) per wave
  int cls_vec = 4;      // floats per lane per load
  int scan_sub = 8;     // count tiles (waves) per count block
  int gen_tile = 512;   // generate tile words
  int gen_xcd = 0;      // 1: each XCD takes a contiguous range of generate tiles
};

std::mutex g_tune_mu;
Tuning g_tune_default;
std::map<std::tuple<int64_t, int64_t, int64_t>, Tuning> g_tune_cache;

Tuning tuning_for(int64_t nx, int64_t ny, int64_t nz) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune_cache.find(std::make_tuple(nx, ny, nz));
  return it == g_tune_cache.end() ? g_tune_default : it->second;
}

bool valid_tuning(const Tuning& t) {
  const bool cls = (t.cls_vec == 1 && (t.cls_unroll == 8 || t.cls_unroll == 16 || t.cls_unroll == 32)) ||
                   (t.cls_vec == 4 && (t.cls_unroll == 2 || t.cls_unroll == 4 || t.cls_unroll == 8));
  const bool scan = t.scan_sub == 1 || t.scan_sub == 2 || t.scan_sub == 4 || t.scan_sub == 8;
  const bool gen = (t.gen_tile == 64 || t.gen_tile == 128 || t.gen_tile == 256 ||
                    t.gen_tile == 512 || t.gen_tile == 1024) &&
                   (t.gen_xcd == 0 || t.gen_xcd == 1);
  return cls && scan && gen;
}

// ---------------------------------------------------------------------------
// Per-cell outputs of the count step: case byte, NumFaces[case] and the
// exclusive triangle offset.  A wavefront takes kCellWords (8) consecutive
// row words: one load instruction brings all their neighbour words (lane
// 8u+v holds neighbour v of word u), each word then reads its eight from
// the lanes and writes its 64 cells (lane = cell).
constexpr int kCellWords = 8;
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

__global__ __launch_bounds__(256) void k_cell_outputs(VolGeom g, const uint64_t* __restrict__ bits,
                                                      const uint32_t* __restrict__ base_t,
                                                      uint8_t* __restrict__ cases,
                                                      int32_t* __restrict__ count,
                                                      int32_t* __restrict__ offsets) {
  __shared__ uint8_t s_nf[256];  // NumFaces per case
  s_nf[threadIdx.x] = (uint8_t)(c_case_table[threadIdx.x] >> 60);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t w0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * kCellWords;
  if (w0 >= g.nwords) return;  // wave-uniform
  const uint64_t* b = bits + g.word_off;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  uint64_t mine = 0ull;  // neighbour (lane & 7) of word w0 + (lane >> 3)
  {
    const int64_t w = w0 + (lane >> 3);
    const int v = lane & 7;
    if (w < g.nwords) {
      const WordPos p = word_pos(g, w);
      if (p.has_y && p.has_z && (!(v & 1) || p.q + 1 < g.wpr))
        mine = b[w + (v & 1) + ((v >> 1) & 1) * g.wpr + (v >> 2) * plane];
    }
  }
  const int64_t cx = g.nx - 1, cy = g.ny - 1;
#pragma unroll
  for (int u = 0; u < kCellWords; ++u) {
    const int64_t w = w0 + u;
    const WordPos p = word_pos(g, w < g.nwords ? w : w0);
    if (w >= g.nwords || !(p.has_y && p.has_z)) continue;  // wave-uniform: no cells
    uint64_t nb[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) nb[v] = readlane64(mine, 8 * u + v);
    CellPlanes cp;
    cell_planes(p, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], nb[6], nb[7], cp);
    const bool valid = ((p.xmask >> lane) & 1ull) != 0;
    const int cs = cell_case(cp, lane);
    const uint32_t nf = valid ? (uint32_t)s_nf[cs] : 0u;
    const uint32_t incl = offsets ? wave_incl_scan(nf) : 0u;
    if (valid) {
      const int64_t cell = (int64_t)p.q * 64 + lane + cx * (p.j + cy * (int64_t)p.k);
      if (cases) cases[cell] = (uint8_t)cs;
      if (count) count[cell] = (int32_t)nf;
      if (offsets) offsets[cell] = (int32_t)(base_t[g.word_off + w] + incl - nf);
    }
  }
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

__global__ void k_interpolate_vec3(const float* __restrict__ in, const int32_t* __restrict__ first,
                                   const int32_t* __restrict__ second,
                                   const float* __restrict__ weight, int64_t n,
                                   float* __restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int64_t a = 3 * (int64_t)first[u], b = 3 * (int64_t)second[u];
  const float w = weight[u];
#pragma unroll
  for (int d = 0; d < 3; ++d) out[3 * u + d] = in[a + d] + w * (in[b + d] - in[a + d]);
}

__global__ void k_point_gradient(VolGeom g, const float* __restrict__ f, float* __restrict__ grad) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = (int64_t)g.nx * g.ny * g.nz;
  if (p >= P) return;
  const uint32_t row = fast_div((uint32_t)p, g.div_nx), zz = fast_div(row, g.div_ny);
  const int x = (int)((uint32_t)p - row * (uint32_t)g.nx), y = (int)(row - zz * (uint32_t)g.ny),
            z = (int)zz;
  float gr[3];
  point_gradient(g, FieldArray{f}, x, y, z, gr);
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
  set_geom_divs(g);
  return DAX_HIP_OK;
}

// Geometry and plane of a slice; the isovalue is dot(normal, origin) summed
// from component 0 (dax/worklet/Slice.h:44,104; dax/Types.h:611-621).
int make_slice(int64_t nx, int64_t ny, int64_t nz, const float* origin, const float* spacing,
               const float* plane_origin, const float* plane_normal, VolGeom& g,
               FieldPlane& pl) {
  if (!plane_origin || !plane_normal) {
    set_error("dax_hip_mc_slice: plane origin and normal are required");
    return DAX_HIP_BAD_VALUE;
  }
  float iso = plane_normal[0] * plane_origin[0];
  iso += plane_normal[1] * plane_origin[1];
  iso += plane_normal[2] * plane_origin[2];
  const int rc = make_geom(nx, ny, nz, origin, spacing, iso, g);
  if (rc) return rc;
  for (int d = 0; d < 3; ++d) {
    pl.o[d] = g.origin[d];
    pl.s[d] = g.spacing[d];
    pl.po[d] = plane_origin[d];
    pl.pn[d] = plane_normal[d];
  }
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

// Classify row words [w_lo, w_hi) (all when w_hi < 0); `zero` also clears
// the count/scan state.
int launch_bits(const VolGeom& g, const float* field, char* ws, const Workspace& L,
                const Tuning& tu, hipStream_t s, int64_t w_lo = 0, int64_t w_hi = -1,
                bool zero = true) {
  if (w_hi < 0) w_hi = g.nwords;
  uint64_t* bits = (uint64_t*)(ws + L.off_bits);
  unsigned long long* status = (unsigned long long*)(ws + L.off_status);
  const int64_t zero_words = zero ? L.ntiles + (int64_t)(sizeof(Ctrl) / 8) : 0;
  const bool aligned = (g.nx % 64) == 0;
  const bool aligned16 = aligned && (g.field_off % 4) == 0 && ((uintptr_t)field % 16) == 0 &&
                         (w_lo % 4) == 0;
  const int vec = aligned16 ? tu.cls_vec : 1;
  const int unroll = (vec == 1 && tu.cls_vec != 1) ? 16 : tu.cls_unroll;
  const int64_t units = vec == 4 ? ceil_div(w_hi - w_lo, 4) : w_hi - w_lo;  // chunks or words
  const int64_t waves = ceil_div(units, unroll);
  int64_t blocks = std::max<int64_t>(ceil_div(waves * 64, 256), ceil_div(zero_words, 256));
  blocks = std::max<int64_t>(blocks, 1);
  const dim3 grid((unsigned)blocks), block(256);
#define DAXHIP_CLS(U, A) \
  k_classify<U, A><<<grid, block, 0, s>>>(g, field, bits, status, zero_words, w_lo, w_hi)
#define DAXHIP_CLS4(U) \
  k_classify_v4<U><<<grid, block, 0, s>>>(g, field, bits, status, zero_words, w_lo, w_hi)
  if (vec == 4) {
    if (unroll == 2) DAXHIP_CLS4(2);
    else if (unroll == 4) DAXHIP_CLS4(4);
    else DAXHIP_CLS4(8);
  } else if (aligned) {
    if (unroll == 8) DAXHIP_CLS(8, true);
    else if (unroll == 32) DAXHIP_CLS(32, true);
    else DAXHIP_CLS(16, true);
  } else {
    if (unroll == 8) DAXHIP_CLS(8, false);
    else if (unroll == 32) DAXHIP_CLS(32, false);
    else DAXHIP_CLS(16, false);
  }
#undef DAXHIP_CLS
#undef DAXHIP_CLS4
  DAXHIP_LAUNCH_CHECK("k_classify");
  return DAX_HIP_OK;
}

// Generate tile words of a tuning (launch_generate_src's dispatch).
int gen_tile_words(const Tuning& tu) {
  return (tu.gen_tile == 64 || tu.gen_tile == 128 || tu.gen_tile == 256 || tu.gen_tile == 1024)
             ? tu.gen_tile
             : 512;
}

// Where the generate kernel reads the word bases; a count block holds two
// count tiles per wave, with the waves of launch_scan's dispatch (1, 2, 8,
// else 4).
CountSrc count_src(const char* ws, const Workspace& L, const Tuning& tu) {
  char* wsm = const_cast<char*>(ws);
  return CountSrc{(const uint32_t*)(ws + L.off_counts),
                  (const unsigned long long*)(ws + L.off_tsum),
                  (const unsigned long long*)(ws + L.off_tpre), L.n_ctiles,
                  tu.scan_sub == 1 ? 1 : (tu.scan_sub == 2 ? 2 : (tu.scan_sub == 8 ? 4 : 3)),
                  (uint32_t*)(wsm + L.off_base_v), (uint32_t*)(wsm + L.off_base_t)};
}

// Count + block-sum scan over the whole volume (scan_sub waves, i.e.
// 2 * scan_sub count tiles, per count block); `bases` also stores every
// word's bases (calls that stop before the generate kernel, which otherwise
// stores them).
int launch_scan(const VolGeom& g, char* ws, const Workspace& L, const Tuning& tu, hipStream_t s,
                bool bases = false) {
  const uint64_t* bits = (const uint64_t*)(ws + L.off_bits);
  uint32_t* lpre = (uint32_t*)(ws + L.off_counts);
  unsigned long long* tloc = (unsigned long long*)(ws + L.off_tsum);
  unsigned long long* bsum = (unsigned long long*)(ws + L.off_bsum);
  unsigned long long* bpre = (unsigned long long*)(ws + L.off_tpre);
  const int64_t n = L.n_ctiles;
  const int waves = tu.scan_sub;
  // generate classes when a count block covers whole generate tiles
  const int tw = gen_tile_words(tu);
  const int block_words = 2 * (waves == 1 || waves == 2 || waves == 8 ? waves : 4) * kCountTile;
  const TileClasses tc{(uint8_t*)(ws + L.off_gcls), (uint8_t*)(ws + L.off_tcls),
                       tw >= kCountTile && tw <= block_words ? tw : 0};
  const int64_t nblocks = ceil_div(n, 2 * waves);
  if (nblocks > 0) {
    const dim3 grid((unsigned)nblocks), block(64 * waves);
    switch (waves) {
      case 1: k_count<1><<<grid, block, 0, s>>>(g, bits, lpre, tloc, bsum, tc, n); break;
      case 2: k_count<2><<<grid, block, 0, s>>>(g, bits, lpre, tloc, bsum, tc, n); break;
      case 8: k_count<8><<<grid, block, 0, s>>>(g, bits, lpre, tloc, bsum, tc, n); break;
      default: k_count<4><<<grid, block, 0, s>>>(g, bits, lpre, tloc, bsum, tc, n); break;
    }
    DAXHIP_LAUNCH_CHECK("k_count");
  }
  // with the dispatch order of the generate tiles, in units of two tiles a
  // plane apart when tiles split planes evenly and the units fill groups of 8
  const int64_t ng = ceil_div(g.nwords, (int64_t)tw);
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int64_t sp = plane % tw == 0 ? plane / tw : 0;
  constexpr int64_t kUnitG = kOrderUnitTiles;
  const bool units = tc.tw > 0 && sp > 0 && ng % (kUnitG * sp) == 0 && (ng / kUnitG) % 8 == 0 &&
                     ceil_div(ng / kUnitG, (int64_t)kTsumThreads) * kUnitG <= kOrderRegTiles;
  const GenOrder go{tc.gcls, g.nwords, tw, tc.tw > 0, units ? (uint32_t)sp : 0u,
                    (uint32_t*)(ws + L.off_order), (uint32_t*)(ws + L.off_gorder)};
  if (ceil_div(ceil_div(g.nwords, (int64_t)tw), (int64_t)kTsumThreads) > kOrderRegTiles)
    k_scan_tsums<true><<<1, dim3(kTsumThreads), 0, s>>>(bsum, bpre, nblocks, (Ctrl*)(ws + L.off_ctrl), go);
  else
    k_scan_tsums<false><<<1, dim3(kTsumThreads), 0, s>>>(bsum, bpre, nblocks, (Ctrl*)(ws + L.off_ctrl), go);
  DAXHIP_LAUNCH_CHECK("k_scan_tsums");
  if (bases) {
    k_bases<<<dim3((unsigned)ceil_div(g.nwords + 1, 256)), dim3(256), 0, s>>>(
        g, count_src(ws, L, tu));
    DAXHIP_LAUNCH_CHECK("k_bases");
  }
  return DAX_HIP_OK;
}

// Generate tiles of TW words [t_lo, t_hi) (all when t_hi < 0).
template <int TW, class Src>
int launch_generate_tw(const VolGeom& g, const Src& field, const char* ws, const Workspace& L,
                       unsigned flags, const OutPtrs& o, int64_t cap_points, int64_t cap_tris,
                       int64_t* counts, const Tuning& tu, hipStream_t s, int64_t t_lo,
                       int64_t t_hi) {
  const int xcd = tu.gen_xcd;
  const size_t lds = GenLayout<TW>(g.wpr).bytes();
  const int64_t all = ceil_div(g.nwords, (int64_t)TW);
  if (t_hi < 0) t_hi = all;
  const int64_t ntiles = std::max<int64_t>(t_hi - t_lo, 1);
  const CountSrc cs = count_src(ws, L, tu);
  const uint32_t* order =
      (t_lo == 0 && t_hi == all) ? (const uint32_t*)(ws + L.off_order) : nullptr;
  k_generate<TW, Src><<<dim3((unsigned)ntiles), dim3(kScanThreads), lds, s>>>(
      g, field, (const uint64_t*)(ws + L.off_bits), cs, (const Ctrl*)(ws + L.off_ctrl), o, flags,
      cap_points, cap_tris, counts, xcd, t_lo, ntiles, order);
  DAXHIP_LAUNCH_CHECK("k_generate");
  return DAX_HIP_OK;
}

template <class Src>
int launch_generate_src(const VolGeom& g, const Src& field, const char* ws, const Workspace& L,
                        unsigned flags, const OutPtrs& o, int64_t cap_points, int64_t cap_tris,
                        int64_t* counts, const Tuning& tu, hipStream_t s, int64_t t_lo = 0,
                        int64_t t_hi = -1) {
  switch (tu.gen_tile) {
    case 64:
      return launch_generate_tw<64>(g, field, ws, L, flags, o, cap_points, cap_tris, counts, tu,
                                    s, t_lo, t_hi);
    case 128:
      return launch_generate_tw<128>(g, field, ws, L, flags, o, cap_points, cap_tris, counts,
                                     tu, s, t_lo, t_hi);
    case 256:
      return launch_generate_tw<256>(g, field, ws, L, flags, o, cap_points, cap_tris, counts,
                                     tu, s, t_lo, t_hi);
    case 1024:
      return launch_generate_tw<1024>(g, field, ws, L, flags, o, cap_points, cap_tris, counts,
                                      tu, s, t_lo, t_hi);
    default:
      return launch_generate_tw<512>(g, field, ws, L, flags, o, cap_points, cap_tris, counts,
                                     tu, s, t_lo, t_hi);
  }
}

int launch_generate(const VolGeom& g, const float* field, const char* ws, const Workspace& L,
                    unsigned flags, const OutPtrs& o, int64_t cap_points, int64_t cap_tris,
                    int64_t* counts, const Tuning& tu, hipStream_t s, int64_t t_lo = 0,
                    int64_t t_hi = -1) {
  return launch_generate_src(g, FieldArray{field + g.field_off}, ws, L, flags, o, cap_points,
                             cap_tris, counts, tu, s, t_lo, t_hi);
}

// Slice classify over the whole volume (clears the count/scan state).
int launch_bits_plane(const VolGeom& g, const FieldPlane& pl, char* ws, const Workspace& L,
                      hipStream_t s) {
  const int64_t zero_words = L.ntiles + (int64_t)(sizeof(Ctrl) / 8);
  const int64_t waves = ceil_div((int64_t)g.ny * g.nz, kPlaneRowsPerWave);
  const int64_t blocks =
      std::max<int64_t>(std::max(ceil_div(waves * 64, 256), ceil_div(zero_words, 256)), 1);
  k_classify_plane<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
      g, pl, (uint64_t*)(ws + L.off_bits), (unsigned long long*)(ws + L.off_status), zero_words);
  DAXHIP_LAUNCH_CHECK("k_classify_plane");
  return DAX_HIP_OK;
}

// The three kernels of a forward over the whole volume.
int launch_forward(const VolGeom& g, const float* field, char* ws, const Workspace& L,
                   unsigned flags, const OutPtrs& o, int64_t cap_points, int64_t cap_tris,
                   int64_t* counts, const Tuning& tu, hipStream_t s) {
  int rc;
  if ((rc = launch_bits(g, field, ws, L, tu, s))) return rc;
  if ((rc = launch_scan(g, ws, L, tu, s))) return rc;
  return launch_generate(g, field, ws, L, flags, o, cap_points, cap_tris, counts, tu, s);
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
  return launch_bits(g, d_field, (char*)d_workspace, L, tuning_for(nx, ny, nz),
                     (hipStream_t)stream);
}

int dax_hip_mc_scan(int64_t nx, int64_t ny, int64_t nz, void* d_workspace, size_t workspace_bytes,
                    void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, nullptr, 0.0f, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  return launch_scan(g, (char*)d_workspace, L, tuning_for(nx, ny, nz), (hipStream_t)stream);
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
                         d_counts, tuning_for(nx, ny, nz), (hipStream_t)stream);
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
  const Tuning tu = tuning_for(nx, ny, nz);
  if ((rc = launch_bits(g, d_field, ws, L, tu, s))) return rc;
  if ((rc = launch_scan(g, ws, L, tu, s, true))) return rc;
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
  const Tuning tu = tuning_for(nx, ny, nz);
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, d_interp_first, d_interp_second,
            d_interp_weight};
  return launch_forward(g, d_field, ws, L, flags, o, cap_points, cap_tris, d_counts, tu, s);
}

// ---- forward without allocation -------------------------------------------
// dax_hip_mc_forward with its workspace in a file-scope device array, for
// callers that cannot allocate (code under graph capture, serving loops):
// grids whose workspace fits kNoAllocScratchBytes (48 MiB: up to 512^3
// points).  Every GPU the library runs on holds the array until the process
// ends.  The forward clears what it needs cleared (its classify zeroes the
// scan's status and control words); calls that overlap on one GPU share the
// array and race.
constexpr size_t kNoAllocScratchBytes = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocDevices] = {};  // device address of `scratch`, per device

// The current device's copy of `scratch`, looked up once per device (with
// that device current); host code uses this address, never the array's name.
int scratch_on_device(void** ptr) {
  int dev = 0;
  DAXHIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error("dax_hip_mc_forward_noalloc: device %d beyond the scratch table", dev);
    return DAX_HIP_BAD_VALUE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return check_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return DAX_HIP_OK;
}
}  // namespace

int dax_hip_mc_forward_noalloc(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                               const float* h_origin3, const float* h_spacing3, float iso,
                               unsigned flags, float* d_vertices, float* d_normals,
                               int32_t* d_conn, uint32_t* d_edges, int32_t* d_interp_first,
                               int32_t* d_interp_second, float* d_interp_weight,
                               int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                               void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, h_origin3, h_spacing3, iso, g);
  if (rc) return rc;
  const Workspace L = workspace_layout(g.nwords);
  if (L.bytes > kNoAllocScratchBytes) {
    set_error("dax_hip_mc_forward_noalloc: a %lld x %lld x %lld grid needs %zu bytes of scratch, "
              "%zu are available",
              (long long)nx, (long long)ny, (long long)nz, L.bytes, kNoAllocScratchBytes);
    return DAX_HIP_BAD_VALUE;
  }
  if ((rc = check_outputs(g, d_field, flags, d_vertices, d_conn, d_edges, d_counts))) return rc;
  void* ws = nullptr;
  if ((rc = scratch_on_device(&ws))) return rc;
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, d_interp_first, d_interp_second,
            d_interp_weight};
  return launch_forward(g, d_field, (char*)ws, L, flags, o, cap_points, cap_tris, d_counts,
                        tuning_for(nx, ny, nz), (hipStream_t)stream);
}

// ---- slice: marching cubes of the plane distance, computed in the kernels
// (SliceCount / SliceGenerate, dax/worklet/Slice.h:28-173, on a uniform
// grid; no field in memory).
int dax_hip_mc_slice_classify(int64_t nx, int64_t ny, int64_t nz, const float* h_origin3,
                              const float* h_spacing3, const float* h_plane_origin3,
                              const float* h_plane_normal3, void* d_workspace,
                              size_t workspace_bytes, int64_t* d_counts, void* stream) {
  VolGeom g;
  FieldPlane pl;
  int rc = make_slice(nx, ny, nz, h_origin3, h_spacing3, h_plane_origin3, h_plane_normal3, g, pl);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  if ((rc = launch_bits_plane(g, pl, ws, L, s))) return rc;
  if ((rc = launch_scan(g, ws, L, tuning_for(nx, ny, nz), s, true))) return rc;
  if (d_counts) {
    k_write_counts<<<1, 1, 0, s>>>((const Ctrl*)(ws + L.off_ctrl), d_counts, DAX_HIP_MC_DEDUP);
    DAXHIP_LAUNCH_CHECK("k_write_counts");
  }
  return DAX_HIP_OK;
}

int dax_hip_mc_slice(int64_t nx, int64_t ny, int64_t nz, const float* h_origin3,
                     const float* h_spacing3, const float* h_plane_origin3,
                     const float* h_plane_normal3, unsigned flags, void* d_workspace,
                     size_t workspace_bytes, float* d_vertices, float* d_normals, int32_t* d_conn,
                     uint32_t* d_edges, int32_t* d_interp_first, int32_t* d_interp_second,
                     float* d_interp_weight, int64_t cap_points, int64_t cap_tris,
                     int64_t* d_counts, void* stream) {
  VolGeom g;
  FieldPlane pl;
  int rc = make_slice(nx, ny, nz, h_origin3, h_spacing3, h_plane_origin3, h_plane_normal3, g, pl);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  static const float kNoField = 0.0f;  // the slice reads no field
  if ((rc = check_outputs(g, &kNoField, flags, d_vertices, d_conn, d_edges, d_counts))) return rc;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  const Tuning tu = tuning_for(nx, ny, nz);
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, d_interp_first, d_interp_second,
            d_interp_weight};
  if ((rc = launch_bits_plane(g, pl, ws, L, s))) return rc;
  if ((rc = launch_scan(g, ws, L, tu, s))) return rc;
  return launch_generate_src(g, pl, ws, L, flags, o, cap_points, cap_tris, d_counts, tu, s);
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
  const int64_t blocks = ceil_div(ceil_div(g.nwords, (int64_t)kCellWords) * 64, 256);
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

int dax_hip_mc_interpolate_vec3(const float* d_in, const int32_t* d_first, const int32_t* d_second,
                                const float* d_weight, int64_t n, float* d_out, void* stream) {
  if (n <= 0) return DAX_HIP_OK;
  k_interpolate_vec3<<<dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream>>>(
      d_in, d_first, d_second, d_weight, n, d_out);
  DAXHIP_LAUNCH_CHECK("k_interpolate_vec3");
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

// ---------------------------------------------------------------------------
// Tuning and autotuning
namespace {

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  float ms() {
    float m = 0.0f;
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&m, a, b);
    return m;
  }
};

}  // namespace

extern "C" {

int dax_hip_mc_set_tuning(int64_t nx, int64_t ny, int64_t nz, const int* h_params5) {
  if (!h_params5) {
    set_error("dax_hip_mc_set_tuning: NULL parameters");
    return DAX_HIP_BAD_VALUE;
  }
  Tuning t;
  t.cls_unroll = h_params5[0];
  t.cls_vec = h_params5[1];
  t.scan_sub = h_params5[2];
  t.gen_tile = h_params5[3];
  t.gen_xcd = h_params5[4];
  if (!valid_tuning(t)) {
    set_error("dax_hip_mc_set_tuning: invalid parameters {%d,%d,%d,%d,%d}", t.cls_unroll,
              t.cls_vec, t.scan_sub, t.gen_tile, t.gen_xcd);
    return DAX_HIP_BAD_VALUE;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (nx == 0 && ny == 0 && nz == 0)
    g_tune_default = t;
  else
    g_tune_cache[std::make_tuple(nx, ny, nz)] = t;
  return DAX_HIP_OK;
}

int dax_hip_mc_get_tuning(int64_t nx, int64_t ny, int64_t nz, int* h_params5) {
  const Tuning t = (nx == 0 && ny == 0 && nz == 0) ? g_tune_default : tuning_for(nx, ny, nz);
  h_params5[0] = t.cls_unroll;
  h_params5[1] = t.cls_vec;
  h_params5[2] = t.scan_sub;
  h_params5[3] = t.gen_tile;
  h_params5[4] = t.gen_xcd;
  return DAX_HIP_OK;
}

int dax_hip_mc_tuning_save(const char* path) {
  FILE* fp = fopen(path, "w");
  if (!fp) {
    set_error("dax_hip_mc_tuning_save: cannot open %s", path);
    return DAX_HIP_BAD_VALUE;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  fprintf(fp, "# nx ny nz classify_unroll classify_vec scan_sub gen_tile gen_xcd\n");
  for (const auto& kv : g_tune_cache)
    fprintf(fp, "%lld %lld %lld %d %d %d %d %d\n", (long long)std::get<0>(kv.first),
            (long long)std::get<1>(kv.first), (long long)std::get<2>(kv.first), kv.second.cls_unroll,
            kv.second.cls_vec, kv.second.scan_sub, kv.second.gen_tile, kv.second.gen_xcd);
  fclose(fp);
  return DAX_HIP_OK;
}

int dax_hip_mc_tuning_load(const char* path) {
  FILE* fp = fopen(path, "r");
  if (!fp) {
    set_error("dax_hip_mc_tuning_load: cannot open %s", path);
    return DAX_HIP_BAD_VALUE;
  }
  char line[256];
  while (fgets(line, sizeof(line), fp)) {
    if (line[0] == '#') continue;
    long long x, y, z;
    Tuning t;
    if (sscanf(line, "%lld %lld %lld %d %d %d %d %d", &x, &y, &z, &t.cls_unroll, &t.cls_vec,
               &t.scan_sub, &t.gen_tile, &t.gen_xcd) == 8 &&
        valid_tuning(t)) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune_cache[std::make_tuple((int64_t)x, (int64_t)y, (int64_t)z)] = t;
    }
  }
  fclose(fp);
  return DAX_HIP_OK;
}

int dax_hip_mc_autotune(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                        unsigned flags, void* d_workspace, size_t workspace_bytes,
                        float* d_vertices, float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                        int64_t cap_points, int64_t cap_tris, int64_t* d_counts, int reps,
                        int* h_best5, float* h_best_ms3, void* stream) {
  VolGeom g;
  int rc = make_geom(nx, ny, nz, nullptr, nullptr, iso, g);
  if (rc) return rc;
  Workspace L;
  if ((rc = check_ws(g, workspace_bytes, L, d_workspace))) return rc;
  if ((rc = check_outputs(g, d_field, flags, d_vertices, d_conn, d_edges, d_counts))) return rc;
  if (reps < 1) reps = 1;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, nullptr, nullptr, nullptr};
  Tuning best = tuning_for(nx, ny, nz);
  EventPair ev;

  // classify variants
  std::vector<std::pair<int, int>> cls = {{8, 1}, {16, 1}, {32, 1}};
  if (nx % 64 == 0) cls.insert(cls.end(), {{2, 4}, {4, 4}, {8, 4}});
  float best_ms[3] = {1e30f, 1e30f, 1e30f};
  for (auto uv : cls) {
    Tuning t = best;
    t.cls_unroll = uv.first;
    t.cls_vec = uv.second;
    if ((rc = launch_bits(g, d_field, ws, L, t, s))) return rc;  // warm
    DAXHIP_CHECK(hipEventRecord(ev.a, s));
    for (int r = 0; r < reps; ++r)
      if ((rc = launch_bits(g, d_field, ws, L, t, s))) return rc;
    DAXHIP_CHECK(hipEventRecord(ev.b, s));
    const float ms = ev.ms() / reps;
    if (ms < best_ms[0]) {
      best_ms[0] = ms;
      best.cls_unroll = t.cls_unroll;
      best.cls_vec = t.cls_vec;
    }
  }
  // count/scan and generate variants together: the count block size decides
  // whether the count kernel classifies the generate tiles, and the tile
  // size the dispatch order the scan block builds (and its cost).  Each
  // repetition follows a classify, as in the forward; events split the
  // scan's time from generate's.
  EventPair ev2;
  const Tuning base = best;
  for (int sub : {1, 2, 4, 8}) {
    for (int tv = 0; tv < 10; ++tv) {
      Tuning t = base;
      t.scan_sub = sub;
      t.gen_tile = 64 << (tv % 5);
      t.gen_xcd = tv / 5;
      float ts = 0.0f, tg = 0.0f;
      for (int r = 0; r < reps; ++r) {
        if ((rc = launch_bits(g, d_field, ws, L, base, s))) return rc;
        DAXHIP_CHECK(hipEventRecord(ev.a, s));
        if ((rc = launch_scan(g, ws, L, t, s))) return rc;
        DAXHIP_CHECK(hipEventRecord(ev2.a, s));
        if ((rc = launch_generate(g, d_field, ws, L, flags, o, cap_points, cap_tris, d_counts, t,
                                  s)))
          return rc;
        DAXHIP_CHECK(hipEventRecord(ev.b, s));
        DAXHIP_CHECK(hipEventSynchronize(ev.b));
        float m1 = 0.0f, m2 = 0.0f;
        DAXHIP_CHECK(hipEventElapsedTime(&m1, ev.a, ev2.a));
        DAXHIP_CHECK(hipEventElapsedTime(&m2, ev2.a, ev.b));
        ts += m1;
        tg += m2;
      }
      ts /= reps;
      tg /= reps;
      if (ts + tg < best_ms[1] + best_ms[2]) {
        best_ms[1] = ts;
        best_ms[2] = tg;
        best.scan_sub = sub;
        best.gen_tile = t.gen_tile;
        best.gen_xcd = t.gen_xcd;
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[std::make_tuple(nx, ny, nz)] = best;
  }
  if (h_best5) {
    h_best5[0] = best.cls_unroll;
    h_best5[1] = best.cls_vec;
    h_best5[2] = best.scan_sub;
    h_best5[3] = best.gen_tile;
    h_best5[4] = best.gen_xcd;
  }
  if (h_best_ms3)
    for (int i = 0; i < 3; ++i) h_best_ms3[i] = best_ms[i];
  return DAX_HIP_OK;
}

}  // extern "C"

