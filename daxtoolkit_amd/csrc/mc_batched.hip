// Batched marching cubes: many independent volumes (uneven extents,
// per-volume isovalue, origin and spacing) in one word space.  No reference
// counterpart (Dax runs one grid per dispatcher call).
//
// Layout: the volumes' fields are concatenated (volume v at field_off[v]);
// their row words are concatenated in volume order (word_off[v]).  Scan and
// generate tiles never straddle a volume; tile tables map tile -> (volume,
// first local word).  One global exclusive scan over all words gives every
// word its vertex/triangle base; subtracting the base of the volume's first
// word makes the numbering per volume -- a segmented scan without segment
// flags.  Outputs are concatenated in volume order; connectivity is local to
// each volume; per-volume {first triangle, first point} offsets are written
// beside the counts.
//
// Each phase is one persistent launch (grid = a few blocks per CU, looping
// over tiles); the four launches of a forward are graph-capturable (no host
// read-back, scan state cleared by a kernel).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mc_adjoint.h"
#include "mc_kernels.h"

namespace daxhip {
namespace mc {

struct TileRef {
  int32_t vol;
  int32_t w0;  // first row word, local to the volume
};

// Tile tables are built at prepare: 1024-word scan tiles, 256-word
// generate (and classify) tiles.  The launch shapes over them are tuned.
constexpr int kBatchScanTile = 1024;
constexpr int kBatchGenTile = 256;

// Launch shapes of the batch forward, picked by dax_hip_mc_batch_autotune
// and cached per batch shape class: classify unroll (row words per wave
// step) and persistent blocks per CU, threads per scan tile (the tile's
// 1024 words over 256, 512 or 1024 threads), generate blocks per CU and
// its XCD-contiguous tile order.
struct BatchTuning {
  int32_t cls_unroll, cls_blocks, scan_threads, gen_blocks, gen_xcd, pad;
};
constexpr BatchTuning kBatchTuningDefault{16, 4, 256, 4, 1, 0};
bool valid_batch_tuning(const BatchTuning& t) {
  return (t.cls_unroll == 8 || t.cls_unroll == 16 || t.cls_unroll == 32) &&
         (t.cls_blocks == 2 || t.cls_blocks == 4 || t.cls_blocks == 8) &&
         (t.scan_threads == 256 || t.scan_threads == 512 || t.scan_threads == 1024) &&
         (t.gen_blocks == 2 || t.gen_blocks == 4 || t.gen_blocks == 8) &&
         (t.gen_xcd == 0 || t.gen_xcd == 1);
}

struct BatchLayout {
  size_t off_vols, off_scan_tiles, off_gen_tiles, off_bits, off_base_v, off_base_t, off_status,
      off_ctrl, bytes;
  int64_t nvols, nwords, nscan, ngen;
};

BatchLayout batch_layout(int64_t nvols, const int64_t* dims) {
  BatchLayout L;
  L.nvols = nvols;
  L.nwords = 0;
  L.nscan = 0;
  L.ngen = 0;
  for (int64_t v = 0; v < nvols; ++v) {
    const int64_t w = ceil_div(dims[3 * v], 64) * dims[3 * v + 1] * dims[3 * v + 2];
    L.nwords += w;
    L.nscan += ceil_div(w, kBatchScanTile);
    L.ngen += ceil_div(w, kBatchGenTile);
  }
  size_t off = 0;
  L.off_vols = off;
  off = align_up(off + (size_t)nvols * sizeof(VolGeom), 256);
  L.off_scan_tiles = off;
  off = align_up(off + (size_t)L.nscan * sizeof(TileRef), 256);
  L.off_gen_tiles = off;
  off = align_up(off + (size_t)L.ngen * sizeof(TileRef), 256);
  L.off_bits = off;
  off = align_up(off + (size_t)(L.nwords + 1) * 8, 256);
  L.off_base_v = off;
  off = align_up(off + (size_t)(L.nwords + 1) * 4, 256);
  L.off_base_t = off;
  off = align_up(off + (size_t)(L.nwords + 1) * 4, 256);
  L.off_status = off;
  off += (size_t)L.nscan * 8;
  L.off_ctrl = off;
  off = align_up(off + sizeof(Ctrl), 256);
  L.bytes = off;
  return L;
}

// Header at the start of the workspace: the layout, so launches need only
// the workspace pointer after prepare.
struct BatchHeader {
  int64_t magic, nvols, nwords, nscan, ngen, max_wpr;
  size_t off_vols, off_scan_tiles, off_gen_tiles, off_bits, off_base_v, off_base_t, off_status,
      off_ctrl, bytes;
  BatchTuning tune;
};
constexpr int64_t kBatchMagic = 0x4441584842415443LL;  // "DAXHBATC"
constexpr size_t kBatchHeaderBytes = 256;
static_assert(sizeof(BatchHeader) <= kBatchHeaderBytes, "header fits its reserved bytes");

template <int UNROLL>
__global__ __launch_bounds__(256) void k_batch_classify(const VolGeom* __restrict__ vols,
                                                        const TileRef* __restrict__ tiles,
                                                        int64_t ntiles,
                                                        const float* __restrict__ fields,
                                                        uint64_t* __restrict__ bits,
                                                        unsigned long long* __restrict__ zero_buf,
                                                        int64_t zero_words) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gtid; i < zero_words; i += nthreads) zero_buf[i] = 0ull;
  const int wv = threadIdx.x >> 6;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const VolGeom g = vols[tr.vol];
    const int64_t w_end = min((int64_t)tr.w0 + kBatchGenTile, g.nwords);
    for (int64_t w0 = tr.w0 + (int64_t)wv * UNROLL; w0 < w_end;
         w0 += (int64_t)(blockDim.x >> 6) * UNROLL)
      classify_words<UNROLL, false>(g, fields + g.field_off, bits + g.word_off, w0, w_end);
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_batch_count_scan(
    const VolGeom* __restrict__ vols, const TileRef* __restrict__ tiles, int64_t ntiles,
    const uint64_t* __restrict__ bits, uint32_t* __restrict__ base_v,
    uint32_t* __restrict__ base_t, unsigned long long* __restrict__ status,
    Ctrl* __restrict__ ctrl) {
  __shared__ ScanShared sh;
  constexpr int WPT = kBatchScanTile / NT;
  __shared__ uint32_t s_pk[ScanTileLds<NT, WPT>::kSize];
  __shared__ int64_t s_tile;
  while (true) {
    __syncthreads();  // s_tile and sh reused across iterations
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(&ctrl->tile_counter, 1u);
    __syncthreads();
    const int64_t t = s_tile;
    if (t >= ntiles) break;
    const TileRef tr = tiles[t];
    const VolGeom g = vols[tr.vol];
    count_scan_tile<NT, WPT>(g, bits + g.word_off, base_v, base_t, status, ctrl, t, tr.w0, sh,
                             s_pk);
  }
}

__global__ __launch_bounds__(kScanThreads) void k_batch_generate(
    const VolGeom* __restrict__ vols, const TileRef* __restrict__ tiles, int64_t ntiles,
    const float* __restrict__ fields, const uint64_t* __restrict__ bits,
    const uint32_t* __restrict__ base_v, const uint32_t* __restrict__ base_t,
    const Ctrl* __restrict__ ctrl, OutPtrs out, unsigned flags, int64_t cap_points,
    int64_t cap_tris, int64_t* __restrict__ counts, int xcd) {
  extern __shared__ uint64_t s_dyn[];
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
  for (int64_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const TileRef tr = tiles[xcd ? xcd_tile(b, ntiles) : b];
    const VolGeom g = vols[tr.vol];
    const int64_t vbv = base_v[g.word_off], vbt = base_t[g.word_off];
    generate_tile<kBatchGenTile>(g, FieldArray{fields + g.field_off}, bits + g.word_off,
                                 base_v + g.word_off, base_t + g.word_off, tr.w0, out, dedup,
                                 normals, vbv, vbt, s_dyn);
    __syncthreads();  // LDS reused by the next tile
  }
}

// Per-volume {first triangle, first point} in the concatenated outputs.
__global__ void k_batch_offsets(const VolGeom* __restrict__ vols, int64_t nvols,
                                const uint32_t* __restrict__ base_v,
                                const uint32_t* __restrict__ base_t, unsigned flags,
                                int64_t* __restrict__ vol_offsets) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvols) return;
  const int64_t w = vols[v].word_off;
  const bool dedup = (flags & DAX_HIP_MC_DEDUP) != 0;
  vol_offsets[2 * v + 0] = base_t[w];
  vol_offsets[2 * v + 1] = dedup ? (int64_t)base_v[w] : 3 * (int64_t)base_t[w];
}

// ---- batched adjoint --------------------------------------------------------
// Same tile table as generate: a block takes a tile of row words, first the
// adjoint records of the tile's vertices (global ids base_v[w0] ..
// base_v[w_end]), then, in a second launch (a block per tile), dL/dfield of
// the tile's points; tiles whose endpoint records do not fit are listed and
// redone as spans by a third launch.
__global__ __launch_bounds__(256) void k_batch_vertex_adjoint(
    const VolGeom* __restrict__ vols, const TileRef* __restrict__ tiles, int64_t ntiles,
    const float* __restrict__ fields, const uint32_t* __restrict__ base_v,
    const uint32_t* __restrict__ edges, const float* __restrict__ dldx,
    const float* __restrict__ dldn, const int64_t* __restrict__ counts, int64_t n_points,
    VertexAdj* __restrict__ out, uint32_t* __restrict__ list) {
  if (blockIdx.x == 0 && threadIdx.x == 0) list[0] = 0u;
  const int64_t U = counts ? min(counts[1], n_points) : n_points;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const VolGeom g = vols[tr.vol];
    const int64_t w_end = min((int64_t)tr.w0 + kBatchGenTile, g.nwords);
    const int64_t u0 = base_v[g.word_off + tr.w0];
    const int64_t u1 = min((int64_t)base_v[g.word_off + w_end], U);
    for (int64_t u = u0 + threadIdx.x; u < u1; u += blockDim.x)
      vertex_adjoint(g, fields, edges, dldx, dldn, u, out);
  }
}

__global__ __launch_bounds__(kGatherThreads) void k_batch_point_gather(
    const VolGeom* __restrict__ vols, const TileRef* __restrict__ tiles,
    const uint64_t* __restrict__ bits, const uint32_t* __restrict__ base_v,
    const VertexAdj* __restrict__ adj, const int64_t* __restrict__ counts, int64_t n_points,
    float* __restrict__ dldf, uint32_t* __restrict__ list) {
  static_assert(kGatherTile == kBatchGenTile, "gather tiles are the generate tiles");
  __shared__ GatherShared sh;
  extern __shared__ __align__(16) uint64_t s_dyn[];
  const int64_t n_adj = counts ? min(counts[1], n_points) : n_points;
  const TileRef tr = tiles[blockIdx.x];
  const VolGeom g = vols[tr.vol];
  const int64_t w_end = min((int64_t)tr.w0 + kBatchGenTile, g.nwords);
  gather_tile(g, bits + g.word_off, base_v + g.word_off, adj, n_adj, dldf + g.field_off, tr.w0,
              w_end, sh, s_dyn, list, gather_list_capacity(n_points), blockIdx.x, 0);
}

__global__ __launch_bounds__(kGatherThreads) void k_batch_point_gather_spans(
    const VolGeom* __restrict__ vols, const TileRef* __restrict__ tiles,
    const uint64_t* __restrict__ bits, const uint32_t* __restrict__ base_v,
    const VertexAdj* __restrict__ adj, const int64_t* __restrict__ counts, int64_t n_points,
    float* __restrict__ dldf, const uint32_t* __restrict__ list) {
  __shared__ GatherShared sh;
  extern __shared__ __align__(16) uint64_t s_dyn[];
  const int64_t n_adj = counts ? min(counts[1], n_points) : n_points;
  const int64_t n_items = min((int64_t)list[0], gather_list_capacity(n_points));
  const int parts = gather_item_parts(n_items, gridDim.x / 2);
  for (int64_t it = blockIdx.x; it < n_items * parts; it += gridDim.x) {
    const uint32_t item = list[1 + it / parts];
    const TileRef tr = tiles[item >> 8];
    const VolGeom g = vols[tr.vol];
    int64_t w0, w_end;
    gather_item_span(item, (int)(it % parts), parts, tr.w0, g.nwords, w0, w_end);
    if (w0 < w_end)
      gather_span<true>(g, bits + g.word_off, base_v + g.word_off, adj, n_adj,
                        dldf + g.field_off, w0, w_end, sh, s_dyn);
    __syncthreads();  // LDS reused by the next span
  }
}

}  // namespace mc
}  // namespace daxhip

using namespace daxhip;
using namespace daxhip::mc;

namespace {

int persistent_blocks(int per_cu) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus * per_cu;
}

// Tuning cache per batch shape class: (volumes, ceil(log2(row words)),
// widest row in words).
std::mutex g_btune_mu;
std::map<std::tuple<int64_t, int, int64_t>, BatchTuning> g_btune_cache;

std::tuple<int64_t, int, int64_t> batch_class(const BatchHeader& h) {
  int lg = 0;
  while ((1LL << lg) < h.nwords) ++lg;
  return std::make_tuple(h.nvols, lg, h.max_wpr);
}

// Device pointers of a prepared workspace.
struct BatchPtrs {
  const VolGeom* vols;
  const TileRef* scan_tiles;
  const TileRef* gen_tiles;
  uint64_t* bits;
  uint32_t* bv;
  uint32_t* bt;
  unsigned long long* status;
  Ctrl* ctrl;
};

BatchPtrs batch_ptrs(const BatchHeader& h, void* d_workspace) {
  char* body = (char*)d_workspace + kBatchHeaderBytes;
  return BatchPtrs{(const VolGeom*)(body + h.off_vols), (const TileRef*)(body + h.off_scan_tiles),
                   (const TileRef*)(body + h.off_gen_tiles), (uint64_t*)(body + h.off_bits),
                   (uint32_t*)(body + h.off_base_v), (uint32_t*)(body + h.off_base_t),
                   (unsigned long long*)(body + h.off_status), (Ctrl*)(body + h.off_ctrl)};
}

// Classify (which also clears the scan's status and control words).
int batch_classify(const BatchHeader& h, const BatchPtrs& p, const float* d_fields,
                   const BatchTuning& t, hipStream_t s) {
  const int64_t zero_words = h.nscan + (int64_t)(sizeof(Ctrl) / 8);
  const dim3 grid((unsigned)persistent_blocks(t.cls_blocks)), blk(256);
  switch (t.cls_unroll) {
    case 8:
      k_batch_classify<8><<<grid, blk, 0, s>>>(p.vols, p.gen_tiles, h.ngen, d_fields, p.bits,
                                               p.status, zero_words);
      break;
    case 32:
      k_batch_classify<32><<<grid, blk, 0, s>>>(p.vols, p.gen_tiles, h.ngen, d_fields, p.bits,
                                                p.status, zero_words);
      break;
    default:
      k_batch_classify<16><<<grid, blk, 0, s>>>(p.vols, p.gen_tiles, h.ngen, d_fields, p.bits,
                                                p.status, zero_words);
  }
  DAXHIP_LAUNCH_CHECK("k_batch_classify");
  return DAX_HIP_OK;
}

int batch_scan(const BatchHeader& h, const BatchPtrs& p, const BatchTuning& t, hipStream_t s) {
  const dim3 grid((unsigned)std::min<int64_t>(persistent_blocks(4), h.nscan));
  switch (t.scan_threads) {
    case 512:
      k_batch_count_scan<512><<<grid, dim3(512), 0, s>>>(p.vols, p.scan_tiles, h.nscan, p.bits,
                                                         p.bv, p.bt, p.status, p.ctrl);
      break;
    case 1024:
      k_batch_count_scan<1024><<<grid, dim3(1024), 0, s>>>(p.vols, p.scan_tiles, h.nscan, p.bits,
                                                           p.bv, p.bt, p.status, p.ctrl);
      break;
    default:
      k_batch_count_scan<256><<<grid, dim3(256), 0, s>>>(p.vols, p.scan_tiles, h.nscan, p.bits,
                                                         p.bv, p.bt, p.status, p.ctrl);
  }
  DAXHIP_LAUNCH_CHECK("k_batch_count_scan");
  return DAX_HIP_OK;
}

int batch_generate(const BatchHeader& h, const BatchPtrs& p, const float* d_fields,
                   unsigned flags, const OutPtrs& o, int64_t cap_points, int64_t cap_tris,
                   int64_t* d_counts, const BatchTuning& t, hipStream_t s) {
  const size_t lds = GenLayout<kBatchGenTile>((int)h.max_wpr).bytes();
  k_batch_generate<<<dim3((unsigned)std::min<int64_t>(persistent_blocks(t.gen_blocks), h.ngen)),
                     dim3(kScanThreads), lds, s>>>(p.vols, p.gen_tiles, h.ngen, d_fields, p.bits,
                                                   p.bv, p.bt, p.ctrl, o, flags, cap_points,
                                                   cap_tris, d_counts, t.gen_xcd);
  DAXHIP_LAUNCH_CHECK("k_batch_generate");
  return DAX_HIP_OK;
}

}  // namespace

extern "C" {

size_t dax_hip_mc_batch_workspace_bytes(int64_t n_volumes, const int64_t* h_dims3) {
  if (n_volumes <= 0 || !h_dims3) return 0;
  return kBatchHeaderBytes + batch_layout(n_volumes, h_dims3).bytes;
}

// Writes the volume and tile tables into the workspace (synchronous upload).
int dax_hip_mc_batch_prepare(int64_t n_volumes, const int64_t* h_dims3,
                             const int64_t* h_field_offsets, const float* h_isos,
                             const float* h_origins3, const float* h_spacings3, void* d_workspace,
                             size_t workspace_bytes, void* h_header_out, void* stream) {
  if (n_volumes <= 0 || n_volumes > 0x7fffffffLL || !h_dims3 || !h_field_offsets || !h_isos) {
    set_error("dax_hip_mc_batch_prepare: bad volume description");
    return DAX_HIP_BAD_VALUE;
  }
  const BatchLayout L = batch_layout(n_volumes, h_dims3);
  if (!d_workspace || workspace_bytes < kBatchHeaderBytes + L.bytes) {
    set_error("dax_hip_mc_batch_prepare: workspace of %zu bytes, %zu needed", workspace_bytes,
              kBatchHeaderBytes + L.bytes);
    return DAX_HIP_BAD_VALUE;
  }
  std::vector<VolGeom> vols((size_t)n_volumes);
  std::vector<TileRef> scan_tiles, gen_tiles;
  scan_tiles.reserve((size_t)L.nscan);
  gen_tiles.reserve((size_t)L.ngen);
  int64_t word_off = 0;
  for (int64_t v = 0; v < n_volumes; ++v) {
    const int64_t nx = h_dims3[3 * v], ny = h_dims3[3 * v + 1], nz = h_dims3[3 * v + 2];
    if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz > (1LL << 30) || ceil_div(nx, 64) > kMaxWpr) {
      set_error("dax_hip_mc_batch_prepare: volume %lld has unsupported dims %lld x %lld x %lld",
                (long long)v, (long long)nx, (long long)ny, (long long)nz);
      return DAX_HIP_BAD_VALUE;
    }
    VolGeom& g = vols[(size_t)v];
    g.nx = (int32_t)nx;
    g.ny = (int32_t)ny;
    g.nz = (int32_t)nz;
    g.wpr = (int32_t)ceil_div(nx, 64);
    g.iso = h_isos[v];
    for (int d = 0; d < 3; ++d) {
      g.origin[d] = h_origins3 ? h_origins3[3 * v + d] : 0.0f;
      g.spacing[d] = h_spacings3 ? h_spacings3[3 * v + d] : 1.0f;
    }
    g.field_off = h_field_offsets[v];
    g.word_off = word_off;
    g.nwords = (int64_t)g.wpr * ny * nz;
    set_geom_divs(g);
    for (int64_t w = 0; w < g.nwords; w += kBatchScanTile) scan_tiles.push_back({(int32_t)v, (int32_t)w});
    for (int64_t w = 0; w < g.nwords; w += kBatchGenTile) gen_tiles.push_back({(int32_t)v, (int32_t)w});
    word_off += g.nwords;
  }
  BatchHeader h;
  h.magic = kBatchMagic;
  h.nvols = n_volumes;
  h.nwords = L.nwords;
  h.nscan = L.nscan;
  h.ngen = L.ngen;
  h.max_wpr = 1;
  for (const VolGeom& g : vols) h.max_wpr = std::max<int64_t>(h.max_wpr, g.wpr);
  h.off_vols = L.off_vols;
  h.off_scan_tiles = L.off_scan_tiles;
  h.off_gen_tiles = L.off_gen_tiles;
  h.off_bits = L.off_bits;
  h.off_base_v = L.off_base_v;
  h.off_base_t = L.off_base_t;
  h.off_status = L.off_status;
  h.off_ctrl = L.off_ctrl;
  h.bytes = L.bytes;
  h.tune = kBatchTuningDefault;
  {
    std::lock_guard<std::mutex> lk(g_btune_mu);
    const auto it = g_btune_cache.find(batch_class(h));
    if (it != g_btune_cache.end()) h.tune = it->second;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  char* body = ws + kBatchHeaderBytes;
  DAXHIP_CHECK(hipMemcpyAsync(ws, &h, sizeof(h), hipMemcpyHostToDevice, s));
  DAXHIP_CHECK(hipMemcpyAsync(body + L.off_vols, vols.data(), vols.size() * sizeof(VolGeom),
                              hipMemcpyHostToDevice, s));
  DAXHIP_CHECK(hipMemcpyAsync(body + L.off_scan_tiles, scan_tiles.data(),
                              scan_tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice, s));
  DAXHIP_CHECK(hipMemcpyAsync(body + L.off_gen_tiles, gen_tiles.data(),
                              gen_tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  if (h_header_out) memcpy(h_header_out, &h, sizeof(h));
  return DAX_HIP_OK;
}

size_t dax_hip_mc_batch_header_bytes(void) { return sizeof(BatchHeader); }

// The batch forward: four persistent launches, graph-capturable.  h_header is
// the header returned by prepare.  d_counts {T, U} totals; d_vol_offsets
// [n][2] {first triangle, first point} of each volume.
int dax_hip_mc_batch_forward(const float* d_fields, const void* h_header, void* d_workspace,
                             size_t workspace_bytes, unsigned flags, float* d_vertices,
                             float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                             int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                             int64_t* d_vol_offsets, void* stream) {
  const BatchHeader* h = (const BatchHeader*)h_header;
  if (!h || h->magic != kBatchMagic || !d_workspace ||
      workspace_bytes < kBatchHeaderBytes + h->bytes) {
    set_error("dax_hip_mc_batch_forward: header/workspace not prepared");
    return DAX_HIP_BAD_VALUE;
  }
  if (!d_fields || !d_vertices || !d_conn || !d_counts) {
    set_error("dax_hip_mc_batch_forward: fields, vertices, connectivity and counts are required");
    return DAX_HIP_BAD_VALUE;
  }
  if (d_edges && !(flags & DAX_HIP_MC_DEDUP)) {
    set_error("dax_hip_mc_batch_forward: the compact edge list needs dedup");
    return DAX_HIP_BAD_VALUE;
  }
  hipStream_t s = (hipStream_t)stream;
  const BatchPtrs p = batch_ptrs(*h, d_workspace);
  const BatchTuning t = valid_batch_tuning(h->tune) ? h->tune : kBatchTuningDefault;
  int rc;
  if ((rc = batch_classify(*h, p, d_fields, t, s))) return rc;
  if ((rc = batch_scan(*h, p, t, s))) return rc;
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, nullptr, nullptr, nullptr};
  if ((rc = batch_generate(*h, p, d_fields, flags, o, cap_points, cap_tris, d_counts, t, s)))
    return rc;
  if (d_vol_offsets) {
    k_batch_offsets<<<dim3((unsigned)ceil_div(h->nvols, 256)), dim3(256), 0, s>>>(
        p.vols, h->nvols, p.bv, p.bt, flags, d_vol_offsets);
    DAXHIP_LAUNCH_CHECK("k_batch_offsets");
  }
  return DAX_HIP_OK;
}

// Launch shapes of a prepared batch: {classify unroll, classify blocks per
// CU, scan threads per tile, generate blocks per CU, generate XCD order}.
int dax_hip_mc_batch_set_tuning(void* h_header, const int* h_params5) {
  BatchHeader* h = (BatchHeader*)h_header;
  if (!h || h->magic != kBatchMagic || !h_params5) {
    set_error("dax_hip_mc_batch_set_tuning: header not prepared or NULL parameters");
    return DAX_HIP_BAD_VALUE;
  }
  const BatchTuning t{h_params5[0], h_params5[1], h_params5[2], h_params5[3], h_params5[4], 0};
  if (!valid_batch_tuning(t)) {
    set_error("dax_hip_mc_batch_set_tuning: invalid parameters {%d,%d,%d,%d,%d}", t.cls_unroll,
              t.cls_blocks, t.scan_threads, t.gen_blocks, t.gen_xcd);
    return DAX_HIP_BAD_VALUE;
  }
  h->tune = t;
  return DAX_HIP_OK;
}

int dax_hip_mc_batch_get_tuning(const void* h_header, int* h_params5) {
  const BatchHeader* h = (const BatchHeader*)h_header;
  if (!h || h->magic != kBatchMagic || !h_params5) {
    set_error("dax_hip_mc_batch_get_tuning: header not prepared or NULL parameters");
    return DAX_HIP_BAD_VALUE;
  }
  const BatchTuning t = valid_batch_tuning(h->tune) ? h->tune : kBatchTuningDefault;
  h_params5[0] = t.cls_unroll;
  h_params5[1] = t.cls_blocks;
  h_params5[2] = t.scan_threads;
  h_params5[3] = t.gen_blocks;
  h_params5[4] = t.gen_xcd;
  return DAX_HIP_OK;
}

// Times every launch shape of each phase on this batch (reps launches each,
// HIP events on `stream`; synchronous), keeps the fastest in the header and
// in the cache for the batch's shape class (later prepares of that class
// start from it).  h_best5: the choice; h_best_ms3: its phase times.
int dax_hip_mc_batch_autotune(const float* d_fields, void* h_header, void* d_workspace,
                              size_t workspace_bytes, unsigned flags, float* d_vertices,
                              float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                              int64_t cap_points, int64_t cap_tris, int64_t* d_counts, int reps,
                              int* h_best5, float* h_best_ms3, void* stream) {
  BatchHeader* h = (BatchHeader*)h_header;
  if (!h || h->magic != kBatchMagic || !d_workspace ||
      workspace_bytes < kBatchHeaderBytes + h->bytes || !d_fields || !d_vertices || !d_conn ||
      !d_counts) {
    set_error("dax_hip_mc_batch_autotune: header/workspace not prepared or outputs missing");
    return DAX_HIP_BAD_VALUE;
  }
  if (reps < 1) reps = 1;
  hipStream_t s = (hipStream_t)stream;
  const BatchPtrs p = batch_ptrs(*h, d_workspace);
  OutPtrs o{d_vertices, d_normals, d_conn, d_edges, nullptr, nullptr, nullptr};
  BatchTuning best = valid_batch_tuning(h->tune) ? h->tune : kBatchTuningDefault;
  float best_ms[3] = {1e30f, 1e30f, 1e30f};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DAXHIP_CHECK(hipEventCreate(&e0));
  DAXHIP_CHECK(hipEventCreate(&e1));
  int rc = DAX_HIP_OK;
  auto elapsed = [&]() -> float {
    float m = 0.0f;
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&m, e0, e1);
    return m;
  };
  // classify
  for (int u : {8, 16, 32})
    for (int bpc : {2, 4, 8}) {
      BatchTuning t = best;
      t.cls_unroll = u;
      t.cls_blocks = bpc;
      if ((rc = batch_classify(*h, p, d_fields, t, s))) goto done;  // warm
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps; ++r)
        if ((rc = batch_classify(*h, p, d_fields, t, s))) goto done;
      (void)hipEventRecord(e1, s);
      const float ms = elapsed() / reps;
      if (ms < best_ms[0]) {
        best_ms[0] = ms;
        best.cls_unroll = u;
        best.cls_blocks = bpc;
      }
    }
  // scan: each timed launch follows a classify that resets its state
  for (int nt : {256, 512, 1024}) {
    BatchTuning t = best;
    t.scan_threads = nt;
    float total = 0.0f;
    for (int r = 0; r < reps; ++r) {
      if ((rc = batch_classify(*h, p, d_fields, best, s))) goto done;
      (void)hipEventRecord(e0, s);
      if ((rc = batch_scan(*h, p, t, s))) goto done;
      (void)hipEventRecord(e1, s);
      total += elapsed();
    }
    if (total / reps < best_ms[1]) {
      best_ms[1] = total / reps;
      best.scan_threads = nt;
    }
  }
  // generate over one classify + scan
  if ((rc = batch_classify(*h, p, d_fields, best, s))) goto done;
  if ((rc = batch_scan(*h, p, best, s))) goto done;
  for (int bpc : {2, 4, 8})
    for (int xcd : {0, 1}) {
      BatchTuning t = best;
      t.gen_blocks = bpc;
      t.gen_xcd = xcd;
      if ((rc = batch_generate(*h, p, d_fields, flags, o, cap_points, cap_tris, d_counts, t, s)))
        goto done;
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps; ++r)
        if ((rc = batch_generate(*h, p, d_fields, flags, o, cap_points, cap_tris, d_counts, t, s)))
          goto done;
      (void)hipEventRecord(e1, s);
      const float ms = elapsed() / reps;
      if (ms < best_ms[2]) {
        best_ms[2] = ms;
        best.gen_blocks = bpc;
        best.gen_xcd = xcd;
      }
    }
  h->tune = best;
  {
    std::lock_guard<std::mutex> lk(g_btune_mu);
    g_btune_cache[batch_class(*h)] = best;
  }
  if (h_best5) {
    h_best5[0] = best.cls_unroll;
    h_best5[1] = best.cls_blocks;
    h_best5[2] = best.scan_threads;
    h_best5[3] = best.gen_blocks;
    h_best5[4] = best.gen_xcd;
  }
  if (h_best_ms3)
    for (int i = 0; i < 3; ++i) h_best_ms3[i] = best_ms[i];
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Batched adjoint: dL/dfield of every volume (at the fields' offsets) from
// dL/dvertex and dL/dnormal of the concatenated outputs of the last batch
// forward on this workspace.  d_scratch: dax_hip_mc_backward_scratch_bytes
// (n_points).
int dax_hip_mc_batch_backward(const float* d_fields, const void* h_header,
                              const void* d_workspace, size_t workspace_bytes,
                              const uint32_t* d_edges, const int64_t* d_counts, int64_t n_points,
                              const float* d_dl_dvertices, const float* d_dl_dnormals,
                              void* d_scratch, size_t scratch_bytes, float* d_dl_dfields,
                              void* stream) {
  const BatchHeader* h = (const BatchHeader*)h_header;
  if (!h || h->magic != kBatchMagic || !d_workspace ||
      workspace_bytes < kBatchHeaderBytes + h->bytes) {
    set_error("dax_hip_mc_batch_backward: header/workspace not prepared");
    return DAX_HIP_BAD_VALUE;
  }
  if (h->ngen >= (1 << 24)) {
    set_error("dax_hip_mc_batch_backward: %lld tiles exceed the 2^24 the adjoint lists",
              (long long)h->ngen);
    return DAX_HIP_BAD_VALUE;
  }
  if (h->max_wpr > kMaxGatherWpr) {
    set_error("dax_hip_mc_batch_backward: rows of %lld points exceed the %d the adjoint stages",
              (long long)h->max_wpr * 64, kMaxGatherWpr * 64);
    return DAX_HIP_BAD_VALUE;
  }
  const size_t need = n_points > 0 ? adjoint_scratch_bytes(n_points) : 0;
  if (n_points < 0 || scratch_bytes < need ||
      (n_points > 0 && (!d_scratch || !d_edges || !d_dl_dvertices || !d_dl_dnormals)) ||
      !d_fields || !d_dl_dfields) {
    set_error("dax_hip_mc_batch_backward: missing buffers or scratch smaller than %zu bytes",
              need);
    return DAX_HIP_BAD_VALUE;
  }
  hipStream_t s = (hipStream_t)stream;
  const char* body = (const char*)d_workspace + kBatchHeaderBytes;
  const VolGeom* vols = (const VolGeom*)(body + h->off_vols);
  const TileRef* gen_tiles = (const TileRef*)(body + h->off_gen_tiles);
  const uint64_t* bits = (const uint64_t*)(body + h->off_bits);
  const uint32_t* bv = (const uint32_t*)(body + h->off_base_v);
  const int grid = (int)std::min<int64_t>(persistent_blocks(4), h->ngen);
  uint32_t* list =
      n_points > 0 ? (uint32_t*)((char*)d_scratch + gather_list_offset(n_points)) : nullptr;
  const size_t lds = gather_lds_bytes((int)h->max_wpr);
  if (n_points > 0) {
    k_batch_vertex_adjoint<<<dim3(grid), dim3(256), 0, s>>>(vols, gen_tiles, h->ngen, d_fields,
                                                            bv, d_edges, d_dl_dvertices,
                                                            d_dl_dnormals, d_counts, n_points,
                                                            (VertexAdj*)d_scratch, list);
    DAXHIP_LAUNCH_CHECK("k_batch_vertex_adjoint");
  }
  k_batch_point_gather<<<dim3((unsigned)h->ngen), dim3(kGatherThreads), lds, s>>>(
      vols, gen_tiles, bits, bv, (const VertexAdj*)d_scratch, d_counts, n_points, d_dl_dfields,
      list);
  DAXHIP_LAUNCH_CHECK("k_batch_point_gather");
  if (n_points > 0) {
    k_batch_point_gather_spans<<<dim3((unsigned)persistent_blocks(3)), dim3(kGatherThreads), lds,
                                 s>>>(vols, gen_tiles, bits, bv, (const VertexAdj*)d_scratch,
                                      d_counts, n_points, d_dl_dfields, list);
    DAXHIP_LAUNCH_CHECK("k_batch_point_gather_spans");
  }
  return DAX_HIP_OK;
}

// Capture of any sequence of this library's launches on `stream` (e.g. a
// batch forward followed by its backward) into one executable graph.
int dax_hip_graph_begin_capture(void* stream) {
  if (!stream) {
    set_error("dax_hip_graph_begin_capture: capture needs a non-null stream");
    return DAX_HIP_BAD_VALUE;
  }
  DAXHIP_CHECK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return DAX_HIP_OK;
}

int dax_hip_graph_end_capture(void* stream, void** h_graph_exec) {
  hipGraph_t graph = nullptr;
  DAXHIP_CHECK(hipStreamEndCapture((hipStream_t)stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
  *h_graph_exec = exec;
  return DAX_HIP_OK;
}

// Baseline of the batched path: the same volumes, one single-volume forward
// (three launches) and, when upstream gradients are given, one single-volume
// backward (two launches) per volume, written at the offsets of the batched
// outputs (h_vol_offsets, as the batched forward reports them).
int dax_hip_mc_volume_loop(int64_t n_volumes, const int64_t* h_dims3,
                           const int64_t* h_field_offsets, const float* h_isos,
                           const float* h_origins3, const float* h_spacings3,
                           const float* d_fields, unsigned flags, void* d_workspace,
                           size_t workspace_bytes, const int64_t* h_vol_offsets,
                           int64_t cap_points, int64_t cap_tris, float* d_vertices,
                           float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                           int64_t* d_counts_scratch, const float* d_dl_dvertices,
                           const float* d_dl_dnormals, void* d_scratch, size_t scratch_bytes,
                           float* d_dl_dfields, void* stream) {
  if (n_volumes <= 0 || !h_dims3 || !h_field_offsets || !h_isos || !h_vol_offsets ||
      !d_counts_scratch) {
    set_error("dax_hip_mc_volume_loop: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  const bool backward = d_dl_dfields != nullptr;
  for (int64_t v = 0; v < n_volumes; ++v) {
    const int64_t t0 = h_vol_offsets[2 * v], p0 = h_vol_offsets[2 * v + 1];
    const int64_t t1 = v + 1 < n_volumes ? h_vol_offsets[2 * v + 2] : cap_tris;
    const int64_t p1 = v + 1 < n_volumes ? h_vol_offsets[2 * v + 3] : cap_points;
    if (t0 < 0 || p0 < 0 || t1 < t0 || p1 < p0 || t1 > cap_tris || p1 > cap_points) {
      set_error("dax_hip_mc_volume_loop: volume %lld offsets outside the outputs", (long long)v);
      return DAX_HIP_BAD_VALUE;
    }
    const int64_t* d = h_dims3 + 3 * v;
    const float* o = h_origins3 ? h_origins3 + 3 * v : nullptr;
    const float* sp = h_spacings3 ? h_spacings3 + 3 * v : nullptr;
    const float* f = d_fields + h_field_offsets[v];
    int rc = dax_hip_mc_forward(f, d[0], d[1], d[2], o, sp, h_isos[v], flags, d_workspace,
                                workspace_bytes, d_vertices + 3 * p0,
                                d_normals ? d_normals + 3 * p0 : nullptr, d_conn + 3 * t0,
                                d_edges ? d_edges + p0 : nullptr, nullptr, nullptr, nullptr,
                                p1 - p0, t1 - t0, d_counts_scratch, stream);
    if (rc) return rc;
    if (backward) {
      rc = dax_hip_mc_backward(f, d[0], d[1], d[2], o, sp, h_isos[v], d_workspace,
                               workspace_bytes, d_edges ? d_edges + p0 : nullptr,
                               d_counts_scratch, p1 - p0,
                               d_dl_dvertices ? d_dl_dvertices + 3 * p0 : nullptr,
                               d_dl_dnormals ? d_dl_dnormals + 3 * p0 : nullptr,
                               d_scratch, scratch_bytes,
                               d_dl_dfields + h_field_offsets[v], stream);
      if (rc) return rc;
    }
  }
  return DAX_HIP_OK;
}

// ---- graphs: capture a launch sequence once, replay it -------------------
// Captures dax_hip_mc_batch_forward(args) on `stream` into an executable
// graph (h_graph_exec receives a hipGraphExec_t).
int dax_hip_mc_batch_graph_create(const float* d_fields, const void* h_header, void* d_workspace,
                                  size_t workspace_bytes, unsigned flags, float* d_vertices,
                                  float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                                  int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                                  int64_t* d_vol_offsets, void* stream, void** h_graph_exec) {
  hipStream_t s = (hipStream_t)stream;
  if (!s) {
    set_error("dax_hip_mc_batch_graph_create: capture needs a non-null stream");
    return DAX_HIP_BAD_VALUE;
  }
  DAXHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = dax_hip_mc_batch_forward(d_fields, h_header, d_workspace, workspace_bytes, flags,
                                    d_vertices, d_normals, d_conn, d_edges, cap_points, cap_tris,
                                    d_counts, d_vol_offsets, stream);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
  *h_graph_exec = exec;
  return DAX_HIP_OK;
}

int dax_hip_graph_launch(void* graph_exec, void* stream) {
  DAXHIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_graph_destroy(void* graph_exec) {
  if (graph_exec) DAXHIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return DAX_HIP_OK;
}

}  // extern "C"
