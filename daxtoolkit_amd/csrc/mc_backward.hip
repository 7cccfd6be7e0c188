// Adjoint of the marching-cubes forward (no reference counterpart): the
// single-volume launches and C-ABI.  The math and the per-vertex / per-point
// device code are in mc_adjoint.h.
#include "mc_adjoint.h"

#include <atomic>

namespace daxhip {
namespace mc {

// One thread per vertex; block 0 also empties the list of dense tiles and
// writes the gather tiles' dispatch order (those with active edges of their
// own first, from the forward count kernel's classes; the other blocks'
// records hide its time).
constexpr int kVertexAdjThreads = 256;
__global__ __launch_bounds__(kVertexAdjThreads) void k_vertex_adjoint(
    VolGeom g, const float* __restrict__ field, const uint32_t* __restrict__ edges,
    const float* __restrict__ dldx, const float* __restrict__ dldn,
    const int64_t* __restrict__ counts, int64_t n_points, VertexAdj* __restrict__ out,
    uint32_t* __restrict__ list, const uint8_t* __restrict__ tcls, uint32_t* __restrict__ gorder) {
  // XCD-contiguous vertex ranges: vertices are in edge order, so the field
  // lines of neighbouring stencils stay in one XCD's L2
  const int64_t u = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (u == 0) list[0] = 0u;
  if (blockIdx.x == 0) {  // uniform
    __shared__ uint32_t s_a[kVertexAdjThreads / 64], s_b[kVertexAdjThreads / 64];
    const uint32_t ngt = (uint32_t)ceil_div(g.nwords, (int64_t)kGatherTile);
    const uint32_t K = order_chunk(ngt, kVertexAdjThreads);
    const uint64_t reg = load_classes(tcls, threadIdx.x * K, K, ngt);
    write_order<true>(reg, tcls, K, ngt, gorder, (uint32_t)kGatherTile, s_a, s_b);
  }
  const int64_t U = counts ? min(counts[1], n_points) : n_points;
  if (u >= U) return;
  vertex_adjoint(g, field, edges, dldx, dldn, u, out);
}

// One tile of kGatherTile row words per block, in the dispatch order the
// vertex-record launch wrote (tiles with active edges of their own first, so
// the long tiles start early and the quiet ones fill the tail), else (no
// vertices: the forward left the order cleared) in tile order.  A tile whose
// endpoint records do not fit runs in spans or goes to the list (without
// records, list == nullptr, every point receives 0).
__global__ __launch_bounds__(kGatherThreads) void k_point_gather(
    VolGeom g, const uint64_t* __restrict__ bits, const uint32_t* __restrict__ base_v,
    const VertexAdj* __restrict__ adj, const int64_t* __restrict__ counts, int64_t n_points,
    float* __restrict__ dldf, uint32_t* __restrict__ list, const uint32_t* __restrict__ order,
    int span_lg) {
  __shared__ GatherShared sh;
  extern __shared__ __align__(16) uint64_t s_dyn[];
  const uint32_t tag = order[0], ord = order[1 + blockIdx.x] & 0x7fffffffu;
  const int64_t tile =
      (tag == (uint32_t)kGatherTile && (int64_t)ord < (int64_t)gridDim.x) ? (int64_t)ord
                                                                            : (int64_t)blockIdx.x;
  const int64_t n_adj = counts ? min(counts[1], n_points) : n_points;
  const int64_t w0 = tile * kGatherTile, w_end = min(w0 + kGatherTile, g.nwords);
  gather_tile(g, bits + g.word_off, base_v + g.word_off, adj, n_adj, dldf + g.field_off, w0,
              w_end, sh, s_dyn, list, gather_list_capacity(n_points), (uint32_t)tile, span_lg);
}

// The listed span items (persistent blocks).
__global__ __launch_bounds__(kGatherThreads) void k_point_gather_spans(
    VolGeom g, const uint64_t* __restrict__ bits, const uint32_t* __restrict__ base_v,
    const VertexAdj* __restrict__ adj, const int64_t* __restrict__ counts, int64_t n_points,
    float* __restrict__ dldf, const uint32_t* __restrict__ list) {
  __shared__ GatherShared sh;
  extern __shared__ __align__(16) uint64_t s_dyn[];
  const int64_t n_adj = counts ? min(counts[1], n_points) : n_points;
  const int64_t n_items = min((int64_t)list[0], gather_list_capacity(n_points));
  const int parts = gather_item_parts(n_items, gridDim.x / 2);
  for (int64_t it = blockIdx.x; it < n_items * parts; it += gridDim.x) {
    const uint32_t item = list[1 + it / parts];
    int64_t w0, w_end;
    gather_item_span(item, (int)(it % parts), parts, (int64_t)(item >> 8) * kGatherTile,
                     g.nwords, w0, w_end);
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

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}


// Dense tiles run their spans in their own block when the gather holds at
// least this many tiles per CU (by default 16: 1024^3, 256 per CU, 1.55 +
// 0.47 ms -> 1.92 + 0.01 ms; 512^3, 32 per CU, forward + backward 0.369 ->
// 0.359 ms); with fewer rounds of tiles such a block is the critical path
// (256^3, 4 per CU: 49 + 29 us -> 88 + 15 us).
std::atomic<int64_t> g_span_tiles_per_cu{16};

int gather_tile_span_lg(int64_t n_tiles, int cus) {
  return n_tiles >= g_span_tiles_per_cu.load(std::memory_order_relaxed) * (int64_t)cus
             ? kTileSpanLg
             : 0;
}

}  // namespace

extern "C" {

int dax_hip_mc_backward_set_span_threshold(int64_t tiles_per_cu) {
  if (tiles_per_cu < 0) {
    set_error("dax_hip_mc_backward_set_span_threshold: negative threshold %lld",
              (long long)tiles_per_cu);
    return DAX_HIP_BAD_VALUE;
  }
  g_span_tiles_per_cu.store(tiles_per_cu, std::memory_order_relaxed);
  return DAX_HIP_OK;
}

size_t dax_hip_mc_backward_scratch_bytes(int64_t n_points) {
  return n_points > 0 ? adjoint_scratch_bytes(n_points) : 0;
}

int dax_hip_mc_backward(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                        const float* h_origin3, const float* h_spacing3, float iso,
                        void* d_workspace, size_t workspace_bytes, const uint32_t* d_edges,
                        const int64_t* d_counts, int64_t n_points, const float* d_dl_dvertices,
                        const float* d_dl_dnormals, void* d_scratch, size_t scratch_bytes,
                        float* d_dl_dfield, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz > (1LL << 30) ||
      (nx + 63) / 64 > kMaxGatherWpr) {
    set_error("dax_hip_mc_backward: unsupported grid %lld x %lld x %lld", (long long)nx,
              (long long)ny, (long long)nz);
    return DAX_HIP_BAD_VALUE;
  }
  VolGeom g;
  g.nx = (int32_t)nx;
  g.ny = (int32_t)ny;
  g.nz = (int32_t)nz;
  g.wpr = (int32_t)ceil_div(nx, 64);
  g.iso = iso;
  for (int d = 0; d < 3; ++d) {
    g.origin[d] = h_origin3 ? h_origin3[d] : 0.0f;
    g.spacing[d] = h_spacing3 ? h_spacing3[d] : 1.0f;
  }
  g.field_off = 0;
  g.word_off = 0;
  g.nwords = (int64_t)g.wpr * ny * nz;
  set_geom_divs(g);
  const Workspace L = workspace_layout(g.nwords);
  if (!d_workspace || workspace_bytes < L.bytes) {
    set_error("dax_hip_mc_backward: workspace of %zu bytes, %zu needed", workspace_bytes, L.bytes);
    return DAX_HIP_BAD_VALUE;
  }
  if (n_points < 0 || scratch_bytes < dax_hip_mc_backward_scratch_bytes(n_points) ||
      (n_points > 0 && (!d_scratch || !d_edges || !d_dl_dvertices || !d_dl_dnormals)) ||
      !d_field || !d_dl_dfield) {
    set_error("dax_hip_mc_backward: missing buffers or scratch smaller than %zu bytes",
              dax_hip_mc_backward_scratch_bytes(n_points));
    return DAX_HIP_BAD_VALUE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_workspace;
  const uint64_t* bits = (const uint64_t*)(ws + L.off_bits);
  const uint32_t* bv = (const uint32_t*)(ws + L.off_base_v);
  uint32_t* list =
      n_points > 0 ? (uint32_t*)((char*)d_scratch + gather_list_offset(n_points)) : nullptr;
  const int cus = device_cus();
  if (n_points > 0) {
    k_vertex_adjoint<<<dim3((unsigned)ceil_div(n_points, kVertexAdjThreads)),
                       dim3(kVertexAdjThreads), 0, s>>>(
        g, d_field, d_edges, d_dl_dvertices, d_dl_dnormals, d_counts, n_points,
        (VertexAdj*)d_scratch, list, (const uint8_t*)(ws + L.off_tcls),
        (uint32_t*)(ws + L.off_gorder));
    DAXHIP_LAUNCH_CHECK("k_vertex_adjoint");
  }
  k_point_gather<<<dim3((unsigned)ceil_div(g.nwords, kGatherTile)), dim3(kGatherThreads),
                   gather_lds_bytes(g.wpr), s>>>(g, bits, bv, (const VertexAdj*)d_scratch,
                                                 d_counts, n_points, d_dl_dfield, list,
                                                 (const uint32_t*)(ws + L.off_gorder),
                                                 gather_tile_span_lg(ceil_div(g.nwords, kGatherTile), cus));
  DAXHIP_LAUNCH_CHECK("k_point_gather");
  if (n_points > 0) {
    // persistent blocks: three per CU, as many as the LDS allows
    k_point_gather_spans<<<dim3((unsigned)(3 * cus)), dim3(kGatherThreads),
                           gather_lds_bytes(g.wpr), s>>>(g, bits, bv, (const VertexAdj*)d_scratch,
                                                         d_counts, n_points, d_dl_dfield, list);
    DAXHIP_LAUNCH_CHECK("k_point_gather_spans");
  }
  return DAX_HIP_OK;
}

}  // extern "C"
