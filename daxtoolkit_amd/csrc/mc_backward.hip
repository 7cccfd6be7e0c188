// Adjoint of the marching-cubes forward (no reference counterpart): the
// single-volume launches and C-ABI.  The math and the per-vertex / per-point
// device code are in mc_adjoint.h.
#include "mc_adjoint.h"

namespace daxhip {
namespace mc {

__global__ __launch_bounds__(256) void k_vertex_adjoint(VolGeom g, const float* __restrict__ field,
                                                        const uint32_t* __restrict__ edges,
                                                        const float* __restrict__ dldx,
                                                        const float* __restrict__ dldn,
                                                        const int64_t* __restrict__ counts,
                                                        int64_t n_points,
                                                        VertexAdj* __restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t U = counts ? min(counts[1], n_points) : n_points;
  if (u >= U) return;
  vertex_adjoint(g, field, edges, dldx, dldn, u, out);
}

// One tile of kGatherTile row words per block, in tile order (round-robin
// over the XCDs: a contiguous range per XCD leaves the XCDs that hold dense
// surface working long after the others).
__global__ __launch_bounds__(kGatherThreads) void k_point_gather(
    VolGeom g, const uint64_t* __restrict__ bits, const uint32_t* __restrict__ base_v,
    const VertexAdj* __restrict__ adj, const int64_t* __restrict__ counts, int64_t n_points,
    float* __restrict__ dldf) {
  __shared__ GatherShared sh;
  extern __shared__ __align__(16) uint64_t s_dyn[];
  const int64_t tile = blockIdx.x;
  const int64_t n_adj = counts ? min(counts[1], n_points) : n_points;
  const int64_t w0 = tile * kGatherTile;
  gather_tile(g, bits + g.word_off, base_v + g.word_off, adj, n_adj, dldf + g.field_off, w0,
              min(w0 + kGatherTile, g.nwords), sh, s_dyn);
}

}  // namespace mc
}  // namespace daxhip

using namespace daxhip;
using namespace daxhip::mc;

extern "C" {

size_t dax_hip_mc_backward_scratch_bytes(int64_t n_points) {
  return n_points > 0 ? (size_t)n_points * sizeof(VertexAdj) : 0;
}

int dax_hip_mc_backward(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                        const float* h_origin3, const float* h_spacing3, float iso,
                        const void* d_workspace, size_t workspace_bytes, const uint32_t* d_edges,
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
  const char* ws = (const char*)d_workspace;
  if (n_points > 0) {
    k_vertex_adjoint<<<dim3((unsigned)ceil_div(n_points, 256)), dim3(256), 0, s>>>(
        g, d_field, d_edges, d_dl_dvertices, d_dl_dnormals, d_counts, n_points,
        (VertexAdj*)d_scratch);
    DAXHIP_LAUNCH_CHECK("k_vertex_adjoint");
  }
  k_point_gather<<<dim3((unsigned)ceil_div(g.nwords, kGatherTile)), dim3(kGatherThreads),
                   gather_lds_bytes(g.wpr), s>>>(
      g, (const uint64_t*)(ws + L.off_bits), (const uint32_t*)(ws + L.off_base_v),
      (const VertexAdj*)d_scratch, d_counts, n_points, d_dl_dfield);
  DAXHIP_LAUNCH_CHECK("k_point_gather");
  return DAX_HIP_OK;
}

}  // extern "C"
