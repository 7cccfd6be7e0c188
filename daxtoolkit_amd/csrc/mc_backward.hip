// Adjoint of the marching-cubes forward (no reference counterpart).
//
// Forward, per output vertex u on the active edge (A, B = A + e_axis):
//   t   = (iso - fA) / (fB - fA)
//   x_u = cA + t (cB - cA)                       (Lerp of the corner coordinates)
//   m   = gA + t (gB - gA),  n_u = m / |m|       (g: central-difference gradient)
// Topology (which edges are active) is held piecewise constant.  Given
// dL/dx_u and dL/dn_u the adjoint is
//   dL/dm = (dL/dn - n (n . dL/dn)) / |m|
//   dL/dt = dL/dx . (cB - cA) + dL/dm . (gB - gA)
//   dL/dfA += dL/dt * (-(1-t)/(fB-fA)),   dL/dfB += dL/dt * (-t/(fB-fA))
//   dL/dgA  = (1-t) dL/dm,                dL/dgB  = t dL/dm
// and each dL/dg_p[d] flows to the stencil points of g_p[d].
//
// Two kernels, deterministic (no atomics; every sum has a fixed order):
//   k_vertex_adjoint : one thread per vertex, recomputes t, gA, gB from the
//                      field and the compact edge list (the interpolants are
//                      not stored by the forward), writes 8 floats per vertex
//                      (sA, sB, GA[3], GB[3])
//   k_point_gather   : one wavefront per row word, lane = point q.  Lanes
//                      farther than L1 distance 2 from any inside/outside
//                      change write 0; the others sum over the active edges
//                      incident to q and to its 6 neighbours.
#include "mc_common.h"

namespace daxhip {
namespace mc {

struct VertexAdj {
  float sA, sB, gA[3], gB[3];
};
static_assert(sizeof(VertexAdj) == 32, "32-byte vertex records");

__device__ __forceinline__ void grad_at(const VolGeom& g, const float* f, int x, int y, int z,
                                        float out[3]) {
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  const int64_t p = x + sy * y + sz * z;
  const int c[3] = {x, y, z};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int64_t st[3] = {1, sy, sz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float h = g.spacing[d];
    if (n[d] < 2)
      out[d] = 0.0f;
    else if (c[d] == 0)
      out[d] = (f[p + st[d]] - f[p]) / h;
    else if (c[d] == n[d] - 1)
      out[d] = (f[p] - f[p - st[d]]) / h;
    else
      out[d] = (f[p + st[d]] - f[p - st[d]]) / (2.0f * h);
  }
}

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
  const float* f = field + g.field_off;
  const uint32_t e = edges[u];
  const int axis = (int)(e >> 30);
  const int64_t A = (int64_t)(e & 0x3fffffffu);
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  const int x = (int)(A % g.nx), y = (int)((A / g.nx) % g.ny), z = (int)(A / sz);
  const int64_t B = A + (axis == 0 ? 1 : (axis == 1 ? sy : sz));
  const float fA = f[A], fB = f[B];
  const float den = fB - fA;
  const float t = (g.iso - fA) / den;
  float gA[3], gB[3];
  grad_at(g, f, x, y, z, gA);
  grad_at(g, f, x + (axis == 0), y + (axis == 1), z + (axis == 2), gB);
  float m[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) m[d] = gA[d] + t * (gB[d] - gA[d]);
  const float len2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
  float dn[3] = {dldn[3 * u + 0], dldn[3 * u + 1], dldn[3 * u + 2]};
  float dm[3] = {0.0f, 0.0f, 0.0f};
  if (len2 > 0.0f) {
    const float len = sqrtf(len2);
    float nv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) nv[d] = m[d] / len;
    const float dot = nv[0] * dn[0] + nv[1] * dn[1] + nv[2] * dn[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) dm[d] = (dn[d] - nv[d] * dot) / len;
  }
  const int ca[3] = {x, y, z};
  float dldt = 0.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float a = g.origin[d] + g.spacing[d] * (float)ca[d];
    const float b = g.origin[d] + g.spacing[d] * (float)(ca[d] + (d == axis ? 1 : 0));
    dldt += dldx[3 * u + d] * (b - a);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) dldt += dm[d] * (gB[d] - gA[d]);
  VertexAdj r;
  r.sA = dldt * (-(1.0f - t) / den);
  r.sB = dldt * (-t / den);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.gA[d] = (1.0f - t) * dm[d];
    r.gB[d] = t * dm[d];
  }
  out[u] = r;
}

// Per-wave staging of the 5x5 rows around the word (dy, dz in -2..2) and the
// words q-1, q, q+1 of each: bits of every point within L1/Linf distance 2.
struct Rows {
  uint64_t w[5][5][3];  // [dz+2][dy+2][dq+1]
};

struct GatherCtx {
  const VolGeom* g;
  const Rows* r;
  const uint32_t* base_v;  // volume's words
  const VertexAdj* adj;
  int q, j, k, lane;
};

// Does point q + (dx, dy, dz) exist?
__device__ __forceinline__ bool exists(const GatherCtx& c, int dx, int dy, int dz) {
  const int x = c.q * 64 + c.lane + dx, y = c.j + dy, z = c.k + dz;
  return x >= 0 && x < c.g->nx && y >= 0 && y < c.g->ny && z >= 0 && z < c.g->nz;
}

__device__ __forceinline__ int bit_at(const GatherCtx& c, int dx, int dy, int dz) {
  const int xl = c.lane + dx;  // -2 .. 65
  const int wi = (xl + 64) >> 6;  // 0, 1, 2
  return (int)((c.r->w[dz + 2][dy + 2][wi] >> (xl & 63)) & 1ull);
}

// Vertex record of the edge with lower endpoint q + (dx,dy,dz) along `axis`,
// or nullptr if that edge is not active.
__device__ __forceinline__ const VertexAdj* edge_vertex(const GatherCtx& c, int dx, int dy, int dz,
                                                        int axis) {
  const int ex = axis == 0, ey = axis == 1, ez = axis == 2;
  if (!exists(c, dx, dy, dz) || !exists(c, dx + ex, dy + ey, dz + ez)) return nullptr;
  if (bit_at(c, dx, dy, dz) == bit_at(c, dx + ex, dy + ey, dz + ez)) return nullptr;
  // rank of the edge within its row word
  const VolGeom& g = *c.g;
  const int xg = c.q * 64 + c.lane + dx;
  const int qq = xg >> 6, bb = xg & 63;
  const int wi = qq - c.q + 1;  // 0..2
  WordPos p;
  p.q = qq;
  p.j = c.j + dy;
  p.k = c.k + dz;
  p.has_y = p.j + 1 < g.ny;
  p.has_z = p.k + 1 < g.nz;
  p.pmask = low_mask(g.nx - qq * 64);
  p.xmask = low_mask(g.nx - 1 - qq * 64);
  const uint64_t A = c.r->w[dz + 2][dy + 2][wi];
  const uint64_t A1 = (wi < 2 && qq + 1 < g.wpr) ? c.r->w[dz + 2][dy + 2][wi + 1] : 0ull;
  const uint64_t B = c.r->w[dz + 2][dy + 3][wi];
  const uint64_t C = c.r->w[dz + 3][dy + 2][wi];
  uint64_t fx, fy, fz;
  edge_flags(p, A, A1, B, C, fx, fy, fz);
  const int64_t w = (int64_t)qq + (int64_t)g.wpr * (p.j + (int64_t)g.ny * p.k);
  const uint32_t vid = c.base_v[w] + (uint32_t)edge_rank(fx, fy, fz, bb, axis);
  return c.adj + vid;
}

// dL/dg_p (3 components) and the t-path sum S_p at p = q + (dx,dy,dz):
// over the 3 edges leaving p and the 3 edges entering p.
__device__ __forceinline__ void point_sums(const GatherCtx& c, int dx, int dy, int dz, float G[3],
                                           float& S) {
  G[0] = G[1] = G[2] = 0.0f;
  S = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const VertexAdj* v = edge_vertex(c, dx, dy, dz, a);
    if (v) {
      S += v->sA;
      G[0] += v->gA[0];
      G[1] += v->gA[1];
      G[2] += v->gA[2];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const VertexAdj* v =
        edge_vertex(c, dx - (a == 0), dy - (a == 1), dz - (a == 2), a);
    if (v) {
      S += v->sB;
      G[0] += v->gB[0];
      G[1] += v->gB[1];
      G[2] += v->gB[2];
    }
  }
}

__global__ __launch_bounds__(256) void k_point_gather(VolGeom g, const uint64_t* __restrict__ bits,
                                                      const uint32_t* __restrict__ base_v,
                                                      const VertexAdj* __restrict__ adj,
                                                      float* __restrict__ dldf) {
  __shared__ Rows s_rows[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const bool has_word = w < g.nwords;
  const WordPos p = word_pos(g, has_word ? w : 0);
  const uint64_t* b = bits + g.word_off;
  Rows& r = s_rows[wv];
  // 75 words, one per lane (lanes 0..74 -> two passes)
  if (has_word) {
    for (int i = lane; i < 75; i += 64) {
      const int dq = i % 3 - 1, dy = (i / 3) % 5 - 2, dz = i / 15 - 2;
      const int qq = p.q + dq, jj = p.j + dy, kk = p.k + dz;
      uint64_t v = 0ull;
      if (qq >= 0 && qq < g.wpr && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz)
        v = b[(int64_t)qq + (int64_t)g.wpr * (jj + (int64_t)g.ny * kk)];
      r.w[dz + 2][dy + 2][dq + 1] = v;
    }
  }
  __syncthreads();  // every wave of the block reaches this barrier
  if (!has_word) return;

  // lanes whose L1 ball of radius 2 holds both inside and outside points
  uint64_t any = 0ull, all = ~0ull;
#pragma unroll
  for (int dz = -2; dz <= 2; ++dz)
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
      const int rad = 2 - (dy < 0 ? -dy : dy) - (dz < 0 ? -dz : dz);
      if (rad < 0) continue;
      const int jj = p.j + dy, kk = p.k + dz;
      if (jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
      const uint64_t m0 = r.w[dz + 2][dy + 2][0], m1 = r.w[dz + 2][dy + 2][1],
                     m2 = r.w[dz + 2][dy + 2][2];
#pragma unroll
      for (int dx = -2; dx <= 2; ++dx) {
        if ((dx < 0 ? -dx : dx) > rad) continue;
        uint64_t v;
        if (dx == 0)
          v = m1;
        else if (dx > 0)
          v = (m1 >> dx) | (m2 << (64 - dx));
        else
          v = (m1 << (-dx)) | (m0 >> (64 + dx));
        any |= v;
        all &= v;
      }
    }
  const bool has_cells = g.nx > 1 && g.ny > 1 && g.nz > 1;  // else no vertices at all
  const uint64_t near = has_cells ? (any & ~all & p.pmask) : 0ull;
  const int x = p.q * 64 + lane;
  if (x >= g.nx) return;
  const int64_t pid = (int64_t)x + (int64_t)g.nx * (p.j + (int64_t)g.ny * p.k);
  float* out = dldf + g.field_off;
  if (!((near >> lane) & 1ull)) {
    out[pid] = 0.0f;
    return;
  }
  GatherCtx c{&g, &r, base_v + g.word_off, adj, p.q, p.j, p.k, lane};
  float G[3], S;
  point_sums(c, 0, 0, 0, G, S);
  float acc = S;
  const int co[3] = {x, p.j, p.k};
  const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float h = g.spacing[d];
    if (n[d] < 2) continue;
    // f_q in g_q[d] (one-sided stencils on the first and last plane)
    if (co[d] == 0)
      acc += G[d] * (-1.0f / h);
    else if (co[d] == n[d] - 1)
      acc += G[d] * (1.0f / h);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float h = g.spacing[d];
    if (n[d] < 2) continue;
    const int ex = d == 0, ey = d == 1, ez = d == 2;
    float Gn[3], Sn;
    if (co[d] >= 1) {  // p = q - e_d: f_q = f(p + e_d)
      point_sums(c, -ex, -ey, -ez, Gn, Sn);
      acc += (co[d] - 1 == 0) ? Gn[d] / h : Gn[d] / (2.0f * h);
    }
    if (co[d] <= n[d] - 2) {  // p = q + e_d: f_q = f(p - e_d)
      point_sums(c, ex, ey, ez, Gn, Sn);
      acc += (co[d] + 1 == n[d] - 1) ? -(Gn[d] / h) : -(Gn[d] / (2.0f * h));
    }
  }
  out[pid] = acc;
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
  if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz > (1LL << 30) || (nx + 63) / 64 > kMaxWpr) {
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
  k_point_gather<<<dim3((unsigned)ceil_div(g.nwords * 64, 256)), dim3(256), 0, s>>>(
      g, (const uint64_t*)(ws + L.off_bits), (const uint32_t*)(ws + L.off_base_v),
      (const VertexAdj*)d_scratch, d_dl_dfield);
  DAXHIP_LAUNCH_CHECK("k_point_gather");
  return DAX_HIP_OK;
}

}  // extern "C"
