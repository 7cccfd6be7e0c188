// Device code of the marching-cubes adjoint, shared by the single-volume
// and the batched backward.
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
//                      (sA, GA[3], sB, GB[3])
//   k_point_gather   : tiles of 256 row words.  A thread per word forms the
//                      mask of points within L1 distance 1 of an active-edge
//                      endpoint (bit-parallel, from the inside-bits); the
//                      other points get 0 with coalesced stores; the
//                      receiving points are compacted and each gets one
//                      thread, which sums over the active edges incident to
//                      it and to its 6 neighbours.
#pragma once

#include "mc_kernels.h"

namespace daxhip {
namespace mc {

// Per-vertex adjoint record: the lower endpoint's half and the upper
// endpoint's half, 16 bytes each.
struct alignas(16) VertexAdj {
  float sA, gA[3], sB, gB[3];
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

// Adjoint record of vertex u (global index into edges/dldx/dldn/out) of the
// volume g; the edge code is local to the volume.
__device__ __forceinline__ void vertex_adjoint(const VolGeom& g, const float* __restrict__ field,
                                               const uint32_t* __restrict__ edges,
                                               const float* __restrict__ dldx,
                                               const float* __restrict__ dldn, int64_t u,
                                               VertexAdj* __restrict__ out) {
  const float* f = field + g.field_off;
  const uint32_t e = edges[u];
  const int axis = (int)(e >> 30);
  const int64_t A = (int64_t)(e & 0x3fffffffu);
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  const int x = (int)(A % g.nx), y = (int)((A / g.nx) % g.ny), z = (int)(A / sz);
  // an edge code outside the grid (edge list not written by a forward that
  // fit its outputs) contributes nothing and is never dereferenced
  if (axis > 2 || A >= sz * g.nz || x + (axis == 0) >= g.nx || y + (axis == 1) >= g.ny ||
      z + (axis == 2) >= g.nz) {
    out[u] = VertexAdj{0.0f, {0.0f, 0.0f, 0.0f}, 0.0f, {0.0f, 0.0f, 0.0f}};
    return;
  }
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

// ---- per-point gather ------------------------------------------------------
// A point receives a non-zero dL/df only within L1 distance 1 of an endpoint
// of an active edge (its own t-terms and the central-difference stencils of
// the endpoint gradients).  The gather of one point reads the inside-bits and
// vertex bases of the volume from global memory (they stay in L2 / MALL).
struct GatherCtx {
  const VolGeom* g;
  const uint64_t* bits;    // volume's inside-bits
  const uint32_t* base_v;  // volume's word bases (global vertex ids)
  const VertexAdj* adj;
  int64_t n_adj;           // records in adj
  int x, j, k;
};

__device__ __forceinline__ uint64_t row_word(const GatherCtx& c, int qq, int jj, int kk) {
  const VolGeom& g = *c.g;
  if (qq < 0 || qq >= g.wpr || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) return 0ull;
  return c.bits[(int64_t)qq + (int64_t)g.wpr * (jj + (int64_t)g.ny * kk)];
}

// Does point (x,j,k) + (dx, dy, dz) exist?
__device__ __forceinline__ bool exists(const GatherCtx& c, int dx, int dy, int dz) {
  const int x = c.x + dx, y = c.j + dy, z = c.k + dz;
  return x >= 0 && x < c.g->nx && y >= 0 && y < c.g->ny && z >= 0 && z < c.g->nz;
}

__device__ __forceinline__ int bit_at(const GatherCtx& c, int dx, int dy, int dz) {
  const int x = c.x + dx;
  return (int)((row_word(c, x >> 6, c.j + dy, c.k + dz) >> (x & 63)) & 1ull);
}

// Vertex record of the edge with lower endpoint (x,j,k) + (dx,dy,dz) along
// `axis`, or nullptr if that edge is not active.
__device__ __forceinline__ const VertexAdj* edge_vertex(const GatherCtx& c, int dx, int dy, int dz,
                                                        int axis) {
  const int ex = axis == 0, ey = axis == 1, ez = axis == 2;
  if (!exists(c, dx, dy, dz) || !exists(c, dx + ex, dy + ey, dz + ez)) return nullptr;
  if (bit_at(c, dx, dy, dz) == bit_at(c, dx + ex, dy + ey, dz + ez)) return nullptr;
  // rank of the edge within its row word
  const VolGeom& g = *c.g;
  const int xg = c.x + dx;
  const int qq = xg >> 6, bb = xg & 63;
  WordPos p;
  p.q = qq;
  p.j = c.j + dy;
  p.k = c.k + dz;
  p.has_y = p.j + 1 < g.ny;
  p.has_z = p.k + 1 < g.nz;
  p.pmask = low_mask(g.nx - qq * 64);
  p.xmask = low_mask(g.nx - 1 - qq * 64);
  const uint64_t A = row_word(c, qq, p.j, p.k);
  const uint64_t A1 = row_word(c, qq + 1, p.j, p.k);
  const uint64_t B = row_word(c, qq, p.j + 1, p.k);
  const uint64_t C = row_word(c, qq, p.j, p.k + 1);
  uint64_t fx, fy, fz;
  edge_flags(p, A, A1, B, C, fx, fy, fz);
  const int64_t w = (int64_t)qq + (int64_t)g.wpr * (p.j + (int64_t)g.ny * p.k);
  const uint32_t vid = c.base_v[w] + (uint32_t)edge_rank(fx, fy, fz, bb, axis);
  return (int64_t)vid < c.n_adj ? c.adj + vid : nullptr;
}

// dL/dg_p (3 components) and the t-path sum S_p at p = (x,j,k) + (dx,dy,dz):
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

// dL/df of the point (x, j, k).
__device__ __forceinline__ float point_adjoint(const GatherCtx& c) {
  const VolGeom& g = *c.g;
  float G[3], S;
  point_sums(c, 0, 0, 0, G, S);
  float acc = S;
  const int co[3] = {c.x, c.j, c.k};
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
  return acc;
}

// Endpoint mask of row word (qq, jj, kk): points with an active incident
// edge (0 outside the grid and in grids without cells).
__device__ __forceinline__ uint64_t endpoint_mask(const VolGeom& g, const uint64_t* __restrict__ b,
                                                  int qq, int jj, int kk) {
  if (qq < 0 || qq >= g.wpr || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) return 0ull;
  if (!(g.nx > 1 && g.ny > 1 && g.nz > 1)) return 0ull;
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int64_t w = (int64_t)qq + (int64_t)g.wpr * jj + plane * kk;
  const uint64_t a = b[w];
  const uint64_t an = qq + 1 < g.wpr ? b[w + 1] : 0ull;
  const int x0 = qq * 64;
  const uint64_t fx = (a ^ shr1(a, an)) & low_mask(g.nx - 1 - x0);  // x -> x+1
  uint64_t e = fx | (fx << 1);
  if (qq > 0) e |= ((b[w - 1] >> 63) ^ a) & 1ull;  // edge entering bit 0
  if (jj + 1 < g.ny) e |= a ^ b[w + g.wpr];
  if (jj >= 1) e |= a ^ b[w - g.wpr];
  if (kk + 1 < g.nz) e |= a ^ b[w + plane];
  if (kk >= 1) e |= a ^ b[w - plane];
  return e & low_mask(g.nx - x0);
}

// Points of word p within L1 distance 1 of an endpoint.
__device__ __forceinline__ uint64_t receiving_mask(const VolGeom& g, const uint64_t* __restrict__ b,
                                                   const WordPos& p) {
  const uint64_t e = endpoint_mask(g, b, p.q, p.j, p.k);
  uint64_t r = e | (e << 1) | (e >> 1);
  r |= endpoint_mask(g, b, p.q - 1, p.j, p.k) >> 63;
  r |= endpoint_mask(g, b, p.q + 1, p.j, p.k) << 63;
  r |= endpoint_mask(g, b, p.q, p.j - 1, p.k) | endpoint_mask(g, b, p.q, p.j + 1, p.k);
  r |= endpoint_mask(g, b, p.q, p.j, p.k - 1) | endpoint_mask(g, b, p.q, p.j, p.k + 1);
  return r & low_mask(g.nx - p.q * 64);
}

// LDS of one gather tile of up to kGatherTile words.
constexpr int kGatherTile = 256;
struct GatherShared {
  uint64_t recv[kGatherTile];
  uint32_t first[kGatherTile + 1];  // exclusive scan of popcount(recv)
  uint32_t scan[16];
};

// dL/df of the points of words [w0, w_end) of a volume (w_end - w0 <=
// kGatherTile, blockDim.x == kGatherTile): receiving masks (thread per word),
// zeros for the other points (a wave per word, coalesced), then one thread per
// receiving point.  Every thread of the block calls this.
__device__ __forceinline__ void gather_tile(const VolGeom& g, const uint64_t* __restrict__ b,
                                            const uint32_t* __restrict__ base_v,
                                            const VertexAdj* __restrict__ adj, int64_t n_adj,
                                            float* __restrict__ out, int64_t w0, int64_t w_end,
                                            GatherShared& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nw = (int)(w_end - w0);
  uint64_t r = 0ull;
  if (tid < nw) r = receiving_mask(g, b, word_pos(g, w0 + tid));
  sh.recv[tid] = r;
  uint32_t total;
  const uint32_t off = block_excl_scan((uint32_t)popc(r), total, sh.scan);
  if (tid < nw) sh.first[tid] = off;
  if (tid == 0) sh.first[nw] = total;
  __syncthreads();
  for (int i = wv; i < nw; i += kGatherTile / 64) {
    const WordPos p = word_pos(g, w0 + i);
    const int x = p.q * 64 + lane;
    if (x < g.nx && !((sh.recv[i] >> lane) & 1ull))
      out[(int64_t)x + (int64_t)g.nx * (p.j + (int64_t)g.ny * p.k)] = 0.0f;
  }
  for (uint32_t e = (uint32_t)tid; e < total; e += kGatherTile) {
    const int i = find_word(sh.first, nw, e);
    const int bit = select_bit(sh.recv[i], (int)(e - sh.first[i]));
    const WordPos p = word_pos(g, w0 + i);
    const GatherCtx c{&g, b, base_v, adj, n_adj, p.q * 64 + bit, p.j, p.k};
    out[(int64_t)c.x + (int64_t)g.nx * (p.j + (int64_t)g.ny * p.k)] = point_adjoint(c);
  }
}

}  // namespace mc
}  // namespace daxhip
