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
  // branch-free loads (point_gradient): the six values of a gradient, and
  // both gradients of a vertex, are in flight together
  point_gradient(g, FieldArray{f}, x, y, z, out);
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
  const uint32_t A32 = (uint32_t)A, row = fast_div(A32, g.div_nx);  // A < 2^30
  const uint32_t zz = fast_div(row, g.div_ny);
  const int x = (int)(A32 - row * (uint32_t)g.nx), y = (int)(row - zz * (uint32_t)g.ny),
            z = (int)zz;
  // an edge code outside the grid (edge list not written by a forward that
  // fit its outputs) contributes nothing and is never dereferenced
  if (axis > 2 || A >= sz * g.nz || x + (axis == 0) >= g.nx || y + (axis == 1) >= g.ny ||
      z + (axis == 2) >= g.nz) {
    out[u] = VertexAdj{0.0f, {0.0f, 0.0f, 0.0f}, 0.0f, {0.0f, 0.0f, 0.0f}};
    return;
  }
  const int64_t B = A + (axis == 0 ? 1 : (axis == 1 ? sy : sz));
  // the upstream gradients do not depend on the field: load them first
  const float dn[3] = {dldn[3 * u + 0], dldn[3 * u + 1], dldn[3 * u + 2]};
  const float dx[3] = {dldx[3 * u + 0], dldx[3 * u + 1], dldx[3 * u + 2]};
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
    dldt += dx[d] * (b - a);
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
// the endpoint gradients).  A gather tile of up to kGatherTile consecutive row
// words stages, for the planes k-2..k+1 around it, the active-edge flags
// (fx, fy, fz: the same flags the forward numbers its vertices with) and the
// vertex bases of the words within two rows (and one word) of the tile:
// segment dz holds words [w0 + dz*plane - 2*wpr - 1, w_end + dz*plane +
// 2*wpr + 1).  Every lower endpoint of an edge a tile point reads lies at
// |dq| <= 1, -2 <= dy, dz <= 1 from its own word, inside its segment.
constexpr int kGatherTile = kGatherTileWords;  // row words per tile
constexpr int kGatherThreads = 512;  // threads per tile
// Endpoint records a tile keeps (4 per thread, held in registers until the
// flags are no longer needed, then written over them).
constexpr int kRecPerThread = 4;
constexpr int kMaxEndpoints = kRecPerThread * kGatherThreads;

__host__ __device__ constexpr int gather_seg(int wpr) { return kGatherTile + 4 * wpr + 2; }
// G-words: the tile's words with one row above and below in its plane, and
// the tile's words in the planes below and above.
__host__ __device__ constexpr int gather_gwords(int nw, int wpr) { return 3 * nw + 2 * wpr; }
// Dynamic LDS of a gather tile: the staged flags and vertex bases (later the
// endpoint records), the G-word endpoint masks and their exclusive scan, and
// the G-word of each endpoint slot.
__host__ __device__ constexpr size_t gather_flags_bytes(int wpr) {
  const size_t flags = (size_t)4 * gather_seg(wpr) * (3 * sizeof(uint64_t) + sizeof(uint32_t));
  const size_t recs = (size_t)kMaxEndpoints * 16;
  return flags > recs ? flags : recs;
}
__host__ __device__ constexpr size_t gather_lds_bytes(int wpr) {
  const size_t ng = (size_t)gather_gwords(kGatherTile, wpr);
  return gather_flags_bytes(wpr) + ng * sizeof(uint64_t) + (ng + 1) * sizeof(uint32_t) +
         (size_t)kMaxEndpoints * sizeof(uint16_t);
}

// Part `part` (0..3) of the test whether the tile at word w0 stages any
// active edge: segment dz = part - 2 holds one (the vertex bases count the
// active edges of a word range).  A tile is quiet when no part holds one.
__device__ __forceinline__ bool gather_segment_active(const VolGeom& g,
                                                      const uint32_t* __restrict__ base_v,
                                                      int64_t w0, int part) {
  const int64_t plane = (int64_t)g.wpr * g.ny;
  const int64_t ws0 = w0 + (part - 2) * plane - 2 * g.wpr - 1;
  const int64_t ws = max(ws0, (int64_t)0);
  const int64_t we = min(ws0 + gather_seg(g.wpr), g.nwords);
  return we > ws && base_v[we] != base_v[ws];
}

struct Stage {
  const uint64_t* fx;  // LDS [4][seg] each
  const uint64_t* fy;
  const uint64_t* fz;
  const uint32_t* bv;
  int seg;
  int64_t w0;          // tile's first word
};

// A point's view of the staged words: its word at (q, j, k) lies dzp planes
// from the tile's plane, at tile-plane word w (its own word minus dzp*plane).
struct GatherCtx {
  const VolGeom* g;
  const Stage* st;
  const VertexAdj* adj;
  int64_t n_adj;  // records in adj
  int64_t w;
  int dzp;
  int q, x, j, k;

  // LDS index of word (q+dq, j+dy, k+dz); dzp + dz in -2..1
  __device__ __forceinline__ int idx(int dq, int dy, int dz) const {
    return (dzp + dz + 2) * st->seg + (int)(w - st->w0) + dq + dy * g->wpr + 2 * g->wpr + 1;
  }
};

// Vertex id of the edge with lower endpoint (x,j,k) + (dx,dy,dz) along
// `axis`, or -1 if that edge is not active (LDS only).
__device__ __forceinline__ int64_t edge_vid(const GatherCtx& c, int dx, int dy, int dz, int axis) {
  const VolGeom& g = *c.g;
  const int xl = c.x + dx, jl = c.j + dy, kl = c.k + dz;
  if (xl < 0 || xl >= g.nx || jl < 0 || jl >= g.ny || kl < 0 || kl >= g.nz) return -1;
  const int qq = xl >> 6, bb = xl & 63;
  const int i = c.idx(qq - c.q, dy, dz);
  const uint64_t fa = (axis == 0 ? c.st->fx : (axis == 1 ? c.st->fy : c.st->fz))[i];
  if (!((fa >> bb) & 1ull)) return -1;
  const uint64_t fx = axis == 0 ? fa : c.st->fx[i];
  const uint64_t fy = axis == 1 ? fa : c.st->fy[i];
  const uint64_t fz = axis == 2 ? fa : c.st->fz[i];
  const int64_t vid = (int64_t)c.st->bv[i] + edge_rank(fx, fy, fz, bb, axis);
  return vid < c.n_adj ? vid : -1;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// (S_p, dL/dg_p) at p = (x,j,k) + (dx,dy,dz): over the 3 edges leaving p
// (lower halves of their records) and the 3 edges entering p (upper halves),
// summed in that order.
__device__ __forceinline__ f32x4 point_sums(const GatherCtx& c, int dx, int dy, int dz) {
  int64_t vid[6];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    vid[a] = edge_vid(c, dx, dy, dz, a);
    vid[3 + a] = edge_vid(c, dx - (a == 0), dy - (a == 1), dz - (a == 2), a);
  }
  const f32x4* rec = reinterpret_cast<const f32x4*>(c.adj);
  f32x4 r[6];
  // unconditional loads (an inactive edge reads record 0 and is dropped
  // below), so the six are in flight together instead of one conditional
  // load and wait at a time; every record index is below n_adj
  if (c.n_adj > 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) r[i] = rec[vid[i] >= 0 ? 2 * vid[i] + (i < 3 ? 0 : 1) : 0];
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) r[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  }
  f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};  // S, G[0..2]
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (vid[i] >= 0) {
      acc.x += r[i].x;
      acc.y += r[i].y;
      acc.z += r[i].z;
      acc.w += r[i].w;
    }
  }
  return acc;
}

// dL/df of the point q = (x, j, k) from (S, G) of q and of its 6 neighbours:
// at(d, s) returns (S_p, G_p) of p = q + s e_d (s = 0: q itself).
template <class At>
__device__ __forceinline__ float combine_point(const VolGeom& g, int x, int j, int k, At at) {
  const f32x4 own = at(0, 0);
  const float G[3] = {own.y, own.z, own.w};
  float acc = own.x;
  const int co[3] = {x, j, k};
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
    if (co[d] >= 1) {  // p = q - e_d: f_q = f(p + e_d)
      const f32x4 r = at(d, -1);
      const float Gd = d == 0 ? r.y : (d == 1 ? r.z : r.w);
      acc += (co[d] - 1 == 0) ? Gd / h : Gd / (2.0f * h);
    }
    if (co[d] <= n[d] - 2) {  // p = q + e_d: f_q = f(p - e_d)
      const f32x4 r = at(d, 1);
      const float Gd = d == 0 ? r.y : (d == 1 ? r.z : r.w);
      acc += (co[d] + 1 == n[d] - 1) ? -(Gd / h) : -(Gd / (2.0f * h));
    }
  }
  return acc;
}

// dL/df of the point (x, j, k), every sum from the staged flags.
__device__ __forceinline__ float point_adjoint(const GatherCtx& c) {
  return combine_point(*c.g, c.x, c.j, c.k, [&](int d, int s) {
    return point_sums(c, d == 0 ? s : 0, d == 1 ? s : 0, d == 2 ? s : 0);
  });
}

// Endpoint mask of row word (q+dq, j+dy, k+dz) of the context's word
// (dzp + dz in -1..1): points with an active incident edge, from the staged
// flags (0 outside the grid; the flags are 0 in grids without cells).
__device__ __forceinline__ uint64_t endpoint_mask(const GatherCtx& c, int dq, int dy, int dz) {
  const VolGeom& g = *c.g;
  const int qq = c.q + dq, jj = c.j + dy, kk = c.k + dz;
  if (qq < 0 || qq >= g.wpr || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) return 0ull;
  const Stage& st = *c.st;
  const int i = c.idx(dq, dy, dz);
  const uint64_t fx = st.fx[i];
  uint64_t e = fx | (fx << 1) | st.fy[i] | st.fz[i];
  if (qq > 0) e |= st.fx[i - 1] >> 63;  // x-edge entering bit 0
  if (jj > 0) e |= st.fy[i - g.wpr];     // y-edges entering the row
  if (kk > 0) e |= st.fz[i - st.seg];    // z-edges entering the row
  return e;
}

// Points of the context's word within L1 distance 1 of an endpoint.
__device__ __forceinline__ uint64_t receiving_mask(const GatherCtx& c) {
  const uint64_t e = endpoint_mask(c, 0, 0, 0);
  uint64_t r = e | (e << 1) | (e >> 1);
  r |= endpoint_mask(c, -1, 0, 0) >> 63;
  r |= endpoint_mask(c, 1, 0, 0) << 63;
  r |= endpoint_mask(c, 0, -1, 0) | endpoint_mask(c, 0, 1, 0);
  r |= endpoint_mask(c, 0, 0, -1) | endpoint_mask(c, 0, 0, 1);
  return r & low_mask(c.g->nx - c.q * 64);
}

// Zeros over the points of row words [w0, w_end) of a volume: one
// contiguous run of the field (16-byte stores between scalar head and tail).
__device__ __forceinline__ void zero_run(const VolGeom& g, float* __restrict__ out, int64_t w0,
                                         int64_t w_end, int tid, int nt) {
  const WordPos pa = word_pos(g, w0), pb = word_pos(g, w_end - 1);
  const int64_t p0 = (int64_t)pa.q * 64 + (int64_t)g.nx * (pa.j + (int64_t)g.ny * pa.k);
  const int64_t p1 = min((int64_t)pb.q * 64 + 64, (int64_t)g.nx) +
                     (int64_t)g.nx * (pb.j + (int64_t)g.ny * pb.k);
  float* o = out + p0;
  const int64_t n = p1 - p0;
  const int64_t head = min(n, (int64_t)(((16 - (reinterpret_cast<uintptr_t>(o) & 15)) & 15) / 4));
  for (int64_t t = tid; t < head; t += nt) o[t] = 0.0f;
  f32x4* o4 = reinterpret_cast<f32x4*>(o + head);
  const int64_t n4 = (n - head) >> 2;
  for (int64_t t = tid; t < n4; t += nt) o4[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t t = head + 4 * n4 + tid; t < n; t += nt) o[t] = 0.0f;
}

struct GatherShared {
  uint64_t recv[kGatherTile];
  uint32_t first[kGatherTile + 1];  // exclusive scan of popcount(recv)
  uint32_t scan[16];
};

// Widest rows the gather stages (nx <= 15872): its LDS stays within a CU's
// 160 KiB.
constexpr int kMaxGatherWpr = 248;
static_assert(gather_lds_bytes(kMaxGatherWpr) + sizeof(GatherShared) <= 160 * 1024,
              "gather tile LDS");

// Staging of one flag slot: the word's inside-bits and those it needs.
struct SlotLoad {
  uint64_t A, A1, B, C;
  uint32_t bv;
  WordPos p;
  bool in;
};

// dL/df of the points of row words [w0, w_end) of a volume (a span of at
// most kGatherTile words, blockDim.x == kGatherThreads).  Every thread of
// the block calls this; `lds` holds gather_lds_bytes(g.wpr).
//   1. stage flags and vertex bases of the planes k-2..k+1
//   2. receiving masks of the span's words (thread per word)
//   3. endpoint masks of the G-words, restricted to the points some
//      receiving point's stencil reads; block scan -> compact endpoint slots
//   4. (S, G) of every such endpoint (thread per endpoint, 6 edge records)
//   5. zeros for the other points (coalesced) and dL/df of every receiving
//      point from the 7 (S, G) of its stencil; no barrier follows, so these
//      stores drain while the next tiles run
// A span with more than kMaxEndpoints endpoints writes nothing and returns
// its endpoint count (kDirect false), or computes every receiving point's
// stencil sums directly from the flags instead of 3-4 (kDirect true).
// Returns 0 when its points are written.
template <bool kDirect>
__device__ __forceinline__ uint32_t gather_span(const VolGeom& g, const uint64_t* __restrict__ b,
                                                const uint32_t* __restrict__ base_v,
                                                const VertexAdj* __restrict__ adj, int64_t n_adj,
                                                float* __restrict__ out, int64_t w0,
                                                int64_t w_end, GatherShared& sh, void* lds) {
  // tid is opaque to the compiler, so what is derived from it is not hoisted
  // out of a caller's loop over spans (it would stay live across the loop,
  // about 30 more VGPRs)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wv = tid >> 6;
  const int nw = (int)(w_end - w0);
  const int wpr = g.wpr;
  const int seg = nw + 4 * wpr + 2;  // staged words per plane
  const int64_t plane = (int64_t)wpr * g.ny;

  // 1. flags and vertex bases: every load of a round is issued before any
  //    is used
  uint64_t* s_fx = reinterpret_cast<uint64_t*>(lds);
  uint64_t* s_fy = s_fx + 4 * seg;
  uint64_t* s_fz = s_fy + 4 * seg;
  uint32_t* s_bv = reinterpret_cast<uint32_t*>(s_fz + 4 * seg);
  const int ng = gather_gwords(nw, wpr);
  uint64_t* s_gm =
      reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(lds) + gather_flags_bytes(wpr));
  uint32_t* s_gp = reinterpret_cast<uint32_t*>(s_gm + gather_gwords(kGatherTile, wpr));
  uint16_t* s_eg = reinterpret_cast<uint16_t*>(s_gp + gather_gwords(kGatherTile, wpr) + 1);
  constexpr int kSlots = 3;
  for (int base = 0; base < 4 * seg; base += kSlots * kGatherThreads) {
    SlotLoad L[kSlots];
#pragma unroll
    for (int u = 0; u < kSlots; ++u) {
      const int i = base + u * kGatherThreads + tid;
      const int dz = i / seg - 2;
      const int64_t w = w0 + dz * plane - 2 * wpr - 1 + (i - (dz + 2) * seg);
      L[u].in = i < 4 * seg && w >= 0 && w < g.nwords;
      L[u].A = L[u].A1 = L[u].B = L[u].C = 0ull;
      L[u].bv = 0u;
      if (L[u].in) {
        L[u].p = word_pos(g, w);
        L[u].A = b[w];
        if (L[u].p.q + 1 < wpr) L[u].A1 = b[w + 1];
        if (L[u].p.has_y) L[u].B = b[w + wpr];
        if (L[u].p.has_z) L[u].C = b[w + plane];
        L[u].bv = base_v[w];
      }
    }
#pragma unroll
    for (int u = 0; u < kSlots; ++u) {
      const int i = base + u * kGatherThreads + tid;
      if (i < 4 * seg) {
        uint64_t fx = 0ull, fy = 0ull, fz = 0ull;
        if (L[u].in) edge_flags(L[u].p, L[u].A, L[u].A1, L[u].B, L[u].C, fx, fy, fz);
        s_fx[i] = fx;
        s_fy[i] = fy;
        s_fz[i] = fz;
        s_bv[i] = L[u].bv;
      }
    }
  }
  __syncthreads();
  const Stage st{s_fx, s_fy, s_fz, s_bv, seg, w0};

  // 2. receiving masks
  uint64_t r = 0ull;
  if (tid < nw) {
    const WordPos p = word_pos(g, w0 + tid);
    const GatherCtx c{&g, &st, adj, n_adj, w0 + tid, 0, p.q, p.q * 64, p.j, p.k};
    r = receiving_mask(c);
    sh.recv[tid] = r;
  }
  uint32_t total;
  const uint32_t off = block_excl_scan((uint32_t)popc(r), total, sh.scan);
  if (tid < nw) sh.first[tid] = off;
  if (tid == 0) sh.first[nw] = total;
  __syncthreads();

  // 3. endpoint masks of the G-words: gi in [0, nw + 2 wpr) is word
  //    w0 - wpr + gi of the tile's plane; then nw words of the plane below,
  //    then nw of the plane above.
  auto recv_of = [&](int64_t w) -> uint64_t {  // receiving mask of a tile word, else 0
    return (w >= w0 && w < w_end) ? sh.recv[w - w0] : 0ull;
  };
  uint32_t n_end = 0;
  if (total > 0) {
    const int ch = (ng + kGatherThreads - 1) / kGatherThreads;
    uint32_t cnt = 0;
    for (int u = 0; u < ch; ++u) {
      const int gi = tid * ch + u;
      if (gi >= ng) break;
      int dzp;
      int64_t wt;  // tile-plane word
      if (gi < nw + 2 * wpr) {
        dzp = 0;
        wt = w0 - wpr + gi;
      } else if (gi < 2 * nw + 2 * wpr) {
        dzp = -1;
        wt = w0 + (gi - nw - 2 * wpr);
      } else {
        dzp = 1;
        wt = w0 + (gi - 2 * nw - 2 * wpr);
      }
      const int64_t wa = wt + dzp * plane;  // the word itself
      uint64_t m = 0ull;
      if (wa >= 0 && wa < g.nwords) {
        const WordPos p = word_pos(g, wa);
        uint64_t need;
        if (dzp == 0) {
          const uint64_t rw = recv_of(wa);
          need = rw | (rw << 1) | (rw >> 1);
          if (p.q > 0) need |= recv_of(wa - 1) >> 63;
          if (p.q + 1 < wpr) need |= recv_of(wa + 1) << 63;
          if (p.j > 0) need |= recv_of(wa - wpr);
          if (p.j + 1 < g.ny) need |= recv_of(wa + wpr);
        } else {
          need = recv_of(wt);
        }
        if (need) {
          const GatherCtx c{&g, &st, adj, n_adj, wt, dzp, p.q, p.q * 64, p.j, p.k};
          m = endpoint_mask(c, 0, 0, 0) & need;
        }
      }
      s_gm[gi] = m;
      cnt += (uint32_t)popc(m);
    }
    uint32_t run = block_excl_scan(cnt, n_end, sh.scan);
    for (int u = 0; u < ch; ++u) {
      const int gi = tid * ch + u;
      if (gi >= ng) break;
      s_gp[gi] = run;
      const uint32_t c = (uint32_t)popc(s_gm[gi]);
      for (uint32_t t = run; t < run + c && t < (uint32_t)kMaxEndpoints; ++t)
        s_eg[t] = (uint16_t)gi;  // G-word of endpoint slot t
      run += c;
    }
    if (tid == 0) s_gp[ng] = n_end;
    __syncthreads();
  }
  const bool shared_sums = n_end <= (uint32_t)kMaxEndpoints;  // block-uniform
  if (!kDirect && !shared_sums) return n_end;
  f32x4* s_rec = reinterpret_cast<f32x4*>(lds);
  if (total > 0 && shared_sums) {
    // 4. (S, G) per endpoint, kept in registers until the flags are done
    f32x4 rr[kRecPerThread];
#pragma unroll
    for (int u = 0; u < kRecPerThread; ++u) {
      const uint32_t e = (uint32_t)tid + (uint32_t)u * kGatherThreads;
      rr[u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      if (e < n_end) {
        const int gi = s_eg[e];
        const int bit = select_bit(s_gm[gi], (int)(e - s_gp[gi]));
        const int dzp = gi < nw + 2 * wpr ? 0 : (gi < 2 * nw + 2 * wpr ? -1 : 1);
        const int64_t wt = dzp == 0 ? w0 - wpr + gi
                                    : w0 + (gi - nw - 2 * wpr - (dzp == 1 ? nw : 0));
        const WordPos p = word_pos(g, wt + dzp * plane);
        const GatherCtx c{&g, &st, adj, n_adj, wt, dzp, p.q, p.q * 64 + bit, p.j, p.k};
        rr[u] = point_sums(c, 0, 0, 0);
      }
    }
    __syncthreads();  // the flags region now holds the endpoint records
#pragma unroll
    for (int u = 0; u < kRecPerThread; ++u) {
      const uint32_t e = (uint32_t)tid + (uint32_t)u * kGatherThreads;
      if (e < n_end) s_rec[e] = rr[u];
    }
    __syncthreads();
  }

  // 5. zeros for the points that receive nothing (four per lane where rows
  //    are 16-byte aligned: 16 lanes per word, 4 words per wave step) ...
  if ((g.nx & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0) {
    const int sub = lane >> 4, x4 = (lane & 15) * 4;
    for (int i0 = wv * 4; i0 < nw; i0 += kGatherThreads / 16) {
      const int i = i0 + sub;
      if (i >= nw) continue;
      const WordPos p = word_pos(g, w0 + i);
      const int x = p.q * 64 + x4;
      if (x >= g.nx) continue;
      const uint32_t rb = (uint32_t)(sh.recv[i] >> x4) & 0xfu;
      float* o = out + (int64_t)x + (int64_t)g.nx * (p.j + (int64_t)g.ny * p.k);
      if (rb == 0u) {
        *reinterpret_cast<f32x4*>(o) = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (!((rb >> t) & 1u)) o[t] = 0.0f;
      }
    }
  } else {
    for (int i = wv; i < nw; i += kGatherThreads / 64) {
      const WordPos p = word_pos(g, w0 + i);
      const int x = p.q * 64 + lane;
      if (x < g.nx && !((sh.recv[i] >> lane) & 1ull))
        out[(int64_t)x + (int64_t)g.nx * (p.j + (int64_t)g.ny * p.k)] = 0.0f;
    }
  }
  // ... and dL/df of the receiving points
  for (uint32_t e = (uint32_t)tid; e < total; e += kGatherThreads) {
    const int i = find_word(sh.first, nw, e);
    const int bit = select_bit(sh.recv[i], (int)(e - sh.first[i]));
    const WordPos p = word_pos(g, w0 + i);
    const int x = p.q * 64 + bit;
    float v = 0.0f;
    if (shared_sums) {
      const int g0 = i + wpr;  // G-word of the point's own word
      auto rec = [&](int gi, int bb) -> f32x4 {
        const uint64_t m = s_gm[gi];
        if (!((m >> bb) & 1ull)) return f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        return s_rec[s_gp[gi] + (uint32_t)popc(m & ((1ull << bb) - 1ull))];
      };
      v = combine_point(g, x, p.j, p.k, [&](int d, int s) -> f32x4 {
        if (s == 0) return rec(g0, bit);
        if (d == 0) {
          const int xb = bit + s;  // -1..64
          return rec(g0 + (xb < 0 ? -1 : (xb > 63 ? 1 : 0)), xb & 63);
        }
        if (d == 1) return rec(g0 + s * wpr, bit);
        return rec((s < 0 ? nw + 2 * wpr : 2 * nw + 2 * wpr) + i, bit);
      });
    } else if (kDirect) {
      const GatherCtx c{&g, &st, adj, n_adj, w0 + i, 0, p.q, x, p.j, p.k};
      v = point_adjoint(c);
    }
    out[(int64_t)x + (int64_t)g.nx * (p.j + (int64_t)g.ny * p.k)] = v;
  }
  return 0;
}

// Tiles whose endpoints exceed kMaxEndpoints are redone in spans of
// kGatherTile >> lg words, lg in 1..4 the smallest with at most
// kSpanEndpoints << lg endpoints (2..16 spans).  A list item is
// (tile << 8) | (lg << 4) | span.  With many tiles per CU, a tile with
// lg <= kTileSpanLg runs its spans itself, one after another, and lists
// only a span that still does not fit: its block then takes a few tiles'
// time among many rounds of tiles, and the span pass (a launch later, its
// spans staged again) is spared.  Otherwise, and for denser tiles, the tile
// lists all its spans, and a span pass runs the items.  The listed word ranges are disjoint and a
// point is an endpoint slot of at most 5 of them (its own, the ranges of
// the words one row before and after it and one plane below and above),
// so with U vertices (at most 2U endpoints) the slots of listed ranges sum
// to at most 10U.  A listed tile with n slots gives 2^lg < 2n / 1536 items
// and a listed span with n > 2048 slots one, so the items number fewer
// than 20U / 1536 < U / 32 + 16.
constexpr int kSpanEndpoints = 1536;
constexpr int kTileSpanLg = 2;
__host__ __device__ constexpr int64_t gather_list_capacity(int64_t n_points) {
  return n_points / 32 + 16;
}
// Scratch of the adjoint: the vertex records, then the list (count, items).
__host__ __device__ constexpr size_t gather_list_offset(int64_t n_points) {
  return ((size_t)(n_points > 0 ? n_points : 0) * sizeof(VertexAdj) + 15) & ~(size_t)15;
}
__host__ __device__ constexpr size_t adjoint_scratch_bytes(int64_t n_points) {
  return gather_list_offset(n_points) + (size_t)(1 + gather_list_capacity(n_points)) * 4;
}

__device__ __forceinline__ int gather_span_lg(uint32_t n_end) {
  int lg = 1;
  while (lg < 4 && (uint32_t)(kSpanEndpoints << lg) < n_end) ++lg;
  return lg;
}

// Appends items (tile, lg, s0 .. s0 + n - 1) (one thread of the block).
__device__ __forceinline__ void gather_list_push(uint32_t* __restrict__ list, int64_t cap,
                                                 uint32_t tile, int lg, uint32_t s0, uint32_t n) {
  const uint32_t i = atomicAdd(&list[0], n);
  for (uint32_t s = 0; s < n; ++s)
    if (i + s < cap) list[1 + i + s] = (tile << 8) | ((uint32_t)lg << 4) | (s0 + s);
}

// A gather tile [w0, w_end) with list id `tile`: zeros only when no active
// edge lies among its staged words, else the tile as one span, or (too many
// endpoints) as the spans above, run here when lg <= max_lg (0 or
// kTileSpanLg).  Without a list (no vertex records) a tile with active edges
// is zeroed.
__device__ __forceinline__ void gather_tile(const VolGeom& g, const uint64_t* __restrict__ b,
                                            const uint32_t* __restrict__ base_v,
                                            const VertexAdj* __restrict__ adj, int64_t n_adj,
                                            float* __restrict__ out, int64_t w0, int64_t w_end,
                                            GatherShared& sh, void* lds,
                                            uint32_t* __restrict__ list, int64_t cap,
                                            uint32_t tile, int max_lg) {
  const int tid = threadIdx.x;
  if (!__syncthreads_or(tid < 4 && gather_segment_active(g, base_v, w0, tid))) {
    zero_run(g, out, w0, w_end, tid, kGatherThreads);
    return;
  }
  // block-uniform counts are read as scalars, so the span bounds stay scalars
  const uint32_t n_end = __builtin_amdgcn_readfirstlane(
      gather_span<false>(g, b, base_v, adj, n_adj, out, w0, w_end, sh, lds));
  if (!n_end) return;
  if (!list) {
    zero_run(g, out, w0, w_end, tid, kGatherThreads);
    return;
  }
  const int lg = gather_span_lg(n_end);
  if (lg > max_lg) {
    if (tid == 0) gather_list_push(list, cap, tile, lg, 0u, 1u << lg);
    return;
  }
  const int len = kGatherTile >> lg;
  for (int s = 0; s < (1 << lg); ++s) {
    const int64_t a = w0 + (int64_t)s * len, e = min(a + len, w_end);
    if (a >= e) break;
    __syncthreads();  // LDS reused by this span
    const uint32_t n_span = __builtin_amdgcn_readfirstlane(
        gather_span<false>(g, b, base_v, adj, n_adj, out, a, e, sh, lds));
    if (n_span && tid == 0) gather_list_push(list, cap, tile, lg, (uint32_t)s, 1u);
  }
}

// The span pass splits every item into `parts` equal pieces, the largest
// power of two (up to 16) that keeps the pieces within 2 * half_blocks: a
// short list (a few dense tiles) then runs as many small spans, whose
// latency is shorter, and a long one as the items themselves.  The span
// kernels pass half their grid, so the pieces fit one round: a piece is a
// chain of dependent phases (about 11 us at 8 words), and a second round
// costs more than longer pieces save.  (Written this way, the span kernel
// compiles to 76 VGPRs, three blocks per CU; `n * parts * 2 <= blocks`
// gives 93.)
__device__ __forceinline__ int gather_item_parts(int64_t n_items, int64_t half_blocks) {
  int parts = 1;
  while (parts < 16 && n_items * parts * 2 <= 2 * half_blocks) parts *= 2;
  return parts;
}

// Word range [w0, w_end) of piece `part` of `parts` of a span item within
// the tile starting at word tile_w0 of a volume with nwords words (spans
// are at least 16 words, pieces at least one).
__device__ __forceinline__ void gather_item_span(uint32_t item, int part, int parts,
                                                 int64_t tile_w0, int64_t nwords, int64_t& w0,
                                                 int64_t& w_end) {
  const int lg = (int)((item >> 4) & 15u), s = (int)(item & 15u);
  const int len = kGatherTile >> lg, piece = max(len / parts, 1);
  w0 = tile_w0 + (int64_t)s * len + (int64_t)part * piece;
  w_end = min(min(w0 + piece, tile_w0 + (int64_t)(s + 1) * len), nwords);
}

}  // namespace mc
}  // namespace daxhip
