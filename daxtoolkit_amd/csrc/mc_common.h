// Marching-cubes data layout shared by the forward, backward and batched
// kernels.
//
// Inside-bits: one bit per grid point, bit = (f > iso) (strict, as
// GetHexahedronClassification, dax/worklet/MarchingCubes.h:36-48).  Bits are
// stored per x-row in 64-bit "row words": word q of row (j,k) holds points
// x = 64q .. 64q+63 (bit x%64), padding bits beyond nx are 0.  The words of a
// volume are in flat (k, j, q) order, so word index w = q + wpr*(j + ny*k)
// with wpr = ceil(nx/64) words per row.  A row word is the unit of every
// count: its active edges (points of the word, axes x<y<z) and its active
// cells (cells whose lowest corner is in the word).  The exclusive scan over
// words of those two counts gives, for every word, the first output vertex
// id and the first output triangle id -- the order of the reference's output
// (vertices sorted by (lower point id, upper point id), triangles by cell then
// visit index, DispatcherGenerateInterpolatedCells.h:180-290).
#pragma once

#include "common.h"
#include "mc_case_table.h"

namespace daxhip {
namespace mc {

constexpr int kTileWords = 1024;     // row words per scan tile
constexpr int kCountTile = 128;      // row words per count tile: its prefixes fit 16 bits
constexpr int kScanThreads = 256;    // threads of the count/scan and generate blocks
constexpr int kWordsPerThread = kTileWords / kScanThreads;
constexpr int kMaxWpr = 256;         // nx <= 16384
constexpr int kGatherTileWords = 256;  // row words per tile of the adjoint's gather

// Geometry of one volume in the (possibly batched) word space.
// Division by an invariant d >= 1 for numerators below 2^31: q = (hi32(n*m)
// + n) >> l with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, section 4).
struct FastDiv {
  uint32_t m;
  int32_t l;
};

inline FastDiv make_fast_div(uint32_t d) {
  FastDiv f;
  f.l = 0;
  while ((1ull << f.l) < d) ++f.l;
  f.m = (uint32_t)((((1ull << 32) * ((1ull << f.l) - d)) / d) + 1);
  return f;
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, FastDiv f) {
  return (uint32_t)(((uint64_t)__umulhi(n, f.m) + n) >> f.l);
}

struct VolGeom {
  int32_t nx, ny, nz, wpr;
  float iso;
  float origin[3];
  float spacing[3];
  int64_t field_off;  // first field element of the volume
  int64_t word_off;   // first row word of the volume
  int64_t nwords;     // wpr*ny*nz
  FastDiv div_wpr, div_ny, div_nx;
  // Spacings that are powers of two (bit d): then x / (2h) == x * (rcp_h/2)
  // and x / h == x * rcp_h exactly in IEEE arithmetic.
  uint32_t pow2_h;
  float rcp_h[3];
};

// Host: fill the division constants after nx, ny, wpr and spacing are set.
inline void set_geom_divs(VolGeom& g) {
  g.div_wpr = make_fast_div((uint32_t)g.wpr);
  g.div_ny = make_fast_div((uint32_t)g.ny);
  g.div_nx = make_fast_div((uint32_t)g.nx);
  g.pow2_h = 0;
  for (int d = 0; d < 3; ++d) {
    const float h = g.spacing[d];
    int e = 0;
    const float m = frexpf(h, &e);  // h = m * 2^e, m in [0.5, 1)
    g.rcp_h[d] = 1.0f;
    if (h > 0.0f && m == 0.5f && e >= -120 && e <= 120) {  // h = 2^(e-1), far from the range limits
      g.pow2_h |= 1u << d;
      g.rcp_h[d] = ldexpf(1.0f, 1 - e);
    }
  }
}

// Edge e of a cell with lowest corner (i,j,k): its lower endpoint is
// (i+dx, j+dy, k+dz) and it runs along `axis`.  Derived from the vertex pairs
// of MarchingCubes.h:131-135 and the corner order of
// TopologyUniform.h:139-159; the lower endpoint is always the smaller point
// id (MarchingCubes.h:148-157).  Packed as dx | dy<<1 | dz<<2 | axis<<3.
__host__ __device__ constexpr uint8_t edge_code(int e) {
  return (uint8_t)(e == 0 ? (0 | 0 << 1 | 0 << 2 | 0 << 3)
                 : e == 1 ? (1 | 0 << 1 | 0 << 2 | 1 << 3)
                 : e == 2 ? (0 | 1 << 1 | 0 << 2 | 0 << 3)
                 : e == 3 ? (0 | 0 << 1 | 0 << 2 | 1 << 3)
                 : e == 4 ? (0 | 0 << 1 | 1 << 2 | 0 << 3)
                 : e == 5 ? (1 | 0 << 1 | 1 << 2 | 1 << 3)
                 : e == 6 ? (0 | 1 << 1 | 1 << 2 | 0 << 3)
                 : e == 7 ? (0 | 0 << 1 | 1 << 2 | 1 << 3)
                 : e == 8 ? (0 | 0 << 1 | 0 << 2 | 2 << 3)
                 : e == 9 ? (1 | 0 << 1 | 0 << 2 | 2 << 3)
                 : e == 10 ? (1 | 1 << 1 | 0 << 2 | 2 << 3)
                           : (0 | 1 << 1 | 0 << 2 | 2 << 3));
}

__device__ __forceinline__ uint64_t shr1(uint64_t a, uint64_t a_next) {
  return (a >> 1) | (a_next << 63);
}

// Mask of the first n bits (n in [0,64]).
__device__ __forceinline__ uint64_t low_mask(int n) {
  return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1ull));
}

__device__ __forceinline__ int popc(uint64_t x) { return __popcll(x); }

// Per-word geometry: row coordinates and validity masks.
struct WordPos {
  int q, j, k;
  bool has_y, has_z;   // a +y / +z neighbour row exists
  uint64_t pmask;      // bits that are grid points
  uint64_t xmask;      // bits whose +x neighbour exists (x-edges, cells in x)
};

// lw < 2^31: a volume holds at most 2^30 points and wpr <= nx, so its word
// count wpr*ny*nz is at most 2^30 (32-bit division is ~4x cheaper).
__device__ __forceinline__ WordPos word_pos(const VolGeom& g, int64_t lw) {
  WordPos p;
  const uint32_t w = (uint32_t)lw;
  const uint32_t row = fast_div(w, g.div_wpr);
  const uint32_t k = fast_div(row, g.div_ny);
  p.q = (int)(w - row * (uint32_t)g.wpr);
  p.j = (int)(row - k * (uint32_t)g.ny);
  p.k = (int)k;
  p.has_y = p.j + 1 < g.ny;
  p.has_z = p.k + 1 < g.nz;
  const int x0 = p.q * 64;
  // A grid without cells (fewer than 2 points along an axis) has no output:
  // the reference returns before generating when the triangle count is 0
  // (DispatcherGenerateInterpolatedCells.h:190-194), so no edge counts.
  const bool has_cells = g.nx > 1 && g.ny > 1 && g.nz > 1;
  p.pmask = has_cells ? low_mask(g.nx - x0) : 0ull;
  p.xmask = has_cells ? low_mask(g.nx - 1 - x0) : 0ull;
  return p;
}

// Active-edge flags of a word from its row word A, the next word A1 in x,
// and the rows B (+y) and C (+z).
__device__ __forceinline__ void edge_flags(const WordPos& p, uint64_t A, uint64_t A1, uint64_t B,
                                           uint64_t C, uint64_t& fx, uint64_t& fy, uint64_t& fz) {
  fx = (A ^ shr1(A, A1)) & p.xmask;
  fy = p.has_y ? ((A ^ B) & p.pmask) : 0ull;
  fz = p.has_z ? ((A ^ C) & p.pmask) : 0ull;
}

// The eight corner bit-planes of the 64 cells of a word (corner order of
// TopologyUniform.h:139-159) and the mask of cells with mixed corners.
struct CellPlanes {
  uint64_t c[8];
  uint64_t active;
};

__device__ __forceinline__ void cell_planes(const WordPos& p, uint64_t A, uint64_t A1, uint64_t B,
                                            uint64_t B1, uint64_t C, uint64_t C1, uint64_t D,
                                            uint64_t D1, CellPlanes& cp) {
  cp.c[0] = A;
  cp.c[1] = shr1(A, A1);
  cp.c[2] = shr1(B, B1);
  cp.c[3] = B;
  cp.c[4] = C;
  cp.c[5] = shr1(C, C1);
  cp.c[6] = shr1(D, D1);
  cp.c[7] = D;
  uint64_t all = ~0ull, any = 0ull;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    all &= cp.c[v];
    any |= cp.c[v];
  }
  cp.active = (p.has_y && p.has_z) ? ((any & ~all) & p.xmask) : 0ull;
}

// Triangles of a word's cells: NumFaces[case] summed over the active cells,
// 64 cells at a time.  For every case with mixed corners
//   NumFaces = V - 2F - 4H,
// V = inside corners, F = cube faces whose four corners are inside, H = 1
// when exactly two corners are outside and they are the ends of a body
// diagonal (the table's two-triangle cases 125, 190, 215, 235).  Checked
// against MarchingCubesTable.h:28-45 for all 254 mixed cases.  Faces in the
// corner order of TopologyUniform.h:139-159: z = 0 {0,1,2,3}, z = 1
// {4,5,6,7}, y = 0 {0,1,5,4}, y = 1 {3,2,6,7}, x = 0 {0,3,7,4}, x = 1
// {1,2,6,5}; body diagonals (0,6), (1,7), (2,4), (3,5).
__device__ __forceinline__ uint32_t word_triangles(const CellPlanes& cp) {
  const uint64_t a = cp.active;
  const uint64_t* c = cp.c;
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) v += (uint32_t)popc(c[i] & a);
  const uint64_t p01 = c[0] & c[1], p23 = c[2] & c[3], p45 = c[4] & c[5], p67 = c[6] & c[7];
  const uint32_t f = (uint32_t)(popc(p01 & p23 & a) + popc(p45 & p67 & a) + popc(p01 & p45 & a) +
                                popc(p23 & p67 & a) + popc(c[0] & c[3] & c[4] & c[7] & a) +
                                popc(c[1] & c[2] & c[5] & c[6] & a));
  const uint64_t h = (p23 & p45 & ((c[1] & ~c[0] & c[7] & ~c[6]) | (c[0] & ~c[1] & c[6] & ~c[7]))) |
                     (p01 & p67 & ((c[3] & ~c[2] & c[5] & ~c[4]) | (c[2] & ~c[3] & c[4] & ~c[5])));
  return v - 2u * f - 4u * (uint32_t)popc(h & a);
}

__device__ __forceinline__ int cell_case(const CellPlanes& cp, int i) {
  int c = 0;
#pragma unroll
  for (int v = 0; v < 8; ++v) c |= (int)((cp.c[v] >> i) & 1ull) << v;
  return c;
}

// Rank of the edge (bit i, axis a) among the active edges of a word.
__device__ __forceinline__ int edge_rank(uint64_t fx, uint64_t fy, uint64_t fz, int i, int a) {
  const uint64_t lo = (1ull << i) - 1ull;
  int r = popc(fx & lo) + popc(fy & lo) + popc(fz & lo);
  if (a > 0) r += (int)((fx >> i) & 1ull);
  if (a > 1) r += (int)((fy >> i) & 1ull);
  return r;
}

// Inverse of edge_rank: the k-th active edge of a word in (bit, axis x<y<z)
// order (k < popc(fx) + popc(fy) + popc(fz)): the 32-bit half that holds
// it, then halving steps, each counting the three flag sets below the split.
__device__ __forceinline__ void edge_select(uint64_t fx, uint64_t fy, uint64_t fz, int k, int& bit,
                                            int& axis) {
  const int nlo = __popc((uint32_t)fx) + __popc((uint32_t)fy) + __popc((uint32_t)fz);
  const bool up = k >= nlo;
  uint32_t x = up ? (uint32_t)(fx >> 32) : (uint32_t)fx;
  uint32_t y = up ? (uint32_t)(fy >> 32) : (uint32_t)fy;
  uint32_t z = up ? (uint32_t)(fz >> 32) : (uint32_t)fz;
  int r = up ? k - nlo : k, pos = up ? 32 : 0;
#pragma unroll
  for (int w = 16; w > 0; w >>= 1) {
    const uint32_t m = (1u << w) - 1u;
    const int c = __popc(x & m) + __popc(y & m) + __popc(z & m);
    const bool hi = r >= c;
    r = hi ? r - c : r;
    x = hi ? x >> w : x;
    y = hi ? y >> w : y;
    z = hi ? z >> w : z;
    pos += hi ? w : 0;
  }
  // bit `pos` holds its set edges x, y, z in that order; r < their count
  const int bx = (int)(x & 1u), by = (int)(y & 1u);
  bit = pos;
  axis = r < bx ? 0 : (r < bx + by ? 1 : 2);
}

// Decoupled look-back tile status: flag in bits 62-63 (0 = not ready,
// 1 = aggregate, 2 = inclusive prefix), triangle count in bits 31-61, vertex
// count in bits 0-30.  Totals above 2^31-1 overflow dax::Id anyway; the
// 64-bit totals accumulated beside the scan detect that case.
constexpr uint64_t kStatusAgg = 1ull;
constexpr uint64_t kStatusPrefix = 2ull;

__device__ __forceinline__ uint64_t pack_status(uint32_t v, uint32_t t, uint64_t flag) {
  return (flag << 62) | ((uint64_t)(t & 0x7fffffffu) << 31) | (uint64_t)(v & 0x7fffffffu);
}

// Control block at the end of the workspace.
struct Ctrl {
  uint32_t tile_counter;
  uint32_t pad;
  unsigned long long total_v;  // 64-bit totals: vertices (active edges), triangles
  unsigned long long total_t;
};

// Workspace layout for a word space of `nwords` row words.  off_counts
// holds every word's exclusive vertex / triangle prefix within its count
// tile of kCountTile words (v | t << 16: at most 128 * 192 and 127 * 320,
// both below 2^16); off_tsum every count tile's exclusive prefix within its
// count block (v | t << 32); off_bsum the count blocks' sums and off_tpre
// their exclusive prefixes, the totals after the last.  A word's bases are
// block prefix + tile prefix + local prefix.  off_gcls / off_tcls: the
// dispatch class of every generate / gather tile (a byte each, from the
// count kernel); off_order: the generate tile size the order was built for
// (0: none), then the generate tiles in dispatch order; off_gorder: the
// same for the adjoint's gather tiles.
struct Workspace {
  size_t off_bits, off_base_v, off_base_t, off_status, off_ctrl, off_counts, off_tsum, off_bsum,
      off_tpre, off_gcls, off_tcls, off_order, off_gorder, bytes;
  int64_t nwords, ntiles, n_ctiles;
};

inline Workspace workspace_layout(int64_t nwords) {
  Workspace w;
  w.nwords = nwords;
  w.ntiles = ceil_div(nwords, kTileWords);
  size_t off = 0;
  w.off_bits = off;
  off = align_up(off + (size_t)(nwords + 1) * 8, 256);
  w.off_base_v = off;
  off = align_up(off + (size_t)(nwords + 1) * 4, 256);
  w.off_base_t = off;
  off = align_up(off + (size_t)(nwords + 1) * 4, 256);
  w.off_status = off;  // status then ctrl: one contiguous block zeroed per run
  off = off + (size_t)w.ntiles * 8;
  w.off_ctrl = off;
  off = align_up(off + sizeof(Ctrl), 256);
  w.off_counts = off;  // packed per-word counts
  off = align_up(off + (size_t)w.ntiles * kTileWords * 4, 256);
  w.n_ctiles = ceil_div(nwords, kCountTile);
  w.off_tsum = off;  // per count tile: exclusive prefix within its block
  off = align_up(off + (size_t)(w.n_ctiles + 1) * 8, 256);
  w.off_bsum = off;  // per count block (at least one tile each): sums
  off = align_up(off + (size_t)(w.n_ctiles + 1) * 8, 256);
  w.off_tpre = off;  // per count block: exclusive prefixes, then the totals
  off = align_up(off + (size_t)(w.n_ctiles + 1) * 8, 256);
  w.off_gcls = off;  // per generate tile (at least kCountTile words): dispatch class
  off = align_up(off + (size_t)w.n_ctiles + 16, 256);
  w.off_tcls = off;  // per gather tile: dispatch class
  off = align_up(off + (size_t)ceil_div(nwords, kGatherTileWords) + 16, 256);
  w.off_order = off;  // generate tiles, heaviest first: the tile size, then the tiles
  off = align_up(off + (size_t)(w.n_ctiles + 2) * 4, 256);
  w.off_gorder = off;  // the adjoint's gather tiles, those with output first
  off = align_up(off + (size_t)(ceil_div(nwords, kGatherTileWords) + 2) * 4, 256);
  w.bytes = off;
  return w;
}

}  // namespace mc
}  // namespace daxhip
