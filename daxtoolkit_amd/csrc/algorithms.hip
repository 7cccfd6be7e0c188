// DeviceAdapterAlgorithm<DeviceAdapterTagHip> primitives for plain numeric
// arrays (the marching-cubes path has its own fused scan in mc_kernels.h).
//
// Reference contract: dax/cont/internal/DeviceAdapterAlgorithm.h:44-331;
// semantics follow the Serial adapter (DeviceAdapterAlgorithmSerial.h:46-199)
// and DeviceAdapterAlgorithmGeneral.h (Unique :904-927, StreamCompact
// :769-830, LowerBounds/UpperBounds :258-308 and :1044-1092).
//
//   scan        reduce-then-scan (tile sums, one-block scan of the sums,
//               tile scans with the prefix): deterministic, also for floats
//   bounds      one thread per query, binary search
//   sort        stable LSD radix sort, 8-bit digits; per-wave digit ranks
//               from ballots over the digit bits (no atomics -> stable)
//   sort by key radix sort of (key, index), then a gather of the values
//   unique      keep-flags + scan + scatter
//   compact     stencil flags + scan + scatter
#include <algorithm>
#include <type_traits>

#include "algorithms.h"

namespace daxhip {
namespace alg {

// ---------------------------------------------------------------------------
// copy
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

__global__ void k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int copy_bytes(void* d_dst, const void* d_src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return DAX_HIP_OK;
  const uintptr_t a = (uintptr_t)d_src, b = (uintptr_t)d_dst;
  if ((a % 16) == 0 && (b % 16) == 0) {
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 > 0) {
      const int64_t blocks = std::min<int64_t>(ceil_div(n16, 256), 256 * 32);
      k_copy16<<<dim3((unsigned)blocks), dim3(256), 0, s>>>((const uint4*)d_src, (uint4*)d_dst, n16);
      DAXHIP_LAUNCH_CHECK("k_copy16");
    }
    const int64_t tail = (int64_t)(bytes - (size_t)n16 * 16);
    if (tail > 0) {
      k_copy1<<<1, 256, 0, s>>>((const uint8_t*)d_src + n16 * 16, (uint8_t*)d_dst + n16 * 16, tail);
      DAXHIP_LAUNCH_CHECK("k_copy1");
    }
  } else {
    k_copy1<<<dim3((unsigned)ceil_div((int64_t)bytes, 256)), dim3(256), 0, s>>>(
        (const uint8_t*)d_src, (uint8_t*)d_dst, (int64_t)bytes);
    DAXHIP_LAUNCH_CHECK("k_copy1");
  }
  return DAX_HIP_OK;
}

// ---------------------------------------------------------------------------
// block-wide helpers (256 threads)
constexpr int kNT = 256;

template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)u, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(u >> 32), d, 64);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __shfl_up(__builtin_bit_cast(int, v), d, 64));
  }
}

// Inclusive wave scan in lane order (fixed association for floats).
template <typename T>
__device__ __forceinline__ T wave_scan_incl(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = shfl_up(v, d);
    if (lane >= d) v = o + v;
  }
  return v;
}

// Exclusive block scan of one value per thread (NT threads); returns the
// block total in `total`.
template <typename T, int NT = kNT>
__device__ __forceinline__ T block_scan_excl(T v, T& total, T* s_wave /* [NT / 64] */) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const T inc = wave_scan_incl(v);
  if (lane == 63) s_wave[wv] = inc;
  __syncthreads();
  T off = T(0);
  T tot = T(0);
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) {
    if (q < wv) off = off + s_wave[q];
    tot = tot + s_wave[q];
  }
  __syncthreads();
  total = tot;
  return off + (inc - v);
}

// ---------------------------------------------------------------------------
// scan: reduce-then-scan over tiles of kNT threads x scan_ipt<T>() items; each
// thread owns 64 contiguous bytes, read and written as 16-byte vectors when
// the arrays are 16-byte aligned.
template <typename T>
__host__ __device__ constexpr int scan_ipt() {
  return 64 / (int)sizeof(T);
}

template <typename T>
__device__ __forceinline__ void load_chunk(const T* __restrict__ in, int64_t base, int64_t n,
                                           bool vec, T* vals) {
  constexpr int IPT = scan_ipt<T>();
  if (vec && base + IPT <= n) {
    const int4* p = (const int4*)(in + base);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int4 v = p[q];
      __builtin_memcpy(&vals[q * IPT / 4], &v, 16);
    }
  } else {
#pragma unroll
    for (int i = 0; i < IPT; ++i) vals[i] = base + i < n ? in[base + i] : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void store_chunk(T* __restrict__ out, int64_t base, int64_t n, bool vec,
                                            const T* vals) {
  constexpr int IPT = scan_ipt<T>();
  if (vec && base + IPT <= n) {
    int4* p = (int4*)(out + base);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int4 v;
      __builtin_memcpy(&v, &vals[q * IPT / 4], 16);
      p[q] = v;
    }
  } else {
#pragma unroll
    for (int i = 0; i < IPT; ++i)
      if (base + i < n) out[base + i] = vals[i];
  }
}

template <typename T>
__global__ __launch_bounds__(kNT) void k_tile_sums(const T* __restrict__ in, int64_t n, bool vec,
                                                   T* __restrict__ sums) {
  constexpr int IPT = scan_ipt<T>();
  __shared__ T s_wave[kNT / 64];
  const int64_t base = ((int64_t)blockIdx.x * kNT + threadIdx.x) * IPT;
  T vals[IPT];
  load_chunk(in, base, n, vec, vals);
  T v = T(0);
#pragma unroll
  for (int i = 0; i < IPT; ++i) v = v + vals[i];
  T total;
  (void)block_scan_excl(v, total, s_wave);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the n tile sums in one block of kSumsNT threads: chunks
// of kSumsNT x 16 bytes in order, each thread scanning 16 contiguous bytes
// (coalesced), the chunk's total carried to the next; the grand total goes
// to *grand.
constexpr int kSumsNT = 1024;
template <typename T>
__global__ __launch_bounds__(kSumsNT) void k_scan_sums(T* __restrict__ sums, int64_t n,
                                                       T* __restrict__ grand) {
  constexpr int V = 16 / (int)sizeof(T);
  __shared__ T s_wave[kSumsNT / 64];
  T carry = T(0);
  for (int64_t c0 = 0; c0 < n; c0 += (int64_t)kSumsNT * V) {
    const int64_t b = c0 + (int64_t)threadIdx.x * V;
    T v[V];
    T t = T(0);
#pragma unroll
    for (int q = 0; q < V; ++q) {
      v[q] = b + q < n ? sums[b + q] : T(0);
      t = t + v[q];
    }
    T total;
    T run = carry + block_scan_excl<T, kSumsNT>(t, total, s_wave);
#pragma unroll
    for (int q = 0; q < V; ++q) {
      if (b + q < n) sums[b + q] = run;
      run = run + v[q];
    }
    carry = carry + total;
  }
  if (threadIdx.x == 0) *grand = carry;
}

template <typename T>
__global__ __launch_bounds__(kNT) void k_tile_scan(const T* __restrict__ in, int64_t n,
                                                   const T* __restrict__ prefix, bool vec,
                                                   T* __restrict__ out, int inclusive) {
  constexpr int IPT = scan_ipt<T>();
  __shared__ T s_wave[kNT / 64];
  const int64_t base = ((int64_t)blockIdx.x * kNT + threadIdx.x) * IPT;
  T vals[IPT];
  load_chunk(in, base, n, vec, vals);
  T v = T(0);
#pragma unroll
  for (int i = 0; i < IPT; ++i) v = v + vals[i];
  T total;
  T run = prefix[blockIdx.x] + block_scan_excl(v, total, s_wave);
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const T x = vals[i];
    if (inclusive) {
      run = run + x;
      vals[i] = run;
    } else {
      vals[i] = run;
      run = run + x;
    }
  }
  store_chunk(out, base, n, vec, vals);
}

template <typename T>
int scan(const T* in, T* out, int64_t n, int inclusive, T* h_total, hipStream_t s) {
  if (n <= 0) {
    if (h_total) *h_total = T(0);
    return DAX_HIP_OK;
  }
  const int64_t tile = (int64_t)kNT * scan_ipt<T>();
  const int64_t tiles = ceil_div(n, tile);
  const bool vec = (((uintptr_t)in | (uintptr_t)out) & 15u) == 0;
  TempBuf tmp(s);
  int rc = tmp.alloc((size_t)(tiles + 1) * sizeof(T));
  if (rc) return rc;
  T* sums = (T*)tmp.p;
  k_tile_sums<T><<<dim3((unsigned)tiles), dim3(kNT), 0, s>>>(in, n, vec, sums);
  DAXHIP_LAUNCH_CHECK("k_tile_sums");
  k_scan_sums<T><<<1, kSumsNT, 0, s>>>(sums, tiles, sums + tiles);
  DAXHIP_LAUNCH_CHECK("k_scan_sums");
  k_tile_scan<T><<<dim3((unsigned)tiles), dim3(kNT), 0, s>>>(in, n, sums, vec, out, inclusive);
  DAXHIP_LAUNCH_CHECK("k_tile_scan");
  if (!h_total) return DAX_HIP_OK;  // stream-ordered, no host round trip
  T total;
  DAXHIP_CHECK(hipMemcpyAsync(&total, sums + tiles, sizeof(T), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  *h_total = total;
  return DAX_HIP_OK;
}

// ---------------------------------------------------------------------------
// bounds
template <typename T, bool UPPER, bool DESC>
__global__ void k_bounds(const T* __restrict__ sorted, int64_t n, const T* __restrict__ vals,
                         int64_t m, int32_t* __restrict__ out32, int64_t* __restrict__ out64) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const T v = vals[q];
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const T x = sorted[mid];
    // lower bound: first x with !(x < v); upper bound: first x with v < x
    // (with "<" replaced by ">" for a descending order)
    const bool go_right = UPPER ? (DESC ? !(v > x) : !(v < x)) : (DESC ? (x > v) : (x < v));
    if (go_right)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (out32) out32[q] = (int32_t)lo;
  if (out64) out64[q] = lo;
}

// ---------------------------------------------------------------------------
// radix sort: keys are mapped to unsigned integers whose order is the key
// order (sign flip for signed ints, IEEE trick for floats; descending order
// complements the mapped bits).
template <typename T>
struct RadixKey;
template <>
struct RadixKey<int32_t> {
  typedef uint32_t U;
  __device__ static U enc(int32_t k) { return (uint32_t)k ^ 0x80000000u; }
  __device__ static int32_t dec(U u) { return (int32_t)(u ^ 0x80000000u); }
};
template <>
struct RadixKey<uint32_t> {
  typedef uint32_t U;
  __device__ static U enc(uint32_t k) { return k; }
  __device__ static uint32_t dec(U u) { return u; }
};
template <>
struct RadixKey<int64_t> {
  typedef uint64_t U;
  __device__ static U enc(int64_t k) { return (uint64_t)k ^ 0x8000000000000000ull; }
  __device__ static int64_t dec(U u) { return (int64_t)(u ^ 0x8000000000000000ull); }
};
template <>
struct RadixKey<uint64_t> {
  typedef uint64_t U;
  __device__ static U enc(uint64_t k) { return k; }
  __device__ static uint64_t dec(U u) { return u; }
};
template <>
struct RadixKey<float> {
  typedef uint32_t U;
  __device__ static U enc(float k) {
    const uint32_t b = __builtin_bit_cast(uint32_t, k);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
  __device__ static float dec(U u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __builtin_bit_cast(float, b);
  }
};
template <>
struct RadixKey<double> {
  typedef uint64_t U;
  __device__ static U enc(double k) {
    const uint64_t b = __builtin_bit_cast(uint64_t, k);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  }
  __device__ static double dec(U u) {
    const uint64_t b = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    return __builtin_bit_cast(double, b);
  }
};

// Items per thread per sort tile: 16 (4096-key tiles) for large inputs, 4
// (1024-key tiles) below kSmallSort keys, so that small sorts still spread
// over the CUs.
constexpr int64_t kSmallSort = 1 << 22;

template <typename T>
__global__ void k_encode(const T* __restrict__ in, typename RadixKey<T>::U* __restrict__ out,
                         int32_t* __restrict__ idx, int64_t n, int desc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename RadixKey<T>::U u = RadixKey<T>::enc(in[i]);
  out[i] = desc ? ~u : u;
  if (idx) idx[i] = (int32_t)i;
}

template <typename T>
__global__ void k_decode(const typename RadixKey<T>::U* __restrict__ in, T* __restrict__ out,
                         int64_t n, int desc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const typename RadixKey<T>::U u = in[i];
  out[i] = RadixKey<T>::dec(desc ? ~u : u);
}

// The sort tile of block L.  Workgroups are dispatched round-robin over the
// 8 XCDs, so block L takes tile (L % 8) * (N8 / 8) + L / 8 (N8: the tile
// count rounded down to a multiple of 8; the rest keep their index), and
// each XCD works through a contiguous run of tiles: the histogram words of
// consecutive tiles, which share cache lines, fill whole lines in one L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t ntiles) {
  const int64_t n8 = ntiles & ~(int64_t)7;
  const int64_t b = blockIdx.x;
  return b < n8 ? (b % 8) * (n8 / 8) + b / 8 : b;
}

// Per-tile digit histograms, digit-major: hist[digit * ntiles + tile]; a
// block writes one word to each of the 256 rows.
template <typename U, int kSortIPT>
__global__ __launch_bounds__(kNT) void k_radix_hist(const U* __restrict__ keys, int64_t n, int shift,
                                                    uint32_t* __restrict__ hist, int64_t ntiles) {
  constexpr int kSortTile = kNT * kSortIPT;
  __shared__ uint32_t s_h[256];
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t tile = xcd_tile(ntiles);
  const int64_t base = tile * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortIPT; ++i) {
    const int64_t k = base + (int64_t)i * kNT + threadIdx.x;
    if (k < n) atomicAdd(&s_h[(uint32_t)(keys[k] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + tile] = s_h[threadIdx.x];
}

// Stable scatter: the tile is processed in rounds of 256 keys in index order;
// within a round each wave ranks its 64 keys per digit with ballots, waves
// are ordered by a per-round digit-count scan in LDS.
template <typename U, bool WITH_IDX, int kSortIPT>
__global__ __launch_bounds__(kNT) void k_radix_scatter(const U* __restrict__ keys_in,
                                                       const int32_t* __restrict__ idx_in,
                                                       int64_t n, int shift,
                                                       const uint32_t* __restrict__ offsets,
                                                       int64_t ntiles, U* __restrict__ keys_out,
                                                       int32_t* __restrict__ idx_out) {
  constexpr int kSortTile = kNT * kSortIPT;
  __shared__ uint32_t s_base[256];          // running output offset per digit
  __shared__ uint32_t s_wcnt[kNT / 64][256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  s_base[threadIdx.x] = offsets[(int64_t)threadIdx.x * ntiles + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortIPT; ++r) {
    const int64_t k = base + (int64_t)r * kNT + threadIdx.x;
    const bool valid = k < n;
    const U key = valid ? keys_in[k] : U(0);
    const uint32_t d = (uint32_t)(key >> shift) & 0xffu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const uint32_t cnt = (uint32_t)__popcll(peers);
    for (int i = threadIdx.x; i < (kNT / 64) * 256; i += kNT) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    // the first lane of each digit group records the wave's count
    if (valid && rank == 0) s_wcnt[wv][d] = cnt;
    __syncthreads();
    if (valid) {
      uint32_t off = s_base[d];
      for (int q = 0; q < wv; ++q) off += s_wcnt[q][d];
      const uint32_t pos = off + rank;
      if (pos < (uint64_t)n) {  // always true when the histograms are consistent
        keys_out[pos] = key;
        if (WITH_IDX) idx_out[pos] = idx_in[k];
      }
    }
    __syncthreads();
    // advance the digit bases by this round's counts
    {
      const int dd = threadIdx.x;
      uint32_t add = 0;
      for (int q = 0; q < kNT / 64; ++q) add += s_wcnt[q][dd];
      s_base[dd] += add;
    }
    __syncthreads();
  }
}

// LSD passes over the low `bits` bits of the encoded keys (8 bits a pass);
// the result is in `keys` (the buffers are swapped after every pass).
template <typename U, int kSortIPT>
int radix_passes(U*& keys, U*& keys_alt, int32_t*& idx, int32_t*& idx_alt, int64_t n, int bits,
                 hipStream_t s) {
  constexpr int kSortTile = kNT * kSortIPT;
  const int64_t ntiles = ceil_div(n, (int64_t)kSortTile);
  TempBuf th(s);
  int rc = th.alloc((size_t)ntiles * 256 * sizeof(uint32_t));
  if (rc) return rc;
  uint32_t* hist = (uint32_t*)th.p;
  for (int shift = 0; shift < bits; shift += 8) {
    k_radix_hist<U, kSortIPT><<<dim3((unsigned)ntiles), dim3(kNT), 0, s>>>(keys, n, shift, hist,
                                                                          ntiles);
    DAXHIP_LAUNCH_CHECK("k_radix_hist");
    if ((rc = scan<uint32_t>(hist, hist, ntiles * 256, 0, nullptr, s))) return rc;
    if (idx)
      k_radix_scatter<U, true, kSortIPT><<<dim3((unsigned)ntiles), dim3(kNT), 0, s>>>(
          keys, idx, n, shift, hist, ntiles, keys_alt, idx_alt);
    else
      k_radix_scatter<U, false, kSortIPT><<<dim3((unsigned)ntiles), dim3(kNT), 0, s>>>(
          keys, nullptr, n, shift, hist, ntiles, keys_alt, nullptr);
    DAXHIP_LAUNCH_CHECK("k_radix_scatter");
    std::swap(keys, keys_alt);
    std::swap(idx, idx_alt);
  }
  return DAX_HIP_OK;
}

template <typename U>
int radix_sort_encoded(U*& keys, U*& keys_alt, int32_t*& idx, int32_t*& idx_alt, int64_t n,
                       hipStream_t s, int bits = 8 * (int)sizeof(U)) {
  return n < kSmallSort ? radix_passes<U, 4>(keys, keys_alt, idx, idx_alt, n, bits, s)
                        : radix_passes<U, 16>(keys, keys_alt, idx, idx_alt, n, bits, s);
}

template <typename T>
int sort(T* d, int64_t n, int desc, int32_t* perm_out, hipStream_t s) {
  if (n <= 1) {
    if (perm_out && n == 1) {
      const int32_t z = 0;
      DAXHIP_CHECK(hipMemcpyAsync(perm_out, &z, 4, hipMemcpyHostToDevice, s));
      DAXHIP_CHECK(hipStreamSynchronize(s));
    }
    return DAX_HIP_OK;
  }
  if (n > 0x7fffffffLL) {
    set_error("dax_hip_sort: more than 2^31-1 keys");
    return DAX_HIP_BAD_VALUE;
  }
  typedef typename RadixKey<T>::U U;
  TempBuf tk(s);
  const bool with_idx = perm_out != nullptr;
  int rc = tk.alloc((size_t)n * (2 * sizeof(U) + (with_idx ? 2 * sizeof(int32_t) : 0)));
  if (rc) return rc;
  U* k0 = (U*)tk.p;
  U* k1 = k0 + n;
  int32_t* i0 = with_idx ? (int32_t*)(k1 + n) : nullptr;
  int32_t* i1 = with_idx ? i0 + n : nullptr;
  const dim3 g((unsigned)ceil_div(n, 256)), b(256);
  k_encode<T><<<g, b, 0, s>>>(d, k0, i0, n, desc);
  DAXHIP_LAUNCH_CHECK("k_encode");
  if ((rc = radix_sort_encoded<U>(k0, k1, i0, i1, n, s))) return rc;
  k_decode<T><<<g, b, 0, s>>>(k0, d, n, desc);
  DAXHIP_LAUNCH_CHECK("k_decode");
  if (with_idx) DAXHIP_CHECK(hipMemcpyAsync(perm_out, i0, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  return DAX_HIP_OK;
}

int scan_exclusive_i32_async(const int32_t* d_in, int32_t* d_out, int64_t n, hipStream_t s) {
  return scan<int32_t>(d_in, d_out, n, 0, nullptr, s);
}

// Ascending, stream-ordered sort of unsigned keys that are below 2^bits.
int sort_u64_bits(uint64_t* d, int64_t n, int bits, hipStream_t s, int32_t* d_vals) {
  if (n <= 1 || bits <= 0) return DAX_HIP_OK;
  if (n > 0x7fffffffLL || bits > 64) {
    set_error("sort_u64_bits: more than 2^31-1 keys or more than 64 bits");
    return DAX_HIP_BAD_VALUE;
  }
  TempBuf tk(s), tv(s);
  int rc = tk.alloc((size_t)n * sizeof(uint64_t));
  if (rc) return rc;
  if (d_vals && (rc = tv.alloc((size_t)n * sizeof(int32_t)))) return rc;
  uint64_t* k0 = d;
  uint64_t* k1 = (uint64_t*)tk.p;
  int32_t* i0 = d_vals;
  int32_t* i1 = d_vals ? (int32_t*)tv.p : nullptr;
  if ((rc = radix_sort_encoded<uint64_t>(k0, k1, i0, i1, n, s, bits))) return rc;
  if (k0 != d && (rc = copy_bytes(d, k0, (size_t)n * sizeof(uint64_t), s))) return rc;
  if (d_vals && i0 != d_vals) return copy_bytes(d_vals, i0, (size_t)n * sizeof(int32_t), s);
  return DAX_HIP_OK;
}

// gather of fixed-size records: out[i] = in[perm[i]]
__global__ void k_gather_bytes(const uint8_t* __restrict__ in, const int32_t* __restrict__ perm,
                               int64_t n, int64_t rec, uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t* src = in + (int64_t)perm[i] * rec;
  uint8_t* dst = out + i * rec;
  if (rec % 4 == 0) {
    for (int64_t b = 0; b < rec; b += 4) *(uint32_t*)(dst + b) = *(const uint32_t*)(src + b);
  } else {
    for (int64_t b = 0; b < rec; ++b) dst[b] = src[b];
  }
}

template <typename T>
int sort_by_key(T* keys, void* values, size_t value_bytes, int64_t n, int desc, hipStream_t s) {
  if (n <= 1) return DAX_HIP_OK;
  TempBuf tp(s);
  int rc = tp.alloc((size_t)n * (4 + value_bytes));
  if (rc) return rc;
  int32_t* perm = (int32_t*)tp.p;
  if ((rc = sort<T>(keys, n, desc, perm, s))) return rc;
  uint8_t* tmpv = (uint8_t*)(perm + n);
  k_gather_bytes<<<dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s>>>(
      (const uint8_t*)values, perm, n, (int64_t)value_bytes, tmpv);
  DAXHIP_LAUNCH_CHECK("k_gather_bytes");
  if ((rc = copy_bytes(values, tmpv, (size_t)n * value_bytes, s))) return rc;
  DAXHIP_CHECK(hipStreamSynchronize(s));
  return DAX_HIP_OK;
}

// ---------------------------------------------------------------------------
// flags + scan + scatter (unique, stream compact)
template <typename T>
__global__ void k_unique_flags(const T* __restrict__ v, int64_t n, int32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || !(v[i] == v[i - 1])) ? 1 : 0;
}

// stencil element types: 0 = i32, 1 = i64, 2 = f32, 3 = u8, 4 = f64
__global__ void k_stencil_flags(const void* __restrict__ st, int type, int64_t n,
                                int32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool keep;
  switch (type) {
    case 0: keep = ((const int32_t*)st)[i] != 0; break;
    case 1: keep = ((const int64_t*)st)[i] != 0; break;
    case 2: keep = ((const float*)st)[i] != 0.0f; break;
    case 3: keep = ((const uint8_t*)st)[i] != 0; break;
    default: keep = ((const double*)st)[i] != 0.0; break;
  }
  flags[i] = keep ? 1 : 0;
}

__global__ void k_scatter_flagged(const uint8_t* __restrict__ in, int64_t rec,
                                  const int32_t* __restrict__ flags,
                                  const int32_t* __restrict__ pos, int64_t n,
                                  uint8_t* __restrict__ out, int32_t* __restrict__ idx32,
                                  int64_t* __restrict__ idx64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const int64_t p = pos[i];
  if (p < 0 || p >= n) return;  // positions come from an exclusive scan of 0/1 flags
  if (in)
    for (int64_t b = 0; b < rec; ++b) out[p * rec + b] = in[i * rec + b];
  if (idx32) idx32[p] = (int32_t)i;
  if (idx64) idx64[p] = i;
}

// Compact `in` (records of rec bytes) or indices by flags; returns the count.
int compact(const void* in, int64_t rec, int32_t* flags, int64_t n, void* out, int32_t* idx32,
            int64_t* idx64, int64_t* h_count, hipStream_t s) {
  TempBuf tp(s);
  int rc = tp.alloc((size_t)n * 4);
  if (rc) return rc;
  int32_t* pos = (int32_t*)tp.p;
  int32_t total = 0;
  if ((rc = scan<int32_t>(flags, pos, n, 0, &total, s))) return rc;
  k_scatter_flagged<<<dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s>>>(
      (const uint8_t*)in, rec, flags, pos, n, (uint8_t*)out, idx32, idx64);
  DAXHIP_LAUNCH_CHECK("k_scatter_flagged");
  DAXHIP_CHECK(hipStreamSynchronize(s));
  *h_count = total;
  return DAX_HIP_OK;
}

template <typename T>
int unique(T* v, int64_t n, int64_t* h_new_n, hipStream_t s) {
  if (n <= 1) {
    *h_new_n = n;
    return DAX_HIP_OK;
  }
  TempBuf tp(s);
  int rc = tp.alloc((size_t)n * (4 + sizeof(T)));
  if (rc) return rc;
  int32_t* flags = (int32_t*)tp.p;
  T* tmp = (T*)(flags + n);
  k_unique_flags<T><<<dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s>>>(v, n, flags);
  DAXHIP_LAUNCH_CHECK("k_unique_flags");
  int64_t cnt = 0;
  if ((rc = compact(v, sizeof(T), flags, n, tmp, nullptr, nullptr, &cnt, s))) return rc;
  if ((rc = copy_bytes(v, tmp, (size_t)cnt * sizeof(T), s))) return rc;
  DAXHIP_CHECK(hipStreamSynchronize(s));
  *h_new_n = cnt;
  return DAX_HIP_OK;
}

}  // namespace alg
}  // namespace daxhip

using namespace daxhip;
using namespace daxhip::alg;

#define DAXHIP_S(stream) ((hipStream_t)(stream))

namespace daxhip {
// Magnitude (dax/worklet/Magnitude.h:26-33): Sqrt(dot(v, v)) with the dot
// product summed from component 0 (dax/math/VectorAnalysis.h:95-130,
// dax/Types.h:611-621), one thread per vector.
__global__ __launch_bounds__(256) void k_magnitude_vec3(const float* __restrict__ in, int64_t n,
                                                        float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
  float d = x * x;
  d = d + y * y;
  d = d + z * z;
  out[i] = sqrtf(d);
}

// Sum of n int32 values in 64 bits (integer addition: the order of the
// per-wave atomics does not change the result).
// Sum of n int32 values into *out (64-bit): 16-byte loads, four in flight
// per thread, the unaligned head and the tail by scalar loads.
__global__ __launch_bounds__(256) void k_sum_i32(const int32_t* __restrict__ in, int64_t n,
                                                 unsigned long long* __restrict__ out) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t head = min(n, (int64_t)(((16 - (reinterpret_cast<uintptr_t>(in) & 15)) & 15) / 4));
  const int4* in4 = reinterpret_cast<const int4*>(in + head);
  const int64_t n4 = (n - head) >> 2;
  long long s = 0;
  if (tid < head) s += in[tid];
  for (int64_t t = head + 4 * n4 + tid; t < n; t += stride) s += in[t];
  int64_t i = tid;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    int4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = in4[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) s += (long long)v[u].x + v[u].y + v[u].z + v[u].w;
  }
  for (; i < n4; i += stride) {
    const int4 v = in4[i];
    s += (long long)v.x + v.y + v.z + v.w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, (unsigned long long)s);
}
}  // namespace daxhip

extern "C" {

int dax_hip_reduce_sum_i32(const int32_t* d_in, int64_t n, int64_t* h_total, void* stream) {
  if (n < 0 || (n > 0 && !d_in) || !h_total) {
    daxhip::set_error("dax_hip_reduce_sum_i32: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  hipStream_t s = (hipStream_t)stream;
  daxhip::TempBuf acc(s);
  int rc;
  if ((rc = acc.alloc(sizeof(unsigned long long)))) return rc;
  DAXHIP_CHECK(hipMemsetAsync(acc.p, 0, sizeof(unsigned long long), s));
  if (n > 0) {
    const int64_t blocks = std::min<int64_t>(daxhip::ceil_div(n, 1024), 4096);
    daxhip::k_sum_i32<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        d_in, n, (unsigned long long*)acc.p);
    DAXHIP_LAUNCH_CHECK("k_sum_i32");
  }
  unsigned long long t = 0;
  DAXHIP_CHECK(hipMemcpyAsync(&t, acc.p, sizeof(t), hipMemcpyDeviceToHost, s));
  DAXHIP_CHECK(hipStreamSynchronize(s));
  *h_total = (int64_t)t;
  return DAX_HIP_OK;
}

int dax_hip_magnitude_vec3(const float* d_in, int64_t n, float* d_out, void* stream) {
  if (n < 0 || (n > 0 && (!d_in || !d_out))) {
    daxhip::set_error("dax_hip_magnitude_vec3: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  if (n == 0) return DAX_HIP_OK;
  daxhip::k_magnitude_vec3<<<dim3((unsigned)daxhip::ceil_div(n, 256)), dim3(256), 0,
                             (hipStream_t)stream>>>(d_in, n, d_out);
  DAXHIP_LAUNCH_CHECK("k_magnitude_vec3");
  return DAX_HIP_OK;
}

int dax_hip_copy(void* d_dst, const void* d_src, size_t bytes, void* stream) {
  if (bytes == 0) return DAX_HIP_OK;
  if (!d_dst || !d_src) {
    set_error("dax_hip_copy: NULL pointer");
    return DAX_HIP_BAD_VALUE;
  }
  return copy_bytes(d_dst, d_src, bytes, DAXHIP_S(stream));
}

// The public scans are synchronous whether or not the total is asked for.
#define DAXHIP_SCAN(SUF, T)                                                                   \
  int dax_hip_scan_inclusive_##SUF(const T* d_in, T* d_out, int64_t n, T* h_total,            \
                                   void* stream) {                                            \
    T total = T(0);                                                                           \
    const int rc = scan<T>(d_in, d_out, n, 1, &total, DAXHIP_S(stream));                     \
    if (!rc && h_total) *h_total = total;                                                     \
    return rc;                                                                                \
  }                                                                                           \
  int dax_hip_scan_exclusive_##SUF(const T* d_in, T* d_out, int64_t n, T* h_total,            \
                                   void* stream) {                                            \
    T total = T(0);                                                                           \
    const int rc = scan<T>(d_in, d_out, n, 0, &total, DAXHIP_S(stream));                     \
    if (!rc && h_total) *h_total = total;                                                     \
    return rc;                                                                                \
  }
DAXHIP_SCAN(i32, int32_t)
DAXHIP_SCAN(i64, int64_t)
DAXHIP_SCAN(f32, float)
DAXHIP_SCAN(f64, double)
#undef DAXHIP_SCAN

#define DAXHIP_BOUNDS(SUF, T)                                                                    \
  int dax_hip_lower_bounds_##SUF(const T* d_sorted, int64_t n, const T* d_values, int64_t m,     \
                                 int descending, int32_t* d_out32, int64_t* d_out64,             \
                                 void* stream) {                                                 \
    if (m <= 0) return DAX_HIP_OK;                                                               \
    const dim3 g((unsigned)ceil_div(m, 256)), b(256);                                           \
    if (descending)                                                                              \
      k_bounds<T, false, true><<<g, b, 0, DAXHIP_S(stream)>>>(d_sorted, n, d_values, m, d_out32, \
                                                               d_out64);                         \
    else                                                                                         \
      k_bounds<T, false, false><<<g, b, 0, DAXHIP_S(stream)>>>(d_sorted, n, d_values, m,        \
                                                                d_out32, d_out64);               \
    DAXHIP_LAUNCH_CHECK("k_bounds");                                                             \
    return DAX_HIP_OK;                                                                           \
  }                                                                                              \
  int dax_hip_upper_bounds_##SUF(const T* d_sorted, int64_t n, const T* d_values, int64_t m,     \
                                 int descending, int32_t* d_out32, int64_t* d_out64,             \
                                 void* stream) {                                                 \
    if (m <= 0) return DAX_HIP_OK;                                                               \
    const dim3 g((unsigned)ceil_div(m, 256)), b(256);                                           \
    if (descending)                                                                              \
      k_bounds<T, true, true><<<g, b, 0, DAXHIP_S(stream)>>>(d_sorted, n, d_values, m, d_out32,  \
                                                              d_out64);                          \
    else                                                                                         \
      k_bounds<T, true, false><<<g, b, 0, DAXHIP_S(stream)>>>(d_sorted, n, d_values, m, d_out32, \
                                                               d_out64);                         \
    DAXHIP_LAUNCH_CHECK("k_bounds");                                                             \
    return DAX_HIP_OK;                                                                           \
  }
DAXHIP_BOUNDS(i32, int32_t)
DAXHIP_BOUNDS(i64, int64_t)
DAXHIP_BOUNDS(f32, float)
DAXHIP_BOUNDS(f64, double)
#undef DAXHIP_BOUNDS

#define DAXHIP_SORT(SUF, T)                                                                     \
  int dax_hip_sort_##SUF(T* d_values, int64_t n, int descending, void* stream) {                \
    return sort<T>(d_values, n, descending, nullptr, DAXHIP_S(stream));                         \
  }                                                                                             \
  int dax_hip_sort_by_key_##SUF(T* d_keys, void* d_values, size_t value_bytes, int64_t n,       \
                                int descending, void* stream) {                                 \
    return sort_by_key<T>(d_keys, d_values, value_bytes, n, descending, DAXHIP_S(stream));      \
  }                                                                                             \
  int dax_hip_unique_##SUF(T* d_values, int64_t n, int64_t* h_new_n, void* stream) {           \
    return unique<T>(d_values, n, h_new_n, DAXHIP_S(stream));                                   \
  }
DAXHIP_SORT(i32, int32_t)
DAXHIP_SORT(i64, int64_t)
DAXHIP_SORT(u32, uint32_t)
DAXHIP_SORT(u64, uint64_t)
DAXHIP_SORT(f32, float)
DAXHIP_SORT(f64, double)
#undef DAXHIP_SORT

int dax_hip_stream_compact(const void* d_in, size_t value_bytes, const void* d_stencil,
                           int stencil_type, int64_t n, void* d_out, int32_t* d_indices32,
                           int64_t* d_indices64, int64_t* h_count, void* stream) {
  if (n <= 0) {
    *h_count = 0;
    return DAX_HIP_OK;
  }
  if (stencil_type < 0 || stencil_type > 4 || !d_stencil || (d_in && !d_out)) {
    set_error("dax_hip_stream_compact: bad arguments");
    return DAX_HIP_BAD_VALUE;
  }
  hipStream_t s = DAXHIP_S(stream);
  TempBuf tf(s);
  int rc = tf.alloc((size_t)n * 4);
  if (rc) return rc;
  k_stencil_flags<<<dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s>>>(d_stencil, stencil_type, n,
                                                                         (int32_t*)tf.p);
  DAXHIP_LAUNCH_CHECK("k_stencil_flags");
  return compact(d_in, (int64_t)value_bytes, (int32_t*)tf.p, n, d_out, d_indices32, d_indices64,
                 h_count, s);
}

}  // extern "C"
