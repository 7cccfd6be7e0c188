// Shared helpers of the gfx950 library: status codes, the thread-local error
// string behind dax_hip_last_error(), and wave64 primitives (DPP scans,
// ballots) used by the scan, compaction and marching-cubes kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/dax_hip.h"

namespace daxhip {

void set_error(const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

// Temporaries of the library's own launches: stream-ordered allocations from
// a per-device memory pool that keeps its memory between calls (no device
// synchronisation on free, no fresh mapping per call).
int temp_alloc(void** p, size_t bytes, hipStream_t s);
void temp_free(void* p, hipStream_t s);

// RAII holder of one temporary, released in stream order.
struct TempBuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  explicit TempBuf(hipStream_t stream) : s(stream) {}
  ~TempBuf() { temp_free(p, s); }
  TempBuf(const TempBuf&) = delete;
  TempBuf& operator=(const TempBuf&) = delete;
  int alloc(size_t bytes) { return temp_alloc(&p, bytes, s); }
};

#define DAXHIP_CHECK(expr)                                  \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ::daxhip::check_hip(_e, #expr); \
  } while (0)

#define DAXHIP_LAUNCH_CHECK(what)                            \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) return ::daxhip::check_hip(_e, what); \
  } while (0)

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP row shifts
// and the GFX9 row broadcasts (row_shr:1/2/4/8, row_bcast:15, row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return v;
}

// Value of lane 63 broadcast to the whole wave.
__device__ __forceinline__ uint32_t wave_last(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return wave_last(wave_incl_scan(v)); }

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  // two 32-bit halves scanned separately, recombined with the carry of the
  // low half accumulated in 64 bits
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  // sum of lows can exceed 32 bits: split further into 16-bit pieces
  uint32_t l0 = lo & 0xffffu, l1 = lo >> 16;
  uint64_t s = (uint64_t)wave_sum(l0) + ((uint64_t)wave_sum(l1) << 16) +
               ((uint64_t)wave_sum(hi) << 32);
  return s;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__host__ __device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

__host__ __device__ __forceinline__ size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace daxhip
