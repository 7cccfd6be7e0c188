// Internal entry points of the adapter algorithms (algorithms.hip) used by
// other parts of the library.
#pragma once
#include "common.h"

namespace daxhip {
namespace alg {

// Ascending, stream-ordered LSD radix sort of n unsigned keys that are all
// below 2^bits (ceil(bits / 8) passes); d_vals, when given, holds n int32
// values that move with their keys (stable).
int sort_u64_bits(uint64_t* d_keys, int64_t n, int bits, hipStream_t s,
                  int32_t* d_vals = nullptr);

// Exclusive scan of n int32 values (in place allowed), stream-ordered: no
// total is read back.
int scan_exclusive_i32_async(const int32_t* d_in, int32_t* d_out, int64_t n, hipStream_t s);

}  // namespace alg
}  // namespace daxhip
