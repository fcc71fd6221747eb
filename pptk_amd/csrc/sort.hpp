#pragma once
#include "common.hpp"

namespace pk {

// Stable LSD radix sort of (key, value) pairs on key bits [0, end_bit),
// in place. Synchronizes `s` before returning.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, hipStream_t s);

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total.
// Synchronizes `s` before returning.
template <typename Tin, typename Tout>
void exclusive_scan(const Tin* in, Tout* out, int64_t n, hipStream_t s);

}  // namespace pk
