#pragma once
#include "common.hpp"

namespace pk {

// Preallocated temporaries of radix_sort_pairs_async for up to n_cap keys.
struct SortWorkspace {
  DevBuf kalt, valt, hist, offs, ssum;
  int64_t n_cap = 0;
  void reserve(int64_t n);
};

// As radix_sort_pairs, with no allocation and no synchronisation (safe to
// record into a hip graph).
void radix_sort_pairs_async(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit,
                            SortWorkspace& ws, hipStream_t s);

// In-place exclusive scan of n uint32 values by one block; data[n] receives
// the total. Asynchronous.
void scan_u32_inplace_async(uint32_t* data, int64_t n, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on key bits [0, end_bit),
// in place. Synchronizes `s` before returning.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, hipStream_t s);

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total.
// Synchronizes `s` before returning.
template <typename Tin, typename Tout>
void exclusive_scan(const Tin* in, Tout* out, int64_t n, hipStream_t s);

}  // namespace pk
