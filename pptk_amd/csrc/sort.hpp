#pragma once
#include "common.hpp"

namespace pk {

// Temporaries of radix_sort_pairs_async for up to n_cap keys: either owned
// pool buffers (reserve) or a slice of caller memory (bind).
struct SortWorkspace {
  uint64_t* kalt = nullptr;
  uint32_t* valt = nullptr;
  uint32_t* hist = nullptr;
  uint32_t* offs = nullptr;
  uint32_t* ssum = nullptr;
  int64_t n_cap = 0;
  // Allocates owned buffers for n keys (no-op if n <= n_cap).
  void reserve(int64_t n);
  // Bytes of caller memory bind() needs for n keys.
  static size_t bytes_for(int64_t n);
  // Points the workspace at `base` (256-byte aligned, bytes_for(n) bytes),
  // dropping any owned buffers.
  void bind(void* base, int64_t n);
  // Device bytes held (owned or bound).
  size_t bytes() const { return n_cap > 0 ? bytes_for(n_cap) : 0; }

 private:
  DevBuf owned_;
};

// As radix_sort_pairs, with no allocation and no synchronisation (safe to
// record into a hip graph). With keys_out false only the values are sorted
// in place; the keys are left in an unspecified order (an odd number of
// passes then copies one array back instead of two).
void radix_sort_pairs_async(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit,
                            SortWorkspace& ws, hipStream_t s, bool keys_out = true);

// In-place exclusive scan of n uint32 values by one block; data[n] receives
// the total. Asynchronous.
void scan_u32_inplace_async(uint32_t* data, int64_t n, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on key bits [0, end_bit),
// in place; asynchronous on `s` (its temporaries return to the pool fenced
// on s).
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, hipStream_t s);

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total;
// asynchronous on `s`.
template <typename Tin, typename Tout>
void exclusive_scan(const Tin* in, Tout* out, int64_t n, hipStream_t s);

}  // namespace pk
