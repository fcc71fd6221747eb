#pragma once
#include "tree.hpp"

namespace pk {

// Query set resolved to device pointers (see pptk_queries in pptk_amd.h).
struct PreparedQueries {
  int mode = 0;        // QueryMode
  int64_t m = 0;       // threads to launch
  int64_t nrows = 0;   // output rows
  DevInput list;
  const int32_t* dlist = nullptr;
  DevBuf conv;
  const void* dpts = nullptr;
  const void* src = nullptr;   // the caller's query array (order cache key)
  // Processing order (thread t serves query order[t]) so that neighbouring
  // lanes get nearby queries; null = input order.
  std::shared_ptr<DevBuf> order_buf;
  const int32_t* order = nullptr;
};

// Device memory a call may use for its temporaries instead of allocating
// (the allocation-free entry point): the query order and its sort, and the
// heaps of k > 32 searches.
struct QueryScratch {
  void* base = nullptr;  // 256-byte aligned
  size_t bytes = 0;
};

// With `scratch`, the query order lives in it and the tree's order cache is
// neither read nor written; queries must already be in device memory.
template <typename D>
void prepare_queries(const pptk_tree* t, const pptk_queries* q, hipStream_t s, PreparedQueries& pq,
                     const QueryScratch* scratch = nullptr);

// Asynchronous device fill of n elements with v.
template <typename T>
void fill_async(T* p, int64_t n, T v, hipStream_t s);

}  // namespace pk
