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

template <typename D>
void prepare_queries(const pptk_tree* t, const pptk_queries* q, hipStream_t s, PreparedQueries& pq);

// Asynchronous device fill of n elements with v.
template <typename T>
void fill_async(T* p, int64_t n, T v, hipStream_t s);

}  // namespace pk
