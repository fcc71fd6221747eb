// TEST INFRASTRUCTURE ONLY. C entry points around pptk's own k-d tree
// (pointkd::KdTree from pptk/kdtree/src/kdtree.h, compiled from the
// reference sources where they lie, without TBB: the reference then builds
// serially, kdtree-impl.h:797-801). The batch loop over queries is
// parallelised here with OpenMP, standing in for the reference's
// tbb::parallel_for (kdtree-impl.h:916-923, 963-970, 1014-1020, 1057-1063);
// each query calls the reference's single-query method unchanged.
//
// Built by oracle/Makefile into oracle/_ref/libpptk_ref.so when the
// reference tree is present; used by tests and the CPU baseline only.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "kdtree.h"

namespace {

template <typename T>
struct RefTree {
  pointkd::KdTree<T, 3>* tree;
};

template <typename T>
void* build(const T* pts, int n, int leaf) {
  pointkd::BuildParams bp;
  bp.num_proc = 1;
  if (leaf > 0) bp.max_leaf_size = leaf;
  return new pointkd::KdTree<T, 3>(pts, n, bp);
}

template <typename T>
void knn(void* h, int mode, const T* qpts, int m, int k, double r, int* out_idx, int* out_cnt,
         int nthreads) {
  auto* t = (pointkd::KdTree<T, 3>*)h;
  typedef typename pointkd::KdTree<T, 3>::DistT DistT;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int i = 0; i < m; ++i) {
    pointkd::Indices res;
    if (mode == 0) t->KNearestNeighborsSelf(res, i, k, (DistT)r);
    else t->KNearestNeighbors(res, qpts + (size_t)i * 3, k, (DistT)r);
    int* row = out_idx + (size_t)i * k;
    for (int s = 0; s < k; ++s) row[s] = s < (int)res.size() ? res[s] : -1;
    out_cnt[i] = (int)res.size();
  }
}

template <typename T>
void rnear(void* h, int mode, const T* qpts, int m, double r, long long* offsets, int* out_idx,
           int nthreads) {
  auto* t = (pointkd::KdTree<T, 3>*)h;
  typedef typename pointkd::KdTree<T, 3>::DistT DistT;
  const bool fill = out_idx != nullptr;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int i = 0; i < m; ++i) {
    pointkd::Indices res;
    if (mode == 0) t->RNearNeighborsSelf(res, i, (DistT)r);
    else t->RNearNeighbors(res, qpts + (size_t)i * 3, (DistT)r);
    std::sort(res.begin(), res.end());
    if (fill) std::copy(res.begin(), res.end(), out_idx + offsets[i]);
    else offsets[i + 1] = (long long)res.size();
  }
  if (!fill) {
    offsets[0] = 0;
    for (int i = 0; i < m; ++i) offsets[i + 1] += offsets[i];
  }
}

}  // namespace

extern "C" {
void* ref_build_f32(const float* pts, int n, int leaf) { return build<float>(pts, n, leaf); }
void* ref_build_f64(const double* pts, int n, int leaf) { return build<double>(pts, n, leaf); }
void ref_free_f32(void* h) { delete (pointkd::KdTree<float, 3>*)h; }
void ref_free_f64(void* h) { delete (pointkd::KdTree<double, 3>*)h; }
int ref_num_points_f32(void* h) { return ((pointkd::KdTree<float, 3>*)h)->num_points(); }
int ref_num_points_f64(void* h) { return ((pointkd::KdTree<double, 3>*)h)->num_points(); }
// mode 0: self queries on original indices 0..m-1; mode 1: query points.
void ref_knn_f32(void* h, int mode, const float* q, int m, int k, double r, int* oi, int* oc,
                 int nt) {
  knn<float>(h, mode, q, m, k, r, oi, oc, nt);
}
void ref_knn_f64(void* h, int mode, const double* q, int m, int k, double r, int* oi, int* oc,
                 int nt) {
  knn<double>(h, mode, q, m, k, r, oi, oc, nt);
}
void ref_rnear_f32(void* h, int mode, const float* q, int m, double r, long long* off, int* oi,
                   int nt) {
  rnear<float>(h, mode, q, m, r, off, oi, nt);
}
void ref_rnear_f64(void* h, int mode, const double* q, int m, double r, long long* off, int* oi,
                   int nt) {
  rnear<double>(h, mode, q, m, r, off, oi, nt);
}
}
