// Neighbourhood PCA shared by the normals kernels: the centred FP64
// covariance of a neighbourhood, the 3x3 eigensolve with the FP64 fallback
// rule, and output writers.
//
// Reference: estimate_normals<T> (pptk/processing/estimate_normals/
// estimate_normals.cpp:64-111): X -= colwise mean, C = X^T X / n,
// Eigen::SelfAdjointEigenSolver, eigenvalues ascending.
#pragma once
#include "eigen3.hpp"
#include "traverse.hpp"

namespace pk {

constexpr double kGapThreshold = 1e-3;

template <typename D>
__device__ __forceinline__ D qnan() { return (D)__builtin_nan(""); }

template <typename D, int NE>
__device__ __forceinline__ void write_nan(int64_t row, D* evecs, D* evals, int32_t* sizes) {
  const D nn = qnan<D>();
  if (evecs)
#pragma unroll
    for (int i = 0; i < 3 * NE; ++i) evecs[row * 3 * NE + i] = nn;
  if (evals)
#pragma unroll
    for (int i = 0; i < NE; ++i) evals[row * NE + i] = nn;
  if (sizes) sizes[row] = 0;
}

template <typename D, typename F, int NE>
__device__ __forceinline__ void write_eig(int64_t row, const F (&w)[3], const F (&v)[3][3],
                                          int n, D* evecs, D* evals, int32_t* sizes) {
  if (evecs) {
#pragma unroll
    for (int j = 0; j < NE; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) evecs[row * 3 * NE + j * 3 + c] = (D)v[c][j];
  }
  if (evals) {
#pragma unroll
    for (int j = 0; j < NE; ++j) evals[row * NE + j] = (D)w[j];
  }
  if (sizes) sizes[row] = n;
}

// Covariance of the neighbourhood enumerated by `each(f)` (f receives tree
// positions), centred, divided by n, in FP64 whatever the point type: the
// offsets from q, their mean, then the centred products, each step rounded
// in FP64 (the library is built without contraction, so no step is fused).
// On this hardware FP64 adds and multiplies issue at the FP32 rate, so this
// is cheaper than compensated FP32 sums and exact to FP64 rounding.
template <typename D, typename Each>
__device__ __forceinline__ void knn_covariance(const TreeView<D, 3>& tv, const D* q, int n,
                                               Each each, double (&cd)[3][3]) {
  double s[3] = {0.0, 0.0, 0.0};
  each([&](int32_t p) {
    const PointRec<D, 3> r = tv.pts[p];
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] = s[a] + ((double)r.c[a] - (double)q[a]);
  });
  double mu[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) mu[a] = s[a] / (double)n;
  double c[6] = {0, 0, 0, 0, 0, 0};
  each([&](int32_t p) {
    const PointRec<D, 3> r = tv.pts[p];
    double e[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) e[a] = ((double)r.c[a] - (double)q[a]) - mu[a];
    c[0] = c[0] + e[0] * e[0];
    c[1] = c[1] + e[0] * e[1];
    c[2] = c[2] + e[0] * e[2];
    c[3] = c[3] + e[1] * e[1];
    c[4] = c[4] + e[1] * e[2];
    c[5] = c[5] + e[2] * e[2];
  });
  const double inv = 1.0 / (double)n;
  cd[0][0] = c[0] * inv; cd[0][1] = c[1] * inv; cd[0][2] = c[2] * inv;
  cd[1][0] = c[1] * inv; cd[1][1] = c[3] * inv; cd[1][2] = c[4] * inv;
  cd[2][0] = c[2] * inv; cd[2][1] = c[4] * inv; cd[2][2] = c[5] * inv;
}

// True if the FP32 eigenvalues (ascending) leave a gap below the threshold
// that the written eigenvectors depend on: (l1 - l0) / l2 for the normal,
// also (l2 - l1) / l2 when all three eigenvectors are written.
template <int NE>
__device__ __forceinline__ bool needs_fp64(const float (&w)[3]) {
  const float l2 = w[2];
  bool fb = !(l2 > 0.f) || ((double)(w[1] - w[0]) < kGapThreshold * (double)l2);
  if (NE == 3) fb = fb || ((double)(w[2] - w[1]) < kGapThreshold * (double)l2);
  return fb;
}

// PCA of one row whose neighbourhood `each` enumerates (tree positions, the
// same order every time it is called): the FP64 covariance, then
//   * PPTK_PREC_FP64: the FP64 Jacobi solve;
//   * PPTK_PREC_MIXED: the FP32 Jacobi solve of the covariance rounded to
//     float, and, for a row whose FP32 eigengap is below the threshold, the
//     FP64 solve of the same FP64 covariance. A gap g amplifies a covariance
//     error e into an eigenvector error of about e / g; such rows equal the
//     FP64 mode's rows bit for bit.
// Returns true if the FP64 solve was taken in the mixed mode.
template <typename D, int PREC, int NE, typename Each>
__device__ __forceinline__ bool pca_row(const TreeView<D, 3>& tv, const D* q, int n, Each each,
                                        int64_t row, D* evecs, D* evals, int32_t* sizes) {
  double cd[3][3];
  knn_covariance<D>(tv, q, n, each, cd);
  if constexpr (PREC == PPTK_PREC_MIXED) {
    float cf[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) cf[i][j] = (float)cd[i][j];
    float w[3], v[3][3];
    eig_sym3<float>(cf, w, v);
    if (!needs_fp64<NE>(w)) {
      write_eig<D, float, NE>(row, w, v, n, evecs, evals, sizes);
      return false;
    }
  }
  double w[3], v[3][3];
  eig_sym3<double>(cd, w, v);
  write_eig<D, double, NE>(row, w, v, n, evecs, evals, sizes);
  return PREC == PPTK_PREC_MIXED;
}

__device__ __forceinline__ void count_fallbacks(bool fb, unsigned long long* nfb) {
  const uint64_t b = __ballot(fb);
  const uint64_t act = __ballot(true);
  if (b && (int)(threadIdx.x & 63) == __builtin_ctzll(act))
    atomicAdd(nfb, (unsigned long long)__popcll(b));
}

}  // namespace pk
