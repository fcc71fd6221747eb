// Neighbourhood PCA shared by the normals kernels: double-float sums, the
// centred covariance of a neighbourhood, the 3x3 eigensolve with the FP64
// fallback rule, and output writers.
//
// Reference: estimate_normals<T> (pptk/processing/estimate_normals/
// estimate_normals.cpp:64-111): X -= colwise mean, C = X^T X / n,
// Eigen::SelfAdjointEigenSolver, eigenvalues ascending.
#pragma once
#include "eigen3.hpp"
#include "traverse.hpp"

namespace pk {

constexpr double kGapThreshold = 1e-3;

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Double-float running sum.
struct DF {
  float hi, lo;
  __device__ __forceinline__ void zero() { hi = 0.f; lo = 0.f; }
  __device__ __forceinline__ void add(float x) {
    float s, e;
    two_sum(hi, x, s, e);
    hi = s;
    lo = lo + e;
  }
  __device__ __forceinline__ void add_prod(float x, float y) {
    const float p = x * y;
    const float pe = fmaf(x, y, -p);  // exact product error
    add(p);
    lo = lo + pe;
  }
  __device__ __forceinline__ double value() const { return (double)hi + (double)lo; }
};

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
// positions), centred, divided by n. Two passes: mean of the offsets from
// q, then the centred products.
template <typename D, int PREC, typename Each>
__device__ __forceinline__ void knn_covariance(const TreeView<D, 3>& tv, const D* q, int n,
                                               Each each, double (&cd)[3][3], float (&cf)[3][3]) {
  if constexpr (PREC == PPTK_PREC_MIXED) {
    DF s[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a].zero();
    each([&](int32_t p) {
      const PointRec<D, 3> r = tv.pts[p];
#pragma unroll
      for (int a = 0; a < 3; ++a) s[a].add((float)(r.c[a] - q[a]));
    });
    float mu[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) mu[a] = (float)(s[a].value() / (double)n);
    DF c[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i].zero();
    each([&](int32_t p) {
      const PointRec<D, 3> r = tv.pts[p];
      float e[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) e[a] = (float)(r.c[a] - q[a]) - mu[a];
      c[0].add_prod(e[0], e[0]);
      c[1].add_prod(e[0], e[1]);
      c[2].add_prod(e[0], e[2]);
      c[3].add_prod(e[1], e[1]);
      c[4].add_prod(e[1], e[2]);
      c[5].add_prod(e[2], e[2]);
    });
    const double inv = 1.0 / (double)n;
    const double v[6] = {c[0].value() * inv, c[1].value() * inv, c[2].value() * inv,
                         c[3].value() * inv, c[4].value() * inv, c[5].value() * inv};
    cd[0][0] = v[0]; cd[0][1] = v[1]; cd[0][2] = v[2];
    cd[1][0] = v[1]; cd[1][1] = v[3]; cd[1][2] = v[4];
    cd[2][0] = v[2]; cd[2][1] = v[4]; cd[2][2] = v[5];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) cf[i][j] = (float)cd[i][j];
  } else {
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
    (void)cf;
  }
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
// same order every time it is called). PPTK_PREC_FP64: FP64 covariance and
// solve. PPTK_PREC_MIXED: FP32 covariance centred on q with compensated
// sums and an FP32 solve; a row whose FP32 eigengap is below the threshold
// recomputes its covariance in FP64 from the same neighbours and solves in
// FP64. The FP32-centred covariance carries ~1e-7 relative rounding (each
// centred offset is rounded to float), and a gap g amplifies a covariance
// error e into an eigenvector error of about e / g, so the FP64 solve alone
// would not give such rows FP64 accuracy; recomputed, they equal the FP64
// mode's rows bit for bit. Returns true if the fallback was taken.
template <typename D, int PREC, int NE, typename Each>
__device__ __forceinline__ bool pca_row(const TreeView<D, 3>& tv, const D* q, int n, Each each,
                                        int64_t row, D* evecs, D* evals, int32_t* sizes) {
  double cd[3][3];
  float cf[3][3];
  if constexpr (PREC == PPTK_PREC_MIXED) {
    knn_covariance<D, PPTK_PREC_MIXED>(tv, q, n, each, cd, cf);
    float w[3], v[3][3];
    eig_sym3<float>(cf, w, v);
    if (!needs_fp64<NE>(w)) {
      write_eig<D, float, NE>(row, w, v, n, evecs, evals, sizes);
      return false;
    }
  }
  knn_covariance<D, PPTK_PREC_FP64>(tv, q, n, each, cd, cf);
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
