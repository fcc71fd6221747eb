// Symmetric 3x3 eigensolver: cyclic Jacobi (Golub & Van Loan, Matrix
// Computations, section 8.5: cyclic-by-row sweeps of the 2x2 symmetric Schur
// step), using Rutishauser's update form a' = a -+ s*(b +- tau*a),
// tau = s/(1+c), for the off-pivot entries. Each rotation zeroes its pivot
// exactly; sweeps continue until the off-diagonal mass vanishes relative to
// the diagonal. Eigenvectors come out orthonormal to working precision even
// for repeated eigenvalues, which the closed-form (trigonometric) solution
// does not guarantee.
//
// Output convention: eigenvalues ascending, column j of V is the j-th
// eigenvector, each eigenvector signed so its largest-magnitude component is
// positive (first such component on ties). For a diagonal matrix V is a
// permutation of the identity, so the eigenvector of the largest eigenvalue
// of diag(l, 0, 0) is +e_x (pptk/tests/test_estimate_normals.py:46,87).
#pragma once
#include <hip/hip_runtime.h>

namespace pk {

template <typename F> __device__ __forceinline__ F fabs_(F x) { return x < F(0) ? -x : x; }
// Rotation parameters in float use the hardware square root, reciprocal and
// reciprocal square root (about 1 ulp each): a rotation's angle only steers
// convergence, and c*c + s*s stays 1 to rounding, so the float solve keeps
// its accuracy without IEEE division sequences. Double stays exact.
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_(double x) { return 1.0 / x; }
__device__ __forceinline__ float rsq_(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq_(double x) { return 1.0 / sqrt(x); }

template <typename F, int P, int Q>
__device__ __forceinline__ void jacobi_rotate(F (&a)[3][3], F (&v)[3][3]) {
  constexpr int R = 3 - P - Q;
  const F apq = a[P][Q];
  if (apq == F(0)) return;
  const F theta = (a[Q][Q] - a[P][P]) * rcp_(F(2) * apq);
  const F at = fabs_(theta);
  F t;
  if (at > F(1e15)) {
    t = F(0.5) * rcp_(theta);
  } else {
    t = rcp_(at + sqrt_(theta * theta + F(1)));
    if (theta < F(0)) t = -t;
  }
  const F c = rsq_(t * t + F(1));
  const F s = t * c;
  const F tau = s * rcp_(F(1) + c);
  a[P][P] = a[P][P] - t * apq;
  a[Q][Q] = a[Q][Q] + t * apq;
  a[P][Q] = F(0);
  a[Q][P] = F(0);
  const F g = a[R][P], h = a[R][Q];
  const F rp = g - s * (h + g * tau);
  const F rq = h + s * (g - h * tau);
  a[R][P] = rp; a[P][R] = rp;
  a[R][Q] = rq; a[Q][R] = rq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const F vg = v[i][P], vh = v[i][Q];
    v[i][P] = vg - s * (vh + vg * tau);
    v[i][Q] = vh + s * (vg - vh * tau);
  }
}

// a: symmetric input (destroyed). w: eigenvalues ascending. v: columns =
// eigenvectors.
template <typename F>
__device__ __forceinline__ void eig_sym3(F (&a)[3][3], F (&w)[3], F (&v)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) v[i][j] = i == j ? F(1) : F(0);
  for (int sweep = 0; sweep < 12; ++sweep) {
    const F off = fabs_(a[0][1]) + fabs_(a[0][2]) + fabs_(a[1][2]);
    const F diag = fabs_(a[0][0]) + fabs_(a[1][1]) + fabs_(a[2][2]);
    if (!(off > diag * F(1e-30))) break;  // also exits on NaN
    jacobi_rotate<F, 0, 1>(a, v);
    jacobi_rotate<F, 0, 2>(a, v);
    jacobi_rotate<F, 1, 2>(a, v);
  }
  w[0] = a[0][0]; w[1] = a[1][1]; w[2] = a[2][2];
  // Stable ascending sort of (w, column) pairs.
  auto cswap = [&](int i, int j) {
    if (w[i] > w[j]) {
      const F tw = w[i]; w[i] = w[j]; w[j] = tw;
#pragma unroll
      for (int r = 0; r < 3; ++r) { const F tv = v[r][i]; v[r][i] = v[r][j]; v[r][j] = tv; }
    }
  };
  cswap(0, 1);
  cswap(1, 2);
  cswap(0, 1);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    int imax = 0;
    F amax = fabs_(v[0][j]);
#pragma unroll
    for (int r = 1; r < 3; ++r) {
      const F x = fabs_(v[r][j]);
      if (x > amax) { amax = x; imax = r; }
    }
    F sel = v[0][j];
    if (imax == 1) sel = v[1][j];
    if (imax == 2) sel = v[2][j];
    if (sel < F(0)) {
#pragma unroll
      for (int r = 0; r < 3; ++r) v[r][j] = -v[r][j];
    }
  }
}

}  // namespace pk
