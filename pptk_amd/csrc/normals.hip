// Fused PCA normals: one lane per output row runs the neighbourhood search on
// the tree (kNN including the point itself, or r-near for k == -1), forms the
// covariance of the centred neighbourhood, C = X^T X / n, and solves the 3x3
// symmetric eigenproblem in registers.
//
// Reference: estimate_normals<T> (pptk/processing/estimate_normals/
// estimate_normals.cpp:64-111): neighbourhood via KNearestNeighbors(q, k,
// (T)r) or RNearNeighbors(q, (T)r), X -= colwise mean, C = X^T X / n,
// Eigen::SelfAdjointEigenSolver, eigenvalues ascending.
//
// Precision modes (both form the covariance in FP64: neighbour offsets
// e = p - q relative to the query point, their mean, the centred products;
// FP64 adds and multiplies issue at the FP32 rate on this hardware):
//   MIXED : FP32 Jacobi of the covariance rounded to float; rows whose
//           eigengap (l1 - l0) / l2 < 1e-3 (or, with all eigenvectors,
//           (l2 - l1) / l2 < 1e-3) are re-solved in FP64 from the same
//           covariance, so they equal the FP64 mode's rows.
//   FP64  : Jacobi in double.
#include <algorithm>
#include <type_traits>

#include "coop.hpp"
#include "pca.hpp"
#include "query.hpp"
#include "traverse.hpp"
#include "tree.hpp"

namespace pk {

namespace {

constexpr int kBlock = 256;

// Resolves row `t`: tree-order mode (rows == nullptr, t = tree position) or
// explicit rows (t = output row, rows[t] = original index).
template <typename D>
__device__ __forceinline__ void resolve_row(const TreeView<D, 3>& tv, const int32_t* rows, int64_t t,
                                            D* q, int64_t& row, int32_t& pos, bool& ok) {
  ok = true;
  pos = -1;
  if (!rows) {
    if (tv.n_dev && t >= *tv.n_dev) {
      ok = false;
      row = -1;
      return;
    }
    const PointRec<D, 3> r = tv.pts[t];
    row = r.orig;
    pos = (int32_t)t;
#pragma unroll
    for (int d = 0; d < 3; ++d) q[d] = r.c[d];
  } else {
    row = t;
    const int32_t o = rows[t];
    const int32_t p = (o >= 0 && (int64_t)o < tv.n_input) ? tv.fwd[o] : -1;
    if (p < 0) {
      ok = false;
    } else {
      pos = p;
      const PointRec<D, 3> r = tv.pts[p];
#pragma unroll
      for (int d = 0; d < 3; ++d) q[d] = r.c[d];
    }
  }
}

template <typename D, int K, int PREC, int NE>
__global__ __launch_bounds__(kBlock) void normals_knn_kernel(TreeView<D, 3> tv,
                                                             const int32_t* __restrict__ rows,
                                                             int64_t m, int k, D r2,
                                                             D* __restrict__ evecs,
                                                             D* __restrict__ evals,
                                                             int32_t* __restrict__ sizes,
                                                             unsigned long long* __restrict__ nfb) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  D q[3];
  int64_t row;
  int32_t pos;
  bool ok;
  resolve_row<D>(tv, rows, t, q, row, pos, ok);
  if (row < 0) return;
  // Positions are not carried through the top-k insertions: after the
  // search each neighbour's tree position is read from the forward map.
  TopK<D, K, false> st;
  st.init(k);
  if (ok) knn_search_from_leaf<D, 3>(tv, pos, q, r2, -1, k, st);
  int n = 0;
#pragma unroll
  for (int s = 0; s < K; ++s) n += (s >= K - k && st.found(s)) ? 1 : 0;
  bool fb = false;
  if (n == 0) {
    write_nan<D, NE>(row, evecs, evals, sizes);
  } else {
    auto each = [&](auto f) {
#pragma unroll
      for (int s = 0; s < K; ++s)
        if (s >= K - k && st.found(s)) f(tv.fwd[st.idx(s)]);
    };
    fb = pca_row<D, PREC, NE>(tv, q, n, each, row, evecs, evals, sizes);
  }
  count_fallbacks(fb, nfb);
}

// Tree-order rows (rows == nullptr) of a float tree: the same as
// normals_knn_kernel with the wave-cooperative search (coop.hpp). Lanes past
// the end stay in the wave as inactive lanes: the search needs every lane.
template <int K, int PREC, int NE>
__global__ __launch_bounds__(kBlock, K <= 16 ? 8 : 1) void normals_coop_kernel(TreeView<float, 3> tv, int64_t m,
                                                              int k, float r2,
                                                              float* __restrict__ evecs,
                                                              float* __restrict__ evals,
                                                              int32_t* __restrict__ sizes,
                                                              unsigned long long* __restrict__ nfb,
                                                              int xrun) {
  __shared__ uint64_t lds[kBlock / 64 * coop_wave_bytes<K, 3, K>() / 8];
  // the wave's first position and LDS area are scalars; a lane's position
  // is recomputed after the search rather than held through it
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t t0 = xcd_block(blockIdx.x, gridDim.x, xrun) * kBlock + wave * 64;
  int64_t t = t0 + __lane_id();
  float q[3] = {0.0f, 0.0f, 0.0f};
  bool active = false;
  if (t < m && !(tv.n_dev && t >= *tv.n_dev)) {
    const PointRec<float, 3> r = tv.pts[t];
    active = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) q[d] = r.c[d];
  }
  TopK<float, K, false> st;
  st.init(k);
  uint64_t* wl = lds + wave * (coop_wave_bytes<K, 3, K>() / 8);
  knn_coop_self<3, K>(tv, active, (int32_t)t, -1, q, r2, k, false, st, wl);
  if (!active) return;
  t = t0 + __lane_id();
  const int32_t row = tv.orig[t];
  // The neighbours' tree positions, in key order, go to this lane's LDS
  // buffer: the keys are dead before the covariance's two passes.
  int32_t* pos = coop_positions<3>(wl);
  int n = 0;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (s >= K - k && st.found(s)) {
      pos[n * 64] = tv.fwd[st.idx(s)];
      ++n;
    }
  }
  bool fb = false;
  if (n == 0) {
    write_nan<float, NE>(row, evecs, evals, sizes);
  } else {
    auto each = [&](auto f) {
      for (int j = 0; j < n; ++j) f(pos[j * 64]);
    };
    fb = pca_row<float, PREC, NE>(tv, q, n, each, row, evecs, evals, sizes);
  }
  count_fallbacks(fb, nfb);
}

// The lane's index in its wave from an asm block, which the compiler cannot
// merge with an earlier lane id: a value rebuilt from it after a search is
// not kept in a register (or spilled) through the search.
__device__ __forceinline__ int32_t fresh_lane_id() {
  int32_t v;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(v));
  return v;
}

// Self-query kNN rows and normals of a float tree from ONE search per point
// (pptk_knn_normals). The wave-cooperative search finds the kk = max(k, nk)
// nearest OTHER points within the larger of the two bounds, in (d, index)
// order. The kNN row is the first k of them with d < dmax^2: the k nearest
// with d < dmax^2, as pptk_knn writes for PPTK_Q_ALL (rows 0 .. m-1 by
// original index). The normals neighbourhood, the nk nearest including the
// point itself with d < r^2 as normals_coop_kernel searches, is that list
// with the point merged in at its (0, own index) rank, cut to d < r^2 and to
// nk entries: the same set in the same order, so the same sums. nk others
// (not nk - 1) are needed when nk or more points coincide with the point at
// lower indices: then the point itself is not among its own nk nearest.
template <int K, int PREC, int NE>
__global__ __launch_bounds__(kBlock, K <= 16 ? 8 : 1) void knn_normals_coop_kernel(
    TreeView<float, 3> tv, int64_t m, int k, int kk, int nk, float r2s, float r2k, float r2n,
    int32_t* __restrict__ out_idx, int32_t* __restrict__ out_cnt, float* __restrict__ out_dist,
    float* __restrict__ evecs, float* __restrict__ evals, int32_t* __restrict__ sizes,
    unsigned long long* __restrict__ nfb, int xrun) {
  __shared__ uint64_t lds[kBlock / 64 * coop_wave_bytes<K, 3, K>() / 8];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // tree positions fit int32; the lane's own is rebuilt from the wave's
  // first after the search instead of being held through it
  const int32_t t0 = (int32_t)xcd_block(blockIdx.x, gridDim.x, xrun) * kBlock + wave * 64;
  const int32_t t = t0 + (int32_t)__lane_id();
  float q[3] = {0.0f, 0.0f, 0.0f};
  int32_t self = -1;
  if (t < m) {
    const PointRec<float, 3> r = tv.pts[t];
    self = r.orig;
#pragma unroll
    for (int d = 0; d < 3; ++d) q[d] = r.c[d];
  }
  TopK<float, K, false> st;
  st.init(kk);
  uint64_t* wl = lds + wave * (coop_wave_bytes<K, 3, K>() / 8);
  // the row-broadcast reductions would spill at K 16 (as in knn_coop_kernel)
  knn_coop_self<3, K, K != 16>(tv, self >= 0, t, self, q, r2s, kk, true, st, wl);
  if (self < 0) return;
  const int32_t me = t0 + fresh_lane_id();
  if (self < m) {
    int32_t c = 0;
    int32_t* orow = out_idx + (int64_t)self * k;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const int j = s - (K - kk);
      if (j >= 0 && j < k) {
        const float ds = st.dist(s);
        const bool f = st.found(s) && ds < r2k;
        orow[j] = f ? st.idx(s) : -1;
        if (out_dist) out_dist[(int64_t)self * k + j] = f ? ds : pos_inf<float>();
        c += f ? 1 : 0;
      }
    }
    out_cnt[self] = c;
  }
  // neighbourhood positions in order, the point itself at its rank, staged in
  // this lane's LDS buffer (free after the search)
  int32_t* pos = coop_positions<3>(wl);
  bool pending = 0.0f < r2n;
  int n = 0;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (s >= K - kk && st.found(s) && st.dist(s) < r2n) {
      if (pending && key_less(0.0f, self, st.dist(s), st.idx(s))) {
        if (n < nk) pos[n * 64] = me;
        n += n < nk ? 1 : 0;
        pending = false;
      }
      if (n < nk) pos[n * 64] = tv.fwd[st.idx(s)];
      n += n < nk ? 1 : 0;
    }
  }
  if (pending && n < nk) {
    pos[n * 64] = me;
    ++n;
  }
  bool fb = false;
  if (n == 0) {
    write_nan<float, NE>(self, evecs, evals, sizes);
  } else {
    auto each = [&](auto f) {
      for (int j = 0; j < n; ++j) f(pos[j * 64]);
    };
    fb = pca_row<float, PREC, NE>(tv, q, n, each, self, evecs, evals, sizes);
  }
  count_fallbacks(fb, nfb);
}

template <typename D, int PREC, int NE>
__global__ __launch_bounds__(kBlock) void normals_heap_kernel(TreeView<D, 3> tv,
                                                              const int32_t* __restrict__ rows,
                                                              int64_t t0, int64_t t1, int k, D r2,
                                                              D* __restrict__ ws_d,
                                                              int32_t* __restrict__ ws_o,
                                                              int32_t* __restrict__ ws_p,
                                                              D* __restrict__ evecs,
                                                              D* __restrict__ evals,
                                                              int32_t* __restrict__ sizes,
                                                              unsigned long long* __restrict__ nfb) {
  const int64_t t = t0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= t1) return;
  D q[3];
  int64_t row;
  int32_t pos;
  bool ok;
  resolve_row<D>(tv, rows, t, q, row, pos, ok);
  if (row < 0) return;
  const int64_t w0 = (t - t0) * k;
  HeapK<D> st;
  st.init(ws_d + w0, ws_o + w0, ws_p + w0, k);
  if (ok) knn_search_from_leaf<D, 3>(tv, pos, q, r2, -1, k, st);
  const int n = st.size;
  bool fb = false;
  if (n == 0) {
    write_nan<D, NE>(row, evecs, evals, sizes);
  } else {
    const int32_t* hp = ws_p + w0;
    auto each = [&](auto f) {
      for (int s = 0; s < n; ++s) f(hp[s]);
    };
    fb = pca_row<D, PREC, NE>(tv, q, n, each, row, evecs, evals, sizes);
  }
  count_fallbacks(fb, nfb);
}

// k == -1: neighbourhood = every point with d < r*r (RNearNeighbors,
// estimate_normals.cpp:79-80). One pass of shifted sums around q, combined
// as C = S2/n - mu mu^T.
template <typename D, int PREC, int NE>
__global__ __launch_bounds__(kBlock) void normals_rnear_kernel(TreeView<D, 3> tv,
                                                               const int32_t* __restrict__ rows,
                                                               int64_t m, D r2,
                                                               D* __restrict__ evecs,
                                                               D* __restrict__ evals,
                                                               int32_t* __restrict__ sizes,
                                                               unsigned long long* __restrict__ nfb) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  D q[3];
  int64_t row;
  int32_t pos;
  bool ok;
  resolve_row<D>(tv, rows, t, q, row, pos, ok);
  if (row < 0) return;
  // Shifted sums around q over the ball, in FP64; returns n.
  auto sums = [&](double (&s1)[3], double (&s2)[6]) -> int64_t {
    int64_t n = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) s1[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) s2[i] = 0.0;
    if (!ok) return 0;
    auto visit = [&](const PointRec<D, 3>& r) {
      ++n;
      double e[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) { e[a] = (double)r.c[a] - (double)q[a]; s1[a] = s1[a] + e[a]; }
      s2[0] = s2[0] + e[0] * e[0]; s2[1] = s2[1] + e[0] * e[1]; s2[2] = s2[2] + e[0] * e[2];
      s2[3] = s2[3] + e[1] * e[1]; s2[4] = s2[4] + e[1] * e[2]; s2[5] = s2[5] + e[2] * e[2];
    };
    rnear_visit<D, 3>(tv, root_of<D, 3>(tv, pos), q, r2, visit);
    return n;
  };
  auto covariance = [](int64_t n, const double (&s1)[3], const double (&s2)[6],
                       double (&cd)[3][3]) {
    const double inv = 1.0 / (double)n;
    const double mu[3] = {s1[0] * inv, s1[1] * inv, s1[2] * inv};
    cd[0][0] = s2[0] * inv - mu[0] * mu[0];
    cd[0][1] = s2[1] * inv - mu[0] * mu[1];
    cd[0][2] = s2[2] * inv - mu[0] * mu[2];
    cd[1][1] = s2[3] * inv - mu[1] * mu[1];
    cd[1][2] = s2[4] * inv - mu[1] * mu[2];
    cd[2][2] = s2[5] * inv - mu[2] * mu[2];
    cd[1][0] = cd[0][1]; cd[2][0] = cd[0][2]; cd[2][1] = cd[1][2];
  };
  double s1[3], s2[6], cd[3][3];
  const int64_t n = sums(s1, s2);
  bool fb = false;
  if (n == 0) {
    write_nan<D, NE>(row, evecs, evals, sizes);
  } else {
    covariance(n, s1, s2, cd);
    bool solved = false;
    if constexpr (PREC == PPTK_PREC_MIXED) {
      float cf[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) cf[i][j] = (float)cd[i][j];
      float w[3], v[3][3];
      eig_sym3<float>(cf, w, v);
      if (!needs_fp64<NE>(w)) {
        write_eig<D, float, NE>(row, w, v, (int)n, evecs, evals, sizes);
        solved = true;
      } else {
        fb = true;  // below the gap threshold: the FP64 solve of the same covariance
      }
    }
    if (!solved) {
      double w[3], v[3][3];
      eig_sym3<double>(cd, w, v);
      write_eig<D, double, NE>(row, w, v, (int)n, evecs, evals, sizes);
    }
  }
  count_fallbacks(fb, nfb);
}

template <typename D, int NE>
__global__ __launch_bounds__(kBlock) void nan_rows_kernel(const int32_t* __restrict__ fwd,
                                                          int64_t nrows, D* __restrict__ evecs,
                                                          D* __restrict__ evals,
                                                          int32_t* __restrict__ sizes) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrows) return;
  if (fwd[r] < 0) write_nan<D, NE>(r, evecs, evals, sizes);
}

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, kBlock)); }

template <typename D, int PREC, int NE>
void normals_launch(const pptk_tree* t, int k, D r2, const int32_t* drows, int64_t m, D* evecs,
                    D* evals, int32_t* sizes, unsigned long long* nfb, hipStream_t s) {
  const TreeView<D, 3> tv = tree_view<D, 3>(t);
  if (m == 0) return;
  if (k == -1) {
    normals_rnear_kernel<D, PREC, NE><<<grid_of(m), kBlock, 0, s>>>(tv, drows, m, r2, evecs,
                                                                     evals, sizes, nfb);
    PK_LAUNCH_CHECK();
  } else if (k <= 32) {
    auto launch = [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if constexpr (std::is_same<D, float>::value) {
        if (!drows) {
          normals_coop_kernel<K, PREC, NE><<<grid_of(m), kBlock, 0, s>>>(tv, m, k, r2, evecs,
                                                                         evals, sizes, nfb,
                                                                         kXcdRun);
          PK_LAUNCH_CHECK();
          return;
        }
      }
      normals_knn_kernel<D, K, PREC, NE><<<grid_of(m), kBlock, 0, s>>>(tv, drows, m, k, r2, evecs,
                                                                        evals, sizes, nfb);
      PK_LAUNCH_CHECK();
    };
    if (k <= 4) launch(std::integral_constant<int, 4>());
    else if (k <= 8) launch(std::integral_constant<int, 8>());
    else if (k <= 16) launch(std::integral_constant<int, 16>());
    else launch(std::integral_constant<int, 32>());
  } else {
    const int64_t per = (int64_t)k * (sizeof(D) + 8);
    const int64_t chunk = std::max<int64_t>(kBlock, ((int64_t)256 << 20) / per);
    const int64_t c = std::min(chunk, m);
    DevBuf wd((size_t)c * k * sizeof(D)), wo((size_t)c * k * 4), wp((size_t)c * k * 4);
    for (int64_t t0 = 0; t0 < m; t0 += chunk) {
      const int64_t t1 = std::min(m, t0 + chunk);
      normals_heap_kernel<D, PREC, NE><<<grid_of(t1 - t0), kBlock, 0, s>>>(
          tv, drows, t0, t1, k, r2, wd.as<D>(), wo.as<int32_t>(), wp.as<int32_t>(), evecs, evals,
          sizes, nfb);
      PK_LAUNCH_CHECK();
    }
    PK_HIP(hipStreamSynchronize(s));
  }
}

template <typename D>
void normals_run(const pptk_tree* t, int k, double r, const int32_t* rows, int64_t m_rows,
                 int num_eigen, pptk_precision prec, void* out_evecs, void* out_evals,
                 int32_t* out_sizes, int64_t* out_num_fp64, pptk_mem mem, hipStream_t s) {
  const bool tree_order = rows == nullptr;
  const int64_t nrows = tree_order ? t->n_input : m_rows;
  if (!tree_order && mem == PPTK_MEM_HOST) {
    for (int64_t i = 0; i < m_rows; ++i)
      if (rows[i] < 0 || rows[i] >= t->n_input)
        fail(PPTK_E_RUNTIME, "%lld-th subsample index %d is outside of [0,%lld]", (long long)i,
             (int)rows[i], (long long)t->n_input - 1);
  }
  DevInput drow;
  if (!tree_order) stage_input(drow, rows, (size_t)m_rows * 4, mem, s);
  DevOutput oe, ov, os;
  oe.init(out_evecs, (size_t)nrows * 3 * num_eigen * sizeof(D), mem);
  ov.init(out_evals, (size_t)nrows * num_eigen * sizeof(D), mem);
  os.init(out_sizes, (size_t)nrows * 4, mem);
  DevBuf nfb(8);
  PK_HIP(hipMemsetAsync(nfb.get(), 0, 8, s));
  D* evecs = (D*)oe.dptr;
  D* evals = (D*)ov.dptr;
  int32_t* sizes = (int32_t*)os.dptr;
  const D rr = (D)r;
  const D r2 = rr * rr;
  const int64_t m = tree_order ? t->n_valid : m_rows;
  const bool all_nan = k < -1;  // KNearestNeighbors with k < 1 returns nothing
  if (nrows > 0 && (all_nan || (tree_order && t->n_valid < t->n_input))) {
    // Rows without a valid point (or every row, for k < -1) are NaN / size 0.
    const int32_t* fwd = t->fwd.as<const int32_t>();
    if (all_nan) {
      if (evecs) fill_async<D>(evecs, nrows * 3 * num_eigen, qnan_host<D>(), s);
      if (evals) fill_async<D>(evals, nrows * num_eigen, qnan_host<D>(), s);
      if (sizes) fill_async<int32_t>(sizes, nrows, 0, s);
    } else if (num_eigen == 1) {
      nan_rows_kernel<D, 1><<<grid_of(nrows), kBlock, 0, s>>>(fwd, nrows, evecs, evals, sizes);
      PK_LAUNCH_CHECK();
    } else {
      nan_rows_kernel<D, 3><<<grid_of(nrows), kBlock, 0, s>>>(fwd, nrows, evecs, evals, sizes);
      PK_LAUNCH_CHECK();
    }
  }
  if (!all_nan) {
    const int32_t* dr = tree_order ? nullptr : (const int32_t*)drow.ptr;
    unsigned long long* c = nfb.as<unsigned long long>();
    const bool mixed = prec == PPTK_PREC_MIXED;
    if (num_eigen == 1) {
      if (mixed) normals_launch<D, PPTK_PREC_MIXED, 1>(t, k, r2, dr, m, evecs, evals, sizes, c, s);
      else normals_launch<D, PPTK_PREC_FP64, 1>(t, k, r2, dr, m, evecs, evals, sizes, c, s);
    } else {
      if (mixed) normals_launch<D, PPTK_PREC_MIXED, 3>(t, k, r2, dr, m, evecs, evals, sizes, c, s);
      else normals_launch<D, PPTK_PREC_FP64, 3>(t, k, r2, dr, m, evecs, evals, sizes, c, s);
    }
  }
  oe.finish(s);
  ov.finish(s);
  os.finish(s);
  unsigned long long nf = 0;
  PK_HIP(hipMemcpyAsync(&nf, nfb.get(), 8, hipMemcpyDeviceToHost, s));
  PK_HIP(hipStreamSynchronize(s));
  if (out_num_fp64) *out_num_fp64 = (int64_t)nf;
}

// pptk_knn_normals on a float 3-D tree with max(k, nk) <= 32: the outputs of
// pptk_knn(PPTK_Q_ALL) and pptk_normals(rows = NULL) from one launch.
void knn_normals_run(const pptk_tree* t, int k, double dmax, int nk, double r, int num_eigen,
                     pptk_precision prec, int32_t* out_idx, int32_t* out_cnt, void* out_dist,
                     void* out_evecs, void* out_evals, int32_t* out_sizes,
                     int64_t* out_num_fp64, pptk_mem mem, hipStream_t s) {
  const int64_t m = t->n_valid;      // kNN rows, and the points searched from
  const int64_t nrows = t->n_input;  // normals rows
  DevOutput oi, oc, od, oe, ov, os;
  oi.init(out_idx, (size_t)m * k * 4, mem);
  oc.init(out_cnt, (size_t)m * 4, mem);
  od.init(out_dist, (size_t)m * k * 4, mem);
  oe.init(out_evecs, (size_t)nrows * 3 * num_eigen * 4, mem);
  ov.init(out_evals, (size_t)nrows * num_eigen * 4, mem);
  os.init(out_sizes, (size_t)nrows * 4, mem);
  DevBuf nfb(8);
  PK_HIP(hipMemsetAsync(nfb.get(), 0, 8, s));
  float* evecs = (float*)oe.dptr;
  float* evals = (float*)ov.dptr;
  int32_t* sizes = (int32_t*)os.dptr;
  if (m < nrows) {
    // kNN rows of NaN/inf points get no lane (pre-filled empty, as pptk_knn);
    // their normals rows are NaN / size 0 (as pptk_normals)
    if (m > 0) {
      fill_async<int32_t>((int32_t*)oi.dptr, m * k, -1, s);
      fill_async<int32_t>((int32_t*)oc.dptr, m, 0, s);
      if (od.dptr) fill_async<float>((float*)od.dptr, m * k, pos_inf<float>(), s);
    }
    const int32_t* fwd = t->fwd.as<const int32_t>();
    if (num_eigen == 1)
      nan_rows_kernel<float, 1><<<grid_of(nrows), kBlock, 0, s>>>(fwd, nrows, evecs, evals, sizes);
    else
      nan_rows_kernel<float, 3><<<grid_of(nrows), kBlock, 0, s>>>(fwd, nrows, evecs, evals, sizes);
    PK_LAUNCH_CHECK();
  }
  const float rk = (float)dmax, rn = (float)r;
  const float r2k = rk * rk, r2n = rn * rn;
  const float r2s = r2k > r2n ? r2k : r2n;
  const int kk = std::max(k, nk);
  if (m > 0) {
    const TreeView<float, 3> tv = tree_view<float, 3>(t);
    auto launch = [&](auto kc, auto pc, auto ec) {
      constexpr int K = decltype(kc)::value;
      constexpr int P = decltype(pc)::value;
      constexpr int NE = decltype(ec)::value;
      knn_normals_coop_kernel<K, P, NE><<<grid_of(m), kBlock, 0, s>>>(
          tv, m, k, kk, nk, r2s, r2k, r2n, (int32_t*)oi.dptr, (int32_t*)oc.dptr,
          (float*)od.dptr, evecs, evals, sizes, nfb.as<unsigned long long>(),
          kXcdRun);
      PK_LAUNCH_CHECK();
    };
    auto by_k = [&](auto pc, auto ec) {
      if (kk <= 4) launch(std::integral_constant<int, 4>(), pc, ec);
      else if (kk <= 8) launch(std::integral_constant<int, 8>(), pc, ec);
      else if (kk <= 16) launch(std::integral_constant<int, 16>(), pc, ec);
      else launch(std::integral_constant<int, 32>(), pc, ec);
    };
    using Mixed = std::integral_constant<int, PPTK_PREC_MIXED>;
    using F64 = std::integral_constant<int, PPTK_PREC_FP64>;
    using One = std::integral_constant<int, 1>;
    using Three = std::integral_constant<int, 3>;
    const bool mixed = prec == PPTK_PREC_MIXED;
    if (num_eigen == 1) mixed ? by_k(Mixed(), One()) : by_k(F64(), One());
    else mixed ? by_k(Mixed(), Three()) : by_k(F64(), Three());
  }
  oi.finish(s);
  oc.finish(s);
  od.finish(s);
  oe.finish(s);
  ov.finish(s);
  os.finish(s);
  unsigned long long nf = 0;
  PK_HIP(hipMemcpyAsync(&nf, nfb.get(), 8, hipMemcpyDeviceToHost, s));
  PK_HIP(hipStreamSynchronize(s));
  if (out_num_fp64) *out_num_fp64 = (int64_t)nf;
}

}  // namespace
}  // namespace pk

extern "C" pptk_status pptk_normals(const pptk_tree* tree, int k, double r, const int32_t* rows,
                                    int64_t m, int num_eigen, pptk_precision precision,
                                    void* out_evecs, void* out_evals, int32_t* out_sizes,
                                    int64_t* out_num_fp64, pptk_mem mem, void* stream) {
  return pk::guard([&] {
    pk::check_tree(tree);
    if (tree->dim != 3) pk::fail(PPTK_E_VALUE, "normals need 3-D points (tree dim = %d)", tree->dim);
    if (tree->dtype != PPTK_F32 && tree->dtype != PPTK_F64)
      pk::fail(PPTK_E_TYPE, "points must be float32 or float64");
    if (!(r > 0.0)) pk::fail(PPTK_E_VALUE, "r must be positive");
    if (k == 0) pk::fail(PPTK_E_VALUE, "k cannot be zero");
    if (k < 0 && r == __builtin_huge_val()) pk::fail(PPTK_E_VALUE, "invalid combo: r == inf and k < 0");
    if (num_eigen != 1 && num_eigen != 3) pk::fail(PPTK_E_VALUE, "num_eigen must be 1 or 3");
    if (rows && m < 0) pk::fail(PPTK_E_VALUE, "negative subsample size");
    if (!out_evecs) pk::fail(PPTK_E_VALUE, "null eigenvector output");
    pptk_precision prec = precision;
    if (prec == PPTK_PREC_AUTO) prec = tree->dtype == PPTK_F64 ? PPTK_PREC_FP64 : PPTK_PREC_MIXED;
    if (prec != PPTK_PREC_MIXED && prec != PPTK_PREC_FP64)
      pk::fail(PPTK_E_VALUE, "unknown precision mode %d", (int)precision);
    hipStream_t s = pk::pick_stream(stream);
    if (tree->dist == PPTK_F64)
      pk::normals_run<double>(tree, k, r, rows, m, num_eigen, prec, out_evecs, out_evals,
                              out_sizes, out_num_fp64, mem, s);
    else
      pk::normals_run<float>(tree, k, r, rows, m, num_eigen, prec, out_evecs, out_evals,
                             out_sizes, out_num_fp64, mem, s);
  });
}

namespace pk {
pptk_status api_knn(const pptk_tree* tree, const pptk_queries* q, int k, double dmax,
                    int32_t* out_idx, int32_t* out_cnt, void* out_dist, pptk_mem out_mem,
                    void* stream);
}  // namespace pk

extern "C" pptk_status pptk_knn_normals(const pptk_tree* tree, int k, double dmax, int nk,
                                        double r, int num_eigen, pptk_precision precision,
                                        int32_t* out_idx, int32_t* out_cnt, void* out_dist,
                                        void* out_evecs, void* out_evals, int32_t* out_sizes,
                                        int64_t* out_num_fp64, pptk_mem mem, void* stream) {
  bool fused = false;
  const pptk_status st = pk::guard([&] {
    pk::check_tree(tree);
    if (k < 1) pk::fail(PPTK_E_VALUE, "k must be positive (got %d)", k);
    if (!(dmax >= 0.0)) pk::fail(PPTK_E_VALUE, "dmax must be non-negative (got %g)", dmax);
    if (nk < 1) pk::fail(PPTK_E_VALUE, "normals k must be positive (got %d)", nk);
    if (!(r > 0.0)) pk::fail(PPTK_E_VALUE, "r must be positive");
    if (tree->dim != 3) pk::fail(PPTK_E_VALUE, "normals need 3-D points (tree dim = %d)", tree->dim);
    if (tree->dtype != PPTK_F32 && tree->dtype != PPTK_F64)
      pk::fail(PPTK_E_TYPE, "points must be float32 or float64");
    if (num_eigen != 1 && num_eigen != 3) pk::fail(PPTK_E_VALUE, "num_eigen must be 1 or 3");
    if (!out_idx || !out_cnt) pk::fail(PPTK_E_VALUE, "null kNN output");
    if (!out_evecs) pk::fail(PPTK_E_VALUE, "null eigenvector output");
    pptk_precision prec = precision;
    if (prec == PPTK_PREC_AUTO) prec = tree->dtype == PPTK_F64 ? PPTK_PREC_FP64 : PPTK_PREC_MIXED;
    if (prec != PPTK_PREC_MIXED && prec != PPTK_PREC_FP64)
      pk::fail(PPTK_E_VALUE, "unknown precision mode %d", (int)precision);
    if (tree->dist != PPTK_F32 || std::max(k, nk) > 32 || tree->nroots > 1) return;
    fused = true;
    hipStream_t s = pk::pick_stream(stream);
    pk::knn_normals_run(tree, k, dmax, nk, r, num_eigen, prec, out_idx, out_cnt, out_dist,
                        out_evecs, out_evals, out_sizes, out_num_fp64, mem, s);
  });
  if (st != PPTK_OK || fused) return st;
  // the two searches of the separate calls (float64 trees, k or nk > 32)
  pptk_queries all{};
  all.kind = PPTK_Q_ALL;
  const pptk_status sk =
      pk::api_knn(tree, &all, k, dmax, out_idx, out_cnt, out_dist, mem, stream);
  if (sk != PPTK_OK) return sk;
  return pptk_normals(tree, nk, r, nullptr, 0, num_eigen, precision, out_evecs, out_evals,
                      out_sizes, out_num_fp64, mem, stream);
}
