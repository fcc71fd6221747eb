// Drives the oracle's C API (oracle/pptk_oracle.cpp) under AddressSanitizer
// and UndefinedBehaviorSanitizer: every entry point on edge-case inputs
// (empty and one-point clouds, duplicates, NaN/inf rows, k above the point
// count, finite and infinite bounds, leaf sizes 1 to 64, dimensions 2 to 4,
// float and double), each result checked against brute force in this file
// (the same squared-distance formula, ties by original index). Test
// infrastructure: built and run by tests/test_sanitizers_cpu.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <vector>

extern "C" {
void* oracle_build(const void* pts, int n, int dim, int dbl, int leaf, double empty_split);
void oracle_free(void* h);
int oracle_num_valid(void* h);
void oracle_knn(void* h, int mode, const int* list, const void* qpts, int m, int k, double r,
                int* out_idx, int* out_cnt, void* out_dist, int nthreads);
void oracle_rnear(void* h, int mode, const int* list, const void* qpts, int m, double r,
                  long long* offsets, int* out_idx, int nthreads);
void oracle_normals(void* h, int k, double r, const int* rows, int m, double* evecs,
                    double* evals, int* sizes, int nthreads);
long long oracle_csr_first_mismatch(const long long* off_a, const int* idx_a,
                                    const long long* off_b, const int* idx_b, long long m,
                                    int nthreads);
}

static int g_failures = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);  \
      std::printf(__VA_ARGS__);                         \
      std::printf("\n");                                \
      if (++g_failures > 20) std::exit(1);              \
    }                                                   \
  } while (0)

template <typename T>
static T dist2(const T* a, const T* b, int dim) {
  T s = T(0);
  for (int d = 0; d < dim; ++d) {
    const T t = a[d] - b[d];
    s = s + t * t;
  }
  return s;
}

template <typename T>
static bool finite_row(const T* p, int dim) {
  for (int d = 0; d < dim; ++d)
    if (!std::isfinite(p[d])) return false;
  return true;
}

// Brute force: the k smallest (d, index) with d < r*r among valid points,
// `self` (an original index, or -1) excluded.
template <typename T>
static std::vector<int> brute_knn(const std::vector<T>& pts, int n, int dim, const T* q, int k,
                                  T r2, int self) {
  std::vector<std::pair<T, int>> c;
  for (int i = 0; i < n; ++i) {
    if (i == self || !finite_row(&pts[(size_t)i * dim], dim)) continue;
    const T d = dist2(q, &pts[(size_t)i * dim], dim);
    if (d < r2) c.push_back({d, i});
  }
  std::sort(c.begin(), c.end());
  std::vector<int> out;
  for (int i = 0; i < (int)c.size() && i < k; ++i) out.push_back(c[i].second);
  return out;
}

template <typename T>
static void run_case(const std::vector<T>& pts, int n, int dim, int leaf, std::mt19937& rng) {
  const int dbl = sizeof(T) == 8;
  void* h = oracle_build(pts.data(), n, dim, dbl, leaf, 0.2);
  CHECK(h != nullptr, "build");
  std::vector<int> valid;
  for (int i = 0; i < n; ++i)
    if (finite_row(&pts[(size_t)i * dim], dim)) valid.push_back(i);
  const int nv = oracle_num_valid(h);
  CHECK(nv == (int)valid.size(), "num_valid %d vs %d", nv, (int)valid.size());
  const double bounds[2] = {std::numeric_limits<double>::infinity(), 0.35};
  for (double r : bounds) {
    const T rr = (T)r;
    const T r2 = rr * rr;
    for (int k : {1, 5, n + 3}) {
      // mode 0: original indices 0 .. num_valid-1 as self queries
      {
        const int m = nv;
        std::vector<int> idx((size_t)m * k + 1), cnt((size_t)m + 1);
        std::vector<T> dist((size_t)m * k + 1);
        oracle_knn(h, 0, nullptr, nullptr, m, k, r, idx.data(), cnt.data(), dist.data(), 1);
        for (int i = 0; i < m; ++i) {
          const bool ok = finite_row(&pts[(size_t)i * dim], dim);
          const std::vector<int> want =
              ok ? brute_knn(pts, n, dim, &pts[(size_t)i * dim], k, r2, i) : std::vector<int>();
          CHECK(cnt[i] == (int)want.size(), "knn all: row %d count %d vs %d", i, cnt[i],
                (int)want.size());
          for (int s = 0; s < k; ++s) {
            const int w = s < (int)want.size() ? want[s] : -1;
            CHECK(idx[(size_t)i * k + s] == w, "knn all: row %d slot %d", i, s);
          }
        }
      }
      // mode 2: point queries
      {
        const int m = 7;
        std::vector<T> q((size_t)m * dim);
        std::uniform_real_distribution<double> u(-0.2, 1.2);
        for (auto& x : q) x = (T)u(rng);
        std::vector<int> idx((size_t)m * k), cnt(m);
        oracle_knn(h, 2, nullptr, q.data(), m, k, r, idx.data(), cnt.data(), nullptr, 1);
        for (int i = 0; i < m; ++i) {
          const std::vector<int> want = brute_knn(pts, n, dim, &q[(size_t)i * dim], k, r2, -1);
          CHECK(cnt[i] == (int)want.size(), "knn points: row %d", i);
          for (int s = 0; s < (int)want.size(); ++s)
            CHECK(idx[(size_t)i * k + s] == want[s], "knn points: row %d slot %d", i, s);
        }
      }
    }
    // r-near by index list (valid rows, reversed) and by points; sets vs brute force
    if (r < 1.0) {
      std::vector<int> list(valid.rbegin(), valid.rend());
      const int m = (int)list.size();
      std::vector<long long> off((size_t)m + 1, 0);
      oracle_rnear(h, 1, list.data(), nullptr, m, r, off.data(), nullptr, 1);
      std::vector<int> out((size_t)std::max<long long>(off[m], 1));
      oracle_rnear(h, 1, list.data(), nullptr, m, r, off.data(), out.data(), 1);
      std::vector<long long> woff((size_t)m + 1, 0);
      std::vector<int> wout;
      for (int i = 0; i < m; ++i) {
        std::vector<int> w = brute_knn(pts, n, dim, &pts[(size_t)list[i] * dim], n + 1, r2,
                                       list[i]);
        std::sort(w.begin(), w.end());
        wout.insert(wout.end(), w.begin(), w.end());
        woff[i + 1] = (long long)wout.size();
      }
      CHECK(off[m] == woff[m], "rnear total %lld vs %lld", off[m], woff[m]);
      if (off[m] == woff[m]) {
        wout.push_back(0);
        CHECK(oracle_csr_first_mismatch(off.data(), out.data(), woff.data(), wout.data(), m,
                                        1) == -1,
              "rnear sets");
      }
    }
  }
  if (dim == 3) {
    // normals of every input row, k = 8 including the point itself
    std::vector<double> ev((size_t)std::max(n, 1) * 9), ew((size_t)std::max(n, 1) * 3);
    std::vector<int> sz((size_t)std::max(n, 1));
    oracle_normals(h, 8, std::numeric_limits<double>::infinity(), nullptr, n, ev.data(),
                   ew.data(), sz.data(), 1);
    for (int i = 0; i < n; ++i) {
      const bool ok = finite_row(&pts[(size_t)i * dim], dim);
      const int want = ok ? std::min(8, (int)valid.size()) : 0;
      CHECK(sz[i] == want, "normals size row %d: %d vs %d", i, sz[i], want);
      if (sz[i] > 0) {
        CHECK(ew[(size_t)i * 3] <= ew[(size_t)i * 3 + 1] &&
                  ew[(size_t)i * 3 + 1] <= ew[(size_t)i * 3 + 2],
              "eigenvalues ascending, row %d", i);
        double nrm = 0;
        for (int a = 0; a < 3; ++a) nrm += ev[(size_t)i * 9 + a] * ev[(size_t)i * 9 + a];
        CHECK(std::fabs(nrm - 1.0) < 1e-9, "unit eigenvector, row %d", i);
      } else {
        CHECK(std::isnan(ev[(size_t)i * 9]), "NaN row %d", i);
      }
    }
    // radius neighbourhoods (k = -1) on a subsample
    if (!valid.empty()) {
      std::vector<int> rows(valid.begin(), valid.begin() + std::min<size_t>(valid.size(), 9));
      const int m = (int)rows.size();
      std::vector<double> ev2((size_t)m * 9), ew2((size_t)m * 3);
      std::vector<int> sz2(m);
      oracle_normals(h, -1, 0.3, rows.data(), m, ev2.data(), ew2.data(), sz2.data(), 1);
      for (int i = 0; i < m; ++i) {
        const T rr = (T)0.3;
        const int want =
            (int)brute_knn(pts, n, dim, &pts[(size_t)rows[i] * dim], n + 1, rr * rr, -1).size();
        CHECK(sz2[i] == want, "radius normals size row %d", i);
      }
    }
  }
  oracle_free(h);
}

template <typename T>
static void cloud_cases(int dim, std::mt19937& rng) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int n : {0, 1, 2, 17, 300}) {
    for (int leaf : {1, 4, 64}) {
      std::vector<T> pts((size_t)n * dim);
      for (auto& x : pts) x = (T)u(rng);
      run_case(pts, n, dim, leaf, rng);
    }
  }
  // duplicates (each point four times), NaN and inf rows
  {
    const int n = 240;
    std::vector<T> pts((size_t)n * dim);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < dim; ++d) pts[(size_t)i * dim + d] = (T)u(rng);
    for (int i = 0; i < n; i += 4)
      for (int j = 1; j < 4 && i + j < n; ++j)
        for (int d = 0; d < dim; ++d) pts[(size_t)(i + j) * dim + d] = pts[(size_t)i * dim + d];
    for (int i = 5; i < n; i += 37) pts[(size_t)i * dim + (i % dim)] = std::numeric_limits<T>::quiet_NaN();
    pts[(size_t)11 * dim] = std::numeric_limits<T>::infinity();
    run_case(pts, n, dim, 8, rng);
  }
  // every point equal
  {
    const int n = 50;
    std::vector<T> pts((size_t)n * dim, (T)0.5);
    run_case(pts, n, dim, 4, rng);
  }
}

int main() {
  std::mt19937 rng(7);
  for (int dim = 2; dim <= 4; ++dim) {
    cloud_cases<float>(dim, rng);
    cloud_cases<double>(dim, rng);
  }
  if (g_failures) {
    std::printf("oracle_check: %d failures\n", g_failures);
    return 1;
  }
  std::printf("oracle_check: ok\n");
  return 0;
}
