// TEST INFRASTRUCTURE ONLY. CPU restatement of pptk's k-d tree and PCA
// normals, used by tests/, bench.py's cpu_baseline leg and
// __graft_entry__.smoke() as the checker. Never part of the product path.
//
// Restates pptk/kdtree/src/kdtree-impl.h (build :53-329, :773-818; kNN
// :348-429, :873-976; r-near :519-558, :978-1069), box.h:157-199 and
// accumulator.h:12-27, with one deliberate change: the kNN heap orders
// candidates by (squared distance, original index) instead of distance alone
// (kdtree.h:58-64), so ties at the k-th neighbour resolve to the smaller
// original index. This is the results contract of the GPU library; without
// exact distance ties the two orders give identical rows.
//
// Normals (pptk/processing/estimate_normals/estimate_normals.cpp:17-120):
// neighbourhood by kNN including the point itself (or r-near for k == -1),
// covariance of the centred neighbourhood divided by n, computed in double,
// eigen-decomposition by cyclic Jacobi in double. Eigen (the reference's
// solver) is not available here; tests pin these normals against
// numpy.linalg.eigh.
//
// Build: make -C oracle  ->  oracle/liboracle.so (C ABI below, OpenMP).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <vector>

namespace {

template <typename T, int DIM>
struct Box {
  T lo[DIM], hi[DIM];
  void reset() {
    for (int i = 0; i < DIM; ++i) {
      lo[i] = std::numeric_limits<T>::max();
      hi[i] = std::numeric_limits<T>::lowest();
    }
  }
  void add(const T* p) {
    for (int i = 0; i < DIM; ++i) {
      lo[i] = std::min(p[i], lo[i]);
      hi[i] = std::max(p[i], hi[i]);
    }
  }
  bool is_point() const {
    for (int i = 0; i < DIM; ++i)
      if (lo[i] != hi[i]) return false;
    return true;
  }
  double width(int i) const { return (double)hi[i] - (double)lo[i]; }
  T& val(int i, int side) { return side == 0 ? lo[i] : hi[i]; }
};

// box.h:157-189 (MinDist2) and 171-198 (MaxDist2), same operation order.
template <typename T, int DIM>
T min_dist2(const T* q, const Box<T, DIM>& b) {
  T s = 0;
  for (int i = 0; i < DIM; ++i) {
    T dmin = b.lo[i] - q[i], dmax = b.hi[i] - q[i];
    T v = dmin > 0 ? dmin * dmin : (dmax < 0 ? dmax * dmax : T(0));
    s += v;
  }
  return s;
}
template <typename T, int DIM>
T max_dist2(const T* q, const Box<T, DIM>& b) {
  T s = 0;
  for (int i = 0; i < DIM; ++i) {
    T a = b.lo[i] - q[i], c = b.hi[i] - q[i];
    a = a * a;
    c = c * c;
    s += std::max(a, c);
  }
  return s;
}

// kdtree-impl.h:331-339: sum from zero of ((DistT)p_i - (DistT)q_i)^2.
template <typename T, int DIM>
T point_dist2(const T* q, const T* p) {
  T s = 0;
  for (int i = 0; i < DIM; ++i) {
    T t = p[i] - q[i];
    s += t * t;
  }
  return s;
}

struct Node {
  double split_value;  // holds a T value exactly
  int split_dim;
  int split_index;
  Node* left;
  Node* right;
};

template <typename T, int DIM>
struct Tree {
  std::vector<T> pts;          // reordered valid points (AoS)
  std::vector<int> fwd;        // original -> reordered, -1 invalid
  std::vector<int> rev;        // reordered -> original
  Box<T, DIM> bbox;
  Node* root = nullptr;
  int n_input = 0;
  int leaf = 10;
  double empty_split = 0.2;
  std::vector<Node*> all;

  Node* new_node() {
    Node* x = new Node();
    x->left = x->right = nullptr;
    all.push_back(x);
    return x;
  }
  ~Tree() {
    for (Node* x : all) delete x;
  }
  int n() const { return (int)rev.size(); }
};

// kdtree-impl.h:99-112
template <typename T, int DIM>
void largest_gap(const Box<T, DIM>& a, const Box<T, DIM>& b, int& dim, int& side, double& size) {
  dim = -1; side = -1; size = 0.0;
  for (int i = 0; i < DIM; ++i) {
    double g = std::max(0.0, (double)b.hi[i] - (double)a.hi[i]);
    if (g >= size) { dim = i; side = 1; size = g; }
    g = std::max(0.0, (double)a.lo[i] - (double)b.lo[i]);
    if (g >= size) { dim = i; side = 0; size = g; }
  }
}

// kdtree-impl.h:114-158
template <typename T, int DIM>
Node* trim_empty(Tree<T, DIM>& t, Node*& cur, int begin, int end, Box<T, DIM>& node_box,
                 Box<T, DIM>& tight) {
  Node* first = nullptr;
  cur = nullptr;
  for (;;) {
    int gd, gs;
    double gsz;
    largest_gap<T, DIM>(tight, node_box, gd, gs, gsz);
    if (gsz == 0.0) break;
    const double w = node_box.width(gd);
    if (gsz / w <= t.empty_split) break;
    Node* x = t.new_node();
    x->split_dim = gd;
    if (gs == 0) {
      x->split_value = (double)tight.lo[gd];
      x->split_index = begin;
      node_box.lo[gd] = tight.lo[gd];
    } else {
      x->split_value = (double)tight.hi[gd];
      x->split_index = end;
      node_box.hi[gd] = tight.hi[gd];
    }
    if (cur) {
      if (cur->split_index == begin) { cur->left = nullptr; cur->right = x; }
      else { cur->left = x; cur->right = nullptr; }
    }
    cur = x;
    if (!first) first = x;
  }
  return first;
}

// kdtree-impl.h:160-173 (floating-point branch: (b - a)/2 + a for same signs)
template <typename T>
T midpoint(T a, T b) {
  if ((a >= 0 && b >= 0) || (a < 0 && b < 0)) return (b - a) / (T)2 + a;
  return (a + b) / (T)2;
}

// kdtree-impl.h:175-190
template <typename T, int DIM>
void compute_split(int& dim, T& val, const Box<T, DIM>& b) {
  T maxw = 0;
  for (int i = 0; i < DIM; ++i) {
    T w = b.hi[i] - b.lo[i];
    if (w >= maxw) {
      maxw = w;
      dim = i;
      val = b.lo[i] == b.hi[i] ? b.lo[i] : midpoint(b.lo[i], b.hi[i]);
    }
  }
}

// kdtree-impl.h:192-246
template <typename T, int DIM>
int partition(int* idx, int count, int d, T v, const T* P) {
  int left = 0, right = count - 1;
  for (;;) {
    while (left <= right && P[idx[left] * DIM + d] < v) ++left;
    while (left <= right && P[idx[right] * DIM + d] >= v) --right;
    if (left >= right) break;
    std::swap(idx[left], idx[right]);
    ++left; --right;
  }
  const int lim1 = left;
  right = count - 1;
  for (;;) {
    while (left <= right && P[idx[left] * DIM + d] <= v) ++left;
    while (left <= right && P[idx[right] * DIM + d] > v) --right;
    if (left >= right) break;
    std::swap(idx[left], idx[right]);
    ++left; --right;
  }
  const int lim2 = left;
  int nl;
  if (lim1 > count / 2) nl = lim1;
  else if (lim2 < count / 2) nl = lim2;
  else nl = count / 2;
  if (lim1 == count && lim2 == 0) nl = count / 2;
  return nl;
}

// kdtree-impl.h:248-329
template <typename T, int DIM>
Node* build_rec(Tree<T, DIM>& t, int begin, int end, std::vector<int>& idx, Box<T, DIM> node_box,
                const T* P, int nvalid) {
  const int size = end - begin;
  Node* node = nullptr;
  Node* cur = nullptr;
  Box<T, DIM> tight = node_box;
  if (size != nvalid) {
    tight.reset();
    for (int i = begin; i < end; ++i) tight.add(&P[idx[i] * DIM]);
    node = trim_empty<T, DIM>(t, cur, begin, end, node_box, tight);
  }
  if (!tight.is_point() && size > t.leaf) {
    Node* x = t.new_node();
    int d = 0;
    T v = 0;
    compute_split<T, DIM>(d, v, tight);
    x->split_dim = d;
    x->split_value = (double)v;
    x->split_index = begin + partition<T, DIM>(&idx[begin], size, d, v, P);
    if (cur) {
      if (cur->split_index == begin) cur->right = x;
      else cur->left = x;
    }
    cur = x;
    if (!node) node = x;
  }
  if (!cur) return node;
  const int sd = cur->split_dim, si = cur->split_index;
  const T sv = (T)cur->split_value;
  if (begin < si) {
    Box<T, DIM> lb = node_box;
    lb.hi[sd] = sv;
    cur->left = build_rec<T, DIM>(t, begin, si, idx, lb, P, nvalid);
  }
  if (si < end) {
    Box<T, DIM> rb = node_box;
    rb.lo[sd] = sv;
    cur->right = build_rec<T, DIM>(t, si, end, idx, rb, P, nvalid);
  }
  return node;
}

template <typename T, int DIM>
Tree<T, DIM>* build(const T* P, int n, int leaf, double es) {
  Tree<T, DIM>* t = new Tree<T, DIM>();
  t->n_input = n;
  t->leaf = leaf;
  t->empty_split = es;
  std::vector<int> idx;
  for (int i = 0; i < n; ++i) {  // kdtree-impl.h:61-75
    bool ok = true;
    for (int d = 0; d < DIM; ++d)
      if (!std::isfinite(P[(size_t)i * DIM + d])) ok = false;
    if (ok) idx.push_back(i);
  }
  t->bbox.reset();
  for (int i : idx) t->bbox.add(&P[(size_t)i * DIM]);
  const int nv = (int)idx.size();
  t->root = build_rec<T, DIM>(*t, 0, nv, idx, t->bbox, P, nv);
  t->pts.resize((size_t)nv * DIM);
  for (int i = 0; i < nv; ++i)
    for (int d = 0; d < DIM; ++d) t->pts[(size_t)i * DIM + d] = P[(size_t)idx[i] * DIM + d];
  t->fwd.assign(n, -1);
  for (int i = 0; i < nv; ++i) t->fwd[idx[i]] = i;
  t->rev = idx;
  return t;
}

template <typename T>
struct Cand {
  T d;
  int orig;
  int pos;
  bool operator<(const Cand& o) const { return d < o.d || (d == o.d && orig < o.orig); }
};

// kdtree-impl.h:348-429 with the (d, orig) heap order. `excl` is a
// reordered position to skip (self queries), or -1.
template <typename T, int DIM>
void knn_rec(const Tree<T, DIM>& t, std::priority_queue<Cand<T>>& h, Box<T, DIM>& box, int begin,
             int end, const Node* node, const T* q, int k, T r2, int excl) {
  const int np = end - begin;
  if (!node || (np <= k - (int)h.size() && max_dist2<T, DIM>(q, box) < r2)) {
    for (int i = begin; i < end; ++i) {
      if (i == excl) continue;
      const T d = point_dist2<T, DIM>(q, &t.pts[(size_t)i * DIM]);
      if (d >= r2) continue;
      Cand<T> c{d, t.rev[i], i};
      if ((int)h.size() < k) h.push(c);
      else if (c < h.top()) { h.pop(); h.push(c); }
    }
    return;
  }
  const T sv = (T)node->split_value;
  const int sd = node->split_dim;
  const Node *nptr, *fptr;
  int nb, ne, ns, fb, fe, fs;
  if (q[sd] < sv) {
    nptr = node->left; nb = begin; ne = node->split_index; ns = 1;
    fptr = node->right; fb = node->split_index; fe = end; fs = 0;
  } else {
    nptr = node->right; nb = node->split_index; ne = end; ns = 0;
    fptr = node->left; fb = begin; fe = node->split_index; fs = 1;
  }
  if (nb < ne) {
    T tmp = box.val(sd, ns);
    box.val(sd, ns) = sv;
    knn_rec<T, DIM>(t, h, box, nb, ne, nptr, q, k, r2, excl);
    box.val(sd, ns) = tmp;
  }
  if (fb == fe) return;
  T tmp = box.val(sd, fs);
  box.val(sd, fs) = sv;
  const T fd = min_dist2<T, DIM>(q, box);
  // `<=` (not `<`) against the k-th distance: an equal distance with a
  // smaller original index can still enter under the (d, orig) order.
  if (fd < r2 && ((int)h.size() < k || fd <= h.top().d))
    knn_rec<T, DIM>(t, h, box, fb, fe, fptr, q, k, r2, excl);
  box.val(sd, fs) = tmp;
}

template <typename T, int DIM>
int knn_one(const Tree<T, DIM>& t, const T* q, int k, T r, int excl, int* out_idx, T* out_d) {
  if (k < 1 || r < 0) return 0;
  for (int d = 0; d < DIM; ++d)
    if (!std::isfinite(q[d])) return 0;
  Box<T, DIM> box = t.bbox;
  const T r2 = r * r;
  if (t.n() == 0 || min_dist2<T, DIM>(q, box) >= r2) return 0;
  std::priority_queue<Cand<T>> h;
  knn_rec<T, DIM>(t, h, box, 0, t.n(), t.root, q, k, r2, excl);
  const int c = (int)h.size();
  for (int i = c - 1; i >= 0; --i) {
    out_idx[i] = h.top().orig;
    if (out_d) out_d[i] = h.top().d;
    h.pop();
  }
  return c;
}

// kdtree-impl.h:519-558
template <typename T, int DIM>
void rnear_rec(const Tree<T, DIM>& t, std::vector<int>& res, Box<T, DIM>& box, int begin, int end,
               const Node* node, const T* q, T r2, int excl) {
  if (min_dist2<T, DIM>(q, box) >= r2) return;
  if (max_dist2<T, DIM>(q, box) < r2) {
    for (int i = begin; i < end; ++i)
      if (i != excl) res.push_back(t.rev[i]);
  } else if (!node) {
    for (int i = begin; i < end; ++i)
      if (i != excl && point_dist2<T, DIM>(q, &t.pts[(size_t)i * DIM]) < r2) res.push_back(t.rev[i]);
  } else {
    const T sv = (T)node->split_value;
    const int sd = node->split_dim, si = node->split_index;
    if (si > begin) {
      T tmp = box.hi[sd];
      box.hi[sd] = sv;
      rnear_rec<T, DIM>(t, res, box, begin, si, node->left, q, r2, excl);
      box.hi[sd] = tmp;
    }
    if (si < end) {
      T tmp = box.lo[sd];
      box.lo[sd] = sv;
      rnear_rec<T, DIM>(t, res, box, si, end, node->right, q, r2, excl);
      box.lo[sd] = tmp;
    }
  }
}

template <typename T, int DIM>
void rnear_one(const Tree<T, DIM>& t, const T* q, T r, int excl, std::vector<int>& res) {
  res.clear();
  if (r < 0 || t.n() == 0) return;
  for (int d = 0; d < DIM; ++d)
    if (!std::isfinite(q[d])) return;
  Box<T, DIM> box = t.bbox;
  rnear_rec<T, DIM>(t, res, box, 0, t.n(), t.root, q, r * r, excl);
  std::sort(res.begin(), res.end());
}

// ----------------------------------------------------------- eigensolver ----
// Cyclic Jacobi in double (Golub & Van Loan, Matrix Computations, section 8.5).
void jacobi3(double a[3][3], double w[3], double v[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = i == j;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
    double diag = std::fabs(a[0][0]) + std::fabs(a[1][1]) + std::fabs(a[2][2]);
    if (!(off > diag * 1e-30)) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        const double apq = a[p][q];
        if (apq == 0) continue;
        const double th = (a[q][q] - a[p][p]) / (2 * apq);
        double t = std::fabs(th) > 1e15 ? 0.5 / th
                                         : (th < 0 ? -1.0 : 1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        // explicit rotation A <- J^T A J, V <- V J
        double J[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        J[p][p] = c; J[q][q] = c; J[p][q] = s; J[q][p] = -s;
        double tmp[3][3], b[3][3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            tmp[i][j] = 0;
            for (int l = 0; l < 3; ++l) tmp[i][j] += a[i][l] * J[l][j];
          }
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            b[i][j] = 0;
            for (int l = 0; l < 3; ++l) b[i][j] += J[l][i] * tmp[l][j];
          }
        b[p][q] = b[q][p] = 0;
        std::memcpy(a, b, sizeof(b));
        for (int i = 0; i < 3; ++i) {
          const double x = v[i][p], y = v[i][q];
          v[i][p] = c * x - s * y;
          v[i][q] = s * x + c * y;
        }
      }
  }
  for (int i = 0; i < 3; ++i) w[i] = a[i][i];
  auto cs = [&](int i, int j) {
    if (w[i] > w[j]) {
      std::swap(w[i], w[j]);
      for (int r = 0; r < 3; ++r) std::swap(v[r][i], v[r][j]);
    }
  };
  cs(0, 1); cs(1, 2); cs(0, 1);
  for (int j = 0; j < 3; ++j) {
    int im = 0;
    for (int r = 1; r < 3; ++r)
      if (std::fabs(v[r][j]) > std::fabs(v[im][j])) im = r;
    if (v[im][j] < 0)
      for (int r = 0; r < 3; ++r) v[r][j] = -v[r][j];
  }
}

struct Handle {
  int dim;
  int dbl;
  void* tree;
};

template <typename F>
void with_tree(const Handle* h, F&& f) {
  if (h->dbl) {
    if (h->dim == 2) f((Tree<double, 2>*)h->tree);
    else if (h->dim == 3) f((Tree<double, 3>*)h->tree);
    else f((Tree<double, 4>*)h->tree);
  } else {
    if (h->dim == 2) f((Tree<float, 2>*)h->tree);
    else if (h->dim == 3) f((Tree<float, 3>*)h->tree);
    else f((Tree<float, 4>*)h->tree);
  }
}

template <typename Tr> struct TreeTraits;
template <typename T, int DIM> struct TreeTraits<Tree<T, DIM>> {
  using type = T;
  static constexpr int dim = DIM;
};

}  // namespace

extern "C" {

// pts: n x dim in the distance type (float32 if !dbl, float64 if dbl).
void* oracle_build(const void* pts, int n, int dim, int dbl, int leaf, double empty_split) {
  Handle* h = new Handle{dim, dbl, nullptr};
  if (dbl) {
    const double* p = (const double*)pts;
    if (dim == 2) h->tree = build<double, 2>(p, n, leaf, empty_split);
    else if (dim == 3) h->tree = build<double, 3>(p, n, leaf, empty_split);
    else h->tree = build<double, 4>(p, n, leaf, empty_split);
  } else {
    const float* p = (const float*)pts;
    if (dim == 2) h->tree = build<float, 2>(p, n, leaf, empty_split);
    else if (dim == 3) h->tree = build<float, 3>(p, n, leaf, empty_split);
    else h->tree = build<float, 4>(p, n, leaf, empty_split);
  }
  return h;
}

void oracle_free(void* hp) {
  Handle* h = (Handle*)hp;
  if (!h) return;
  with_tree(h, [](auto* t) { delete t; });
  delete h;
}

int oracle_num_valid(void* hp) {
  int n = 0;
  with_tree((Handle*)hp, [&](auto* t) { n = t->n(); });
  return n;
}

// mode 0: all (original indices 0..num_valid-1, self), 1: list (self),
// 2: points (m x dim distance type). out_idx m x k (-1 padded), out_cnt m.
// out_dist: NULL or m x k distance type.
void oracle_knn(void* hp, int mode, const int* list, const void* qpts, int m, int k, double r,
                int* out_idx, int* out_cnt, void* out_dist, int nthreads) {
  with_tree((Handle*)hp, [&](auto* t) {
    using Tr = std::remove_pointer_t<decltype(t)>;
    using T = typename TreeTraits<Tr>::type;
    constexpr int DIM = TreeTraits<Tr>::dim;
    const T rr = (T)r;
    const T* qp = (const T*)qpts;
    T* od = (T*)out_dist;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int i = 0; i < m; ++i) {
      int* row = out_idx + (size_t)i * k;
      T* drow = od ? od + (size_t)i * k : nullptr;
      for (int s = 0; s < k; ++s) {
        row[s] = -1;
        if (drow) drow[s] = std::numeric_limits<T>::infinity();
      }
      int c = 0;
      if (mode == 2) {
        c = knn_one<T, DIM>(*t, qp + (size_t)i * DIM, k, rr, -1, row, drow);
      } else {
        const int o = mode == 0 ? i : list[i];
        const int p = t->fwd[o];
        if (p >= 0) c = knn_one<T, DIM>(*t, &t->pts[(size_t)p * DIM], k, rr, p, row, drow);
      }
      out_cnt[i] = c;
    }
  });
}

// r-near: counts (m) then lists written at offsets (sorted by original index).
void oracle_rnear(void* hp, int mode, const int* list, const void* qpts, int m, double r,
                  long long* offsets, int* out_idx, int nthreads) {
  with_tree((Handle*)hp, [&](auto* t) {
    using Tr = std::remove_pointer_t<decltype(t)>;
    using T = typename TreeTraits<Tr>::type;
    constexpr int DIM = TreeTraits<Tr>::dim;
    const T rr = (T)r;
    const T* qp = (const T*)qpts;
    const bool fill = out_idx != nullptr;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int i = 0; i < m; ++i) {
      std::vector<int> res;
      if (mode == 2) {
        rnear_one<T, DIM>(*t, qp + (size_t)i * DIM, rr, -1, res);
      } else {
        const int o = mode == 0 ? i : list[i];
        const int p = t->fwd[o];
        if (p >= 0) rnear_one<T, DIM>(*t, &t->pts[(size_t)p * DIM], rr, p, res);
      }
      if (fill) std::copy(res.begin(), res.end(), out_idx + offsets[i]);
      else offsets[i + 1] = (long long)res.size();
    }
    if (!fill) {
      offsets[0] = 0;
      for (int i = 0; i < m; ++i) offsets[i + 1] += offsets[i];
    }
  });
}

// PCA normals in double on the oracle tree (3-D only). rows: NULL -> all
// input rows; else m original indices. Outputs in double:
// evecs m x 3 x 3 ([i, j, :] = j-th eigenvector), evals m x 3, sizes m.
void oracle_normals(void* hp, int k, double r, const int* rows, int m, double* evecs,
                    double* evals, int* sizes, int nthreads) {
  Handle* h = (Handle*)hp;
  if (h->dim != 3) return;
  with_tree(h, [&](auto* t) {
    using Tr = std::remove_pointer_t<decltype(t)>;
    using T = typename TreeTraits<Tr>::type;
    constexpr int DIM = TreeTraits<Tr>::dim;
    if constexpr (DIM == 3) {
      const int nrows = rows ? m : t->n_input;
      const T rr = (T)r;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
      for (int i = 0; i < nrows; ++i) {
        const int o = rows ? rows[i] : i;
        const int p = t->fwd[o];
        std::vector<int> nb;
        if (p >= 0 && k != 0 && k >= -1) {
          const T* q = &t->pts[(size_t)p * 3];
          if (k == -1) {
            rnear_one<T, 3>(*t, q, rr, -1, nb);
          } else {
            nb.resize(k);
            const int c = knn_one<T, 3>(*t, q, k, rr, -1, nb.data(), nullptr);
            nb.resize(c);
          }
        }
        const int n = (int)nb.size();
        if (sizes) sizes[i] = n;
        if (n == 0) {
          for (int j = 0; j < 9; ++j) evecs[(size_t)i * 9 + j] = NAN;
          for (int j = 0; j < 3; ++j) evals[(size_t)i * 3 + j] = NAN;
          continue;
        }
        double mu[3] = {0, 0, 0};
        for (int j : nb) {
          const int pj = t->fwd[j];
          for (int a = 0; a < 3; ++a) mu[a] += (double)t->pts[(size_t)pj * 3 + a];
        }
        for (int a = 0; a < 3; ++a) mu[a] /= n;
        double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int j : nb) {
          const int pj = t->fwd[j];
          double e[3];
          for (int a = 0; a < 3; ++a) e[a] = (double)t->pts[(size_t)pj * 3 + a] - mu[a];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) c[a][b] += e[a] * e[b];
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) c[a][b] /= n;
        double w[3], v[3][3];
        jacobi3(c, w, v);
        for (int j = 0; j < 3; ++j) {
          evals[(size_t)i * 3 + j] = w[j];
          for (int a = 0; a < 3; ++a) evecs[(size_t)i * 9 + j * 3 + a] = v[a][j];
        }
      }
    }
  });
}

// Checker utility: whether two CSR neighbour lists hold the same set in every
// row (rows compared after sorting copies; neither input is modified).
// Returns -1 if every row matches, else the smallest differing row.
long long oracle_csr_first_mismatch(const long long* off_a, const int* idx_a,
                                    const long long* off_b, const int* idx_b, long long m,
                                    int nthreads) {
  long long first = -1;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  {
    std::vector<int> a, b;
    long long mine = -1;
#pragma omp for schedule(dynamic, 1024)
    for (long long i = 0; i < m; ++i) {
      if (mine >= 0 && i > mine) continue;
      const long long na = off_a[i + 1] - off_a[i], nb = off_b[i + 1] - off_b[i];
      bool same = na == nb;
      if (same && na > 0) {
        a.assign(idx_a + off_a[i], idx_a + off_a[i + 1]);
        b.assign(idx_b + off_b[i], idx_b + off_b[i + 1]);
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        same = a == b;
      }
      if (!same && (mine < 0 || i < mine)) mine = i;
    }
#pragma omp critical
    if (mine >= 0 && (first < 0 || mine < first)) first = mine;
  }
  return first;
}

}  // extern "C"
