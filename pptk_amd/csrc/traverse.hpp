// Per-lane tree traversal shared by the kNN, r-near and normals kernels.
//
// Nodes are Morton prefix cells holding contiguous tree positions [a, b);
// children come in adjacent pairs (child, child + 1) and a node's sibling is
// id + 1 for an odd id, id - 1 for an even id. A lane walks the tree without
// a stack: it keeps its node, its depth and one bit per level recording which
// child it entered first (two 64-bit words: depth stays below 64 key bits plus
// 31 levels of equal-code halving). Backtracking reads the sibling from that
// bit and re-tests it against the current k-th bound.
//
// Neighbour order is the key (d, original index): smaller squared distance
// first, ties broken by the smaller original index. A node is entered while
// its box lower bound is < r*r and <= the current k-th distance (<=, because
// a point at exactly the k-th distance with a smaller index still enters).
// Reference semantics: pointkd::impl::KNearestNeighborsHelper
// (pptk/kdtree/src/kdtree-impl.h:348-429) and RNearNeighborsHelper
// (kdtree-impl.h:519-558).
#pragma once
#include "common.hpp"

namespace pk {

enum QueryMode : int { kQueryAll = 0, kQueryList = 1, kQueryPoints = 2 };

template <typename D>
__device__ __forceinline__ bool key_less(D a, int32_t ao, D b, int32_t bo) {
  return a < b || (a == b && ao < bo);
}

// Sorted top-k in registers, K slots, runtime k <= K. Slots [0, K-k) hold
// -inf sentinels that never move, so the k-th entry is always slot K-1 and no
// register array is indexed dynamically.
template <typename D, int K, bool POS>
struct TopK {
  D d[K];
  int32_t o[K];
  int32_t p[POS ? K : 1];

  __device__ __forceinline__ void init(int k) {
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const bool pad = s < K - k;
      d[s] = pad ? -pos_inf<D>() : pos_inf<D>();
      o[s] = pad ? INT32_MIN : INT32_MAX;
      if constexpr (POS) p[s] = -1;
    }
  }
  __device__ __forceinline__ D bound() const { return d[K - 1]; }
  // slot s holds a neighbour
  __device__ __forceinline__ bool found(int s) const { return d[s] < pos_inf<D>(); }
  __device__ __forceinline__ D dist(int s) const { return d[s]; }
  __device__ __forceinline__ int32_t idx(int s) const { return o[s]; }
  __device__ __forceinline__ bool beats(D dd, int32_t oo) const {
    return key_less(dd, oo, d[K - 1], o[K - 1]);
  }
  // Precondition: beats(dd, oo).
  __device__ __forceinline__ void insert(D dd, int32_t oo, int32_t pp) {
#pragma unroll
    for (int s = K - 1; s > 0; --s) {
      const bool sh = key_less(dd, oo, d[s - 1], o[s - 1]);
      const bool here = key_less(dd, oo, d[s], o[s]);
      d[s] = sh ? d[s - 1] : (here ? dd : d[s]);
      o[s] = sh ? o[s - 1] : (here ? oo : o[s]);
      if constexpr (POS) p[s] = sh ? p[s - 1] : (here ? pp : p[s]);
    }
    const bool h0 = key_less(dd, oo, d[0], o[0]);
    d[0] = h0 ? dd : d[0];
    o[0] = h0 ? oo : o[0];
    if constexpr (POS) p[0] = h0 ? pp : p[0];
  }
};

// Unsigned 64-bit min / max of two packed float keys (below) in one
// instruction. Every key read as an IEEE double is non-negative and finite:
// its top 32 bits are a squared distance's float bits (sign 0, at most +inf =
// 0x7f800000, so the double exponent stays below 0x7ff), and non-negative
// doubles order as their bit patterns. Denormal doubles (distances below
// 2^-126) are exact because f64 denormals are preserved. Written in asm
// because the builtins first canonicalise each operand (one more
// instruction each).
__device__ __forceinline__ uint64_t key_min(uint64_t a, uint64_t b) {
  uint64_t r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ uint64_t key_max(uint64_t a, uint64_t b) {
  uint64_t r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Lane L's value of x, as a wave-uniform value.
__device__ __forceinline__ int32_t rl_i32(int32_t x, int L) {
  return __builtin_amdgcn_readlane(x, L);
}
__device__ __forceinline__ int64_t rl_i64(int64_t x, int L) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)x, L);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)((uint64_t)x >> 32), L);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float rl(float x, int L) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), L));
}
__device__ __forceinline__ double rl(double x, int L) {
  return __longlong_as_double(rl_i64(__double_as_longlong(x), L));
}

// Float specialisation: (d, o) packed into one 64-bit key, d's IEEE bits
// above the original index. Squared distances are +0 or positive, so the
// unsigned key order equals the (d, o) order and each slot update is one
// 64-bit compare. Sentinel slots hold key 0 (an inserted key 0 lands right
// after them); empty slots hold (+inf, 0xffffffff).
template <int K, bool POS>
struct TopK<float, K, POS> {
  static constexpr uint64_t kEmpty = (0x7f800000ull << 32) | 0xffffffffull;
  uint64_t key[K];
  int32_t p[POS ? K : 1];
  __device__ __forceinline__ float dist(int s) const {
    return __uint_as_float((uint32_t)(key[s] >> 32));
  }
  __device__ __forceinline__ int32_t idx(int s) const { return (int32_t)(uint32_t)key[s]; }

  static __device__ __forceinline__ uint64_t make(float dd, int32_t oo) {
    return ((uint64_t)__float_as_uint(dd) << 32) | (uint32_t)oo;
  }
  __device__ __forceinline__ void init(int k) {
#pragma unroll
    for (int s = 0; s < K; ++s) {
      key[s] = s < K - k ? 0ull : kEmpty;
      if constexpr (POS) p[s] = -1;
    }
  }
  __device__ __forceinline__ float bound() const {
    return __uint_as_float((uint32_t)(key[K - 1] >> 32));
  }
  // slot s holds a neighbour: a finite distance and an index other than -1
  // (a search may start its slots at (bound, -1) instead of empty)
  __device__ __forceinline__ bool found(int s) const {
    return ((uint32_t)(key[s] >> 32) < 0x7f800000u) & ((uint32_t)key[s] != 0xffffffffu);
  }
  __device__ __forceinline__ bool beats(float dd, int32_t oo) const { return make(dd, oo) < key[K - 1]; }

  __device__ __forceinline__ void insert(float dd, int32_t oo, int32_t pp) {
    const uint64_t kk = make(dd, oo);
    if constexpr (POS) {
#pragma unroll
      for (int s = K - 1; s > 0; --s) {
        const bool sh = kk < key[s - 1];
        const bool here = kk < key[s];
        key[s] = sh ? key[s - 1] : (here ? kk : key[s]);
        p[s] = sh ? p[s - 1] : (here ? pp : p[s]);
      }
      const bool h0 = kk < key[0];
      key[0] = h0 ? kk : key[0];
      p[0] = h0 ? pp : p[0];
    } else {
      // Sorted insert as a min/max network: slot s keeps the smaller of its
      // key and max(slot s-1, kk) -- its own key below the insertion point,
      // kk at it, the key of slot s-1 (shifted up) above it.
      (void)pp;
#pragma unroll
      for (int s = K - 1; s > 0; --s) key[s] = key_min(key[s], key_max(key[s - 1], kk));
      key[0] = key_min(key[0], kk);
    }
  }
};

// Binary max-heap on (d, o) in global memory, for k above the register
// limit. `hd`, `ho`, `hp` point at this query's k-slot row.
template <typename D>
struct HeapK {
  D* hd;
  int32_t* ho;
  int32_t* hp;  // may be null
  int k;
  int size;
  D top_d;
  int32_t top_o;

  __device__ __forceinline__ void init(D* d_, int32_t* o_, int32_t* p_, int k_) {
    hd = d_; ho = o_; hp = p_; k = k_; size = 0;
    top_d = pos_inf<D>(); top_o = INT32_MAX;
  }
  __device__ __forceinline__ D bound() const { return size < k ? pos_inf<D>() : top_d; }
  __device__ __forceinline__ bool beats(D dd, int32_t oo) const {
    return size < k || key_less(dd, oo, top_d, top_o);
  }
  __device__ __forceinline__ void swap_(int a, int b) {
    const D td = hd[a]; hd[a] = hd[b]; hd[b] = td;
    const int32_t to = ho[a]; ho[a] = ho[b]; ho[b] = to;
    if (hp) { const int32_t tp = hp[a]; hp[a] = hp[b]; hp[b] = tp; }
  }
  __device__ void sift_down(int i, int n) {
    for (;;) {
      const int l = 2 * i + 1;
      if (l >= n) return;
      int c = l;
      if (l + 1 < n && key_less(hd[l], ho[l], hd[l + 1], ho[l + 1])) c = l + 1;
      if (!key_less(hd[i], ho[i], hd[c], ho[c])) return;
      swap_(i, c);
      i = c;
    }
  }
  __device__ void insert(D dd, int32_t oo, int32_t pp) {
    if (size < k) {
      int i = size++;
      hd[i] = dd; ho[i] = oo;
      if (hp) hp[i] = pp;
      while (i > 0) {
        const int par = (i - 1) >> 1;
        if (!key_less(hd[par], ho[par], hd[i], ho[i])) break;
        swap_(par, i);
        i = par;
      }
    } else {
      hd[0] = dd; ho[0] = oo;
      if (hp) hp[0] = pp;
      sift_down(0, size);
    }
    top_d = hd[0];
    top_o = ho[0];
  }
  // Heapsort in place: slots [0, size) end up ascending.
  __device__ void finish() {
    for (int e = size - 1; e > 0; --e) {
      swap_(0, e);
      sift_down(0, e);
    }
  }
};

// Largest value below x (x >= 0): d < x  <=>  d <= below(x) for every
// non-NaN d, so the radius test and the k-th bound test become one compare.
__device__ __forceinline__ float below(float x) {
  return x > 0.f ? __uint_as_float(__float_as_uint(x) - 1u) : -1.f;
}
__device__ __forceinline__ double below(double x) {
  return x > 0.0 ? __longlong_as_double(__double_as_longlong(x) - 1ll) : -1.0;
}

template <typename D, int DIM, typename State>
__device__ __forceinline__ void knn_scan_leaf(const TreeView<D, DIM>& tv, int32_t a, int32_t b,
                                              const D* q, D r2, int32_t excl, State& st,
                                              D cap = pos_inf<D>()) {
  // Leaves hold at most 256 points; scan in chunks of 64 so the candidate
  // mask fits one 64-bit word. Each chunk first marks the points that pass
  // d <= min(k-th bound, below(r*r)) -- exactly d < r*r and d <= k-th --
  // four loads in flight at a time (eight cost too many registers, and the
  // lost occupancy outweighs the saved instructions), then inserts the
  // marked ones.
  const D r2b = below(r2);
  for (int32_t c0 = a; c0 < b; c0 += 64) {
    const int32_t cnt = b - c0 > 64 ? 64 : b - c0;
    const D kb = st.bound();
    D thr = kb < r2b ? kb : r2b;
    thr = cap < thr ? cap : thr;
    const PointRec<D, DIM>* base = tv.pts + c0;
    uint64_t mask = 0;
    for (int32_t j = 0; j < cnt; j += 4) {
      uint32_t grp = 0;
      if (j + 4 <= cnt) {
        // full group: four independent loads at fixed offsets
        const PointRec<D, DIM>* g = base + j;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const D dd = dist2<D, DIM>(q, g[u].c);
          grp |= dd <= thr ? (1u << u) : 0u;
        }
      } else {
        // tail: indices clamped to the last point, extra bits cleared
        const int32_t last = cnt - 1;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int32_t i = j + u < last ? j + u : last;
          const D dd = dist2<D, DIM>(q, base[i].c);
          grp |= dd <= thr ? (1u << u) : 0u;
        }
        grp &= (1u << (cnt - j)) - 1u;
      }
      mask |= (uint64_t)grp << j;
    }
    const int32_t ex = excl - c0;
    if (ex >= 0 && ex < cnt) mask &= ~(1ull << ex);
    while (mask) {
      const int s = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      const PointRec<D, DIM> r = tv.pts[c0 + s];
      const D dd = dist2<D, DIM>(q, r.c);
      if (st.beats(dd, r.orig)) st.insert(dd, r.orig, c0 + s);
    }
  }
}

__device__ __forceinline__ void set_flip(uint64_t& lo, uint64_t& hi, int depth, bool v) {
  if (depth < 64) {
    const uint64_t bit = 1ull << depth;
    lo = v ? (lo | bit) : (lo & ~bit);
  } else {
    const uint64_t bit = 1ull << (depth - 64);
    hi = v ? (hi | bit) : (hi & ~bit);
  }
}
__device__ __forceinline__ bool get_flip(uint64_t lo, uint64_t hi, int depth) {
  return depth < 64 ? ((lo >> depth) & 1ull) : ((hi >> (depth - 64)) & 1ull);
}

// k-nearest search inside the subtree rooted at `start`, near child first.
template <typename D, int DIM, typename State>
__device__ void knn_subtree(const TreeView<D, DIM>& tv, int32_t start, const D* q, D r2,
                            int32_t excl, State& st) {
  int32_t node = start;
  int depth = 0;
  uint64_t flo = 0, fhi = 0;
  NodeRec cur = tv.node[start];
  for (;;) {
    bool at_leaf = true;
    while (cur.child >= 0) {
      const int32_t c0 = cur.child;
      const D m0 = box_min_dist2<D, DIM>(tv.box[c0], q);
      const D m1 = box_min_dist2<D, DIM>(tv.box[c0 + 1], q);
      const D kb = st.bound();
      const bool n0 = (m0 < r2) && (m0 <= kb);
      const bool n1 = (m1 < r2) && (m1 <= kb);
      if (!(n0 || n1)) { at_leaf = false; break; }
      const bool right = n1 && (!n0 || m1 < m0);
      set_flip(flo, fhi, depth, right);
      node = c0 + (right ? 1 : 0);
      ++depth;
      cur = tv.node[node];
    }
    if (at_leaf) knn_scan_leaf<D, DIM>(tv, cur.a, cur.b, q, r2, excl, st);
    for (;;) {
      if (depth == 0) return;
      const bool is_right = (node & 1) == 0;
      if (is_right == get_flip(flo, fhi, depth - 1)) {
        const int32_t sib = is_right ? node - 1 : node + 1;
        const D ms = box_min_dist2<D, DIM>(tv.box[sib], q);
        if (ms < r2 && ms <= st.bound()) {
          node = sib;
          cur = tv.node[sib];
          break;
        }
      }
      node = cur.parent;
      cur = tv.node[node];
      --depth;
    }
  }
}

// k-nearest search from the root (queries given as points).
template <typename D, int DIM, typename State>
__device__ void knn_search(const TreeView<D, DIM>& tv, const D* q, D r2, int32_t excl, State& st) {
  if (!(box_min_dist2<D, DIM>(tv.box[0], q) < r2)) return;
  knn_subtree<D, DIM>(tv, 0, q, r2, excl, st);
}

// k-nearest search for a query that is the tree point at position `pos`:
// scan its own leaf, then walk up, searching each ancestor's other child
// while that child's box can still hold a better neighbour. Same result as
// knn_search (every box that can hold a neighbour is still visited); the
// nearest candidates are simply found first.
//
// ksel > 0 is the number of neighbours sought. Before the first scan, the
// ksel points nearest to `pos` in tree order (skipping `excl`) give an upper
// bound on the ksel-th distance -- the ksel-th smallest distance is at most
// the largest distance of any ksel distinct candidates -- so the first leaf
// marks only points that can still make the list.
template <typename D, int DIM, typename State>
__device__ void knn_search_from_leaf(const TreeView<D, DIM>& tv, int32_t pos, const D* q, D r2,
                                     int32_t excl, int ksel, State& st) {
  int32_t node = tv.leafof[pos];
  NodeRec cur = tv.node[node];
  D cap = pos_inf<D>();
  const int32_t nsel = ksel + (excl >= 0 ? 1 : 0);
  if (ksel > 0 && cur.b - cur.a >= nsel) {
    // window of nsel positions inside the leaf that contains pos
    int32_t s0 = pos - nsel / 2;
    s0 = s0 < cur.a ? cur.a : s0;
    s0 = s0 > cur.b - nsel ? cur.b - nsel : s0;
    D mx = D(0);
    for (int32_t j = 0; j < nsel; ++j) {
      const int32_t p = s0 + j;
      if (p == excl) continue;
      const D dd = dist2<D, DIM>(q, tv.pts[p].c);
      mx = dd > mx ? dd : mx;
    }
    cap = mx;
  }
  knn_scan_leaf<D, DIM>(tv, cur.a, cur.b, q, r2, excl, st, cap);
  while (cur.parent >= 0) {
    const int32_t sib = (node & 1) ? node + 1 : node - 1;
    const D ms = box_min_dist2<D, DIM>(tv.box[sib], q);
    if (ms < r2 && ms <= st.bound()) knn_subtree<D, DIM>(tv, sib, q, r2, excl, st);
    node = cur.parent;
    cur = tv.node[node];
  }
}

// Next node after finishing `node`'s subtree in left-to-right order, or -1
// when the walk has left the tree (a root has parent -1).
template <typename D, int DIM>
__device__ __forceinline__ int32_t next_in_order(const TreeView<D, DIM>& tv, int32_t node) {
  // roots are exactly the nodes [0, nroots) (parent -1); every other node
  // is in a child pair that starts at an odd id
  for (;;) {
    if (node < tv.nroots) return -1;
    if (node & 1) return node + 1;  // left child: its sibling is next
    node = tv.node[node].parent;
  }
}

// Root of the tree (cloud) holding tree position `pos`.
template <typename D, int DIM>
__device__ __forceinline__ int32_t root_of(const TreeView<D, DIM>& tv, int32_t pos) {
  if (tv.nroots <= 1) return 0;
  int32_t node = tv.leafof[pos];
  for (;;) {
    const int32_t par = tv.node[node].parent;
    if (par < 0) return node;
    node = par;
  }
}

// Calls visit(record) for every point with d < r*r, in tree order (used by
// the r-near normals; the query point itself is included).
template <typename D, int DIM, typename V>
__device__ void rnear_visit(const TreeView<D, DIM>& tv, int32_t root, const D* q, D r2,
                            V& visit) {
  int32_t node = root;
  for (;;) {
    const BoxRec<D, DIM> b = tv.box[node];
    if (box_min_dist2<D, DIM>(b, q) < r2) {
      const NodeRec nr = tv.node[node];
      if (box_max_dist2<D, DIM>(b, q) < r2) {
        for (int32_t p = nr.a; p < nr.b; ++p) visit(tv.pts[p]);
      } else if (nr.child < 0) {
        for (int32_t p = nr.a; p < nr.b; ++p) {
          const PointRec<D, DIM> r = tv.pts[p];
          if (dist2<D, DIM>(q, r.c) < r2) visit(r);
        }
      } else {
        node = nr.child;
        continue;
      }
    }
    node = next_in_order<D, DIM>(tv, node);
    if (node < 0) break;
  }
}

// Resolves the query of thread `t`. Returns false if the thread has no output
// row (ALL mode, original index >= nrows). Sets ok = false for an invalid
// query (NaN/inf point, index of an invalid row, index out of range), which
// yields an empty result.
template <typename D, int DIM>
__device__ __forceinline__ bool resolve_query(const TreeView<D, DIM>& tv, int mode,
                                              const int32_t* __restrict__ qlist,
                                              const D* __restrict__ qpts,
                                              const int32_t* __restrict__ order, int64_t t,
                                              int64_t nrows, D* q, int32_t& excl, int64_t& row,
                                              bool& ok) {
  ok = true;
  excl = -1;   // for self queries: the query's own tree position
  if (order && mode != kQueryAll) t = order[t];
  if (mode == kQueryAll) {
    if (tv.n_dev && t >= *tv.n_dev) return false;
    const PointRec<D, DIM> r = tv.pts[t];
    row = r.orig;
    if (row >= nrows) return false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[d] = r.c[d];
    excl = (int32_t)t;
  } else if (mode == kQueryList) {
    row = t;
    const int32_t o = qlist[t];
    const int32_t p = (o >= 0 && (int64_t)o < tv.n_input) ? tv.fwd[o] : -1;
    if (p < 0) {
      ok = false;
    } else {
      const PointRec<D, DIM> r = tv.pts[p];
#pragma unroll
      for (int d = 0; d < DIM; ++d) q[d] = r.c[d];
      excl = p;
    }
  } else {
    row = t;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      q[d] = qpts[t * DIM + d];
      ok = ok && finite_val(q[d]);
    }
  }
  return true;
}

}  // namespace pk
