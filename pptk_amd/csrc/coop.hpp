// Wave-cooperative k-nearest search for self queries on float trees.
//
// The 64 lanes of a wave hold 64 consecutive tree positions; each lane's
// query is its own point, and neighbours are ordered by the packed key
// (squared distance bits << 32 | original index), as in TopK<float, K>.
//
// Instead of 64 independent traversals, the wave walks its cloud's tree once:
//   * a node is entered when any lane's query can still improve from it
//     (box lower bound <= that lane's bound, below);
//   * a leaf's points are loaded 64 at a time, one per lane, into LDS, and
//     filtered against the bounding box of the queries whose ball reaches
//     the leaf and the largest bound among them; each surviving point is
//     read back by every lane (one broadcast LDS read), every lane computes
//     its own distance, and a lane whose key is below its bound appends the
//     key to its own buffer in LDS (branch-free: every lane writes its key to
//     the slot past its last one, and only a taken key advances the count);
//   * when any lane's buffer is full, every lane sorts its buffer (bitonic
//     network) and merges it into its sorted top-k (half-cleaner against the
//     reversed buffer, then a bitonic merge), all as 64-bit key min/max.
// A lane's bound is the smaller of its k-th key and its cap: the largest
// distance of the k points nearest to it in tree order within a compact
// Morton cell (any k distinct points bound the k-th distance from above),
// limited to below r*r. Lanes whose cap is far above the wave's median walk
// separately, as do lanes of different clouds.
//
// Same result as knn_search_from_leaf: every node that can hold a key below a
// lane's final k-th key is visited (bounds only shrink), every point of a
// visited leaf that can beat the lane's bound at the time is offered to it,
// and a sorted top-k of the smallest keys is independent of insertion order.
#pragma once
#include "traverse.hpp"

namespace pk {

// Buffer slots per lane for a top-k of K: 8 (4 for K = 4).
template <int K> constexpr int coop_slots() { return K < 8 ? K : 8; }
// A leaf point staged in LDS: the original index first, so that a lane's
// read of it leaves the index in an even register, the low half of the key
// (distance bits, index) it builds.
template <int DIM>
struct alignas(16) CoopRec {
  int32_t orig;
  float c[DIM];
};
// One wave's LDS: a leaf chunk of 64 records, then the buffer area. The
// buffer area holds the lanes' key buffers during the search; afterwards it
// may hold up to NPOS neighbour positions per lane (coop_positions).
template <int DIM> constexpr int coop_chunk_bytes() { return 64 * (int)sizeof(CoopRec<DIM>); }
template <int K, int NPOS> constexpr int coop_buf_bytes() {
  return coop_slots<K>() * 64 * 8 > (NPOS * 64 * 4 + 15) / 16 * 16
             ? coop_slots<K>() * 64 * 8
             : (NPOS * 64 * 4 + 15) / 16 * 16;
}
template <int K, int DIM, int NPOS> constexpr int coop_wave_bytes() {
  return coop_chunk_bytes<DIM>() + coop_buf_bytes<K, NPOS>();
}
// After the search: slot j of this lane's neighbour positions (j < NPOS).
template <int DIM>
__device__ __forceinline__ int32_t* coop_positions(uint64_t* wave_lds) {
  return (int32_t*)((char*)wave_lds + coop_chunk_bytes<DIM>()) + (threadIdx.x & 63);
}

// Compare-exchange of two keys (a <= b after). One asm block, b updated in
// place: the scheduler cannot hoist all the minima of a network stage above
// its maxima, which would keep both generations of the keys live at once.
__device__ __forceinline__ void key_cx(uint64_t& a, uint64_t& b) {
  uint64_t lo;
  asm("v_min_f64 %0, %1, %2\n\tv_max_f64 %1, %1, %2" : "=&v"(lo), "+v"(b) : "v"(a));
  a = lo;
}

// Bitonic sort of M keys, ascending (M a power of two).
template <int M>
__device__ __forceinline__ void key_sort(uint64_t (&v)[M]) {
#pragma unroll
  for (int size = 2; size <= M; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int i = 0; i < M; ++i) {
        const int j = i ^ stride;
        if (j > i) {
          if ((i & size) == 0) key_cx(v[i], v[j]);
          else key_cx(v[j], v[i]);
        }
      }
    }
  }
}

// a (K keys, ascending) <- the K smallest of a and b (M <= K keys, ascending).
// a[i] = min(a[i], b[K-1-i]) (b padded with +inf below) keeps the K smallest
// as a bitonic sequence; the bitonic merge sorts it.
template <int K, int M>
__device__ __forceinline__ void key_merge(uint64_t (&a)[K], const uint64_t (&b)[M]) {
#pragma unroll
  for (int i = K - M; i < K; ++i) a[i] = key_min(a[i], b[K - 1 - i]);
#pragma unroll
  for (int stride = K >> 1; stride > 0; stride >>= 1) {
#pragma unroll
    for (int i = 0; i < K; ++i)
      if ((i & stride) == 0) key_cx(a[i], a[i + stride]);
  }
}

// Floats as integers of the same order (no NaN): the magnitude bits of a
// negative value are flipped. The map is its own inverse.
__device__ __forceinline__ int32_t float_ord(float x) {
  const int32_t b = __float_as_int(x);
  return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float ord_float(int32_t k) {
  return __int_as_float(k ^ ((k >> 31) & 0x7fffffff));
}
// Wave maximum / minimum of x (finite or infinite, never NaN), as a scalar:
// integer max/min of the order-preserving map, within each row of 16 lanes by
// DPP quad permutations and row rotations (each folded into the max/min),
// then across the four rows' first lanes as scalars.
__device__ __forceinline__ float wave_max_f(float x) {
  int32_t k = float_ord(x);
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0xb1, 0xf, 0xf, true));    // quad_perm [1,0,3,2]
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0x4e, 0xf, 0xf, true));    // quad_perm [2,3,0,1]
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0x124, 0xf, 0xf, true));   // row_ror:4
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0x128, 0xf, 0xf, true));   // row_ror:8
  const int32_t r0 = __builtin_amdgcn_readlane(k, 0), r1 = __builtin_amdgcn_readlane(k, 16);
  const int32_t r2 = __builtin_amdgcn_readlane(k, 32), r3 = __builtin_amdgcn_readlane(k, 48);
  return ord_float(max(max(r0, r1), max(r2, r3)));
}
// The same maxima and minima, carried across the rows by row broadcasts
// (lane 15 into rows 1 and 3, then lane 31 into rows 2 and 3, each folded
// into the max/min) and read from lane 63: fewer instructions, one more
// vector register held until the read.
__device__ __forceinline__ float wave_max_f_bcast(float x) {
  int32_t k = float_ord(x);
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0xb1, 0xf, 0xf, true));
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0x4e, 0xf, 0xf, true));
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0x124, 0xf, 0xf, true));
  k = max(k, __builtin_amdgcn_mov_dpp(k, 0x128, 0xf, 0xf, true));
  k = max(k, __builtin_amdgcn_update_dpp(INT32_MIN, k, 0x142, 0xa, 0xf, false));  // row_bcast:15
  k = max(k, __builtin_amdgcn_update_dpp(INT32_MIN, k, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return ord_float(__builtin_amdgcn_readlane(k, 63));
}
__device__ __forceinline__ float wave_min_f_bcast(float x) {
  int32_t k = float_ord(x);
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0xb1, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0x4e, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0x124, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0x128, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_update_dpp(INT32_MAX, k, 0x142, 0xa, 0xf, false));
  k = min(k, __builtin_amdgcn_update_dpp(INT32_MAX, k, 0x143, 0xc, 0xf, false));
  return ord_float(__builtin_amdgcn_readlane(k, 63));
}
__device__ __forceinline__ float wave_min_f(float x) {
  int32_t k = float_ord(x);
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0xb1, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0x4e, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0x124, 0xf, 0xf, true));
  k = min(k, __builtin_amdgcn_mov_dpp(k, 0x128, 0xf, 0xf, true));
  const int32_t r0 = __builtin_amdgcn_readlane(k, 0), r1 = __builtin_amdgcn_readlane(k, 16);
  const int32_t r2 = __builtin_amdgcn_readlane(k, 32), r3 = __builtin_amdgcn_readlane(k, 48);
  return ord_float(min(min(r0, r1), min(r2, r3)));
}

// Number of set bits of mask in the lanes below this one (v_mbcnt).
__device__ __forceinline__ int32_t rank_below(uint64_t mask) {
  return (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                            __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// x as a wave-uniform (scalar) value; x must be equal in every active lane.
__device__ __forceinline__ float uniform_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ float key_dist(uint64_t key) {
  return __uint_as_float((uint32_t)(key >> 32));
}

// Root (cloud) holding tree position pos in a forest whose roots [0, nroots)
// cover consecutive position ranges in root order: the last root whose range
// starts at or before pos (empty roots start where the next one does).
template <int DIM>
__device__ __forceinline__ int32_t root_by_range(const TreeView<float, DIM>& tv, int32_t pos) {
  int32_t lo = 0, hi = tv.nroots - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (tv.node[mid].a <= pos) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int DIM, int K>
struct CoopSearch {
  static constexpr int M = coop_slots<K>();
  const TreeView<float, DIM>& tv;
  TopK<float, K, false>& st;
  uint64_t* lbuf;                // buffers: slot s of lane l at lbuf[s * 64 + l]
  CoopRec<DIM>* chunk;           // the loaded leaf chunk's surviving points
  const float* q;
  int lane;
  int cnt;                       // keys in this lane's buffer
  uint64_t bkey;                 // keys below bkey are taken; 0 = not searching
  float bd;                      // bkey's distance, -1 when not searching

  __device__ __forceinline__ CoopSearch(const TreeView<float, DIM>& tv_, TopK<float, K, false>& st_,
                                        uint64_t* wave_lds, const float* q_)
      : tv(tv_), st(st_), lbuf((uint64_t*)((char*)wave_lds + coop_chunk_bytes<DIM>())),
        chunk((CoopRec<DIM>*)wave_lds),
        q(q_), lane(threadIdx.x & 63), cnt(0), bkey(0), bd(-1.0f) {}

  __device__ __forceinline__ void refresh(bool part) {
    // the cap is in the top-k: its empty slots start at the cap key
    const uint64_t b = st.key[K - 1];
    bkey = part ? b : 0ull;
    bd = part ? key_dist(b) : -1.0f;
  }

  // Sorts every lane's buffer into its top-k and empties the buffers.
  __device__ __forceinline__ void flush(bool part) {
    // a slot at or past cnt gets distance bits +inf: it sorts after every
    // real key and, if it reaches the top-k, reads as no neighbour
    uint64_t b[M];
#pragma unroll
    for (int s = 0; s < M; ++s) {
      const uint64_t v = lbuf[s * 64 + lane];
      const uint32_t hi = s < cnt ? (uint32_t)(v >> 32) : 0x7f800000u;
      b[s] = ((uint64_t)hi << 32) | (uint32_t)v;
    }
    cnt = 0;
    key_sort<M>(b);
    key_merge<K, M>(st.key, b);
    refresh(part);
  }

  // Entry test of a node for the searching lanes (wave-uniform result).
  __device__ __forceinline__ bool need(int32_t node) const {
    const BoxRec<float, DIM> bx = tv.box[node];
    return __ballot(box_min_dist2<float, DIM>(bx, q) <= bd) != 0ull;
  }

  // Advances the walk from `node` (just finished, or a leaf just scanned) to
  // the next leaf to scan; returns false when the walk is over. Touches only
  // wave-uniform state (the top-k is read, never written).
  __device__ __forceinline__ bool next_leaf(int32_t& node, int& depth, uint64_t& flo,
                                            uint64_t& fhi, int32_t mid, bool backtrack) {
    for (;;) {
      if (backtrack) {
        bool found = false;
        while (depth > 0) {
          const bool is_right = (node & 1) == 0;
          if (is_right == get_flip(flo, fhi, depth - 1)) {
            const int32_t sib = is_right ? node - 1 : node + 1;
            if (need(sib)) {
              node = sib;
              found = true;
              break;
            }
          }
          node = tv.node[node].parent;
          --depth;
        }
        if (!found) return false;
      }
      node = __builtin_amdgcn_readfirstlane(node);
      const int32_t child = tv.node[node].child;
      if (child < 0) return true;
      const int32_t split = tv.node[child].b;
      const bool right_first = mid >= split;
      const int32_t first = right_first ? child + 1 : child;
      const int32_t second = right_first ? child : child + 1;
      // flip = "right child first" in both cases: after `second` (entered
      // because `first` is not needed) the backtrack must not test `first`
      set_flip(flo, fhi, depth, right_first);
      if (need(first)) {
        node = first;
        ++depth;
        backtrack = false;
      } else if (need(second)) {
        node = second;
        ++depth;
        backtrack = false;
      } else {
        backtrack = true;
      }
    }
  }
};

// k-nearest search for the tree point at position pos (record index
// self_orig), done by the whole wave. Every lane of the wave must call it
// (inactive lanes with active = false), and lane l must pass the wave's first
// position + l as pos (consecutive positions, inactive lanes included).
// excl_self: the point itself is not a neighbour (kNN self queries);
// otherwise it is one (normals). st must be freshly init(k); wave_lds is this
// wave's coop_wave_bytes<K, DIM, NPOS>() of LDS (any NPOS). ROWBCAST picks
// the row-broadcast form of the reductions at each leaf (fewer instructions,
// one more vector register).
template <int DIM, int K, bool ROWBCAST = true>
__device__ __forceinline__ void knn_coop_self(const TreeView<float, DIM>& tv, bool active,
                                              int32_t pos, int32_t self_orig, const float* q,
                                              float r2, int k, bool excl_self,
                                              TopK<float, K, false>& st, uint64_t* wave_lds) {
  CoopSearch<DIM, K> cs(tv, st, wave_lds, q);
  constexpr int M = CoopSearch<DIM, K>::M;
  // the wave's first position: lane l's position is pos0 + l
  const int32_t pos0 = __builtin_amdgcn_readfirstlane(pos - (int32_t)(threadIdx.x & 63));
  int32_t root = 0;
  if (tv.nroots > 1 && active) root = root_by_range<DIM>(tv, pos);
  // cap: the largest distance of the k points nearest to pos in tree order
  // inside the smallest node around pos that holds enough of them (a compact
  // Morton cell; the point itself excluded for self queries)
  float cap = pos_inf<float>();
  if (active) {
    const int32_t wlen = k + (excl_self ? 1 : 0);
    NodeRec nr = tv.node[tv.leafof[pos]];
    while (nr.b - nr.a < wlen && nr.parent >= 0) nr = tv.node[nr.parent];
    if (nr.b - nr.a >= wlen) {
      int32_t wlo = pos - wlen / 2;
      wlo = wlo > nr.b - wlen ? nr.b - wlen : wlo;
      wlo = wlo < nr.a ? nr.a : wlo;
      float mx = 0.0f;
      for (int32_t j = 0; j < wlen; ++j) {
        const int32_t p = wlo + j;
        if (excl_self && p == pos) continue;
        const float dd = dist2<float, DIM>(q, tv.pts[p].c);
        mx = dd > mx ? dd : mx;
      }
      cap = mx;
    }
  }
  float capd;
  {
    const float r2b = below(r2);
    capd = cap < r2b ? cap : r2b;
    // the cap as a key, (cap, 0xffffffff), fills the empty slots: the top-k's
    // k-th key is then the smaller of the k-th key found and the cap key, and
    // a slot still holding it at the end reads as no neighbour (index -1)
    const uint64_t capkey = ((uint64_t)__float_as_uint(capd) << 32) | 0xffffffffull;
#pragma unroll
    for (int s = 0; s < K; ++s) st.key[s] = key_min(st.key[s], capkey);
  }
  // Lanes whose cap is far above the wave's median (binary exponent more than
  // 3 above the median's: about 8x in squared distance) walk in a segment of
  // their own, so that one isolated point does not widen every lane's walk.
  bool outlier = false;
  {
    const uint32_t e = active ? (__float_as_uint(capd) >> 23) : 511u;
    const int target = (__popcll(__ballot(active)) + 1) >> 1;
    uint32_t lo = 0, hi = 255;
    while (lo < hi) {
      const uint32_t m = (lo + hi) >> 1;
      if (__popcll(__ballot(e <= m)) >= target) hi = m;
      else lo = m + 1;
    }
    outlier = active && e > lo + 3;
  }
  // The inner loop only reads the top-k: it takes the loaded chunk's
  // candidates, loads the next chunk, advances the walk to its next leaf, or
  // starts the walk of the next cloud present in the wave; it stops when a
  // buffer is full or the search is over. The outer loop then merges the
  // buffers: one unconditional merge site keeps the top-k in the same
  // registers.
  uint64_t todo = __ballot(active);
  bool part = false;
  BoxRec<float, DIM> lqb{};       // bounding box of the queries that need the current leaf
  bool inleaf = false;            // this lane needs the current leaf
  float rmax = -1.0f;             // largest bound among them
  int32_t mid = 0, node = 0, depth = 0;
  uint64_t flo = 0, fhi = 0;
  int32_t c0 = 0, cend = 0;       // leaf chunks still to load: [c0, cend)
  int ci = 0, ncand = 0;          // next and number of candidates in the LDS chunk
  bool walking = false, done = false;
  do {
    for (;;) {
      bool full = false;
      if (ci < ncand) {
        do {
          const CoopRec<DIM> cr = cs.chunk[ci];
          ++ci;
          const float dd = dist2<float, DIM>(q, cr.c);
          const uint64_t key = ((uint64_t)__float_as_uint(dd) << 32) | (uint32_t)cr.orig;
          // the own point has the lane's own index (unique in its cloud)
          const bool take = (key < cs.bkey) & !(excl_self & (cr.orig == self_orig));
          // written whether taken or not: slot cnt is past the last key
          cs.lbuf[cs.cnt * 64 + cs.lane] = key;
          cs.cnt += take ? 1 : 0;
          full = __ballot(cs.cnt == M) != 0ull;
        } while ((ci < ncand) & !full);
        if (full) break;
      }
      if (c0 < cend) {
        // the chunk's surviving points are compacted, in tree order, to the
        // front of the LDS chunk
        const int32_t p = c0 + cs.lane;
        const bool valid = p < cend;
        PointRec<float, DIM> rec{};
        if (valid) rec = tv.pts[p];
        const bool near = box_min_dist2<float, DIM>(lqb, rec.c) <= rmax;
        const bool pass = valid & near;
        const uint64_t pm = __ballot(valid) & __ballot(near);
        if (pass) {
          CoopRec<DIM> cr;
          cr.orig = rec.orig;
#pragma unroll
          for (int d = 0; d < DIM; ++d) cr.c[d] = rec.c[d];
          cs.chunk[rank_below(pm)] = cr;
        }
        ncand = __popcll(pm);
        ci = 0;
        c0 += 64;
        continue;
      }
      bool leaf = false;
      if (walking) leaf = cs.next_leaf(node, depth, flo, fhi, mid, true);
      while (!leaf && todo) {
        // start the walk of the next cloud present in the wave
        const int L0 = __builtin_ctzll(todo);
        const int32_t R = rl_i32(root, L0);
        const bool cls = ((__ballot(outlier) >> L0) & 1ull) != 0ull;
        part = active && root == R && outlier == cls;
        const uint64_t pm = __ballot(part);
        todo &= ~pm;
        cs.refresh(part);
        mid = pos0 + ((__builtin_ctzll(pm) + (63 - __builtin_clzll(pm))) >> 1);
        node = R;
        depth = 0;
        flo = fhi = 0;
        walking = cs.need(R);
        if (walking) leaf = cs.next_leaf(node, depth, flo, fhi, mid, false);
      }
      if (!leaf) {
        done = true;
        break;
      }
      walking = true;
      const NodeRec nr = tv.node[node];
      c0 = nr.a;
      cend = nr.b;
      // the leaf's points are filtered against the lanes that need the leaf
      // only: any other lane is farther from every point than its bound
      {
        const BoxRec<float, DIM> lb = tv.box[node];
        inleaf = box_min_dist2<float, DIM>(lb, q) <= cs.bd;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const float lo = inleaf ? q[d] : pos_inf<float>();
          const float hi = inleaf ? q[d] : -pos_inf<float>();
          if constexpr (ROWBCAST) {
            lqb.lo[d] = uniform_f(wave_min_f_bcast(lo));
            lqb.hi[d] = uniform_f(wave_max_f_bcast(hi));
          } else {
            lqb.lo[d] = uniform_f(wave_min_f(lo));
            lqb.hi[d] = uniform_f(wave_max_f(hi));
          }
        }
        if constexpr (ROWBCAST) rmax = uniform_f(wave_max_f_bcast(inleaf ? cs.bd : -1.0f));
        else rmax = uniform_f(wave_max_f(inleaf ? cs.bd : -1.0f));
      }
    }
    cs.flush(part);
    rmax = uniform_f(wave_max_f(inleaf ? cs.bd : -1.0f));
  } while (!done);
}

}  // namespace pk
