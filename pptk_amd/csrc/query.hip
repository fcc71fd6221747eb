// kNN and r-near queries on a device-resident tree.
//
// kNN  (pptk_knn)         : one lane per query, stackless near-first search,
//                           top-k in registers (k <= 32) or in a global-memory
//                           heap (k > 32).
// r-near (pptk_rnear_*)   : count pass -> exclusive scan -> fill pass into
//                           CSR (int64 offsets, int32 original indices).
// Reference entry points: KdTree::KNearestNeighbors[Self] and
// RNearNeighbors[Self] (pptk/kdtree/src/kdtree-impl.h:873-1069).
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "coop.hpp"
#include "morton.hpp"
#include "query.hpp"
#include "sort.hpp"
#include "traverse.hpp"
#include "tree.hpp"

namespace pk {

namespace {

constexpr int kBlock = 256;

template <typename D, int DIM, int K>
__global__ __launch_bounds__(kBlock) void knn_kernel(TreeView<D, DIM> tv, int mode,
                                                     const int32_t* __restrict__ qlist,
                                                     const D* __restrict__ qpts,
                                                     const int32_t* __restrict__ order, int64_t m,
                                                     int64_t nrows, int k, D r2,
                                                     int32_t* __restrict__ out_idx,
                                                     int32_t* __restrict__ out_cnt,
                                                     D* __restrict__ out_dist) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  D q[DIM];
  int32_t excl;
  int64_t row;
  bool ok;
  if (!resolve_query<D, DIM>(tv, mode, qlist, qpts, order, t, nrows, q, excl, row, ok)) return;
  TopK<D, K, false> st;
  st.init(k);
  if (ok) {
    if (excl >= 0) knn_search_from_leaf<D, DIM>(tv, excl, q, r2, excl, k, st);
    else knn_search<D, DIM>(tv, q, r2, excl, st);
  }
  int32_t c = 0;
  int32_t* orow = out_idx + row * k;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (s >= K - k) {
      const D ds = st.dist(s);
      const bool f = st.found(s);
      orow[s - (K - k)] = f ? st.idx(s) : -1;
      if (out_dist) out_dist[row * k + (s - (K - k))] = f ? ds : pos_inf<D>();
      c += f ? 1 : 0;
    }
  }
  out_cnt[row] = c;
}

// kNN self queries over every point of a float tree (ALL mode): one
// wave-cooperative search per wave (coop.hpp). Lanes past the end stay in
// the wave as inactive lanes: the search needs every lane.
template <int DIM, int K>
__global__ __launch_bounds__(kBlock, K <= 16 && DIM <= 3 ? 8 : 1) void knn_coop_kernel(TreeView<float, DIM> tv, int64_t m,
                                                          int64_t nrows, int k, float r2,
                                                          int32_t* __restrict__ out_idx,
                                                          int32_t* __restrict__ out_cnt,
                                                          float* __restrict__ out_dist,
                                                          int xrun) {
  __shared__ uint64_t lds[kBlock / 64 * coop_wave_bytes<K, DIM, 0>() / 8];
  const int64_t t = xcd_block(blockIdx.x, gridDim.x, xrun) * kBlock + threadIdx.x;
  float q[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) q[d] = 0.0f;
  int32_t row = -1;   // original row (int32, like every original index)
  if (t < m && !(tv.n_dev && t >= *tv.n_dev)) {
    const PointRec<float, DIM> r = tv.pts[t];
    if (r.orig < nrows) {
      row = r.orig;
#pragma unroll
      for (int d = 0; d < DIM; ++d) q[d] = r.c[d];
    }
  }
  TopK<float, K, false> st;
  st.init(k);
  // at DIM 3 and K 16 this kernel holds 63 of its 64 vector registers; the
  // row-broadcast reductions would spill there
  knn_coop_self<DIM, K, !(DIM == 3 && K == 16)>(tv, row >= 0, (int32_t)t, row, q, r2, k, true, st,
                        lds + (threadIdx.x >> 6) * (coop_wave_bytes<K, DIM, 0>() / 8));
  if (row < 0) return;
  int32_t c = 0;
  int32_t* orow = out_idx + (int64_t)row * k;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (s >= K - k) {
      const float ds = st.dist(s);
      const bool f = st.found(s);
      orow[s - (K - k)] = f ? st.idx(s) : -1;
      if (out_dist) out_dist[(int64_t)row * k + (s - (K - k))] = f ? ds : pos_inf<float>();
      c += f ? 1 : 0;
    }
  }
  out_cnt[row] = c;
}

template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void knn_large_kernel(TreeView<D, DIM> tv, int mode,
                                                           const int32_t* __restrict__ qlist,
                                                           const D* __restrict__ qpts,
                                                           const int32_t* __restrict__ order,
                                                           int64_t t0,
                                                           int64_t t1, int64_t nrows, int k, D r2,
                                                           int32_t* __restrict__ out_idx,
                                                           int32_t* __restrict__ out_cnt,
                                                           D* __restrict__ ws_d,
                                                           D* __restrict__ out_dist) {
  const int64_t t = t0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= t1) return;
  D q[DIM];
  int32_t excl;
  int64_t row;
  bool ok;
  if (!resolve_query<D, DIM>(tv, mode, qlist, qpts, order, t, nrows, q, excl, row, ok)) return;
  D* hd = ws_d + (t - t0) * k;
  int32_t* ho = out_idx + row * k;
  HeapK<D> st;
  st.init(hd, ho, nullptr, k);
  if (ok) {
    if (excl >= 0) knn_search_from_leaf<D, DIM>(tv, excl, q, r2, excl, k, st);
    else knn_search<D, DIM>(tv, q, r2, excl, st);
  }
  st.finish();
  for (int s = st.size; s < k; ++s) ho[s] = -1;
  out_cnt[row] = st.size;
  if (out_dist)
    for (int s = 0; s < k; ++s) out_dist[row * k + s] = s < st.size ? hd[s] : pos_inf<D>();
}

enum : int { kItemNone = 0, kItemRange = 1, kItemLeaf = 2 };

__device__ __forceinline__ int32_t wave_incl_scan_i32(int32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// Lane whose segment [incl - size, incl) of a wave-wide inclusive prefix
// `incl` contains slot s (0 <= s < total): the smallest lane with incl > s.
__device__ __forceinline__ int owner_of(int32_t incl, int32_t s) {
  int lo = 0, hi = 63;
#pragma unroll
  for (int it = 0; it < 6; ++it) {
    const int mid = (lo + hi) >> 1;
    const int32_t v = __shfl(incl, mid);
    if (v > s) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// r-near, count (FILL = false) or fill (FILL = true).
//
// Each lane walks its query's tree in left-to-right order doing box tests
// only, and stops at the next work item: a node wholly inside the ball (box
// upper bound < r*r, RNearNeighborsHelper, kdtree-impl.h:527-529) or a leaf
// that straddles the ball. Every round, each lane posts one item and the
// wave serves all posted items together: a wave-wide prefix sum over the
// item sizes spreads their elements over the 64 lanes, each lane finds its
// element's owner by binary search over the prefix, and
//   * range elements are copied (fill) or just counted (count: sizes only);
//   * leaf elements are distance-tested against the owner's query; hits are
//     ranked within the owner's segment with a ballot.
// Output order within a query is tree order, identical in both passes.
template <typename D, int DIM, bool FILL>
__global__ __launch_bounds__(kBlock) void rnear_kernel(TreeView<D, DIM> tv, int mode,
                                                       const int32_t* __restrict__ qlist,
                                                       const D* __restrict__ qpts,
                                                       const int32_t* __restrict__ order,
                                                       int64_t m, int64_t nrows, D r2,
                                                       int32_t* __restrict__ counts,
                                                       const int64_t* __restrict__ offsets,
                                                       int64_t total_arg,
                                                       const int64_t* __restrict__ total_dev,
                                                       int32_t* __restrict__ out) {
  // the fill's output length: total_arg, or read here when given on the device
  const int64_t total = total_dev ? (*total_dev > 0 ? *total_dev : 0) : total_arg;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  D q[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) q[d] = D(0);
  int32_t excl = -1;
  int64_t row = 0;
  bool ok = false, has_row = false;
  if (t < m)
    has_row = resolve_query<D, DIM>(tv, mode, qlist, qpts, order, t, nrows, q, excl, row, ok);
  bool active = has_row && ok;
  int64_t base = 0, cap = 0;
  if (FILL && active) {
    // writes stay inside [offsets[row], offsets[row + 1]) clamped to [0, total)
    int64_t e = offsets[row + 1];
    base = offsets[row];
    base = base < 0 ? 0 : (base > total ? total : base);
    e = e < base ? base : (e > total ? total : e);
    cap = e - base;
  }
  int32_t cnt = 0;  // outputs so far for this lane's query
  int32_t node = 0;
  if (active && tv.nroots > 1) node = root_of<D, DIM>(tv, excl);
  // A range item is posted at most kMaxRange elements at a time (the rest is
  // posted in the next rounds), which keeps the wave-wide prefix in int32.
  constexpr int32_t kMaxRange = 1 << 24;
  int32_t rest_a = 0, rest_b = 0;
  // Every lane stays in the loop until the whole wave is done: the item
  // service below is executed by all 64 lanes together.
  while (__ballot(active || rest_b > rest_a)) {
    int kind = kItemNone;
    int32_t ia = 0, ib = 0;
    if (rest_b > rest_a) {
      kind = kItemRange;
      ia = rest_a;
      ib = rest_b - rest_a > kMaxRange ? rest_a + kMaxRange : rest_b;
      rest_a = ib;
    } else if (active) {
      for (;;) {
        const BoxRec<D, DIM> b = tv.box[node];
        if (box_min_dist2<D, DIM>(b, q) < r2) {
          const NodeRec nr = tv.node[node];
          if (box_max_dist2<D, DIM>(b, q) < r2) {
            kind = kItemRange;
            ia = nr.a;
            ib = nr.b;
          } else if (nr.child < 0) {
            kind = kItemLeaf;
            ia = nr.a;
            ib = nr.b;
          } else {
            node = nr.child;
            continue;
          }
        }
        node = next_in_order<D, DIM>(tv, node);
        if (node < 0) active = false;
        if (kind != kItemNone || !active) break;
      }
      if (kind == kItemRange && ib - ia > kMaxRange) {
        rest_a = ia + kMaxRange;
        rest_b = ib;
        ib = rest_a;
      }
    }
    const bool has_ex = excl >= ia && excl < ib;
    // ---- ranges: every element except the query itself is an output ----
    if (kind == kItemRange) cnt += (ib - ia) - (has_ex ? 1 : 0);
    if constexpr (FILL) {
      const int32_t esz = kind == kItemRange ? ib - ia : 0;
      const int32_t incl = wave_incl_scan_i32(esz);
      const int32_t tot = __shfl(incl, 63);
      // this lane's outputs before the current range item
      const int32_t start = kind == kItemRange ? cnt - ((ib - ia) - (has_ex ? 1 : 0)) : cnt;
      for (int32_t s0 = 0; s0 < tot; s0 += 64) {
        const int32_t sl = s0 + lane;
        const int src = owner_of(incl, sl < tot ? sl : tot - 1);
        const int32_t o_incl = __shfl(incl, src);
        const int32_t o_a = __shfl(ia, src);
        const int32_t o_b = __shfl(ib, src);
        const int32_t o_ex = __shfl(excl, src);
        const int32_t o_start = __shfl(start, src);
        const int64_t o_base = __shfl(base, src);
        const int64_t o_cap = __shfl(cap, src);
        if (sl < tot) {
          const int32_t p = o_a + (sl - (o_incl - (o_b - o_a)));
          if (p != o_ex) {
            const bool ex_in = o_ex >= o_a && o_ex < o_b;
            const int64_t o = (int64_t)o_start + (p - o_a) - ((ex_in && p > o_ex) ? 1 : 0);
            if (o < o_cap) out[o_base + o] = tv.orig[p];
          }
        }
      }
    }
    // ---- leaves: test every point against its owner's query ----
    {
      const int32_t lsz = kind == kItemLeaf ? ib - ia : 0;
      const int32_t incl = wave_incl_scan_i32(lsz);
      const int32_t tot = __shfl(incl, 63);
      const int32_t seg0 = incl - lsz;  // this lane's first slot
      int32_t hits_mine = 0;
      for (int32_t s0 = 0; s0 < tot; s0 += 64) {
        const int32_t sl = s0 + lane;
        const int src = owner_of(incl, sl < tot ? sl : tot - 1);
        const int32_t o_seg0 = __shfl(seg0, src);
        const int32_t o_a = __shfl(ia, src);
        const int32_t o_ex = __shfl(excl, src);
        D oq[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) oq[d] = __shfl(q[d], src);
        bool pass = false;
        int32_t og = 0;
        const int32_t p = o_a + (sl - o_seg0);
        if (sl < tot) {
          const PointRec<D, DIM> r = tv.pts[p];
          pass = dist2<D, DIM>(oq, r.c) < r2 && p != o_ex;
          og = r.orig;
        }
        const uint64_t hits = __ballot(pass);
        if constexpr (FILL) {
          // rank among the owner's hits: earlier chunks + earlier lanes of
          // the owner's segment in this chunk
          const int32_t lo_lane = o_seg0 > s0 ? o_seg0 - s0 : 0;
          const uint64_t seg_lt = lt_mask & ~((1ull << lo_lane) - 1ull);
          const int32_t o_done = __shfl(hits_mine, src);
          const int32_t o_cnt = __shfl(cnt, src);
          const int64_t o_base = __shfl(base, src);
          const int64_t o_cap = __shfl(cap, src);
          const int64_t o = (int64_t)o_cnt + o_done + __popcll(hits & seg_lt);
          if (pass && o < o_cap) out[o_base + o] = og;
        }
        // this lane's hits in this chunk (its segment [seg0, incl))
        if (lsz > 0) {
          const int32_t a0 = seg0 > s0 ? seg0 - s0 : 0;
          const int32_t a1 = incl < s0 + 64 ? incl - s0 : 64;
          if (a1 > a0) {
            const uint64_t hi_mask = a1 >= 64 ? ~0ull : ((1ull << a1) - 1ull);
            const uint64_t mine = hi_mask & ~((1ull << a0) - 1ull);
            hits_mine += __popcll(hits & mine);
          }
        }
      }
      cnt += hits_mine;
    }
  }
  if constexpr (!FILL) {
    if (has_row) counts[row] = ok ? cnt : 0;
  }
}

// Event lists of an r-near count pass, replayed by the fill pass of the same
// queries so that the fill does not walk the tree again. An event is a node
// at which some of a count wave's lanes take work: lanes whose ball holds
// the node's box (low 32 bits of the mask: the whole range [a, b) is theirs)
// and lanes whose ball straddles a leaf (high 32 bits: the leaf's points are
// tested one by one). A count wave (pk <= 32 queries) writes its events in
// its traversal order into chunks of 64 taken from a shared pool by one
// atomic add each; its header holds the first chunk and the event count, or
// -1 when the pool ran out (its fill waves then walk the tree). A fill wave
// serves a contiguous subset of a count wave's lanes; it replays the events
// in the recorded order, so every list is written in the same (tree) order
// as by a walk.
constexpr int kEventChunk = 64;
struct RnearEventView {
  int32_t* node = nullptr;        // [nchunks * 64]: event node
  uint64_t* mask = nullptr;       // [nchunks * 64]: inside lanes | straddling lanes << 32
  int32_t* link = nullptr;        // [nchunks]: next chunk of the same wave
  int32_t* hdr = nullptr;         // [2 * count waves]: first chunk, events (-1: not kept)
  uint32_t* head = nullptr;       // chunks taken
  int32_t nchunks = 0;
  int32_t pk = 0;                 // queries per count wave
  const int32_t* same = nullptr;  // fill: nonzero if its queries are the count's
};

// r-near on a single-root tree, one traversal per wave (count: FILL = false,
// fill: FILL = true). A wave takes pk <= 32 consecutive queries in lanes
// 0 .. pk-1 (lanes 32 .. 32+pk-1 hold the same queries for the far-bound
// half of each node test; the others idle); consecutive queries are neighbours on the Morton curve, so their
// balls overlap: the wave walks the tree once, in left-to-right order, and
// visits a node when any lane's ball reaches it. Fewer queries per wave
// split the dense regions' work (thousands of hits per query) over more
// waves, at the price of more node visits per query. wave_order, when
// given, maps launch order to wave index (heaviest waves first).
// Every lane classifies the node against its own ball:
//   * inside (box upper bound < r*r, RNearNeighborsHelper,
//     kdtree-impl.h:527-529): the lane takes the whole range [a, b) and skips
//     the node's subtree (it remembers b; nodes starting before b are done);
//   * straddling: at a leaf, the leaf's points are loaded once (one per lane)
//     and tested against each straddling lane's query in turn, the hits
//     ranked with a ballot; at an inner node the wave descends.
// Each lane's output is in tree order, exactly as rnear_kernel's.
// The count records its events when ev_node is set (pk <= 32); the fill
// replays them when ev_hdr is set, *ev.same is nonzero and its count wave
// kept its list, and walks the tree otherwise.
// EXCL: self queries (tree points): the query's own position is skipped.
// Lane values needed by the whole wave (query, own position, write pointer
// and end) live in LDS and are fetched with one broadcast read; the hit
// tests are branch-free.
template <typename D, int DIM, bool FILL, bool EXCL>
__global__ __launch_bounds__(kBlock) void rnear_packet_kernel(TreeView<D, DIM> tv, int mode,
                                                              const int32_t* __restrict__ qlist,
                                                              const D* __restrict__ qpts,
                                                              const int32_t* __restrict__ order,
                                                              int64_t m, int64_t nrows, D r2,
                                                              int pk, int64_t nw,
                                                              const uint32_t* __restrict__ wave_order,
                                                              int32_t* __restrict__ counts,
                                                              const int64_t* __restrict__ offsets,
                                                              int64_t total_arg,
                                                              const int64_t* __restrict__ total_dev,
                                                              int32_t* __restrict__ out,
                                                              RnearEventView ev,
                                                              int32_t* __restrict__ ev_node,
                                                              uint64_t* __restrict__ ev_mask,
                                                              int32_t* __restrict__ ev_link,
                                                              int32_t* __restrict__ ev_hdr,
                                                              uint32_t* __restrict__ ev_head,
                                                              int xrun) {
  // The event arrays come in as restrict parameters (the same pointers as
  // ev's): the count's stores to them then cannot alias the tree, whose
  // node, box and point reads stay scalar loads.
  // The fill's output length: total_arg, or read here when given on the
  // device (a negative total writes nothing).
  int64_t total = total_arg;
  if constexpr (FILL) {
    if (total_dev) total = *total_dev > 0 ? *total_dev : 0;
  }
  const int lane = threadIdx.x & 63;
  int64_t w = xcd_block(blockIdx.x, gridDim.x, xrun) * (kBlock / 64) + (threadIdx.x >> 6);
  if (wave_order && w < nw) w = wave_order[w];
  // Lanes 0 .. pk-1 serve the wave's queries (pk <= 32). Lane 32 + l holds
  // lane l's query too: in the walk the low half tests each node's near
  // bound and the high half its far bound, in one pass.
  const int ql = lane & 31;
  const int64_t t = w * pk + ql;
  D q[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) q[d] = D(0);
  int32_t excl = -1;
  int64_t row = 0;
  bool ok = false, has_row = false;
  if (ql < pk && t < m)
    has_row = resolve_query<D, DIM>(tv, mode, qlist, qpts, order, t, nrows, q, excl, row, ok);
  if (lane >= 32) has_row = false;
  const bool active = has_row && ok;
  const D sgn = lane < 32 ? D(1) : D(-1);
  // Lane values read by the whole wave live in LDS: the query (and its own
  // position) and, for the fill, the next output slot and the list end; one
  // broadcast read fetches a lane's record. The write pointers are kept only
  // there.
  __shared__ PointRec<D, DIM> qsh[kBlock];
  __shared__ int64_t wsh[FILL ? 2 * kBlock : 2];
  PointRec<D, DIM>* wq = qsh + (threadIdx.x & ~63);
  int64_t* ww = wsh + (FILL ? 2 * (threadIdx.x & ~63) : 0);
  {
    PointRec<D, DIM> qr;
#pragma unroll
    for (int d = 0; d < DIM; ++d) qr.c[d] = q[d];
    qr.orig = excl;
    wq[lane] = qr;
    if constexpr (FILL) {
      // a lane writes only inside [offsets[row], offsets[row + 1]) clamped
      // to the output's [0, total)
      int64_t a = 0, e = 0;
      if (active) {
        a = offsets[row];
        e = offsets[row + 1];
        a = a < 0 ? 0 : (a > total ? total : a);
        e = e < a ? a : (e > total ? total : e);
      }
      ww[2 * lane] = a;
      ww[2 * lane + 1] = e;
    }
  }
  int32_t cnt = 0;  // count: outputs so far for this lane's query
  // wave masks are ballots of bare comparisons combined as scalars (a ballot
  // of a combined per-lane predicate costs two more vector instructions)
  const uint64_t actm = __ballot(active);

  // Output size of a node for this lane (its own position excluded).
  auto range_size = [&](const NodeRec& nr) {
    int32_t size = nr.b - nr.a;
    if constexpr (EXCL) size -= ((excl >= nr.a) & (excl < nr.b)) ? 1 : 0;
    return size;
  };
  // Lane L's next output slot and list end as scalars.
  auto write_window = [&](int L, int64_t& wL, int64_t& eL) {
    const int64_t wv = ww[2 * L];
    const int64_t ev_end = ww[2 * L + 1];
    wL = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(wv >> 32)) << 32) |
                   (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)wv));
    eL = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(ev_end >> 32)) << 32) |
                   (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)ev_end));
  };
  // Fill: copies the range of a node into the list of every lane of insm.
  // The range's indices are held in registers kCopyGroup 64-point chunks at
  // a time; per lane of insm and chunk, the slots are a scalar base plus
  // the lane index, so a chunk is one compare and one store.
  auto copy_range = [&](const NodeRec& nr, uint64_t insm) {
    constexpr int kCopyGroup = 4;
    for (int32_t g0 = nr.a; g0 < nr.b; g0 += 64 * kCopyGroup) {
      int32_t og[kCopyGroup];
#pragma unroll
      for (int c = 0; c < kCopyGroup; ++c) {
        const int32_t p = g0 + 64 * c + lane;
        og[c] = p < nr.b ? tv.orig[p] : 0;
      }
      uint64_t lm = insm;
      while (lm) {
        const int L = __builtin_ctzll(lm);
        lm &= lm - 1ull;
        int64_t wL, eL;
        write_window(L, wL, eL);
        // free slots of lane L's list (at most 2^31 - 1)
        const int64_t fr = eL - wL;
        const int32_t roomL = fr <= 0 ? 0 : (fr >= INT32_MAX ? INT32_MAX : (int32_t)fr);
        int32_t exL = -1;
        bool exin = false;
        if constexpr (EXCL) {
          exL = __builtin_amdgcn_readfirstlane(wq[L].orig);
          exin = (exL >= nr.a) & (exL < nr.b);
        }
#pragma unroll
        for (int c = 0; c < kCopyGroup; ++c) {
          const int32_t s0 = g0 + 64 * c;
          if (s0 >= nr.b) break;
          // list slot of point p = s0 + lane: wL + off + lane, where off is
          // s0 - nr.a, one less past the own position, which is skipped
          int32_t off = s0 - nr.a;
          int32_t d = 64;  // the own position's lane in this chunk
          if constexpr (EXCL) {
            if (exin) {
              d = exL - s0;
              if (d < 0) off -= 1;
            }
          }
          // a lane writes when its point is inside the range (lane < rngc)
          // and its slot inside the list (slot - wL - off < roomc)
          const int32_t rm = roomL - off;
          const int32_t rngc = nr.b - s0 < 64 ? nr.b - s0 : 64;
          const int32_t roomc = rm <= 0 ? 0 : (rm < 64 ? rm : 64);
          int32_t* dst = out + (wL + off);
          if constexpr (EXCL) {
            if ((uint32_t)d < 64u) {
              // lanes past the own position move down one slot
              const int32_t rel = lane - (lane > d ? 1 : 0);
              if ((lane < rngc) & (lane != d) & (rel < roomc)) dst[(uint32_t)rel] = og[c];
              continue;
            }
          }
          const int32_t lim = rngc < roomc ? rngc : roomc;
          if (lane < lim) dst[(uint32_t)lane] = og[c];
        }
      }
    }
  };
  // Tests a leaf's points against each lane of strm: counts the hits, or
  // writes them (fill). The points are held in registers, kGroup 64-point
  // chunks at a time (a whole leaf at bucket sizes up to 64 * kGroup), so a
  // straddling lane's query, write pointer and list end are read once per
  // group; the pointer and end are then scalars, and a hit is one store
  // through a scalar base at the hit's rank.
  auto scan_leaf = [&](const NodeRec& nr, uint64_t strm) {
    constexpr int kGroup = sizeof(PointRec<D, DIM>) <= 16 ? 4 : 2;
    for (int32_t g0 = nr.a; g0 < nr.b; g0 += 64 * kGroup) {
      PointRec<D, DIM> rec[kGroup];
      uint64_t vm[kGroup];
#pragma unroll
      for (int c = 0; c < kGroup; ++c) {
        const int32_t p = g0 + 64 * c + lane;
        const bool valid = p < nr.b;
        rec[c] = PointRec<D, DIM>{};
        if (valid) rec[c] = tv.pts[p];
        vm[c] = __ballot(valid);
      }
      uint64_t sm = strm;
      while (sm) {
        const int L = __builtin_ctzll(sm);
        sm &= sm - 1ull;
        const PointRec<D, DIM> qL = wq[L];
        int32_t exL = -1;
        if constexpr (EXCL) exL = __builtin_amdgcn_readfirstlane(qL.orig);
        int32_t* dst = out;
        uint32_t room = 0;  // free slots left in lane L's list, at most 2^31
        int64_t wL = 0;
        if constexpr (FILL) {
          int64_t eL;
          write_window(L, wL, eL);
          const int64_t free_slots = eL - wL;
          room = free_slots <= 0 ? 0u : (free_slots >= (1ll << 31) ? (1u << 31) : (uint32_t)free_slots);
          dst = out + wL;
        }
        int32_t hl = 0;
#pragma unroll
        for (int c = 0; c < kGroup; ++c) {
          if (vm[c] == 0ull) break;
          // hit masks are ballots of the bare distance test, combined with
          // the chunk's valid lanes (and the own position) as scalar masks
          const bool in = dist2<D, DIM>(qL.c, rec[c].c) < r2;
          uint64_t hits = __ballot(in) & vm[c];
          if constexpr (EXCL) {
            const uint32_t dl = (uint32_t)(exL - (g0 + 64 * c));
            hits &= dl < 64u ? ~(1ull << dl) : ~0ull;
          }
          const uint32_t nh = (uint32_t)__popcll(hits);
          if constexpr (FILL) {
            const uint32_t rk = (uint32_t)rank_below(hits);
            if (__builtin_amdgcn_inverse_ballot_w64(hits) & (rk < room)) dst[rk] = rec[c].orig;
            const uint32_t adv = nh < room ? nh : room;
            dst += adv;
            room -= adv;
          }
          hl += (int32_t)nh;
        }
        if constexpr (FILL) {
          if (lane == 0) ww[2 * L] = wL + hl;
        } else {
          cnt += lane == L ? hl : 0;
        }
      }
    }
  };

  bool walk = true;
  if constexpr (FILL) {
    if (ev_hdr && actm) {
      // this wave's lanes are lanes h0 .. h0 + pk - 1 of count wave cw
      const int64_t t0 = w * pk;
      const int64_t cw = t0 / ev.pk;
      const int h0 = (int)(t0 - cw * ev.pk);
      const int32_t first = ev_hdr[2 * cw];
      const int32_t nev = ev_hdr[2 * cw + 1];
      if (*ev.same != 0 && nev >= 0) {
        walk = false;
        const uint32_t lm = pk >= 32 ? 0xffffffffu : ((1u << pk) - 1u);
        int32_t chunk = first;
        for (int32_t e0 = 0; e0 < nev; e0 += kEventChunk) {
          const bool has = lane < nev - e0;
          const int64_t slot = (int64_t)chunk * kEventChunk + lane;
          const int32_t en = has ? ev_node[slot] : 0;
          const uint64_t em = has ? ev_mask[slot] : 0ull;
          const uint32_t mi = ((uint32_t)em >> h0) & lm;
          const uint32_t ms = ((uint32_t)(em >> 32) >> h0) & lm;
          uint64_t todo = __ballot((mi | ms) != 0u);
          const int32_t next = e0 + kEventChunk < nev ? ev_link[chunk] : -1;
          while (todo) {
            const int j = __builtin_ctzll(todo);
            todo &= todo - 1ull;
            const int32_t node = __builtin_amdgcn_readlane(en, j);
            const uint64_t insm = (uint32_t)__builtin_amdgcn_readlane((int)mi, j);
            const uint64_t strm = (uint32_t)__builtin_amdgcn_readlane((int)ms, j);
            const NodeRec nr = tv.node[node];
            if (insm) {
              copy_range(nr, insm);
              if ((insm >> lane) & 1ull) ww[2 * lane] += range_size(nr);
            }
            if (strm) scan_leaf(nr, strm);
          }
          chunk = next;
        }
      }
    }
  }
  if (walk) {
    // the count's event recording (uniform state; lane j of en/em holds
    // event j of the batch not yet written)
    bool rec = false;
    if constexpr (!FILL) rec = ev_node != nullptr && w < nw;
    int32_t en = 0;
    uint64_t em = 0;
    int32_t e_pos = 0, e_first = -1, e_chunk = -1, e_count = 0;
    auto flush = [&]() {
      int32_t c = 0;
      if (lane == 0) c = (int32_t)atomicAdd(ev_head, 1u);
      c = __builtin_amdgcn_readlane(c, 0);
      if (c < 0 || c >= ev.nchunks) {
        rec = false;  // pool exhausted: the fill walks for this wave
        return;
      }
      const int64_t slot = (int64_t)c * kEventChunk + lane;
      if (lane < e_pos) {
        ev_node[slot] = en;
        ev_mask[slot] = em;
      }
      if (e_chunk >= 0) {
        if (lane == 0) ev_link[e_chunk] = c;
      } else {
        e_first = c;
      }
      e_chunk = c;
      e_count += e_pos;
      e_pos = 0;
    };
    int32_t done_end = INT32_MIN;  // nodes starting before this are taken
    int32_t node = actm ? 0 : -1;
    D sq[DIM];  // sgn * q
#pragma unroll
    for (int d = 0; d < DIM; ++d) sq[d] = sgn * q[d];
    while (node >= 0) {
      node = __builtin_amdgcn_readfirstlane(node);
      const NodeRec nr = tv.node[node];
      const BoxRec<D, DIM> bx = tv.box[node];
      const bool c_new = nr.a >= done_end;
      // low half: the ball reaches the box; high half: the ball holds it
      const uint64_t bm = __ballot(box_near_far_dist2<D, DIM>(bx, sq, sgn) < r2);
      const uint64_t hitm = __ballot(c_new) & bm & actm;
      bool descend = false;
      if (hitm) {
        const uint64_t cmaxm = bm >> 32;
        const uint64_t insm = hitm & cmaxm;
        const uint64_t strm = hitm & ~cmaxm;
        const bool inside = __builtin_amdgcn_inverse_ballot_w64(insm);
        if (inside) done_end = nr.b;
        if constexpr (FILL) {
          if (insm) copy_range(nr, insm);
          if (inside) ww[2 * lane] += range_size(nr);
        } else {
          if (inside) cnt += range_size(nr);
          if (rec) {
            const uint64_t emask = insm | (nr.child < 0 ? strm << 32 : 0ull);
            if (emask) {
              en = lane == e_pos ? node : en;
              em = lane == e_pos ? emask : em;
              if (++e_pos == kEventChunk) flush();
            }
          }
        }
        if (strm) {
          if (nr.child < 0) scan_leaf(nr, strm);
          else descend = true;
        }
      }
      node = descend ? nr.child : next_in_order<D, DIM>(tv, node);
    }
    if constexpr (!FILL) {
      if (ev_node != nullptr && w < nw) {
        if (rec && e_pos > 0) flush();
        if (lane == 0) {
          ev_hdr[2 * w] = e_first;
          ev_hdr[2 * w + 1] = rec ? e_count : -1;
        }
      }
    }
  }
  if constexpr (!FILL) {
    if (has_row) counts[row] = ok ? cnt : 0;
  }
}

// Launch order of the fill's waves, heaviest first: waves bucketed by
// floor(log2(output + 1)) of their output count from the count pass, buckets
// in decreasing order, waves in index order inside a bucket. Three launches,
// one thread per wave in the first and last:
//   rnear_wave_bucket_count_kernel  each wave's bucket, each block's count
//                                   per bucket;
//   scan_u32_inplace_async          one block: every (bucket, block) run's
//                                   start, buckets heaviest first;
//   rnear_wave_order_kernel         each wave's slot: its run's start plus
//                                   its rank among the block's waves of its
//                                   bucket (ballots).
// (One 1,024-thread block for all of it took 61 us on 62.5K waves, plus a
// 17 us cost-key kernel; a 32-bit radix sort of the keys before that took
// 4 passes, 20 launches.)
constexpr int kBucketBlock = 256;
constexpr int kBucketWaves = kBucketBlock / 64;
constexpr int kBuckets = 33;

__device__ __forceinline__ int wave_bucket(int64_t c) {  // 0 .. 32
  const uint64_t v = (uint64_t)(c < 0xffffffffll ? c : 0xffffffffll);
  return 63 - __clzll(v + 1ull);
}

// Calls f(k, m) for each bucket k present among the wave's lanes (b = -1: no
// fill wave), m the lanes holding k; one ballot per bucket present.
template <typename F>
__device__ __forceinline__ void each_bucket(int b, F f) {
  uint64_t todo = __ballot(b >= 0);
  while (todo) {
    const int k = __builtin_amdgcn_readlane(b, __builtin_ctzll(todo));
    const uint64_t m = __ballot(b == k);
    todo &= ~m;
    f(k, m);
  }
}

// Fill wave w serves rows order[w*pk .. w*pk+pk) (pk <= kRnearMaxFillWidth),
// or those positions; for self queries of every point (row_of_pos set)
// row_of_pos[t]. Writes each wave's bucket and the block's count of bucket
// k at cnt[(kBuckets - 1 - k) * nblk + block] (runs in slot order).
constexpr int kRnearMaxFillWidth = 16;
__global__ __launch_bounds__(kBucketBlock) void rnear_wave_bucket_count_kernel(
    const int32_t* __restrict__ order, const int32_t* __restrict__ row_of_pos,
    const int64_t* __restrict__ offsets, int64_t m, int64_t nrows, int pk, int64_t nw,
    uint8_t* __restrict__ bucket_of, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t bc[kBuckets];
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < kBuckets) bc[threadIdx.x] = 0;
  __syncthreads();
  const int64_t w = (int64_t)blockIdx.x * kBucketBlock + threadIdx.x;
  int b = -1;
  if (w < nw) {
    // the wave's rows first, then all their offsets: every read in flight
    int64_t row[kRnearMaxFillWidth];
#pragma unroll
    for (int l = 0; l < kRnearMaxFillWidth; ++l) {
      const int64_t t = w * pk + l;
      row[l] = -1;
      if (l < pk && t < m) row[l] = row_of_pos ? row_of_pos[t] : (order ? order[t] : t);
    }
    int64_t c = 0;
#pragma unroll
    for (int l = 0; l < kRnearMaxFillWidth; ++l) {
      if (row[l] >= 0 && row[l] < nrows) {
        const int64_t n = offsets[row[l] + 1] - offsets[row[l]];
        c += n > 0 ? n : 0;
      }
    }
    b = wave_bucket(c);
    bucket_of[w] = (uint8_t)b;
  }
  each_bucket(b, [&](int k, uint64_t msk) {
    if (lane == 0) atomicAdd(&bc[k], (uint32_t)__popcll(msk));
  });
  __syncthreads();
  if (threadIdx.x < kBuckets)
    cnt[(int64_t)(kBuckets - 1 - threadIdx.x) * gridDim.x + blockIdx.x] = bc[threadIdx.x];
}

__global__ __launch_bounds__(kBucketBlock) void rnear_wave_order_kernel(
    const uint8_t* __restrict__ bucket_of, const uint32_t* __restrict__ start, int64_t nw,
    uint32_t* __restrict__ worder) {
  __shared__ uint32_t wc[kBucketWaves][kBuckets];  // each wave's count per bucket
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t w = (int64_t)blockIdx.x * kBucketBlock + threadIdx.x;
  const int b = w < nw ? (int)bucket_of[w] : -1;
  uint32_t mine = 0;  // lane k: this wave's count of bucket k
  each_bucket(b, [&](int k, uint64_t msk) { mine += lane == k ? (uint32_t)__popcll(msk) : 0u; });
  if (lane < kBuckets) wc[wv][lane] = mine;
  __syncthreads();
  // lane k: the slot of this wave's first fill wave of bucket k
  uint32_t base = 0;
  if (lane < kBuckets) {
    base = start[(int64_t)(kBuckets - 1 - lane) * gridDim.x + blockIdx.x];
    for (int v = 0; v < wv; ++v) base += wc[v][lane];
  }
  each_bucket(b, [&](int k, uint64_t msk) {
    const uint32_t bk = (uint32_t)__builtin_amdgcn_readlane((int)base, k);
    if (b == k) worder[bk + rank_below(msk)] = (uint32_t)w;
  });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fill_kernel(T* __restrict__ p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, kBlock)); }

// Queries per wave of the r-near packet kernels. On the 10M-point lidar
// scene with 1M queries at dmax 0.1 m (median 45, p90 5364 hits per query):
// count 64/32/16/8 -> 2.16/1.85/2.00/2.53 ms; fill 64/32/16/8 -> 4.19/3.44/3.37/4.03 ms.
constexpr int kRnearCountWidth = 32;
constexpr int kRnearFillWidth = 16;
static_assert(kRnearFillWidth <= kRnearMaxFillWidth, "fill waves wider than the wave bucketing reads");
static_assert(kRnearCountWidth <= 32 && kRnearFillWidth <= 32,
              "r-near packets fill the low half of a wave (the high half mirrors it)");

}  // namespace

template <typename T>
void fill_async(T* p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return;
  fill_kernel<T><<<grid_of(n), kBlock, 0, s>>>(p, n, v);
  PK_LAUNCH_CHECK();
}
template void fill_async<int32_t>(int32_t*, int64_t, int32_t, hipStream_t);
template void fill_async<uint32_t>(uint32_t*, int64_t, uint32_t, hipStream_t);
template void fill_async<float>(float*, int64_t, float, hipStream_t);
template void fill_async<double>(double*, int64_t, double, hipStream_t);

namespace {

template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void query_morton_kernel(const D* __restrict__ qpts, int64_t m,
                                                              BuildInfo<D> frame, int shift,
                                                              uint64_t* __restrict__ keys,
                                                              uint32_t* __restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  D c[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) c[d] = qpts[t * DIM + d];
  keys[t] = morton_code<D, DIM>(c, frame) >> shift;
  vals[t] = (uint32_t)t;
}

__global__ __launch_bounds__(kBlock) void query_position_kernel(const int32_t* __restrict__ list,
                                                                const int32_t* __restrict__ fwd,
                                                                int64_t m, int64_t n_input,
                                                                uint64_t* __restrict__ keys,
                                                                uint32_t* __restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  const int32_t o = list[t];
  const int32_t p = (o >= 0 && (int64_t)o < n_input) ? fwd[o] : -1;
  keys[t] = p < 0 ? 0xffffffffull : (uint64_t)(uint32_t)p;
  vals[t] = (uint32_t)t;
}

// Orders point queries along the tree's Morton curve and index queries by
// tree position. Below this many queries the sort costs more than it saves.
constexpr int64_t kOrderMin = 4096;

// Point queries are ordered on the top kQueryOrderBits bits of their Morton
// code (13 to 14 bits per axis in 3-D): finer cells than any wave's extent,
// in 5 radix passes instead of the full code's 8. On the 10M-point lidar
// scene with 1M queries: count 1.41 / 1.39 / 1.38 / 1.58 ms at 48 / 40 / 36
// / 32 bits, the fill 2.06 ms at 48 to 40 bits and 2.08 at 36.
constexpr int kQueryOrderBits = 40;

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Scratch bytes that ordering m queries takes: the order itself, the sort
// keys and the sort's temporaries.
size_t order_scratch_bytes(int64_t m) {
  return align256((size_t)m * 4) + align256((size_t)m * 8) + SortWorkspace::bytes_for(m);
}

// Largest query count whose order fits in `bytes` of scratch.
int64_t max_ordered_queries(size_t bytes) {
  int64_t lo = 0, hi = (int64_t)(bytes / 12) + 1;
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (order_scratch_bytes(mid) <= bytes) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename D>
void order_queries(const pptk_tree* t, PreparedQueries& pq, hipStream_t s,
                   const QueryScratch* scratch) {
  const int64_t m = pq.m;
  if (m < kOrderMin || pq.mode == kQueryAll) return;
  OrderCache& oc = t->order_cache;
  if (!scratch) {
    std::lock_guard<std::mutex> lk(oc.mu);
    if (oc.order && oc.src == pq.src && oc.m == m && oc.mode == pq.mode) {
      pq.order_buf = oc.order;
      pq.order = pq.order_buf->as<const int32_t>();
      return;
    }
  }
  DevBuf keys_owned;
  std::shared_ptr<DevBuf> buf;
  uint32_t* ord = nullptr;
  uint64_t* keys = nullptr;
  SortWorkspace ws;
  if (scratch) {
    if (order_scratch_bytes(m) > scratch->bytes)
      fail(PPTK_E_VALUE,
           "%lld point or index queries need %zu bytes of scratch to be ordered; the "
           "allocation-free entry point has %zu (at most %lld such queries per call)",
           (long long)m, order_scratch_bytes(m), scratch->bytes,
           (long long)max_ordered_queries(scratch->bytes));
    uint8_t* p = (uint8_t*)scratch->base;
    ord = (uint32_t*)p;
    p += align256((size_t)m * 4);
    keys = (uint64_t*)p;
    p += align256((size_t)m * 8);
    ws.bind(p, m);
  } else {
    keys_owned.reset((size_t)m * 8);
    keys = keys_owned.as<uint64_t>();
    buf = std::make_shared<DevBuf>((size_t)m * 4);
    ord = buf->as<uint32_t>();
    ws.reserve(m);
  }
  int bits = 32;
  if (pq.mode == kQueryList) {
    query_position_kernel<<<grid_of(m), kBlock, 0, s>>>(
        pq.dlist, t->fwd.as<const int32_t>(), m, t->n_input, keys, ord);
    PK_LAUNCH_CHECK();
  } else {
    BuildInfo<D> frame;
    frame.n_valid = 0;
    frame.pad = 0;
    for (int d = 0; d < 4; ++d) {
      frame.lo[d] = (D)t->mlo[d];
      frame.hi[d] = (D)t->hi[d];
      frame.scale[d] = (D)t->mscale[d];
    }
    auto launch = [&](auto dimc) {
      constexpr int DIM = decltype(dimc)::value;
      constexpr int full = DIM * MortonBits<DIM>::value;
      constexpr int shift = full > kQueryOrderBits ? full - kQueryOrderBits : 0;
      query_morton_kernel<D, DIM><<<grid_of(m), kBlock, 0, s>>>(
          (const D*)pq.dpts, m, frame, shift, keys, ord);
      PK_LAUNCH_CHECK();
      bits = full - shift;
    };
    if (t->dim == 2) launch(std::integral_constant<int, 2>());
    else if (t->dim == 3) launch(std::integral_constant<int, 3>());
    else launch(std::integral_constant<int, 4>());
  }
  radix_sort_pairs_async(keys, ord, m, bits, ws, s, /*keys_out=*/false);
  pq.order = (const int32_t*)ord;
  if (scratch) return;  // stream order keeps the scratch valid for the queries
  // The cache entry is published only once the order is complete.
  PK_HIP(hipStreamSynchronize(s));
  pq.order_buf = buf;
  std::lock_guard<std::mutex> lk(oc.mu);
  oc.src = pq.src;
  oc.m = m;
  oc.mode = pq.mode;
  oc.order = std::move(buf);
}

}  // namespace

template <typename D>
void prepare_queries(const pptk_tree* t, const pptk_queries* q, hipStream_t s, PreparedQueries& pq,
                     const QueryScratch* scratch) {
  if (!q) fail(PPTK_E_VALUE, "null query set");
  if (scratch && q->kind != PPTK_Q_ALL && q->mem != PPTK_MEM_DEVICE)
    fail(PPTK_E_VALUE, "the allocation-free entry point takes queries in device memory");
  switch (q->kind) {
    case PPTK_Q_ALL:
      pq.mode = kQueryAll;
      pq.m = t->n_valid;
      pq.nrows = t->n_valid;
      break;
    case PPTK_Q_INDICES: {
      pq.mode = kQueryList;
      if (q->m < 0) fail(PPTK_E_VALUE, "negative query count");
      pq.m = pq.nrows = q->m;
      if (q->m > 0 && !q->indices) fail(PPTK_E_VALUE, "null query indices");
      if (q->mem == PPTK_MEM_HOST) {
        for (int64_t i = 0; i < q->m; ++i)
          if (q->indices[i] < 0 || q->indices[i] >= t->n_input)
            fail(PPTK_E_RUNTIME, "%lld-th query index %d is outside of [0,%lld]", (long long)i,
                 (int)q->indices[i], (long long)t->n_input - 1);
      }
      stage_input(pq.list, q->indices, (size_t)q->m * 4, q->mem, s);
      pq.dlist = (const int32_t*)pq.list.ptr;
      pq.src = q->indices;
      break;
    }
    case PPTK_Q_POINTS: {
      if (t->nroots > 1)
        fail(PPTK_E_VALUE, "point queries need a single-cloud tree (this tree has %lld clouds)",
             (long long)t->nroots);
      pq.mode = kQueryPoints;
      if (q->m < 0) fail(PPTK_E_VALUE, "negative query count");
      pq.m = pq.nrows = q->m;
      pq.src = q->points;
      if (q->m > 0 && !q->points) fail(PPTK_E_VALUE, "null query points");
      if (!dtype_valid(q->dtype)) fail(PPTK_E_TYPE, "unsupported query dtype %d", (int)q->dtype);
      const size_t cnt = (size_t)q->m * t->dim;
      if (q->mem == PPTK_MEM_DEVICE) {
        if (q->dtype != t->dist)
          fail(PPTK_E_TYPE, "device query points must have the tree's distance dtype");
        pq.dpts = q->points;
      } else if (cnt > 0) {
        std::vector<D> conv(cnt);
        D* dst = conv.data();
        dispatch_dtype(q->dtype, [&](auto tag) {
          using T = typename decltype(tag)::type;
          const T* src = (const T*)q->points;
          for (size_t i = 0; i < cnt; ++i) dst[i] = (D)src[i];
        });
        pq.conv.reset(cnt * sizeof(D));
        PK_HIP(hipMemcpyAsync(pq.conv.get(), conv.data(), cnt * sizeof(D), hipMemcpyHostToDevice, s));
        PK_HIP(hipStreamSynchronize(s));  // `conv` goes out of scope
        pq.dpts = pq.conv.get();
      }
      break;
    }
    default:
      fail(PPTK_E_VALUE, "unknown query kind %d", (int)q->kind);
  }
  order_queries<D>(t, pq, s, scratch);
}
template void prepare_queries<float>(const pptk_tree*, const pptk_queries*, hipStream_t,
                                     PreparedQueries&, const QueryScratch*);
template void prepare_queries<double>(const pptk_tree*, const pptk_queries*, hipStream_t,
                                      PreparedQueries&, const QueryScratch*);

namespace {

template <typename D, int DIM>
void knn_run(const pptk_tree* t, const pptk_queries* qs, int k, double dmax, int32_t* out_idx,
             int32_t* out_cnt, void* out_dist, pptk_mem out_mem, hipStream_t s,
             const QueryScratch* scratch = nullptr) {
  PreparedQueries pq;
  prepare_queries<D>(t, qs, s, pq, scratch);
  const D r = (D)dmax;
  const D r2 = r * r;
  const int64_t nrows = pq.nrows;
  DevOutput oi, oc, od;
  oi.init(out_idx, (size_t)nrows * k * 4, out_mem);
  oc.init(out_cnt, (size_t)nrows * 4, out_mem);
  od.init(out_dist, (size_t)nrows * k * sizeof(D), out_mem);
  if (nrows == 0) return;
  const TreeView<D, DIM> tv = tree_view<D, DIM>(t);
  if (pq.mode == kQueryAll && t->n_valid < t->n_input) {
    // Rows of NaN/inf points receive no thread; pre-fill them empty.
    fill_async<int32_t>((int32_t*)oi.dptr, nrows * k, -1, s);
    fill_async<int32_t>((int32_t*)oc.dptr, nrows, 0, s);
    if (od.dptr) fill_async<D>((D*)od.dptr, nrows * k, pos_inf<D>(), s);
  }
  const int64_t m = pq.m;
  const D* qp = (const D*)pq.dpts;
  if (m > 0) {
    if (k <= 32) {
      auto launch = [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if constexpr (std::is_same<D, float>::value) {
          if (pq.mode == kQueryAll) {
            knn_coop_kernel<DIM, K><<<grid_of(m), kBlock, 0, s>>>(
                tv, m, nrows, k, r2, (int32_t*)oi.dptr, (int32_t*)oc.dptr, (float*)od.dptr,
                kXcdRun);
            PK_LAUNCH_CHECK();
            return;
          }
        }
        knn_kernel<D, DIM, K><<<grid_of(m), kBlock, 0, s>>>(
            tv, pq.mode, pq.dlist, qp, pq.order, m, nrows, k, r2, (int32_t*)oi.dptr, (int32_t*)oc.dptr,
            (D*)od.dptr);
        PK_LAUNCH_CHECK();
      };
      if (k <= 4) launch(std::integral_constant<int, 4>());
      else if (k <= 8) launch(std::integral_constant<int, 8>());
      else if (k <= 16) launch(std::integral_constant<int, 16>());
      else launch(std::integral_constant<int, 32>());
    } else {
      // Global-memory heaps; bound the distance workspace to ~256 MB, or to
      // the scratch left after the query order.
      int64_t chunk = std::max<int64_t>(kBlock, ((int64_t)256 << 20) / ((int64_t)k * sizeof(D)));
      DevBuf ws;
      D* wsp = nullptr;
      if (scratch) {
        const size_t used = pq.order ? align256((size_t)m * 4) : 0;
        const int64_t fit = (int64_t)((scratch->bytes - used) / ((size_t)k * sizeof(D)));
        chunk = fit / kBlock * kBlock;
        if (chunk < kBlock)
          fail(PPTK_E_VALUE, "k = %d needs more scratch than the allocation-free entry point has",
               k);
        wsp = (D*)((uint8_t*)scratch->base + used);
      } else {
        ws.reset((size_t)std::min(chunk, m) * k * sizeof(D));
        wsp = ws.as<D>();
      }
      for (int64_t t0 = 0; t0 < m; t0 += chunk) {
        const int64_t t1 = std::min(m, t0 + chunk);
        knn_large_kernel<D, DIM><<<grid_of(t1 - t0), kBlock, 0, s>>>(
            tv, pq.mode, pq.dlist, qp, pq.order, t0, t1, nrows, k, r2, (int32_t*)oi.dptr,
            (int32_t*)oc.dptr, wsp, (D*)od.dptr);
        PK_LAUNCH_CHECK();
      }
      PK_HIP(hipStreamSynchronize(s));
    }
  }
  oi.finish(s);
  oc.finish(s);
  od.finish(s);
  PK_HIP(hipStreamSynchronize(s));
}

// Chunks of 64 events per count wave in a fresh pool, and the pool's cap
// (PPTK_RNEAR_EVENT_CHUNKS overrides the cap; 0 keeps no lists).
constexpr double kEventChunksPerWave = 4.0;
constexpr int64_t kEventPoolBytes = (int64_t)1 << 30;
constexpr size_t kEventChunkBytes = (size_t)kEventChunk * 12 + 4;

int64_t event_pool_cap() {
  const char* e = getenv("PPTK_RNEAR_EVENT_CHUNKS");
  if (e && *e) {
    char* end = nullptr;
    const long long v = strtoll(e, &end, 10);
    if (end && *end == '\0' && v >= 0) return (int64_t)v;
  }
  return kEventPoolBytes / (int64_t)kEventChunkBytes;
}

__global__ __launch_bounds__(kBlock) void words_equal_kernel(const uint32_t* __restrict__ a,
                                                             const uint32_t* __restrict__ b,
                                                             int64_t n, int32_t* __restrict__ same) {
  bool diff = false;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    diff |= a[i] != b[i];
  if (diff) *same = 0;
}

template <typename D, int DIM>
void rnear_count_run(const pptk_tree* t, const pptk_queries* qs, double dmax, int64_t* out_offsets,
                     int64_t* out_total, pptk_mem out_mem, hipStream_t s) {
  PreparedQueries pq;
  prepare_queries<D>(t, qs, s, pq);
  const D r = (D)dmax;
  const D r2 = r * r;
  const int64_t nrows = pq.nrows;
  DevOutput oo;
  oo.init(out_offsets, (size_t)(nrows + 1) * 8, out_mem);
  DevBuf counts((size_t)std::max<int64_t>(nrows, 1) * 4);
  if (nrows > 0) PK_HIP(hipMemsetAsync(counts.get(), 0, (size_t)nrows * 4, s));
  const TreeView<D, DIM> tv = tree_view<D, DIM>(t);
  std::shared_ptr<RnearEvents> ent;
  uint32_t used = 0;
  if (pq.m > 0) {
    const int pk = kRnearCountWidth;
    const int64_t nw = ceil_div(pq.m, pk);
    const unsigned grid = (unsigned)std::max<int64_t>(1, ceil_div(nw, kBlock / 64));
    RnearEventView ev;
    if (tv.nroots == 1) {
      // event lists for the fill of the same queries
      double per_wave = kEventChunksPerWave;
      {
        std::lock_guard<std::mutex> lk(t->rnear_cache.mu);
        per_wave = std::max(per_wave, t->rnear_cache.chunks_per_wave);
      }
      const int64_t nchunks = std::min<int64_t>(event_pool_cap(),
                                                (int64_t)(per_wave * (double)nw) + 64);
      if (nchunks > 0) {
        ent = std::make_shared<RnearEvents>();
        ent->src = pq.src;
        ent->m = pq.m;
        ent->mode = pq.mode;
        ent->r = (double)r;
        ent->pk = pk;
        ent->nw = nw;
        ent->nchunks = nchunks;
        ent->order = pq.order ? pq.order_buf : nullptr;
        if (pq.mode == kQueryPoints) ent->qbytes = (size_t)pq.m * DIM * sizeof(D);
        else if (pq.mode == kQueryList) ent->qbytes = (size_t)pq.m * 4;
        if (ent->qbytes) {
          ent->queries.reset(ent->qbytes);
          PK_HIP(hipMemcpyAsync(ent->queries.get(),
                                pq.mode == kQueryPoints ? pq.dpts : (const void*)pq.dlist,
                                ent->qbytes, hipMemcpyDeviceToDevice, s));
        }
        ent->node.reset((size_t)nchunks * kEventChunk * 4);
        ent->mask.reset((size_t)nchunks * kEventChunk * 8);
        ent->link.reset((size_t)nchunks * 4);
        ent->hdr.reset((size_t)nw * 8);
        ent->head.reset(4);
        fill_async<uint32_t>(ent->head.as<uint32_t>(), 1, 0u, s);
        ev.node = ent->node.as<int32_t>();
        ev.mask = ent->mask.as<uint64_t>();
        ev.link = ent->link.as<int32_t>();
        ev.hdr = ent->hdr.as<int32_t>();
        ev.head = ent->head.as<uint32_t>();
        ev.nchunks = (int32_t)nchunks;
        ev.pk = pk;
      }
    }
    if (tv.nroots == 1 && pq.mode == kQueryPoints)
      rnear_packet_kernel<D, DIM, false, false><<<grid, kBlock, 0, s>>>(
          tv, pq.mode, pq.dlist, (const D*)pq.dpts, pq.order, pq.m, nrows, r2, pk, nw, nullptr,
          counts.as<int32_t>(), nullptr, 0, nullptr, nullptr, ev, ev.node, ev.mask, ev.link, ev.hdr, ev.head, kXcdRun);
    else if (tv.nroots == 1)
      rnear_packet_kernel<D, DIM, false, true><<<grid, kBlock, 0, s>>>(
          tv, pq.mode, pq.dlist, (const D*)pq.dpts, pq.order, pq.m, nrows, r2, pk, nw, nullptr,
          counts.as<int32_t>(), nullptr, 0, nullptr, nullptr, ev, ev.node, ev.mask, ev.link, ev.hdr, ev.head, kXcdRun);
    else
      rnear_kernel<D, DIM, false><<<grid_of(pq.m), kBlock, 0, s>>>(
          tv, pq.mode, pq.dlist, (const D*)pq.dpts, pq.order, pq.m, nrows, r2,
          counts.as<int32_t>(), nullptr, 0, nullptr, nullptr);
    PK_LAUNCH_CHECK();
    if (ent) PK_HIP(hipMemcpyAsync(&used, ent->head.get(), 4, hipMemcpyDeviceToHost, s));
  }
  int64_t* doff = (int64_t*)oo.dptr;
  DevBuf tmp;
  if (!doff) {
    tmp.reset((size_t)(nrows + 1) * 8);
    doff = tmp.as<int64_t>();
  }
  exclusive_scan<int32_t, int64_t>(counts.as<int32_t>(), doff, nrows, s);
  int64_t total = 0;
  PK_HIP(hipMemcpyAsync(&total, doff + nrows, 8, hipMemcpyDeviceToHost, s));
  oo.finish(s);
  PK_HIP(hipStreamSynchronize(s));
  if (out_total) *out_total = total;
  if (ent) {
    ent->chunks_used = (int64_t)used;
    // the next count on this tree starts with a pool this count would have
    // filled to 4/5
    std::lock_guard<std::mutex> lk(t->rnear_cache.mu);
    t->rnear_cache.chunks_per_wave = 1.25 * (double)used / (double)ent->nw;
    t->rnear_cache.last = std::move(ent);
  }
}

template <typename D, int DIM>
void rnear_fill_run(const pptk_tree* t, const pptk_queries* qs, double dmax, const int64_t* offsets,
                    int32_t* out_idx, pptk_mem out_mem, hipStream_t s) {
  PreparedQueries pq;
  prepare_queries<D>(t, qs, s, pq);
  const D r = (D)dmax;
  const D r2 = r * r;
  const int64_t nrows = pq.nrows;
  if (!offsets) fail(PPTK_E_VALUE, "null offsets");
  // With offsets and output in device memory the kernels read the total
  // (offsets[nrows]) themselves: no copy to the host and no wait for it.
  const bool dev_total = out_mem == PPTK_MEM_DEVICE && out_idx != nullptr;
  int64_t total = 0;
  if (!dev_total) {
    if (out_mem == PPTK_MEM_HOST) total = offsets[nrows];
    else PK_HIP(hipMemcpyAsync(&total, offsets + nrows, 8, hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    if (total < 0) fail(PPTK_E_VALUE, "negative total in offsets");
  }
  const int64_t* total_dev = dev_total ? offsets + nrows : nullptr;
  const bool work = dev_total || total > 0;
  DevInput doff;
  stage_input(doff, offsets, (size_t)(nrows + 1) * 8, out_mem, s);
  DevOutput oi;
  oi.init(out_idx, (size_t)total * 4, out_mem);
  if (total > 0 && !out_idx) fail(PPTK_E_VALUE, "null output");
  const TreeView<D, DIM> tv = tree_view<D, DIM>(t);
  const int pk = kRnearFillWidth;
  // the event lists of the count of these queries, if it was the last count
  // on this tree (held until the stream has synchronised below)
  std::shared_ptr<RnearEvents> ent;
  if (tv.nroots == 1 && pq.m > 0 && work) {
    std::lock_guard<std::mutex> lk(t->rnear_cache.mu);
    const std::shared_ptr<RnearEvents>& e = t->rnear_cache.last;
    if (e && e->src == pq.src && e->m == pq.m && e->mode == pq.mode && e->r == (double)r &&
        e->pk % pk == 0)
      ent = e;
  }
  // wave order of the fill and the replay check; held until the stream has
  // synchronised below
  DevBuf wkeys, wvals, same;
  RnearEventView ev;
  if (ent) {
    // the count's processing order: a lane of the fill serves the query of
    // the same lane of its count wave
    pq.order = ent->order ? ent->order->as<const int32_t>() : nullptr;
    same.reset(4);
    fill_async<int32_t>(same.as<int32_t>(), 1, 1, s);
    if (ent->qbytes) {
      const int64_t words = (int64_t)(ent->qbytes / 4);
      const unsigned g = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, ceil_div(words, kBlock)));
      words_equal_kernel<<<g, kBlock, 0, s>>>(
          ent->queries.as<const uint32_t>(),
          (const uint32_t*)(pq.mode == kQueryPoints ? pq.dpts : (const void*)pq.dlist), words,
          same.as<int32_t>());
      PK_LAUNCH_CHECK();
    }
    ev.node = ent->node.as<int32_t>();
    ev.mask = ent->mask.as<uint64_t>();
    ev.link = ent->link.as<int32_t>();
    ev.hdr = ent->hdr.as<int32_t>();
    ev.nchunks = (int32_t)ent->nchunks;
    ev.pk = ent->pk;
    ev.same = same.as<const int32_t>();
  }
  if (pq.m > 0 && work) {
    const int64_t nw = ceil_div(pq.m, pk);
    const unsigned grid = (unsigned)std::max<int64_t>(1, ceil_div(nw, kBlock / 64));
    // waves with the most output start first, so that the last waves to
    // start are short ones
    const uint32_t* worder = nullptr;
    if (tv.nroots == 1 && nw > 1) {
      const int64_t nblk = ceil_div(nw, kBucketBlock);
      // run lengths in slot order, scanned in place into run starts (+ total)
      const size_t cbytes = (size_t)(nblk * kBuckets + 1) * 4;
      wkeys.reset(align256((size_t)nw) + cbytes);
      wvals.reset((size_t)nw * 4);
      uint8_t* bof = wkeys.as<uint8_t>();
      uint32_t* bcnt = (uint32_t*)(bof + align256((size_t)nw));
      rnear_wave_bucket_count_kernel<<<(unsigned)nblk, kBucketBlock, 0, s>>>(
          pq.order, pq.mode == kQueryAll ? tv.orig : nullptr, (const int64_t*)doff.ptr, pq.m,
          nrows, pk, nw, bof, bcnt);
      PK_LAUNCH_CHECK();
      scan_u32_inplace_async(bcnt, nblk * kBuckets, s);
      rnear_wave_order_kernel<<<(unsigned)nblk, kBucketBlock, 0, s>>>(bof, bcnt, nw,
                                                                      wvals.as<uint32_t>());
      PK_LAUNCH_CHECK();
      worder = wvals.as<uint32_t>();
    }
    if (tv.nroots == 1 && pq.mode == kQueryPoints)
      rnear_packet_kernel<D, DIM, true, false><<<grid, kBlock, 0, s>>>(
          tv, pq.mode, pq.dlist, (const D*)pq.dpts, pq.order, pq.m, nrows, r2, pk, nw, worder,
          nullptr, (const int64_t*)doff.ptr, total, total_dev, (int32_t*)oi.dptr, ev, ev.node,
          ev.mask, ev.link, ev.hdr, ev.head, 1);
    else if (tv.nroots == 1)
      rnear_packet_kernel<D, DIM, true, true><<<grid, kBlock, 0, s>>>(
          tv, pq.mode, pq.dlist, (const D*)pq.dpts, pq.order, pq.m, nrows, r2, pk, nw, worder,
          nullptr, (const int64_t*)doff.ptr, total, total_dev, (int32_t*)oi.dptr, ev, ev.node,
          ev.mask, ev.link, ev.hdr, ev.head, 1);
    else
      rnear_kernel<D, DIM, true><<<grid_of(pq.m), kBlock, 0, s>>>(
          tv, pq.mode, pq.dlist, (const D*)pq.dpts, pq.order, pq.m, nrows, r2, nullptr,
          (const int64_t*)doff.ptr, total, total_dev, (int32_t*)oi.dptr);
    PK_LAUNCH_CHECK();
  }
  oi.finish(s);
  PK_HIP(hipStreamSynchronize(s));
  if (ent) {
    // one replay per count: the lists are dropped once used
    std::lock_guard<std::mutex> lk(t->rnear_cache.mu);
    if (t->rnear_cache.last == ent) t->rnear_cache.last.reset();
  }
}

void check_dmax(double dmax) {
  if (!(dmax >= 0.0)) fail(PPTK_E_VALUE, "dmax must be non-negative (got %g)", dmax);
}

}  // namespace
}  // namespace pk

namespace pk {
pptk_status api_knn(const pptk_tree* tree, const pptk_queries* q, int k, double dmax,
                                int32_t* out_idx, int32_t* out_cnt, void* out_dist,
                                pptk_mem out_mem, void* stream) {
  return pk::guard([&] {
    pk::check_tree(tree);
    if (k < 1) pk::fail(PPTK_E_VALUE, "k must be positive (got %d)", k);
    pk::check_dmax(dmax);
    hipStream_t s = pk::pick_stream(stream);
    auto run = [&](auto dt, auto dimc) {
      using D = typename decltype(dt)::type;
      constexpr int DIM = decltype(dimc)::value;
      pk::knn_run<D, DIM>(tree, q, k, dmax, out_idx, out_cnt, out_dist, out_mem, s);
    };
    pk::dispatch_tree(tree, run);
  });
}
}  // namespace pk

extern "C" pptk_status pptk_knn(const pptk_tree* tree, const pptk_queries* q, int k, double dmax,
                                int32_t* out_idx, int32_t* out_cnt, void* out_dist,
                                pptk_mem out_mem, void* stream) {
  return pk::api_knn(tree, q, k, dmax, out_idx, out_cnt, out_dist, out_mem, stream);
}

// ---------------------------------------------------- allocation-free kNN ----
//
// pptk_knn_noalloc runs the kNN kernels of pptk_knn with every temporary in
// `scratch`, a file-scope device array each GPU holds from the library's
// load to the process's end: the query order of point and index queries
// (Morton keys, the order and the radix sort's temporaries, ~24.3 bytes per
// query) and the heaps of k > 32 searches. Host code never passes `scratch`
// itself to a kernel: its name is a host-side stand-in, so the device
// address is looked up once per device with hipGetSymbolAddress and kept.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU race and must be serialised by the caller. No part of the array has
// to start at zero (every kernel writes what it later reads).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace pk {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device.
void* scratch_address() {
  int dev = -1;
  PK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(PPTK_E_DEVICE, "device id %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    PK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p || ((uintptr_t)p & 255u) != 0)
      fail(PPTK_E_DEVICE, "scratch symbol address %p is not 256-byte aligned", p);
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

}  // namespace

pptk_status api_knn_noalloc(const pptk_tree* tree, const pptk_queries* q, int k, double dmax,
                            int32_t* out_idx, int32_t* out_cnt, void* out_dist, void* stream) {
  return pk::guard([&] {
    pk::check_tree(tree);
    if (k < 1) pk::fail(PPTK_E_VALUE, "k must be positive (got %d)", k);
    pk::check_dmax(dmax);
    if (!q) pk::fail(PPTK_E_VALUE, "null query set");
    pk::ensure_device();  // the tree's device is current
    const QueryScratch sc{scratch_address(), kNoAllocScratchBytes};
    hipStream_t s = pk::pick_stream(stream);
    auto run = [&](auto dt, auto dimc) {
      using D = typename decltype(dt)::type;
      constexpr int DIM = decltype(dimc)::value;
      pk::knn_run<D, DIM>(tree, q, k, dmax, out_idx, out_cnt, out_dist, PPTK_MEM_DEVICE, s, &sc);
    };
    pk::dispatch_tree(tree, run);
  });
}

}  // namespace pk

extern "C" pptk_status pptk_knn_noalloc(const pptk_tree* tree, const pptk_queries* q, int k,
                                        double dmax, int32_t* out_idx, int32_t* out_cnt,
                                        void* out_dist, void* stream) {
  return pk::api_knn_noalloc(tree, q, k, dmax, out_idx, out_cnt, out_dist, stream);
}

namespace pk {
pptk_status api_rnear_count(const pptk_tree* tree, const pptk_queries* q, double dmax,
                                        int64_t* out_offsets, int64_t* out_total,
                                        pptk_mem out_mem, void* stream) {
  return pk::guard([&] {
    pk::check_tree(tree);
    pk::check_dmax(dmax);
    hipStream_t s = pk::pick_stream(stream);
    auto run = [&](auto dt, auto dimc) {
      using D = typename decltype(dt)::type;
      constexpr int DIM = decltype(dimc)::value;
      pk::rnear_count_run<D, DIM>(tree, q, dmax, out_offsets, out_total, out_mem, s);
    };
    pk::dispatch_tree(tree, run);
  });
}
}  // namespace pk

extern "C" pptk_status pptk_rnear_count(const pptk_tree* tree, const pptk_queries* q, double dmax,
                                        int64_t* out_offsets, int64_t* out_total,
                                        pptk_mem out_mem, void* stream) {
  return pk::api_rnear_count(tree, q, dmax, out_offsets, out_total, out_mem, stream);
}

namespace pk {
pptk_status api_rnear_fill(const pptk_tree* tree, const pptk_queries* q, double dmax,
                                       const int64_t* offsets, int32_t* out_idx, pptk_mem out_mem,
                                       void* stream) {
  return pk::guard([&] {
    pk::check_tree(tree);
    pk::check_dmax(dmax);
    hipStream_t s = pk::pick_stream(stream);
    auto run = [&](auto dt, auto dimc) {
      using D = typename decltype(dt)::type;
      constexpr int DIM = decltype(dimc)::value;
      pk::rnear_fill_run<D, DIM>(tree, q, dmax, offsets, out_idx, out_mem, s);
    };
    pk::dispatch_tree(tree, run);
  });
}
}  // namespace pk

extern "C" pptk_status pptk_rnear_fill(const pptk_tree* tree, const pptk_queries* q, double dmax,
                                       const int64_t* offsets, int32_t* out_idx, pptk_mem out_mem,
                                       void* stream) {
  return pk::api_rnear_fill(tree, q, dmax, offsets, out_idx, out_mem, stream);
}
