// Device-driven level kernels of the top-down tree build, shared by the
// single-tree build and the batched forest build.
//
// A split launch reads its node range [lev_s[l], lev_e[l]) from device
// memory, splits every node of the range that holds more than `leaf` points
// (at the highest key bit where its first and last keys differ, found by a
// search over the sorted keys, or at the middle when all its keys are equal),
// appends the children in adjacent pairs at the shared counter, and its last
// block to finish publishes level l + 1. One launch then computes every
// leaf's box, and the box launches combine child boxes bottom-up over the
// same level table. No host read is needed between levels.
#pragma once
#include "common.hpp"

namespace pk {

constexpr int kLevelBlock = 256;
// lanes a split launch spends on the searches of a narrow level
constexpr int64_t kSplitProbeLanes = 16384;

struct LevelState {
  int32_t counter;    // next free node id
  int32_t overflow;   // node capacity or level bound exceeded
  int32_t pad[2];
  unsigned long long nfb;  // batch: rows that took the FP64 eigensolve
};

// Start of a single-tree build: node 0 = [0, n) with no children, level 0 =
// {node 0}, levels 1..nlev empty, no split launch done, counter = 1. One
// block; replaces host copies and memsets that would each be a queue entry.
static __global__ __launch_bounds__(kLevelBlock) void level_init_kernel(NodeRec* __restrict__ nodes,
                                                            LevelState* __restrict__ st,
                                                            int32_t* __restrict__ lev_s,
                                                            int32_t* __restrict__ lev_e,
                                                            uint32_t* __restrict__ done,
                                                            int nlev, int32_t n_arg,
                                                            const uint32_t* __restrict__ n_dev) {
  // the root's size: n_arg, or read here when given on the device
  const int32_t n = n_dev ? (int32_t)*n_dev : n_arg;
  for (int i = threadIdx.x; i <= nlev; i += kLevelBlock) {
    lev_s[i] = 0;
    lev_e[i] = i == 0 ? 1 : 0;
    if (i < nlev) done[i] = 0u;
  }
  if (threadIdx.x == 0) {
    NodeRec root;
    root.a = 0; root.b = n; root.child = -1; root.parent = -1;
    nodes[0] = root;
    LevelState z;
    z.counter = 1; z.overflow = 0; z.pad[0] = 0; z.pad[1] = 0; z.nfb = 0ull;
    *st = z;
  }
}

// First index p in [lo, hi] whose key has `bit` set, for sorted keys that
// agree above `bit` and with keys[hi] having it set. One lane (g = 1) does a
// binary search. A group of g > 1 lanes (g a power of two <= 64, aligned in
// its wave, converged) probes lo + j * span / g for j < g and keeps the part
// where the bit turns on, so a range of r keys takes about log_g(r) dependent
// loads instead of log_2(r). Every lane of the group returns the same p.
__device__ __forceinline__ int32_t first_set_bit_index(const uint64_t* __restrict__ keys,
                                                       int32_t lo, int32_t hi, int bit,
                                                       int lg, int sub, int gbase) {
  if (lg == 0) {
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if ((keys[mid] >> bit) & 1ull) hi = mid;
      else lo = mid + 1;
    }
    return lo;
  }
  const int g = 1 << lg;
  const uint64_t gmask = g == 64 ? ~0ull : (((1ull << g) - 1ull) << gbase);
  while (lo < hi) {
    const int64_t span = (int64_t)hi - lo;
    // probe j: lo + j * span / g, in [lo, hi) and non-decreasing in j
    const int32_t q = lo + (int32_t)(((int64_t)sub * span) >> lg);
    const uint64_t m = __ballot((keys[q] >> bit) & 1ull) & gmask;
    if (m) {
      const int j = __ffsll((unsigned long long)m) - 1 - gbase;  // first probe with the bit
      // probe j - 1 (a different, lower index) lacks the bit
      const int32_t qj = lo + (int32_t)(((int64_t)j * span) >> lg);
      if (j > 0) lo = lo + (int32_t)(((int64_t)(j - 1) * span) >> lg) + 1;
      hi = qj;
    } else {
      lo = lo + (int32_t)(((int64_t)(g - 1) * span) >> lg) + 1;
    }
  }
  return lo;
}

// Split every node of level l (the same rule as the single-tree build: the
// highest differing key bit, or the middle when all keys are equal). A
// narrow level gives each node a group of up to 64 lanes for the split
// search, as many as keep width * lanes within kSplitProbeLanes (the top and
// bottom levels of a tree are narrow, and their searches would otherwise be
// long chains of dependent loads); a wide level gives each node one lane.
// Only the blocks the level needs take part (the rest return at once); the
// last of them to finish publishes level l+1 = the pairs this launch
// appended.
static __global__ __launch_bounds__(kLevelBlock) void split_level_dev_kernel(const uint64_t* __restrict__ keys,
                                                                 NodeRec* __restrict__ nodes,
                                                                 int32_t* __restrict__ lev_s,
                                                                 int32_t* __restrict__ lev_e,
                                                                 uint32_t* __restrict__ done,
                                                                 int l, int nlev, int leaf,
                                                                 int64_t cap,
                                                                 LevelState* __restrict__ st) {
  const int32_t s = lev_s[l], e = lev_e[l];
  const int64_t width = (int64_t)e - s;
  int lg = 0;
  while (lg < 6 && width * (2 << lg) <= kSplitProbeLanes) ++lg;
  const int g = 1 << lg;
  const int64_t want = width > 0 ? (width * g + kLevelBlock - 1) / kLevelBlock : 1;
  const unsigned used = (unsigned)(want < (int64_t)gridDim.x ? want : (int64_t)gridDim.x);
  if (blockIdx.x >= used) return;
  const int lane = (int)(threadIdx.x & 63);
  const int sub = lane & (g - 1);
  const int gbase = lane - sub;
  const int64_t groups = (int64_t)used * (kLevelBlock / g);
  // every lane of a group takes the same nodes and the same branches
  for (int64_t i = s + ((int64_t)blockIdx.x * kLevelBlock + threadIdx.x) / g; i < e; i += groups) {
    const int32_t a = nodes[i].a, b = nodes[i].b;
    if (b - a <= leaf) continue;  // child stays -1
    const uint64_t x = keys[a] ^ keys[b - 1];
    int32_t split;
    if (x == 0ull) {
      split = a + (b - a) / 2;
    } else {
      const int bit = 63 - __clzll((long long)x);
      split = first_set_bit_index(keys, a + 1, b - 1, bit, lg, sub, gbase);
    }
    if (sub != 0) continue;
    const int c = atomicAdd(&st->counter, 2);
    if ((int64_t)c + 2 > cap || l + 1 >= nlev) {
      st->overflow = 1;
      continue;
    }
    NodeRec lc, rc;
    lc.a = a; lc.b = split; lc.child = -1; lc.parent = (int32_t)i;
    rc.a = split; rc.b = b; rc.child = -1; rc.parent = (int32_t)i;
    nodes[c] = lc;
    nodes[c + 1] = rc;
    nodes[i].child = c;
  }
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&done[l], 1u) == used - 1u;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const int32_t c = atomicAdd(&st->counter, 0);
    lev_s[l + 1] = e;
    lev_e[l + 1] = c < (int32_t)cap ? c : (int32_t)cap;
  }
}

// Boxes and leafof of every leaf, eight lanes per node (nodes [0, nn), nn
// read from *nn_dev when given): the lanes of a leaf read its points eight
// consecutive records at a time and combine their boxes with shuffles.
template <typename D, int DIM>
static __global__ __launch_bounds__(kLevelBlock) void leaf_box_kernel(const PointRec<D, DIM>* __restrict__ rec,
                                                          const NodeRec* __restrict__ nodes,
                                                          BoxRec<D, DIM>* __restrict__ box,
                                                          int32_t* __restrict__ leafof,
                                                          int64_t nn,
                                                          const int32_t* __restrict__ nn_dev) {
  if (nn_dev) nn = *nn_dev;
  const int sub = threadIdx.x & 7;
  const int64_t groups = (int64_t)gridDim.x * (kLevelBlock / 8);
  for (int64_t i = ((int64_t)blockIdx.x * kLevelBlock + threadIdx.x) >> 3; i < nn; i += groups) {
    const NodeRec nr = nodes[i];
    if (nr.child >= 0) continue;  // the same for the group's eight lanes
    D lo[DIM], hi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      lo[d] = pos_inf<D>();
      hi[d] = -pos_inf<D>();
    }
    // four records per lane in flight: a lane's loads do not wait on each
    // other (the last ones of a leaf repeat its last point)
    for (int32_t p0 = nr.a + sub; p0 < nr.b; p0 += 32) {
      PointRec<D, DIM> r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int32_t p = p0 + 8 * u;
        r[u] = rec[p < nr.b ? p : nr.b - 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int32_t p = p0 + 8 * u;
        if (p < nr.b) leafof[p] = (int32_t)i;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          lo[d] = r[u].c[d] < lo[d] ? r[u].c[d] : lo[d];
          hi[d] = r[u].c[d] > hi[d] ? r[u].c[d] : hi[d];
        }
      }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const D l2 = __shfl_xor(lo[d], o);
        const D h2 = __shfl_xor(hi[d], o);
        lo[d] = l2 < lo[d] ? l2 : lo[d];
        hi[d] = h2 > hi[d] ? h2 : hi[d];
      }
    }
    if (sub == 0) {
      BoxRec<D, DIM> b;
#pragma unroll
      for (int d = 0; d < PadDim<DIM>::value; ++d) {
        b.lo[d] = d < DIM ? lo[d < DIM ? d : 0] : D(0);
        b.hi[d] = d < DIM ? hi[d < DIM ? d : 0] : D(0);
      }
      box[i] = b;
    }
  }
}

// Box of inner node i from its children's boxes (leaves keep theirs).
template <typename D, int DIM>
__device__ __forceinline__ void combine_child_boxes(const NodeRec* __restrict__ nodes,
                                                    BoxRec<D, DIM>* __restrict__ box, int64_t i) {
  const int32_t c = nodes[i].child;
  if (c < 0) return;  // leaf boxes come from leaf_box_kernel
  const BoxRec<D, DIM> x = box[c];
  const BoxRec<D, DIM> y = box[c + 1];
  BoxRec<D, DIM> b;
#pragma unroll
  for (int d = 0; d < PadDim<DIM>::value; ++d) {
    b.lo[d] = d < DIM ? (x.lo[d] < y.lo[d] ? x.lo[d] : y.lo[d]) : D(0);
    b.hi[d] = d < DIM ? (x.hi[d] > y.hi[d] ? x.hi[d] : y.hi[d]) : D(0);
  }
  box[i] = b;
}

// Boxes of the inner nodes of level l from their children's boxes; run
// deepest level first.
template <typename D, int DIM>
static __global__ __launch_bounds__(kLevelBlock) void box_level_dev_kernel(const NodeRec* __restrict__ nodes,
                                                               BoxRec<D, DIM>* __restrict__ box,
                                                               const int32_t* __restrict__ lev_s,
                                                               const int32_t* __restrict__ lev_e,
                                                               int l) {
  const int32_t s = lev_s[l], e = lev_e[l];
  const int64_t stride = (int64_t)gridDim.x * kLevelBlock;
  for (int64_t i = s + (int64_t)blockIdx.x * kLevelBlock + threadIdx.x; i < e; i += stride)
    combine_child_boxes<D, DIM>(nodes, box, i);
}

// The same for a run of narrow levels l_hi, l_hi - 1, ..., l_lo in ONE block:
// a barrier between levels takes the place of a launch per level (the block's
// waves share a compute unit, so their box writes are visible to each other
// after the barrier).
constexpr int kBoxRunBlock = 1024;
constexpr int64_t kBoxRunWidth = 4 * kBoxRunBlock;  // widest level a run takes

template <typename D, int DIM>
static __global__ __launch_bounds__(kBoxRunBlock) void box_levels_run_kernel(const NodeRec* __restrict__ nodes,
                                                                 BoxRec<D, DIM>* __restrict__ box,
                                                                 const int32_t* __restrict__ lev_s,
                                                                 const int32_t* __restrict__ lev_e,
                                                                 int l_hi, int l_lo) {
  for (int l = l_hi; l >= l_lo; --l) {
    const int32_t s = lev_s[l], e = lev_e[l];
    for (int64_t i = s + threadIdx.x; i < e; i += kBoxRunBlock)
      combine_child_boxes<D, DIM>(nodes, box, i);
    __syncthreads();
  }
}

}  // namespace pk
