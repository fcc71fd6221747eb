// GPU k-d tree build.
//
// Replaces impl::BuildTree (pptk/kdtree/src/kdtree-impl.h:773-818). The
// reference splits each node at the midpoint of its widest side and recurses
// (kdtree-impl.h:175-329); result sets do not depend on the tree shape
// (pruning bounds are exact lower/upper bounds), so the GPU tree is laid out
// for 64-wide traversal instead:
//   K0 ingest_reduce   validity (no NaN/inf, kdtree-impl.h:53-75) + per-tile
//                      counts and bounding boxes
//   K0 ingest_finalize tile offsets, global box, Morton scale
//   K1 morton          stable compaction of valid rows + Morton code
//   K2 radix sort      (sort.hip) on the 63-bit code
//   K3 gather          tree-order point records, reverse/forward maps
//                      (kdtree-impl.h:803-817)
//   K3 split levels    top-down, one launch per level: a node holding more
//                      than `leaf` points splits at the highest Morton bit in
//                      which its first and last codes differ (binary search
//                      for the first code with that bit set), or at the middle
//                      when all its codes are equal. Every node is a Morton
//                      prefix cell, siblings are disjoint cells, and a leaf is
//                      a cell of at most `leaf` points. Children are allocated
//                      in adjacent pairs into one flat node array.
//   K3 node boxes      bottom-up, one launch per level: tight box of each
//                      leaf's points, union of the children's boxes above.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <set>
#include <vector>

#include "levels.hpp"
#include "morton.hpp"
#include "sort.hpp"
#include "tree.hpp"

namespace pk {

namespace {

constexpr int kBlock = 256;
constexpr int kMortonExtraBits = 7;
constexpr int kPer = 16;                 // rows of kBlock points per tile
constexpr int kTile = kBlock * kPer;     // 4096

template <typename T>
__device__ __forceinline__ bool elem_finite(T v) {
  if constexpr (std::is_floating_point<T>::value) return finite_val(v);
  else return true;
}

template <typename D>
__device__ __forceinline__ D wave_min(D x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { D y = __shfl_xor(x, o); x = y < x ? y : x; }
  return x;
}
template <typename D>
__device__ __forceinline__ D wave_max(D x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { D y = __shfl_xor(x, o); x = y > x ? y : x; }
  return x;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Loads the kPer rows first + e * kBlock (e < kPer) of a tile, zero past n.
template <typename T, int DIM>
__device__ __forceinline__ void load_tile_rows(const T* __restrict__ pts, int64_t n, int64_t first,
                                               T (&v)[kPer][DIM]) {
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = first + (int64_t)e * kBlock;
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[e][d] = i < n ? pts[i * DIM + d] : T(0);
  }
}

template <typename T, typename D, int DIM>
__global__ __launch_bounds__(kBlock) void ingest_reduce_kernel(const T* __restrict__ pts, int64_t n,
                                                               uint32_t* __restrict__ blk_cnt,
                                                               D* __restrict__ blk_lo,
                                                               D* __restrict__ blk_hi) {
  // tile rows of kBlock consecutive points: lane-consecutive loads
  const int64_t first = (int64_t)blockIdx.x * kTile + threadIdx.x;
  uint32_t cnt = 0;
  D lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = pos_inf<D>(); hi[d] = -pos_inf<D>(); }
  // every row's loads in flight before the first use
  T v[kPer][DIM];
  load_tile_rows<T, DIM>(pts, n, first, v);
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = first + (int64_t)e * kBlock;
    bool ok = i < n;
#pragma unroll
    for (int d = 0; d < DIM; ++d) ok = ok & elem_finite(v[e][d]);
    if (ok) {
      ++cnt;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const D c = (D)v[e][d];
        lo[d] = c < lo[d] ? c : lo[d];
        hi[d] = c > hi[d] ? c : hi[d];
      }
    }
  }
  cnt = wave_sum_u32(cnt);
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = wave_min(lo[d]); hi[d] = wave_max(hi[d]); }
  __shared__ uint32_t s_cnt[kBlock / 64];
  __shared__ D s_lo[kBlock / 64][DIM];
  __shared__ D s_hi[kBlock / 64][DIM];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_cnt[w] = cnt;
#pragma unroll
    for (int d = 0; d < DIM; ++d) { s_lo[w][d] = lo[d]; s_hi[w][d] = hi[d]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    D l[DIM], h[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) { l[d] = pos_inf<D>(); h[d] = -pos_inf<D>(); }
    for (int ww = 0; ww < kBlock / 64; ++ww) {
      c += s_cnt[ww];
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        l[d] = s_lo[ww][d] < l[d] ? s_lo[ww][d] : l[d];
        h[d] = s_hi[ww][d] > h[d] ? s_hi[ww][d] : h[d];
      }
    }
    blk_cnt[blockIdx.x] = c;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      blk_lo[(int64_t)blockIdx.x * DIM + d] = l[d];
      blk_hi[(int64_t)blockIdx.x * DIM + d] = h[d];
    }
  }
}

// One block: exclusive scan of tile counts into tile offsets, global box,
// Morton quantisation scale.
template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void ingest_finalize_kernel(const uint32_t* __restrict__ blk_cnt,
                                                                 const D* __restrict__ blk_lo,
                                                                 const D* __restrict__ blk_hi,
                                                                 int64_t nblk,
                                                                 uint32_t* __restrict__ blk_off,
                                                                 BuildInfo<D>* __restrict__ info) {
  __shared__ uint32_t s_wsum[kBlock / 64];
  __shared__ D s_lo[kBlock / 64][DIM];
  __shared__ D s_hi[kBlock / 64][DIM];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint32_t carry = 0;
  D lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = pos_inf<D>(); hi[d] = -pos_inf<D>(); }
  for (int64_t b0 = 0; b0 < nblk; b0 += kBlock) {
    const int64_t b = b0 + threadIdx.x;
    const uint32_t c = b < nblk ? blk_cnt[b] : 0u;
    if (b < nblk) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const D l = blk_lo[b * DIM + d], h = blk_hi[b * DIM + d];
        lo[d] = l < lo[d] ? l : lo[d];
        hi[d] = h > hi[d] ? h : hi[d];
      }
    }
    // block exclusive scan of c
    uint32_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == 63) s_wsum[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
      if (i < w) off += s_wsum[i];
      tot += s_wsum[i];
    }
    if (b < nblk) blk_off[b] = carry + off + incl - c;
    carry += tot;
    __syncthreads();
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = wave_min(lo[d]); hi[d] = wave_max(hi[d]); }
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) { s_lo[w][d] = lo[d]; s_hi[w][d] = hi[d]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    info->n_valid = carry;
    info->pad = 0;
    const D maxq = (D)((1u << MortonBits<DIM>::value) - 1u);
    for (int d = 0; d < 4; ++d) {
      D l = pos_inf<D>(), h = -pos_inf<D>();
      if (d < DIM) {
        for (int ww = 0; ww < kBlock / 64; ++ww) {
          l = s_lo[ww][d] < l ? s_lo[ww][d] : l;
          h = s_hi[ww][d] > h ? s_hi[ww][d] : h;
        }
      } else {
        l = h = D(0);
      }
      info->lo[d] = l;
      info->hi[d] = h;
    }
    // One scale for every axis: cubic Morton cells keep node boxes compact
    // on flat or elongated clouds.
    D ext = D(0);
    for (int d = 0; d < DIM; ++d) {
      const D wdt = info->hi[d] - info->lo[d];
      ext = wdt > ext ? wdt : ext;
    }
    for (int d = 0; d < 4; ++d)
      info->scale[d] = (d < DIM && ext > D(0) && finite_val(ext)) ? maxq / ext : D(0);
  }
}

// Stable compaction of a tile's valid rows and their Morton codes to the
// front of the key array, and of its invalid (NaN/inf) rows, keyed
// `sentinel` (every sorted key bit set), behind all n_valid valid ones: a
// stable sort then leaves every invalid row after every valid row, so the
// host need not know n_valid before sorting. Rows are read lane-consecutively
// (row e of the tile = points [e*kBlock, (e+1)*kBlock) of it); a row's points
// are ranked with wave ballots, and the per-(row, wave) counts are scanned in
// (row, wave) order, which is index order, so the compaction keeps the input
// order.
template <typename T, typename D, int DIM>
__global__ __launch_bounds__(kBlock) void morton_kernel(const T* __restrict__ pts, int64_t n,
                                                        const uint32_t* __restrict__ blk_off,
                                                        const BuildInfo<D>* __restrict__ info_p,
                                                        uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals,
                                                        int kshift, uint64_t sentinel) {
  constexpr int kWaves = kBlock / 64;
  static_assert(kPer * kWaves == 64, "one wave scans the (row, wave) counts");
  __shared__ uint32_t s_off[kPer * kWaves];
  __shared__ uint32_t s_inv[kPer * kWaves];
  const BuildInfo<D> info = *info_p;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const int64_t first = (int64_t)blockIdx.x * kTile + threadIdx.x;
  // the tile's rows stay in registers from the finiteness test to the codes
  T v[kPer][DIM];
  load_tile_rows<T, DIM>(pts, n, first, v);
  uint32_t flags = 0;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = first + (int64_t)e * kBlock;
    bool ok = i < n;
#pragma unroll
    for (int d = 0; d < DIM; ++d) ok = ok & elem_finite(v[e][d]);
    flags |= (uint32_t)ok << e;
    const uint64_t b = __ballot(ok);
    const uint64_t inr = __ballot(i < n);
    if (lane == 0) {
      s_off[e * kWaves + w] = (uint32_t)__popcll(b);
      s_inv[e * kWaves + w] = (uint32_t)__popcll(inr & ~b);
    }
  }
  __syncthreads();
  if (w == 0) {
    // exclusive scans of the kPer * kWaves (= 64) counts, one per lane
    const uint32_t c = s_off[lane], ci = s_inv[lane];
    uint32_t incl = c, incli = ci;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      const uint32_t yi = __shfl_up(incli, o);
      if (lane >= o) {
        incl += y;
        incli += yi;
      }
    }
    s_off[lane] = incl - c;
    s_inv[lane] = incli - ci;
  }
  __syncthreads();
  const uint32_t tile0 = blk_off[blockIdx.x];
  // invalid rows of the earlier tiles (all full): their rows minus their valid ones
  const uint32_t inv0 = info.n_valid + (uint32_t)((int64_t)blockIdx.x * kTile - tile0);
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = first + (int64_t)e * kBlock;
    const bool ok = (flags >> e) & 1u;
    const bool bad = !ok && i < n;
    const uint64_t b = __ballot(ok);
    const uint64_t bb = __ballot(bad);
    if (ok) {
      const uint32_t pos = tile0 + s_off[e * kWaves + w] + (uint32_t)__popcll(b & lt_mask);
      D c[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) c[d] = (D)v[e][d];
      keys[pos] = morton_code<D, DIM>(c, info) >> kshift;
      vals[pos] = (uint32_t)i;
    } else if (bad) {
      const uint32_t pos = inv0 + s_inv[e * kWaves + w] + (uint32_t)__popcll(bb & lt_mask);
      keys[pos] = sentinel;
      vals[pos] = (uint32_t)i;
    }
  }
}

// Tree order from the sorted order of all n rows, in two launches: the row
// maps (positions [0, n_valid) are the valid rows in tree order, the rest the
// invalid rows, whose forward entry is -1), then the records, each input row
// read in input order and written to its position. Scattering the 16-byte
// records costs less than gathering the 12-byte input rows in tree order
// (10M-point lidar scene: build 1.645 -> 1.58 ms; the forward map's 4-byte
// entries are scattered either way).
__global__ __launch_bounds__(kBlock) void row_maps_kernel(const uint32_t* __restrict__ order,
                                                          int64_t n,
                                                          const uint32_t* __restrict__ nv_dev,
                                                          int32_t* __restrict__ orig,
                                                          int32_t* __restrict__ fwd) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const uint32_t i = order[p];
  if (p >= (int64_t)*nv_dev) {
    fwd[i] = -1;
    return;
  }
  orig[p] = (int32_t)i;
  fwd[i] = (int32_t)p;
}

template <typename T, typename D, int DIM>
__global__ __launch_bounds__(kBlock) void scatter_rows_kernel(const T* __restrict__ pts, int64_t n,
                                                              const int32_t* __restrict__ fwd,
                                                              PointRec<D, DIM>* __restrict__ rec) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t p = fwd[i];
  if (p < 0) return;
  PointRec<D, DIM> r;
#pragma unroll
  for (int d = 0; d < DIM; ++d) r.c[d] = (D)pts[i * DIM + d];
  r.orig = (int32_t)i;
  rec[p] = r;
}

unsigned grid_for(int64_t n, int block) {
  const int64_t g = ceil_div(n, block);
  return (unsigned)(g > 0 ? g : 1);
}

template <typename T, typename D, int DIM>
void build_impl(pptk_tree* t, const T* dpts, int64_t n, hipStream_t s) {
  // K0: validity + bounding box
  const int64_t nblk = n > 0 ? ceil_div(n, kTile) : 0;
  DevBuf blk_cnt((size_t)std::max<int64_t>(nblk, 1) * 4);
  DevBuf blk_off((size_t)std::max<int64_t>(nblk, 1) * 4);
  DevBuf blk_lo((size_t)std::max<int64_t>(nblk, 1) * DIM * sizeof(D));
  DevBuf blk_hi((size_t)std::max<int64_t>(nblk, 1) * DIM * sizeof(D));
  DevBuf info_d(sizeof(BuildInfo<D>));
  if (nblk > 0) {
    ingest_reduce_kernel<T, D, DIM><<<(unsigned)nblk, kBlock, 0, s>>>(
        dpts, n, blk_cnt.as<uint32_t>(), blk_lo.as<D>(), blk_hi.as<D>());
    PK_LAUNCH_CHECK();
  }
  ingest_finalize_kernel<D, DIM><<<1, kBlock, 0, s>>>(blk_cnt.as<uint32_t>(), blk_lo.as<D>(),
                                                      blk_hi.as<D>(), nblk,
                                                      blk_off.as<uint32_t>(),
                                                      info_d.as<BuildInfo<D>>());
  PK_LAUNCH_CHECK();
  // The valid-row count and the frame stay on the device until the level
  // table is read back: every launch before that is sized by n (invalid rows
  // sort behind the valid ones) and reads the count where it needs it.
  const uint32_t* nv_dev = &info_d.as<BuildInfo<D>>()->n_valid;

  const int64_t B = t->leaf;
  t->pts.reset((size_t)std::max<int64_t>(n, 1) * sizeof(PointRec<D, DIM>));
  t->orig.reset((size_t)std::max<int64_t>(n, 1) * 4);
  t->fwd.reset((size_t)std::max<int64_t>(n, 1) * 4);

  // A binary tree whose leaves are non-empty has at most 2*nv - 1 <= 2*n - 1 nodes.
  const int64_t cap = std::max<int64_t>(1, 2 * n - 1);
  DevBuf nodes((size_t)cap * sizeof(NodeRec));
  // Level table: every split launch reads its range from device memory and
  // publishes the next level, so the host reads it back only once per batch
  // of levels to see whether the tree is complete: first after the balanced
  // depth log2(nv / B) plus 20 (Morton splits are uneven; a 10M-point lidar
  // scene at 192 points per leaf is 36 levels deep), then every kLevelBatch
  // levels. A launch past the last level only reads an empty
  // range. Depth is bounded by the key bits plus 31 halvings of equal keys.
  constexpr int kLevelBatch = 8;
  const int max_levels = DIM * MortonBits<DIM>::value + 33;
  // level table [lev_s | lev_e] followed by the LevelState, read back in one
  // copy at the end
  const size_t lev_bytes = (size_t)(2 * (max_levels + 1)) * 4;
  const size_t st_off = (lev_bytes + 15) / 16 * 16;
  DevBuf lev(st_off + sizeof(LevelState));
  DevBuf done((size_t)max_levels * 4);
  int32_t* lev_s = lev.as<int32_t>();
  int32_t* lev_e = lev_s + (max_levels + 1);
  LevelState* state_d = (LevelState*)((char*)lev.get() + st_off);
  // root node, state and level table set by one kernel: no host copies
  level_init_kernel<<<1, kLevelBlock, 0, s>>>(nodes.as<NodeRec>(), state_d, lev_s, lev_e,
                                               done.as<uint32_t>(), max_levels, 0, nv_dev);
  PK_LAUNCH_CHECK();
  DevBuf keys, vals;
  if (n > 0) {
    // K1: compaction + Morton codes
    keys.reset((size_t)n * 8);
    vals.reset((size_t)n * 4);
    // Key bits per axis: enough for ~2^kMortonExtraBits more cells per axis
    // than the points need on a surface (ceil(log2 n / DIM) + extra), at most
    // the full code; the low bits are dropped, so fewer radix passes.
    int lg = 0;
    while ((int64_t(1) << lg) < n) ++lg;
    const int kbits = std::min(MortonBits<DIM>::value, (lg + DIM - 1) / DIM + kMortonExtraBits);
    const int kshift = DIM * (MortonBits<DIM>::value - kbits);
    const uint64_t sentinel = (DIM * kbits >= 64) ? ~0ull : ((1ull << (DIM * kbits)) - 1ull);
    morton_kernel<T, D, DIM><<<(unsigned)nblk, kBlock, 0, s>>>(
        dpts, n, blk_off.as<uint32_t>(), info_d.as<BuildInfo<D>>(), keys.as<uint64_t>(),
        vals.as<uint32_t>(), kshift, sentinel);
    PK_LAUNCH_CHECK();
    // K2: sort (stable: the invalid rows, placed behind the valid ones, stay
    // behind any valid key equal to the sentinel)
    radix_sort_pairs(keys.as<uint64_t>(), vals.as<uint32_t>(), n, DIM * kbits, s);
    // K3: row maps, then the records in tree order
    row_maps_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(
        vals.as<uint32_t>(), n, nv_dev, t->orig.as<int32_t>(), t->fwd.as<int32_t>());
    PK_LAUNCH_CHECK();
    scatter_rows_kernel<T, D, DIM><<<grid_for(n, kBlock), kBlock, 0, s>>>(
        dpts, n, t->fwd.as<const int32_t>(), t->pts.as<PointRec<D, DIM>>());
    PK_LAUNCH_CHECK();
  }
  // K3: split level by level until a level creates no children.
  const unsigned split_grid = (unsigned)std::min<int64_t>(
      1024, std::max<int64_t>(kSplitProbeLanes / kLevelBlock, ceil_div(n, (int64_t)B * kBlock)));
  int nlev = 1;  // levels [0, nlev) hold nodes
  // the level table and state come back in one copy per batch of levels
  std::vector<char> tab(st_off + sizeof(LevelState));
  const int32_t* hs = (const int32_t*)tab.data();
  const int32_t* he = hs + (max_levels + 1);
  BuildInfo<D> info;
  auto read_table = [&] {
    PK_HIP(hipMemcpyAsync(tab.data(), lev.get(), tab.size(), hipMemcpyDeviceToHost, s));
    PK_HIP(hipMemcpyAsync(&info, info_d.get(), sizeof(info), hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
  };
  if (n > 0) {
    int balanced = 0;
    while (((int64_t)B << balanced) < n) ++balanced;
    for (int l0 = 0;;) {
      const int l1 = std::min(l0 == 0 ? std::max(kLevelBatch, balanced + 20) : l0 + kLevelBatch,
                              max_levels);
      for (int l = l0; l < l1; ++l) {
        split_level_dev_kernel<<<split_grid, kLevelBlock, 0, s>>>(
            keys.as<uint64_t>(), nodes.as<NodeRec>(), lev_s, lev_e, done.as<uint32_t>(), l,
            max_levels, (int)B, cap, state_d);
        PK_LAUNCH_CHECK();
      }
      read_table();
      if (he[l1] <= hs[l1] || l1 >= max_levels) break;
      l0 = l1;
    }
  } else {
    read_table();
  }
  const int64_t nv = info.n_valid;
  t->n_valid = nv;
  for (int d = 0; d < 4; ++d) {
    t->lo[d] = d < DIM ? (double)info.lo[d] : 0.0;
    t->hi[d] = d < DIM ? (double)info.hi[d] : 0.0;
    t->mlo[d] = (double)info.lo[d];
    t->mscale[d] = (double)info.scale[d];
  }
  LevelState st;
  std::memcpy(&st, tab.data() + st_off, sizeof(st));
  if (st.overflow) fail(PPTK_E_RUNTIME, "tree build exceeded its node or level bound");
  nlev = 0;
  for (int l = 0; l <= max_levels; ++l)
    if (he[l] > hs[l]) nlev = l + 1;
  nlev = std::max(nlev, 1);
  const int64_t nn = st.counter;
  t->nn = nn;
  t->depth = nlev;
  t->nleaves = (nn + 1) / 2;
  t->node.reset((size_t)nn * sizeof(NodeRec));
  PK_HIP(hipMemcpyAsync(t->node.get(), nodes.get(), (size_t)nn * sizeof(NodeRec),
                        hipMemcpyDeviceToDevice, s));
  t->box.reset((size_t)nn * sizeof(BoxRec<D, DIM>));
  t->leafof.reset((size_t)std::max<int64_t>(nv, 1) * 4);
  {
    // leaf boxes (the empty tree's root gets the empty box)
    const unsigned g =
        (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, ceil_div(nn * 8, kLevelBlock)));
    leaf_box_kernel<D, DIM><<<g, kLevelBlock, 0, s>>>(
        t->pts.as<PointRec<D, DIM>>(), t->node.as<NodeRec>(), t->box.as<BoxRec<D, DIM>>(),
        t->leafof.as<int32_t>(), nn, nullptr);
    PK_LAUNCH_CHECK();
  }
  // Inner boxes bottom-up (the deepest level holds only leaves): a wide level
  // takes a launch of its own, a run of narrow levels one single-block launch.
  for (int l = nlev - 2; l >= 0;) {
    const int64_t width = he[l] - hs[l];
    if (width > kBoxRunWidth) {
      const unsigned g =
          (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(width, kLevelBlock)));
      box_level_dev_kernel<D, DIM><<<g, kLevelBlock, 0, s>>>(
          t->node.as<NodeRec>(), t->box.as<BoxRec<D, DIM>>(), lev_s, lev_e, l);
      PK_LAUNCH_CHECK();
      --l;
      continue;
    }
    int lo = l;
    while (lo > 0 && he[lo - 1] - hs[lo - 1] <= kBoxRunWidth) --lo;
    box_levels_run_kernel<D, DIM><<<1, kBoxRunBlock, 0, s>>>(
        t->node.as<NodeRec>(), t->box.as<BoxRec<D, DIM>>(), lev_s, lev_e, l, lo);
    PK_LAUNCH_CHECK();
    l = lo - 1;
  }
  PK_HIP(hipStreamSynchronize(s));  // temporaries return to the pool
}

std::mutex g_trees_mu;
std::set<const void*> g_trees;

}  // namespace

// Measured on MI355X (10M-point lidar scene: build, kNN k = 16 of every
// point, r-near of 1M points at 0.1 m, normals k = 16): the step takes
// 16.0 / 15.4 / 14.8 / 14.6 / 14.5 / 14.3 ms at 64 / 80 / 96 / 112 / 128 /
// 160 points per leaf and stays within noise up to 256. With cheaper leaf
// entries in the cooperative search it takes 14.45 / 14.27 / 14.12 / 14.22 ms
// at 128 / 160 / 192 / 224. Point queries, which scan leaves lane by lane,
// are 8% slower at 192 than at 96.
int default_leaf_size(int dim, bool dbl, int64_t n) {
  (void)dim; (void)dbl; (void)n;
  return 192;
}

void register_tree(pptk_tree* t) {
  std::lock_guard<std::mutex> lk(g_trees_mu);
  g_trees.insert(t);
}

const pptk_tree* check_tree(const pptk_tree* t) {
  if (!t) fail(PPTK_E_TYPE, "null tree");
  std::lock_guard<std::mutex> lk(g_trees_mu);
  if (!g_trees.count(t) || t->magic != pptk_tree::kMagic)
    fail(PPTK_E_TYPE, "object is not a live k-d tree handle");
  return t;
}

}  // namespace pk

namespace pk {
pptk_status api_tree_build(const void* pts, int64_t n, int dim, pptk_dtype dtype,
                                       const pptk_build_params* params, pptk_mem mem,
                                       void* stream, pptk_tree** out) {
  return pk::guard([&] {
    if (!out) pk::fail(PPTK_E_VALUE, "null output handle");
    *out = nullptr;
    if (dim < 2 || dim > 4) pk::fail(PPTK_E_RUNTIME, "Invalid k-d tree dim = %d.", dim);
    if (!pk::dtype_valid(dtype)) pk::fail(PPTK_E_RUNTIME, "Invalid k-d tree type_num = %d.", (int)dtype);
    if (n < 0) pk::fail(PPTK_E_VALUE, "number of points must be non-negative");
    if (n > (1LL << 30)) pk::fail(PPTK_E_VALUE, "at most 2^30 points per tree (int32 node ids)");
    if (n > 0 && !pts) pk::fail(PPTK_E_VALUE, "null points");
    hipStream_t s = pk::pick_stream(stream);
    std::unique_ptr<pptk_tree> t(new pptk_tree());
    t->dim = dim;
    t->dtype = dtype;
    t->dist = dtype == PPTK_F64 ? PPTK_F64 : PPTK_F32;
    t->n_input = n;
    int leaf = params && params->max_leaf_size > 0 ? params->max_leaf_size
                                                   : pk::default_leaf_size(dim, t->dist == PPTK_F64, n);
    t->leaf = std::min(256, std::max(1, leaf));
    pk::DevInput in;
    pk::stage_input(in, pts, (size_t)n * dim * pk::dtype_size(dtype), mem, s);
    pptk_tree* tp = t.get();
    const void* p = in.ptr;
    auto run = [&](auto tag) {
        using T = typename decltype(tag)::type;
        pptk_tree* t = tp;
        if (t->dist == PPTK_F64) {
          switch (t->dim) {
            case 2: pk::build_impl<T, double, 2>(t, (const T*)p, n, s); break;
            case 3: pk::build_impl<T, double, 3>(t, (const T*)p, n, s); break;
            default: pk::build_impl<T, double, 4>(t, (const T*)p, n, s); break;
          }
        } else {
          switch (t->dim) {
            case 2: pk::build_impl<T, float, 2>(t, (const T*)p, n, s); break;
            case 3: pk::build_impl<T, float, 3>(t, (const T*)p, n, s); break;
            default: pk::build_impl<T, float, 4>(t, (const T*)p, n, s); break;
          }
        }
    };
    pk::dispatch_dtype(dtype, run);
    {
      std::lock_guard<std::mutex> lk(pk::g_trees_mu);
      pk::g_trees.insert(t.get());
    }
    *out = t.release();
  });
}
}  // namespace pk

extern "C" pptk_status pptk_tree_build(const void* pts, int64_t n, int dim, pptk_dtype dtype,
                                       const pptk_build_params* params, pptk_mem mem,
                                       void* stream, pptk_tree** out) {
  return pk::api_tree_build(pts, n, dim, dtype, params, mem, stream, out);
}

extern "C" void pptk_tree_free(pptk_tree* tree) {
  if (!tree) return;
  {
    std::lock_guard<std::mutex> lk(pk::g_trees_mu);
    if (!pk::g_trees.erase(tree)) return;
  }
  tree->magic = 0;
  delete tree;
}

extern "C" int pptk_tree_check(const void* p) {
  std::lock_guard<std::mutex> lk(pk::g_trees_mu);
  return p && pk::g_trees.count(p) ? 1 : 0;
}

extern "C" pptk_status pptk_tree_get_info(const pptk_tree* tree, pptk_tree_info* out) {
  return pk::guard([&] {
    pk::check_tree(tree);
    if (!out) pk::fail(PPTK_E_VALUE, "null output");
    out->dim = tree->dim;
    out->dtype = tree->dtype;
    out->dist_dtype = tree->dist;
    out->num_input = tree->n_input;
    out->num_valid = tree->n_valid;
    out->leaf_size = tree->leaf;
    out->num_leaves = tree->nleaves;
    out->num_nodes = tree->nn;
    out->depth = tree->depth;
    for (int d = 0; d < 4; ++d) { out->bbox_lo[d] = tree->lo[d]; out->bbox_hi[d] = tree->hi[d]; }
    out->device_bytes = (int64_t)(tree->pts.bytes() + tree->orig.bytes() + tree->fwd.bytes() +
                                  tree->box.bytes() + tree->node.bytes() +
                                  tree->leafof.bytes());
  });
}

namespace pk {
namespace {
template <typename D, int DIM>
__global__ __launch_bounds__(256) void export_points_kernel(const PointRec<D, DIM>* __restrict__ rec,
                                                            int64_t nv, D* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nv) return;
  const PointRec<D, DIM> r = rec[p];
#pragma unroll
  for (int d = 0; d < DIM; ++d) out[p * DIM + d] = r.c[d];
}
template <typename D, int DIM>
__global__ __launch_bounds__(256) void export_boxes_kernel(const BoxRec<D, DIM>* __restrict__ box,
                                                           int64_t nn, D* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const BoxRec<D, DIM> b = box[i];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    out[i * 2 * DIM + d] = b.lo[d];
    out[i * 2 * DIM + DIM + d] = b.hi[d];
  }
}
}  // namespace
}  // namespace pk

namespace pk {
pptk_status api_tree_export(const pptk_tree* tree, void* points, int32_t* reverse,
                                        int32_t* forward, void* boxes, int32_t* nodes,
                                        pptk_mem mem) {
  return pk::guard([&] {
    pk::check_tree(tree);
    hipStream_t s = pk::lib_stream();
    const int64_t nv = tree->n_valid;
    const int64_t nn = tree->nn;
    const pptk_tree* t = tree;
    auto run = [&](auto dt, auto dimc) {
        using D = typename decltype(dt)::type;
        constexpr int DIM = decltype(dimc)::value;
        if (points && nv > 0) {
          pk::DevOutput o;
          o.init(points, (size_t)nv * DIM * sizeof(D), mem);
          pk::export_points_kernel<D, DIM><<<(unsigned)pk::ceil_div(nv, 256), 256, 0, s>>>(
              t->pts.as<const pk::PointRec<D, DIM>>(), nv, (D*)o.dptr);
          PK_LAUNCH_CHECK();
          o.finish(s);
          PK_HIP(hipStreamSynchronize(s));
        }
        if (boxes) {
          pk::DevOutput o;
          o.init(boxes, (size_t)nn * 2 * DIM * sizeof(D), mem);
          pk::export_boxes_kernel<D, DIM><<<(unsigned)pk::ceil_div(nn, 256), 256, 0, s>>>(
              t->box.as<const pk::BoxRec<D, DIM>>(), nn, (D*)o.dptr);
          PK_LAUNCH_CHECK();
          o.finish(s);
          PK_HIP(hipStreamSynchronize(s));
        }
    };
    pk::dispatch_tree(tree, run);
    const hipMemcpyKind k = mem == PPTK_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (reverse && nv > 0) PK_HIP(hipMemcpyAsync(reverse, tree->orig.get(), (size_t)nv * 4, k, s));
    if (forward && tree->n_input > 0)
      PK_HIP(hipMemcpyAsync(forward, tree->fwd.get(), (size_t)tree->n_input * 4, k, s));
    if (nodes) PK_HIP(hipMemcpyAsync(nodes, tree->node.get(), (size_t)nn * 16, k, s));
    PK_HIP(hipStreamSynchronize(s));
  });
}
}  // namespace pk

extern "C" pptk_status pptk_tree_export(const pptk_tree* tree, void* points, int32_t* reverse,
                                        int32_t* forward, void* boxes, int32_t* nodes,
                                        pptk_mem mem) {
  return pk::api_tree_export(tree, points, reverse, forward, boxes, nodes, mem);
}
