// Batched clouds: a forest build over many concatenated clouds plus the fused
// self-query kNN and PCA normals of every point, with every buffer allocated
// when the plan is made and no host synchronisation inside a run, so that the
// whole run can be captured once into a hip graph and replayed.
//
// Each cloud is the equivalent of one kdtree._build + kdtree._query(None, k)
// + estimate_normals(k) sequence (pptk/kdtree/kdtree_wrapper.cpp:281-304,
// pptk/kdtree/src/kdtree-impl.h:957-976, pptk/processing/estimate_normals/
// estimate_normals.cpp:64-111); the results per cloud are those of the
// single-cloud calls on that cloud alone.
//
// Forest layout (same records as a single tree, see build.hip):
//   * sort key = cloud id above the cloud's Morton code (mb bits per axis on
//     the cloud's own cubic frame, mb from the largest cloud's size); rows with NaN/inf get cloud id C and sort
//     after every valid row, so no compaction pass is needed;
//   * nodes [0, C) are the cloud roots (parent -1); when C is even, node C is
//     an empty placeholder root so that child pairs start at an odd id;
//   * PointRec.orig holds the index LOCAL to the cloud (ties are broken by
//     it, and it is what the kNN rows report); `gorig` maps a tree position
//     to its global input row.
//
// Device-driven level loop: the top-down split and the bottom-up box pass run
// as a fixed number of grid-stride launches (one per possible level; the
// depth is bounded by the Morton bits per cloud plus the halvings that bring
// the largest cloud's equal-key runs to the bucket size).
// Each split launch reads its node range from device memory and its last
// block to finish publishes the next level's range, so the host never reads
// the node counter.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <mutex>
#include <set>
#include <vector>

#include "coop.hpp"
#include "levels.hpp"
#include "morton.hpp"
#include "pca.hpp"
#include "sort.hpp"
#include "traverse.hpp"

struct pptk_batch {
  static constexpr uint64_t kMagic = 0x70707462617463ull;
  uint64_t magic = kMagic;
  int dim = 3;
  pptk_dtype dtype = PPTK_F32;
  int64_t C = 0;      // clouds
  int64_t N = 0;      // total input rows
  int64_t R = 0;      // root slots (C rounded up to odd)
  int64_t cap = 0;    // node capacity
  int64_t ntiles = 0;
  int leaf = 16;
  int mb = 21;        // Morton bits per axis
  int shift = 63;     // bit position of the cloud id in the key
  int end_bit = 63;
  int nlev = 0;       // level launches
  int k = 0, nk = 0, kk = 0;
  double dmax = 0.0;
  int prec = PPTK_PREC_MIXED;
  bool use_graph = false;
  int qblock = 256;
  bool timing = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  double build_ms = 0.0, query_ms = 0.0;
  int launches = 0;
  pk::DevBuf coff;    // int64[C+1] cloud input offsets
  pk::DevBuf tiles;   // TileRec[ntiles]
  pk::DevBuf tbeg;    // int32[C+1] first tile of each cloud
  pk::DevBuf tcnt, tlo, thi;  // per-tile valid count and box
  pk::DevBuf cinfo;   // CloudFrame<D>[C]
  pk::DevBuf vcnt;    // uint32[C+1]: valid counts -> exclusive offsets, [C] = total
  pk::DevBuf keys, vals;
  pk::SortWorkspace ws;
  pk::DevBuf rec, gorig, fwd, leafof, nodes, box;
  pk::DevBuf lev_s, lev_e, done;
  pk::DevBuf state;   // BatchState
  // hip graph of one run, keyed on the buffers it was captured with
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* gkey[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // filled after a run
  int64_t nv = 0, nn = 0, nfb = 0;
  bool ran = false;
};

namespace pk {

namespace {

constexpr int kBlock = 256;
constexpr int kPer = 16;
constexpr int kTile = kBlock * kPer;  // input rows per tile
constexpr int kLevelGrid = 1024;      // blocks of the grid-stride level launches

struct TileRec {
  int64_t start, end;
  int32_t cloud;
  int32_t pad;
};

template <typename D>
struct CloudFrame {
  D lo[4];
  D scale;
};

using BatchState = LevelState;

template <typename T>
__device__ __forceinline__ bool row_valid(const T* p, int dim) {
  bool ok = true;
  for (int d = 0; d < dim; ++d) ok = ok && finite_val(p[d]);
  return ok;
}

template <typename D>
__device__ __forceinline__ D wmin(D x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { const D y = __shfl_xor(x, o); x = y < x ? y : x; }
  return x;
}
template <typename D>
__device__ __forceinline__ D wmax(D x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { const D y = __shfl_xor(x, o); x = y > x ? y : x; }
  return x;
}
__device__ __forceinline__ uint32_t wsum(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__global__ void reset_kernel(int32_t* __restrict__ lev_s, int32_t* __restrict__ lev_e,
                             uint32_t* __restrict__ done, int nlev, BatchState* __restrict__ st) {
  for (int i = threadIdx.x; i <= nlev; i += blockDim.x) {
    lev_s[i] = 0;
    lev_e[i] = 0;
    if (i < nlev) done[i] = 0u;
  }
  if (threadIdx.x == 0) {
    st->counter = 0;
    st->overflow = 0;
    st->nfb = 0ull;
  }
}

// Valid-row count and box of one tile (rows of one cloud).
template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void tile_reduce_kernel(const D* __restrict__ pts,
                                                             const TileRec* __restrict__ tiles,
                                                             uint32_t* __restrict__ tcnt,
                                                             D* __restrict__ tlo,
                                                             D* __restrict__ thi) {
  const TileRec tr = tiles[blockIdx.x];
  uint32_t cnt = 0;
  D lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = pos_inf<D>(); hi[d] = -pos_inf<D>(); }
  for (int64_t i = tr.start + threadIdx.x; i < tr.end; i += kBlock) {
    D v[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[d] = pts[i * DIM + d];
    if (row_valid(v, DIM)) {
      ++cnt;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        lo[d] = v[d] < lo[d] ? v[d] : lo[d];
        hi[d] = v[d] > hi[d] ? v[d] : hi[d];
      }
    }
  }
  cnt = wsum(cnt);
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = wmin(lo[d]); hi[d] = wmax(hi[d]); }
  __shared__ uint32_t s_cnt[kBlock / 64];
  __shared__ D s_lo[kBlock / 64][DIM], s_hi[kBlock / 64][DIM];
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
    tcnt[blockIdx.x] = c;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      tlo[(int64_t)blockIdx.x * DIM + d] = l[d];
      thi[(int64_t)blockIdx.x * DIM + d] = h[d];
    }
  }
}

// One wave per cloud: valid count and cubic Morton frame from the cloud's
// tiles (the same frame rule as the single-tree build: one scale for every
// axis, maxq / widest extent).
template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void cloud_frame_kernel(const int32_t* __restrict__ tbeg,
                                                             const uint32_t* __restrict__ tcnt,
                                                             const D* __restrict__ tlo,
                                                             const D* __restrict__ thi,
                                                             int64_t C, int mb,
                                                             CloudFrame<D>* __restrict__ frame,
                                                             uint32_t* __restrict__ vcnt) {
  const int64_t c = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= C) return;
  uint32_t cnt = 0;
  D lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = pos_inf<D>(); hi[d] = -pos_inf<D>(); }
  for (int32_t t = tbeg[c] + lane; t < tbeg[c + 1]; t += 64) {
    cnt += tcnt[t];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const D l = tlo[(int64_t)t * DIM + d], h = thi[(int64_t)t * DIM + d];
      lo[d] = l < lo[d] ? l : lo[d];
      hi[d] = h > hi[d] ? h : hi[d];
    }
  }
  cnt = wsum(cnt);
#pragma unroll
  for (int d = 0; d < DIM; ++d) { lo[d] = wmin(lo[d]); hi[d] = wmax(hi[d]); }
  if (lane == 0) {
    CloudFrame<D> f;
    D ext = D(0);
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      f.lo[d] = d < DIM ? lo[d] : D(0);
      if (d < DIM && cnt > 0) {
        const D wdt = hi[d] - lo[d];
        ext = wdt > ext ? wdt : ext;
      }
    }
    const D maxq = (D)((1u << mb) - 1u);
    f.scale = (cnt > 0 && ext > D(0) && finite_val(ext)) ? maxq / ext : D(0);
    frame[c] = f;
    vcnt[c] = cnt;
  }
}

template <int DIM>
__device__ __forceinline__ uint64_t interleave(const uint32_t* u, int mb) {
  if constexpr (DIM == 3) {
    return spread3_21(u[0]) | (spread3_21(u[1]) << 1) | (spread3_21(u[2]) << 2);
  } else if constexpr (DIM == 2) {
    return spread2_31(u[0]) | (spread2_31(u[1]) << 1);
  } else {
    uint64_t code = 0;
    for (int b = 0; b < mb; ++b)
#pragma unroll
      for (int d = 0; d < DIM; ++d) code |= (uint64_t)((u[d] >> b) & 1u) << (b * DIM + d);
    return code;
  }
}

// key[i] = cloud << shift | Morton code of row i in its cloud's frame;
// invalid rows get cloud id C. vals[i] = i.
template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void batch_morton_kernel(const D* __restrict__ pts,
                                                              const TileRec* __restrict__ tiles,
                                                              const CloudFrame<D>* __restrict__ frame,
                                                              int64_t C, int mb, int shift,
                                                              uint64_t* __restrict__ keys,
                                                              uint32_t* __restrict__ vals) {
  const TileRec tr = tiles[blockIdx.x];
  const CloudFrame<D> f = frame[tr.cloud];
  const D maxq = (D)((1u << mb) - 1u);
  for (int64_t i = tr.start + threadIdx.x; i < tr.end; i += kBlock) {
    D v[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[d] = pts[i * DIM + d];
    uint64_t key;
    if (row_valid(v, DIM)) {
      // maxq can round up in float (2^28 - 1 -> 2^28): clamp again as an
      // integer so no code spills into the cloud-id bits.
      const uint32_t umax = (1u << mb) - 1u;
      uint32_t u[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        D x = (v[d] - f.lo[d]) * f.scale;
        x = x > D(0) ? x : D(0);
        x = x < maxq ? x : maxq;
        const uint32_t ux = (uint32_t)x;
        u[d] = ux < umax ? ux : umax;
      }
      key = ((uint64_t)tr.cloud << shift) | interleave<DIM>(u, mb);
    } else {
      key = (uint64_t)C << shift;
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
  }
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap so that each XCD serves one contiguous range of logical blocks, i.e.
// one spatially coherent range of tree positions.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t g = gridDim.x, p = blockIdx.x;
  const int64_t g8 = g - (g & 7);
  return p < g8 ? (p & 7) * (g8 >> 3) + (p >> 3) : p;
}

// Tree-order point records. Positions >= num_valid hold the invalid rows.
// Each XCD serves a contiguous range of positions, i.e. of clouds, so the
// rows it gathers stay within a few clouds' input and hit its L2.
template <typename D, int DIM>
__global__ __launch_bounds__(kBlock) void batch_gather_kernel(const D* __restrict__ pts,
                                                              const uint64_t* __restrict__ keys,
                                                              const uint32_t* __restrict__ vals,
                                                              int64_t N, int shift,
                                                              const int64_t* __restrict__ coff,
                                                              const uint32_t* __restrict__ vstart,
                                                              int64_t C,
                                                              PointRec<D, DIM>* __restrict__ rec,
                                                              int32_t* __restrict__ gorig,
                                                              int32_t* __restrict__ fwd) {
  const int64_t p = xcd_block() * kBlock + threadIdx.x;
  if (p >= N) return;
  const uint32_t i = vals[p];
  gorig[p] = (int32_t)i;
  const bool valid = p < (int64_t)vstart[C];
  fwd[i] = valid ? (int32_t)p : -1;
  if (!valid) return;
  const int64_t c = (int64_t)(keys[p] >> shift);
  PointRec<D, DIM> r;
#pragma unroll
  for (int d = 0; d < DIM; ++d) r.c[d] = pts[(int64_t)i * DIM + d];
  r.orig = (int32_t)((int64_t)i - coff[c]);
  rec[p] = r;
}

// Cloud roots [0, C) and the placeholder root; level 0 = [0, R).
__global__ __launch_bounds__(kBlock) void roots_kernel(const uint32_t* __restrict__ vstart,
                                                       int64_t C, int64_t R,
                                                       NodeRec* __restrict__ nodes,
                                                       int32_t* __restrict__ lev_s,
                                                       int32_t* __restrict__ lev_e,
                                                       BatchState* __restrict__ st) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c < R) {
    NodeRec nr;
    nr.a = (int32_t)(c < C ? vstart[c] : vstart[C]);
    nr.b = (int32_t)(c < C ? vstart[c + 1] : vstart[C]);
    nr.child = -1;
    nr.parent = -1;
    nodes[c] = nr;
  }
  if (c == 0) {
    st->counter = (int32_t)R;
    lev_s[0] = 0;
    lev_e[0] = (int32_t)R;
  }
}

// Fused self-query kNN + PCA normals, one lane per tree position. The search
// finds the kk = max(k, nk) nearest other points of the cloud; the kNN row is
// the first k of them. The normals neighbourhood (nk nearest including the
// point itself, as estimate_normals searches) is that list with the point
// merged in at its (0, own index) rank, truncated to nk: the same set, summed
// in the same order, as a separate search would give. nk others (not nk - 1)
// are needed when nk or more points coincide with the point at lower indices:
// then the point itself is not among its own nk nearest.
template <typename D, int DIM, int K, int PREC, bool NORM>
// Float forests with K <= 16 hold 7 waves per SIMD (<= 72 VGPRs): at 64
// VGPRs the top-k merges spill, and without a bound the compiler's budget
// drifts with the code around the search.
__global__ __launch_bounds__(kBlock, (std::is_same<D, float>::value && K <= 16) ? 7 : 1) void batch_query_kernel(
                                                             TreeView<D, DIM> tv, int64_t N,
                                                             const uint32_t* __restrict__ vtotal,
                                                             const int32_t* __restrict__ gorig,
                                                             int k, int kk, int nk, D r2,
                                                             int wave_words,
                                                             int32_t* __restrict__ out_idx,
                                                             int32_t* __restrict__ out_cnt,
                                                             D* __restrict__ out_nrm,
                                                             D* __restrict__ out_eval,
                                                             unsigned long long* __restrict__ nfb) {
  const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
  // Float forests search with the whole wave (coop.hpp): lanes past the end
  // stay in the wave as inactive lanes.
  constexpr bool kCoop = std::is_same<D, float>::value;
  if (!kCoop && t >= N) return;
  const bool valid = t < N && t < (int64_t)*vtotal;
  const int64_t row = t < N ? gorig[t] : -1;
  // Positions are not carried through the top-k insertions: the normals
  // read each neighbour's tree position from the forward map (global row =
  // this point's global row - its cloud-local index + the neighbour's).
  TopK<D, K, false> st;
  st.init(kk);
  D q[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) q[d] = D(0);
  int32_t self = -1;
  if (valid) {
    const PointRec<D, DIM> r = tv.pts[t];
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[d] = r.c[d];
    self = r.orig;
  }
  uint64_t* wl = nullptr;
  if constexpr (kCoop) {
    extern __shared__ uint64_t coop_lds[];
    wl = coop_lds + (threadIdx.x >> 6) * wave_words;
    knn_coop_self<DIM, K>(tv, valid, (int32_t)t, self, q, r2, kk, true, st, wl);
    if (t >= N) return;
  } else {
    if (valid) knn_search_from_leaf<D, DIM>(tv, (int32_t)t, q, r2, (int32_t)t, kk, st);
  }
  if (out_idx) {
    int32_t c = 0;
    int32_t* orow = out_idx + row * k;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const int j = s - (K - kk);
      if (j >= 0 && j < k) {
        const bool f = st.found(s);
        orow[j] = f ? st.idx(s) : -1;
        c += f ? 1 : 0;
      }
    }
    if (out_cnt) out_cnt[row] = c;
  }
  if constexpr (NORM) {
    bool fb = false;
    const bool self_in = valid && D(0) < r2;
    int n = self_in ? 1 : 0;
#pragma unroll
    for (int s = 0; s < K; ++s) n += (s >= K - kk && st.found(s)) ? 1 : 0;
    n = n < nk ? n : nk;
    if (n == 0) {
      write_nan<D, 1>(row, out_nrm, out_eval, nullptr);
    } else {
      // neighbourhood positions in order: the point itself at its rank
      auto merged = [&](auto f) {
        bool pending = self_in;
        int taken = 0;
#pragma unroll
        for (int s = 0; s < K; ++s) {
          if (s >= K - kk && st.found(s)) {
            if (pending && key_less(D(0), self, st.dist(s), st.idx(s))) {
              if (taken < nk) f((int32_t)t);
              ++taken;
              pending = false;
            }
            if (taken < nk) f(tv.fwd[row - self + st.idx(s)]);
            ++taken;
          }
        }
        if (pending && taken < nk) f((int32_t)t);
      };
      if constexpr (kCoop) {
        // staged in this lane's LDS buffer (free after the search), so the
        // keys are dead before the covariance's passes
        int32_t* pl = coop_positions<DIM>(wl);
        int m = 0;
        merged([&](int32_t p) {
          pl[m * 64] = p;
          ++m;
        });
        auto each = [&](auto f) {
          for (int j = 0; j < m; ++j) f(pl[j * 64]);
        };
        fb = pca_row<D, PREC, 1>(tv, q, n, each, row, out_nrm, out_eval, nullptr);
      } else {
        fb = pca_row<D, PREC, 1>(tv, q, n, merged, row, out_nrm, out_eval, nullptr);
      }
    }
    count_fallbacks(fb, nfb);
  }
}

unsigned grid_of(int64_t n, int block) {
  const int64_t g = ceil_div(n, block);
  return (unsigned)(g > 0 ? g : 1);
}

int bits_for(int64_t v) {  // bits needed to represent v (v >= 1)
  int b = 0;
  while ((v >> b) != 0) ++b;
  return b;
}

template <typename D, int DIM>
void enqueue_run(pptk_batch* b, const D* pts, int32_t* out_idx, int32_t* out_cnt, D* out_nrm,
                 D* out_eval, hipStream_t s) {
  int launches = 0;
  auto lc = [&]() { PK_LAUNCH_CHECK(); ++launches; };
  BatchState* st = b->state.as<BatchState>();
  if (b->timing) PK_HIP(hipEventRecord(b->ev[0], s));
  reset_kernel<<<1, kBlock, 0, s>>>(b->lev_s.as<int32_t>(), b->lev_e.as<int32_t>(),
                                    b->done.as<uint32_t>(), b->nlev, st);
  lc();
  if (b->N > 0) {
    tile_reduce_kernel<D, DIM><<<(unsigned)b->ntiles, kBlock, 0, s>>>(
        pts, b->tiles.as<TileRec>(), b->tcnt.as<uint32_t>(), b->tlo.as<D>(), b->thi.as<D>());
    lc();
  }
  cloud_frame_kernel<D, DIM><<<grid_of(b->C, kBlock / 64), kBlock, 0, s>>>(
      b->tbeg.as<int32_t>(), b->tcnt.as<uint32_t>(), b->tlo.as<D>(), b->thi.as<D>(), b->C, b->mb,
      b->cinfo.as<CloudFrame<D>>(), b->vcnt.as<uint32_t>());
  lc();
  scan_u32_inplace_async(b->vcnt.as<uint32_t>(), b->C, s);
  lc();
  if (b->N > 0) {
    batch_morton_kernel<D, DIM><<<(unsigned)b->ntiles, kBlock, 0, s>>>(
        pts, b->tiles.as<TileRec>(), b->cinfo.as<CloudFrame<D>>(), b->C, b->mb, b->shift,
        b->keys.as<uint64_t>(), b->vals.as<uint32_t>());
    lc();
    radix_sort_pairs_async(b->keys.as<uint64_t>(), b->vals.as<uint32_t>(), b->N, b->end_bit, b->ws,
                           s);
    launches += 4 * ((b->end_bit + 7) / 8);  // histogram, two scan launches, scatter
    batch_gather_kernel<D, DIM><<<grid_of(b->N, kBlock), kBlock, 0, s>>>(
        pts, b->keys.as<uint64_t>(), b->vals.as<uint32_t>(), b->N, b->shift,
        b->coff.as<int64_t>(), b->vcnt.as<uint32_t>(), b->C, b->rec.as<PointRec<D, DIM>>(),
        b->gorig.as<int32_t>(), b->fwd.as<int32_t>());
    lc();
  }
  roots_kernel<<<grid_of(b->R, kBlock), kBlock, 0, s>>>(b->vcnt.as<uint32_t>(), b->C, b->R,
                                                       b->nodes.as<NodeRec>(),
                                                       b->lev_s.as<int32_t>(),
                                                       b->lev_e.as<int32_t>(), st);
  lc();
  for (int l = 0; l < b->nlev; ++l) {
    split_level_dev_kernel<<<kLevelGrid, kBlock, 0, s>>>(
        b->keys.as<uint64_t>(), b->nodes.as<NodeRec>(), b->lev_s.as<int32_t>(),
        b->lev_e.as<int32_t>(), b->done.as<uint32_t>(), l, b->nlev, b->leaf, b->cap, st);
    lc();
  }
  leaf_box_kernel<D, DIM><<<4 * kLevelGrid, kBlock, 0, s>>>(
      b->rec.as<PointRec<D, DIM>>(), b->nodes.as<NodeRec>(), b->box.as<BoxRec<D, DIM>>(),
      b->leafof.as<int32_t>(), 0, &st->counter);
  lc();
  for (int l = b->nlev - 1; l >= 0; --l) {
    box_level_dev_kernel<D, DIM><<<kLevelGrid, kBlock, 0, s>>>(
        b->nodes.as<NodeRec>(), b->box.as<BoxRec<D, DIM>>(), b->lev_s.as<int32_t>(),
        b->lev_e.as<int32_t>(), l);
    lc();
  }
  if (b->timing) PK_HIP(hipEventRecord(b->ev[1], s));
  if (b->N > 0 && (b->k > 0 || b->nk > 0)) {
    TreeView<D, DIM> tv;
    tv.pts = b->rec.as<const PointRec<D, DIM>>();
    tv.orig = b->gorig.as<const int32_t>();
    tv.fwd = b->fwd.as<const int32_t>();
    tv.box = b->box.as<const BoxRec<D, DIM>>();
    tv.node = b->nodes.as<const NodeRec>();
    tv.leafof = b->leafof.as<const int32_t>();
    tv.n = (int32_t)b->N;
    tv.leaf = b->leaf;
    tv.nn = (int32_t)b->cap;
    tv.n_input = b->N;
    tv.n_dev = (const int32_t*)(b->vcnt.as<uint32_t>() + b->C);
    tv.nroots = (int32_t)b->R;
    D r2 = pos_inf<D>();
    if (!std::isinf(b->dmax)) {
      const D r = (D)b->dmax;
      r2 = r * r;
    }
    const unsigned grid = grid_of(b->N, b->qblock);
    unsigned long long* nfb = &st->nfb;
    const uint32_t* vtotal = b->vcnt.as<uint32_t>() + b->C;
    const int32_t* gorig = b->gorig.as<int32_t>();
    auto go = [&](auto kc, auto pc, auto nc) {
      constexpr int K = decltype(kc)::value;
      constexpr int P = decltype(pc)::value;
      constexpr bool NM = decltype(nc)::value;
      // per wave: the leaf chunk, then the key buffers or, after the
      // search, the nk neighbour positions of each lane
      const int pos_bytes = (std::max(b->nk, 0) * 64 * 4 + 15) / 16 * 16;
      const int wave_bytes = coop_chunk_bytes<DIM>() +
                             std::max(coop_buf_bytes<K, 0>(), pos_bytes);
      const size_t lds = std::is_same<D, float>::value
                             ? (size_t)(b->qblock / 64) * wave_bytes : 0;
      batch_query_kernel<D, DIM, K, P, NM><<<grid, b->qblock, lds, s>>>(
          tv, b->N, vtotal, gorig, b->k, b->kk, b->nk, r2, wave_bytes / 8, out_idx, out_cnt,
          out_nrm, out_eval, nfb);
    };
    auto by_k = [&](auto pc, auto nc) {
      if (b->kk <= 4) go(std::integral_constant<int, 4>{}, pc, nc);
      else if (b->kk <= 8) go(std::integral_constant<int, 8>{}, pc, nc);
      else if (b->kk <= 16) go(std::integral_constant<int, 16>{}, pc, nc);
      else go(std::integral_constant<int, 32>{}, pc, nc);
    };
    using Mixed = std::integral_constant<int, PPTK_PREC_MIXED>;
    using F64 = std::integral_constant<int, PPTK_PREC_FP64>;
    if constexpr (DIM == 3) {
      if (b->nk > 0) {
        if (b->prec == PPTK_PREC_FP64) by_k(F64{}, std::true_type{});
        else by_k(Mixed{}, std::true_type{});
      } else {
        by_k(Mixed{}, std::false_type{});
      }
    } else {
      by_k(Mixed{}, std::false_type{});
    }
    lc();
  }
  if (b->timing) PK_HIP(hipEventRecord(b->ev[2], s));
  b->launches = launches;
}

template <typename F>
void dispatch_batch(const pptk_batch* b, F&& f) {
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  const bool dbl = b->dtype == PPTK_F64;
  switch (b->dim) {
    case 2: if (dbl) f(Tag<double>{}, I2{}); else f(Tag<float>{}, I2{}); break;
    case 3: if (dbl) f(Tag<double>{}, I3{}); else f(Tag<float>{}, I3{}); break;
    default: if (dbl) f(Tag<double>{}, I4{}); else f(Tag<float>{}, I4{}); break;
  }
}

std::mutex g_batch_mu;
std::set<const void*> g_batches;

pptk_batch* check_batch(const pptk_batch* b) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  if (!b || !g_batches.count(b) || b->magic != pptk_batch::kMagic)
    fail(PPTK_E_TYPE, "object is not a live batch plan");
  return const_cast<pptk_batch*>(b);
}

void drop_graph(pptk_batch* b) {
  // teardown: a failure to free leaves nothing to recover
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (auto& k : b->gkey) k = nullptr;
}

}  // namespace

int default_leaf_size(int dim, bool dbl, int64_t n);

pptk_status api_batch_create(const int64_t* sizes, int64_t C, int dim, pptk_dtype dtype,
                             const pptk_batch_params* prm, pptk_batch** out) {
  return guard([&] {
    if (!out) fail(PPTK_E_VALUE, "null output handle");
    *out = nullptr;
    if (C < 1) fail(PPTK_E_VALUE, "a batch needs at least one cloud");
    if (!sizes) fail(PPTK_E_VALUE, "null cloud sizes");
    if (dim < 2 || dim > 4) fail(PPTK_E_VALUE, "dim must be 2, 3 or 4 (got %d)", dim);
    if (dtype != PPTK_F32 && dtype != PPTK_F64)
      fail(PPTK_E_TYPE, "batched clouds take float32 or float64 points");
    if (!prm) fail(PPTK_E_VALUE, "null batch parameters");
    if (prm->k < 0 || prm->k > 32) fail(PPTK_E_VALUE, "batch k must be in [0, 32]");
    if (prm->normals_k < 0 || prm->normals_k > 32)
      fail(PPTK_E_VALUE, "batch normals_k must be in [0, 32]");
    if (prm->normals_k > 0 && dim != 3) fail(PPTK_E_VALUE, "normals need 3-D points");
    if (!(prm->dmax >= 0.0)) fail(PPTK_E_VALUE, "dmax must be non-negative");
    if (prm->precision < PPTK_PREC_AUTO || prm->precision > PPTK_PREC_FP64)
      fail(PPTK_E_VALUE, "unknown precision mode %d", prm->precision);
    const int qb = prm->query_block == 0 ? 256 : prm->query_block;
    if (qb != 64 && qb != 128 && qb != 256) fail(PPTK_E_VALUE, "query_block must be 64, 128 or 256");
    if (C > (1LL << 24)) fail(PPTK_E_VALUE, "at most 2^24 clouds per batch");
    std::vector<int64_t> off((size_t)C + 1, 0);
    for (int64_t c = 0; c < C; ++c) {
      if (sizes[c] < 0) fail(PPTK_E_VALUE, "negative cloud size");
      off[c + 1] = off[c] + sizes[c];
      if (off[c + 1] > (1LL << 30)) fail(PPTK_E_VALUE, "at most 2^30 points per batch");
    }
    ensure_device();
    std::unique_ptr<pptk_batch> b(new pptk_batch());
    b->dim = dim;
    b->dtype = dtype;
    b->C = C;
    b->N = off[C];
    b->R = C | 1;
    b->cap = b->R + 2 * b->N;
    if (b->cap > (int64_t)INT32_MAX) fail(PPTK_E_VALUE, "too many points for int32 node ids");
    const int leaf = prm->max_leaf_size > 0 ? prm->max_leaf_size
                                            : default_leaf_size(dim, dtype == PPTK_F64, b->N);
    b->leaf = std::min(256, std::max(1, leaf));
    const int cbits = bits_for(C);  // room for ids 0..C (C = invalid rows)
    const int mbmax = dim == 2 ? 31 : dim == 3 ? 21 : 15;
    // Morton bits per axis: enough to give the largest cloud ~32x more cells
    // per axis than its points need on a surface (ceil(log2 n / dim) + 5), so
    // equal codes stay rare; fewer bits mean fewer radix passes and levels.
    // Results do not depend on it (the searches are exact for any tree).
    int64_t maxn = 1;
    for (int64_t c = 0; c < C; ++c) maxn = std::max<int64_t>(maxn, off[c + 1] - off[c]);
    const int need = (bits_for(maxn) + dim - 1) / dim + 5;
    b->mb = std::min({mbmax, (63 - cbits) / dim, need});
    b->shift = dim * b->mb;
    b->end_bit = b->shift + cbits;
    // Levels: along any root-to-leaf path a split either lowers the highest
    // differing key bit (at most `shift` such splits inside one cloud) or
    // halves a node whose keys are all equal, after which only such halvings
    // follow; ceil(log2(maxn / leaf)) of them bring any run to <= leaf. So
    // the depth is at most 1 + shift + that, and no launch goes past it.
    int halvings = 0;
    while (((int64_t)b->leaf << halvings) < maxn) ++halvings;
    b->nlev = b->shift + halvings + 2;
    b->k = prm->k;
    b->nk = prm->normals_k;
    b->kk = std::max(b->k, b->nk);
    b->dmax = prm->dmax;
    b->prec = prm->precision == PPTK_PREC_AUTO
                  ? (dtype == PPTK_F64 ? PPTK_PREC_FP64 : PPTK_PREC_MIXED)
                  : prm->precision;
    b->timing = prm->timing != 0;
    b->use_graph = prm->use_graph != 0 && !b->timing;
    b->qblock = qb;
    if (b->timing) {
      ensure_device();
      for (auto& e : b->ev) PK_HIP(hipEventCreate(&e));
    }

    std::vector<TileRec> tiles;
    std::vector<int32_t> tbeg((size_t)C + 1, 0);
    for (int64_t c = 0; c < C; ++c) {
      tbeg[c] = (int32_t)tiles.size();
      for (int64_t a = off[c]; a < off[c + 1]; a += kTile) {
        TileRec t;
        t.start = a;
        t.end = std::min(off[c + 1], a + (int64_t)kTile);
        t.cloud = (int32_t)c;
        t.pad = 0;
        tiles.push_back(t);
      }
    }
    tbeg[C] = (int32_t)tiles.size();
    b->ntiles = (int64_t)tiles.size();
    const size_t dsz = dtype == PPTK_F64 ? 8 : 4;
    const int64_t N1 = std::max<int64_t>(b->N, 1);
    const int64_t T1 = std::max<int64_t>(b->ntiles, 1);
    b->coff.reset((size_t)(C + 1) * 8);
    b->tiles.reset((size_t)T1 * sizeof(TileRec));
    b->tbeg.reset((size_t)(C + 1) * 4);
    b->tcnt.reset((size_t)T1 * 4);
    b->tlo.reset((size_t)T1 * dim * dsz);
    b->thi.reset((size_t)T1 * dim * dsz);
    b->cinfo.reset((size_t)C * (dtype == PPTK_F64 ? sizeof(CloudFrame<double>)
                                                  : sizeof(CloudFrame<float>)));
    b->vcnt.reset((size_t)(C + 1) * 4);
    b->keys.reset((size_t)N1 * 8);
    b->vals.reset((size_t)N1 * 4);
    b->ws.reserve(N1);
    size_t rec_sz = 0, box_sz = 0;
    dispatch_batch(b.get(), [&](auto dt, auto dc) {
      using D = typename decltype(dt)::type;
      constexpr int DIM = decltype(dc)::value;
      rec_sz = sizeof(PointRec<D, DIM>);
      box_sz = sizeof(BoxRec<D, DIM>);
    });
    b->rec.reset((size_t)N1 * rec_sz);
    b->gorig.reset((size_t)N1 * 4);
    b->fwd.reset((size_t)N1 * 4);
    b->leafof.reset((size_t)N1 * 4);
    b->nodes.reset((size_t)b->cap * sizeof(NodeRec));
    b->box.reset((size_t)b->cap * box_sz);
    b->lev_s.reset((size_t)(b->nlev + 1) * 4);
    b->lev_e.reset((size_t)(b->nlev + 1) * 4);
    b->done.reset((size_t)b->nlev * 4);
    b->state.reset(sizeof(BatchState));
    hipStream_t s = lib_stream();
    PK_HIP(hipMemcpyAsync(b->coff.get(), off.data(), (size_t)(C + 1) * 8, hipMemcpyHostToDevice, s));
    if (!tiles.empty())
      PK_HIP(hipMemcpyAsync(b->tiles.get(), tiles.data(), tiles.size() * sizeof(TileRec),
                            hipMemcpyHostToDevice, s));
    PK_HIP(hipMemcpyAsync(b->tbeg.get(), tbeg.data(), (size_t)(C + 1) * 4, hipMemcpyHostToDevice, s));
    PK_HIP(hipStreamSynchronize(s));
    {
      std::lock_guard<std::mutex> lk(g_batch_mu);
      g_batches.insert(b.get());
    }
    *out = b.release();
  });
}

pptk_status api_batch_run(pptk_batch* bp, const void* pts, int32_t* out_idx, int32_t* out_cnt,
                          void* out_nrm, void* out_eval, void* stream) {
  return guard([&] {
    pptk_batch* b = check_batch(bp);
    if (b->N > 0 && !pts) fail(PPTK_E_VALUE, "null points");
    if (b->k > 0 && b->N > 0 && !out_idx) fail(PPTK_E_VALUE, "null kNN output");
    if (b->nk > 0 && b->N > 0 && !out_nrm) fail(PPTK_E_VALUE, "null normals output");
    if (b->k == 0) { out_idx = nullptr; out_cnt = nullptr; }
    if (b->nk == 0) { out_nrm = nullptr; out_eval = nullptr; }
    hipStream_t s = pick_stream(stream);
    auto enqueue = [&]() {
      dispatch_batch(b, [&](auto dt, auto dc) {
        using D = typename decltype(dt)::type;
        constexpr int DIM = decltype(dc)::value;
        enqueue_run<D, DIM>(b, (const D*)pts, out_idx, out_cnt, (D*)out_nrm, (D*)out_eval, s);
      });
    };
    if (b->use_graph) {
      const void* key[5] = {pts, out_idx, out_cnt, out_nrm, out_eval};
      if (b->exec && !std::equal(key, key + 5, b->gkey)) drop_graph(b);
      if (!b->exec) {
        PK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          enqueue();
        } catch (...) {
          // end the capture before rethrowing; its own errors would hide
          // the one being reported
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        PK_HIP(hipStreamEndCapture(s, &b->graph));
        PK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        std::copy(key, key + 5, b->gkey);
      }
      PK_HIP(hipGraphLaunch(b->exec, s));
    } else {
      enqueue();
    }
    BatchState hs;
    uint32_t nv = 0;
    PK_HIP(hipMemcpyAsync(&hs, b->state.get(), sizeof(hs), hipMemcpyDeviceToHost, s));
    PK_HIP(hipMemcpyAsync(&nv, b->vcnt.as<uint32_t>() + b->C, 4, hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    if (hs.overflow)
      fail(PPTK_E_RUNTIME, "batch build exceeded its node or level bound (internal error)");
    b->nv = nv;
    b->nn = hs.counter;
    b->nfb = (int64_t)hs.nfb;
    b->ran = true;
    if (b->timing) {
      float a = 0.f, q = 0.f;
      PK_HIP(hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
      PK_HIP(hipEventElapsedTime(&q, b->ev[1], b->ev[2]));
      b->build_ms = a;
      b->query_ms = q;
    }
  });
}

pptk_status api_batch_get_info(const pptk_batch* bp, pptk_batch_info* out) {
  return guard([&] {
    pptk_batch* b = check_batch(bp);
    if (!out) fail(PPTK_E_VALUE, "null output");
    out->num_clouds = b->C;
    out->num_points = b->N;
    out->num_valid = b->nv;
    out->num_nodes = b->nn;
    out->leaf_size = b->leaf;
    out->morton_bits = b->mb;
    out->graph_nodes = b->launches;
    out->num_fp64 = b->nfb;
    out->build_ms = b->build_ms;
    out->query_ms = b->query_ms;
    out->depth = 0;
    if (b->ran) {
      std::vector<int32_t> ls(b->nlev + 1), le(b->nlev + 1);
      hipStream_t s = lib_stream();
      PK_HIP(hipMemcpyAsync(ls.data(), b->lev_s.get(), ls.size() * 4, hipMemcpyDeviceToHost, s));
      PK_HIP(hipMemcpyAsync(le.data(), b->lev_e.get(), le.size() * 4, hipMemcpyDeviceToHost, s));
      PK_HIP(hipStreamSynchronize(s));
      for (int l = 0; l <= b->nlev; ++l)
        if (le[l] > ls[l]) out->depth = l + 1;
    }
    out->device_bytes = (int64_t)(b->coff.bytes() + b->tiles.bytes() + b->tbeg.bytes() +
                                  b->tcnt.bytes() + b->tlo.bytes() + b->thi.bytes() +
                                  b->cinfo.bytes() + b->vcnt.bytes() + b->keys.bytes() +
                                  b->vals.bytes() + b->ws.bytes() +
                                  b->rec.bytes() + b->gorig.bytes() + b->fwd.bytes() + b->leafof.bytes() +
                                  b->nodes.bytes() + b->box.bytes() + b->lev_s.bytes() +
                                  b->lev_e.bytes() + b->done.bytes() + b->state.bytes());
  });
}

void api_batch_free(pptk_batch* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    if (!g_batches.erase(b)) return;
  }
  drop_graph(b);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  b->magic = 0;
  delete b;
}

}  // namespace pk

extern "C" pptk_status pptk_batch_create(const int64_t* cloud_sizes, int64_t num_clouds, int dim,
                                         pptk_dtype dtype, const pptk_batch_params* params,
                                         pptk_batch** out) {
  return pk::api_batch_create(cloud_sizes, num_clouds, dim, dtype, params, out);
}

extern "C" pptk_status pptk_batch_run(pptk_batch* b, const void* pts, int32_t* out_idx,
                                      int32_t* out_cnt, void* out_nrm, void* out_eval,
                                      void* stream) {
  return pk::api_batch_run(b, pts, out_idx, out_cnt, out_nrm, out_eval, stream);
}

extern "C" pptk_status pptk_batch_get_info(const pptk_batch* b, pptk_batch_info* out) {
  return pk::api_batch_get_info(b, out);
}

extern "C" void pptk_batch_free(pptk_batch* b) { pk::api_batch_free(b); }
