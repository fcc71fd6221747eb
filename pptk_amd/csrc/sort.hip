// LSD radix sort of (uint64 key, uint32 value) pairs and exclusive prefix
// sums, written for 64-wide wavefronts.
//
// Radix sort, one 8-bit digit per pass:
//   1. radix_hist_kernel    per-tile digit histogram, stored digit-major
//                           (hist[digit * tiles + tile]);
//   2. exclusive scan of the histogram (scan_reduce_kernel: per-chunk sums;
//      scan_down_kernel: each chunk's scan plus the sum of the chunks
//      before it) gives every (digit, tile) its global output offset;
//   3. radix_scatter_kernel ranks keys stably inside the tile with wave
//      ballots (peer mask of equal digits) and per-wave LDS counters, stages
//      the tile in LDS in digit order, and writes each digit's run of the
//      tile contiguously (coalesced, instead of one scattered address per
//      lane).
// A tile is 4096 keys: 4 waves x 16 rows x 64 lanes. Wave w owns the 1024
// consecutive keys [w*1024, w*1024+1024) of the tile and reads them row by
// row, so (wave, row, lane) order equals input order and the sort is stable.
#include "common.hpp"
#include "sort.hpp"

namespace pk {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;  // 4096
constexpr int kRadix = 256;
constexpr int kWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void radix_hist_kernel(const uint64_t* __restrict__ keys,
                                                            int64_t n, int shift,
                                                            uint32_t* __restrict__ hist,
                                                            int tiles) {
  __shared__ uint32_t cnt[kWaves][kRadix];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  for (int i = tid; i < kWaves * kRadix; i += kBlock) (&cnt[0][0])[i] = 0u;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  // every key of the tile requested before the first count (kRadix: none)
  uint32_t dg[kItems];
  if (base + kTile <= n) {
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      dg[j] = (uint32_t)(keys[base + (int64_t)j * kBlock + tid] >> shift) & 255u;
  } else {
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int64_t idx = base + (int64_t)j * kBlock + tid;
      dg[j] = idx < n ? (uint32_t)(keys[idx] >> shift) & 255u : (uint32_t)kRadix;
    }
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j)
    if (dg[j] < (uint32_t)kRadix) atomicAdd(&cnt[w][dg[j]], 1u);
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int ww = 0; ww < kWaves; ++ww) s += cnt[ww][tid];
  hist[(int64_t)tid * tiles + blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void radix_scatter_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n, int shift,
    const uint32_t* __restrict__ offs, int tiles) {
  __shared__ uint32_t cnt[kWaves][kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t dstart[kRadix];
  __shared__ uint32_t wsum[kWaves];
  __shared__ uint64_t skey[kTile];
  __shared__ uint32_t sval[kTile];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  for (int i = tid; i < kWaves * kRadix; i += kBlock) (&cnt[0][0])[i] = 0u;
  // The digit offsets are read before the keys. Read after them, the scatter
  // took the same time at 10M keys but 1.23 -> 1.48 ms per pass at 153M keys
  // (37.5K tiles: a block's 256 offset reads lie 150 KB apart).
  gbase[tid] = offs[(int64_t)tid * tiles + blockIdx.x];
  __syncthreads();

  const int64_t wbase = (int64_t)blockIdx.x * kTile + (int64_t)w * (kItems * 64);
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  uint64_t key[kItems];
  uint32_t val[kItems];
  uint32_t rank[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t idx = wbase + (int64_t)j * 64 + lane;
    const bool ok = idx < n;
    key[j] = ok ? kin[idx] : 0ull;
    val[j] = ok ? vin[idx] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t idx = wbase + (int64_t)j * 64 + lane;
    const bool ok = idx < n;
    const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
    // Lanes holding a valid key with the same digit as this lane.
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
    const uint32_t total = (uint32_t)__popcll(peers);
    const uint32_t c0 = cnt[w][d];
    // The lowest lane of each digit group advances the wave's counter; LDS
    // operations of one wave complete in program order, so the next row
    // reads the updated value.
    if (ok && before == 0u) cnt[w][d] = c0 + total;
    rank[j] = c0 + before;
  }
  __syncthreads();
  uint32_t total;
  {
    // Exclusive prefix over waves for digit `tid` (kBlock == kRadix).
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      const uint32_t c = cnt[ww][tid];
      cnt[ww][tid] = run;
      run += c;
    }
    total = run;  // keys of digit `tid` in this tile
  }
  // Tile-local start of each digit: exclusive prefix of the totals.
  {
    uint32_t incl = total;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww)
      if (ww < w) off += wsum[ww];
    dstart[tid] = off + incl - total;
  }
  __syncthreads();
  // Stage the tile in digit order, then write each digit's run contiguously.
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t idx = wbase + (int64_t)j * 64 + lane;
    if (idx < n) {
      const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
      const uint32_t lp = dstart[d] + cnt[w][d] + rank[j];
      skey[lp] = key[j];
      sval[lp] = val[j];
    }
  }
  __syncthreads();
  const int64_t tbase = (int64_t)blockIdx.x * kTile;
  const int32_t nt = (int32_t)(n - tbase < kTile ? n - tbase : kTile);
  for (int32_t i = tid; i < nt; i += kBlock) {
    const uint64_t k = skey[i];
    const uint32_t d = (uint32_t)(k >> shift) & 255u;
    const uint32_t pos = gbase[d] + (uint32_t)i - dstart[d];
    kout[pos] = k;
    vout[pos] = sval[i];
  }
}

// ------------------------------------------------------------------ scan ----

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;  // 4096

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Block-wide exclusive scan of one value per thread (kScanBlock threads).
// Returns the exclusive prefix; *total receives the block sum.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* total) {
  __shared__ T wsum[kScanBlock / 64];
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  T incl = wave_incl_scan(v);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanBlock / 64; ++i) {
    const T s = wsum[i];
    if (i < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Per-tile sums of `in` (tile = 4096 elements).
template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const Tin* __restrict__ in,
                                                                 int64_t n,
                                                                 Tout* __restrict__ tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  Tout s = 0;
  if (base + kScanTile <= n) {
    // a full tile: every load issued before the first add
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s += (Tout)in[base + (int64_t)j * kScanBlock + threadIdx.x];
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      const int64_t i = base + (int64_t)j * kScanBlock + threadIdx.x;
      if (i < n) s += (Tout)in[i];
    }
  }
  s = wave_sum(s);
  __shared__ Tout ws[kScanBlock / 64];
  if (lane_id() == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Tout t = 0;
#pragma unroll
    for (int i = 0; i < kScanBlock / 64; ++i) t += ws[i];
    tile_sums[blockIdx.x] = t;
  }
}

// In-place exclusive scan of `n` values by one block, tile by tile with a
// running carry. data has room for n+1 values; data[n] receives the total.
template <typename T>
__global__ __launch_bounds__(kScanBlock) void scan_single_kernel(T* __restrict__ data, int64_t n) {
  T carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kScanTile) {
    const int64_t i0 = t0 + (int64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T local = 0;
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) {
      v[e] = (i0 + e < n) ? data[i0 + e] : (T)0;
      local += v[e];
    }
    T total;
    T run = carry + block_excl_scan(local, &total);
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) {
      if (i0 + e < n) data[i0 + e] = run;
      run += v[e];
    }
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) data[n] = carry;
}

// Sum of one value per thread over the block (kScanBlock threads).
template <typename T>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T ws[kScanBlock / 64];
  v = wave_sum(v);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  T t = 0;
#pragma unroll
  for (int i = 0; i < kScanBlock / 64; ++i) t += ws[i];
  __syncthreads();
  return t;
}

// Exclusive scan of each tile plus its tile offset, which the block sums
// from the earlier tiles' sums itself (no separate pass over the sums).
// Each thread owns 16 consecutive elements of the tile. With `total`, the
// last block writes the sum of all n values there.
template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanBlock) void scan_down_kernel(const Tin* __restrict__ in,
                                                               Tout* __restrict__ out, int64_t n,
                                                               const Tout* __restrict__ tile_sums,
                                                               Tout* __restrict__ total) {
  Tout before = 0;
  for (int64_t j = threadIdx.x; j < (int64_t)blockIdx.x; j += kScanBlock) before += tile_sums[j];
  const Tout tile_off = block_sum(before);
  const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  Tout v[kScanItems];
  Tout local = 0;
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    v[e] = (i0 + e < n) ? (Tout)in[i0 + e] : (Tout)0;
    local += v[e];
  }
  Tout tile_total;
  Tout run = tile_off + block_excl_scan(local, &tile_total);
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    if (i0 + e < n) out[i0 + e] = run;
    run += v[e];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = tile_off + tile_total;
}

}  // namespace

template <typename Tin, typename Tout>
void exclusive_scan(const Tin* in, Tout* out, int64_t n, hipStream_t s) {
  // out has n+1 entries; out[n] = total.
  if (n <= 0) {
    PK_HIP(hipMemsetAsync(out, 0, sizeof(Tout), s));
    return;
  }
  const int64_t tiles = ceil_div(n, kScanTile);
  DevBuf sums((size_t)tiles * sizeof(Tout));
  Tout* ts = sums.as<Tout>();
  scan_reduce_kernel<Tin, Tout><<<dim3((unsigned)tiles), dim3(kScanBlock), 0, s>>>(in, n, ts);
  PK_LAUNCH_CHECK();
  scan_down_kernel<Tin, Tout><<<dim3((unsigned)tiles), dim3(kScanBlock), 0, s>>>(in, out, n, ts,
                                                                                 out + n);
  PK_LAUNCH_CHECK();
  // `sums` returns to the pool fenced on s (no other stream gets it before
  // this work is done), so no synchronisation here
}

template void exclusive_scan<int32_t, int64_t>(const int32_t*, int64_t*, int64_t, hipStream_t);
template void exclusive_scan<uint32_t, uint32_t>(const uint32_t*, uint32_t*, int64_t, hipStream_t);

size_t SortWorkspace::bytes_for(int64_t n) {
  const int64_t tiles = ceil_div(std::max<int64_t>(n, 1), kTile);
  const int64_t hlen = (int64_t)kRadix * tiles;
  const int64_t scan_tiles = ceil_div(hlen, kScanTile);
  auto a = [](int64_t b) { return (size_t)((b + 255) / 256 * 256); };
  return a(std::max<int64_t>(n, 1) * 8) + a(std::max<int64_t>(n, 1) * 4) + a(hlen * 4) +
         a((hlen + 1) * 4) + a((scan_tiles + 1) * 4);
}

void SortWorkspace::bind(void* base, int64_t n) {
  const int64_t m = std::max<int64_t>(n, 1);
  const int64_t tiles = ceil_div(m, kTile);
  const int64_t hlen = (int64_t)kRadix * tiles;
  auto a = [](int64_t b) { return (size_t)((b + 255) / 256 * 256); };
  uint8_t* p = (uint8_t*)base;
  kalt = (uint64_t*)p; p += a(m * 8);
  valt = (uint32_t*)p; p += a(m * 4);
  hist = (uint32_t*)p; p += a(hlen * 4);
  offs = (uint32_t*)p; p += a((hlen + 1) * 4);
  ssum = (uint32_t*)p;
  n_cap = m;
}

void SortWorkspace::reserve(int64_t n) {
  if (n <= n_cap) return;
  DevBuf buf(bytes_for(n));  // pool blocks are 256-byte aligned
  bind(buf.get(), n);
  owned_ = std::move(buf);
}

void radix_sort_pairs_async(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit,
                            SortWorkspace& ws, hipStream_t s, bool keys_out) {
  if (n <= 1 || end_bit <= 0) return;
  if (n > (int64_t)0x7fffffff) fail(PPTK_E_VALUE, "radix sort: n=%lld too large", (long long)n);
  if (n > ws.n_cap) fail(PPTK_E_RUNTIME, "radix sort workspace too small");
  const int64_t tiles = ceil_div(n, kTile);
  const int64_t hlen = (int64_t)kRadix * tiles;
  const int64_t scan_tiles = ceil_div(hlen, kScanTile);
  uint64_t* k0 = keys;
  uint32_t* v0 = vals;
  uint64_t* k1 = ws.kalt;
  uint32_t* v1 = ws.valt;
  uint32_t* hist = ws.hist;
  uint32_t* offs = ws.offs;
  uint32_t* ts = ws.ssum;
  int passes = 0;
  for (int shift = 0; shift < end_bit; shift += 8, ++passes) {
    radix_hist_kernel<<<dim3((unsigned)tiles), dim3(kBlock), 0, s>>>(k0, n, shift, hist, (int)tiles);
    PK_LAUNCH_CHECK();
    scan_reduce_kernel<uint32_t, uint32_t>
        <<<dim3((unsigned)scan_tiles), dim3(kScanBlock), 0, s>>>(hist, hlen, ts);
    PK_LAUNCH_CHECK();
    scan_down_kernel<uint32_t, uint32_t><<<dim3((unsigned)scan_tiles), dim3(kScanBlock), 0, s>>>(
        hist, offs, hlen, ts, nullptr);
    PK_LAUNCH_CHECK();
    radix_scatter_kernel<<<dim3((unsigned)tiles), dim3(kBlock), 0, s>>>(k0, v0, k1, v1, n, shift,
                                                                        offs, (int)tiles);
    PK_LAUNCH_CHECK();
    uint64_t* tk = k0; k0 = k1; k1 = tk;
    uint32_t* tv = v0; v0 = v1; v1 = tv;
  }
  if (passes & 1) {
    if (keys_out)
      PK_HIP(hipMemcpyAsync(keys, k0, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    PK_HIP(hipMemcpyAsync(vals, v0, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, hipStream_t s) {
  if (n <= 1 || end_bit <= 0) return;
  SortWorkspace ws;
  ws.reserve(n);
  radix_sort_pairs_async(keys, vals, n, end_bit, ws, s);
  // the workspace returns to the pool fenced on s
}

void scan_u32_inplace_async(uint32_t* data, int64_t n, hipStream_t s) {
  scan_single_kernel<uint32_t><<<1, kScanBlock, 0, s>>>(data, n);
  PK_LAUNCH_CHECK();
}

}  // namespace pk

extern "C" pptk_status pptk_sort_pairs_u64(uint64_t* keys, uint32_t* vals, int64_t n,
                                           int end_bit, pptk_mem mem) {
  return pk::guard([&] {
    if (n < 0) pk::fail(PPTK_E_VALUE, "n must be non-negative");
    if (end_bit < 0 || end_bit > 64) pk::fail(PPTK_E_VALUE, "end_bit must be in [0, 64]");
    if (n == 0) return;
    if (!keys || !vals) pk::fail(PPTK_E_VALUE, "null buffer");
    hipStream_t s = pk::lib_stream();
    pk::DevInput kin, vin;
    pk::stage_input(kin, keys, (size_t)n * 8, mem, s);
    pk::stage_input(vin, vals, (size_t)n * 4, mem, s);
    pk::radix_sort_pairs((uint64_t*)kin.ptr, (uint32_t*)vin.ptr, n, end_bit, s);
    if (mem == PPTK_MEM_HOST) {
      PK_HIP(hipMemcpyAsync(keys, kin.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, s));
      PK_HIP(hipMemcpyAsync(vals, vin.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    }
    PK_HIP(hipStreamSynchronize(s));
  });
}

extern "C" pptk_status pptk_exclusive_scan_i32(const int32_t* in, int64_t* out, int64_t n,
                                               pptk_mem mem) {
  return pk::guard([&] {
    if (n < 0) pk::fail(PPTK_E_VALUE, "n must be non-negative");
    if (!out) pk::fail(PPTK_E_VALUE, "null output");
    if (n > 0 && !in) pk::fail(PPTK_E_VALUE, "null input");
    hipStream_t s = pk::lib_stream();
    pk::DevInput din;
    pk::stage_input(din, in, (size_t)n * 4, mem, s);
    pk::DevOutput dout;
    dout.init(out, (size_t)(n + 1) * 8, mem);
    pk::exclusive_scan<int32_t, int64_t>((const int32_t*)din.ptr, (int64_t*)dout.dptr, n, s);
    dout.finish(s);
    PK_HIP(hipStreamSynchronize(s));
  });
}
