// Internal runtime shared by the library's translation units: error
// reporting, the device allocator, the library stream, element-type helpers
// and the device-side geometry primitives (distance, box bounds).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pptk_amd.h"

namespace pk {

// ---------------------------------------------------------------- errors ----

struct Error {
  pptk_status status;
  std::string msg;
};

[[noreturn]] void fail(pptk_status st, const char* fmt, ...);
void set_last_error(const std::string& msg);
void clear_last_error();

#define PK_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      ::pk::fail(PPTK_E_DEVICE, "%s failed: %s (%s:%d)", #expr,               \
                 hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

#define PK_LAUNCH_CHECK() PK_HIP(hipGetLastError())

// One library call on this thread. Device blocks freed while it is active
// go back to the pool behind a fence: the call's own stream may reuse them at
// once, other streams only after an event recorded on the call's stream at
// its end has completed. So an error that unwinds between a launch and its
// synchronisation never hands an in-flight buffer to another stream. Scopes
// nest; each fences its own blocks.
class CallScope {
 public:
  CallScope();
  ~CallScope();
  CallScope(const CallScope&) = delete;
  CallScope& operator=(const CallScope&) = delete;

 private:
  friend void scope_note_stream(hipStream_t s);
  friend bool scope_defer_free(void* p);
  friend hipStream_t scope_stream();
  CallScope* outer_;
  hipStream_t stream_ = nullptr;
  bool has_stream_ = false;
  bool many_streams_ = false;
  std::vector<void*> unqueued_;   // freed before the call picked a stream
  std::shared_ptr<void> fence_;   // the pool's fence for this call's blocks
};

// Records the stream the active call works on (no-op outside a call).
void scope_note_stream(hipStream_t s);

// Runs `body` and converts pk::Error / std::bad_alloc into a status code.
template <typename F>
pptk_status guard(F&& body) {
  CallScope scope;
  try {
    clear_last_error();
    body();
    return PPTK_OK;
  } catch (const Error& e) {
    set_last_error(e.msg);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return PPTK_E_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return PPTK_E_RUNTIME;
  }
}

// ---------------------------------------------------------------- device ----

// Selects the device, creates the library stream on first use.
void ensure_device();
hipStream_t lib_stream();
inline hipStream_t pick_stream(void* s) {
  if (!s) return lib_stream();
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  scope_note_stream(st);
  return st;
}
int device_cu_count();

// Pooled device allocation. Blocks are cached after free and reused for
// requests of up to 1.25x smaller size; pptk_trim() returns them. A block
// freed inside a library call is fenced on that call's stream (CallScope).
void* dev_alloc(size_t bytes);
void dev_free(void* p);

// RAII device buffer backed by the pool.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { reset(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void reset(size_t bytes) {
    release();
    if (bytes) { p_ = dev_alloc(bytes); n_ = bytes; }
  }
  void release() {
    if (p_) dev_free(p_);
    p_ = nullptr; n_ = 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p_); }
  void* get() const { return p_; }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Copies `bytes` from a host or device source into a fresh device buffer,
// or returns the device pointer itself when the source is already there.
struct DevInput {
  DevBuf owned;
  const void* ptr = nullptr;
};
void stage_input(DevInput& out, const void* src, size_t bytes, pptk_mem mem, hipStream_t s);

// Output buffer: either the caller's device pointer, or a device temporary
// copied back to the caller's host pointer by `finish`.
struct DevOutput {
  DevBuf owned;
  void* dptr = nullptr;
  void* host = nullptr;
  size_t bytes = 0;
  void init(void* user, size_t nbytes, pptk_mem mem);
  void finish(hipStream_t s);
};

// ------------------------------------------------------------ dtype util ----

size_t dtype_size(pptk_dtype t);
bool dtype_valid(int t);
inline bool dist_is_double(pptk_dtype t) { return t == PPTK_F64; }

template <typename T> struct Tag { using type = T; };

// Calls f(Tag<T>{}) with T the C++ type of `t`.
template <typename F>
void dispatch_dtype(pptk_dtype t, F&& f) {
  switch (t) {
    case PPTK_F32: f(Tag<float>{}); break;
    case PPTK_F64: f(Tag<double>{}); break;
    case PPTK_I8: f(Tag<int8_t>{}); break;
    case PPTK_I16: f(Tag<int16_t>{}); break;
    case PPTK_I32: f(Tag<int32_t>{}); break;
    case PPTK_I64: f(Tag<int64_t>{}); break;
    case PPTK_U8: f(Tag<uint8_t>{}); break;
    case PPTK_U16: f(Tag<uint16_t>{}); break;
    case PPTK_U32: f(Tag<uint32_t>{}); break;
    case PPTK_U64: f(Tag<uint64_t>{}); break;
    default: fail(PPTK_E_TYPE, "unsupported dtype code %d", (int)t);
  }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// -------------------------------------------------------- device helpers ----

template <typename D> __host__ __device__ inline D pos_inf();
template <> __host__ __device__ inline float pos_inf<float>() { return __builtin_huge_valf(); }
template <> __host__ __device__ inline double pos_inf<double>() { return __builtin_huge_val(); }

// Point record in tree order: coordinates in the distance type plus the
// original row index, padded to 16-byte multiples so one record is one or
// two vector loads.
template <typename D, int DIM>
struct alignas(16) PointRec {
  D c[DIM];
  int32_t orig;
};

// Node box record: lo[PD], hi[PD] (PD = 4 for 3-D so a box is two float4s).
template <int DIM> struct PadDim { static constexpr int value = DIM == 3 ? 4 : DIM; };
template <typename D, int DIM>
struct alignas(16) BoxRec {
  D lo[PadDim<DIM>::value];
  D hi[PadDim<DIM>::value];
};

// Squared distance exactly as pointkd::PointDistance2 (kdtree-impl.h:331-339):
// running sum from zero, each term (DistT)b - (DistT)a squared. The library is
// compiled with -ffp-contract=off, so no term is fused into an FMA.
template <typename D, int DIM>
__device__ __forceinline__ D dist2(const D* q, const D* p) {
  D s = D(0);
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    D t = p[i] - q[i];
    s = s + t * t;
  }
  return s;
}

// Lower bound of dist2 over a box, same rounding as pointkd::MinDist2
// (pptk/kdtree/src/box.h:157-189). Monotone rounding makes it <= dist2(q,p)
// for every p inside the box, so pruning on it never changes a result set.
template <typename D, int DIM>
__device__ __forceinline__ D box_min_dist2(const BoxRec<D, DIM>& b, const D* q) {
  D s = D(0);
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    // the gap on the side of the box the query lies beyond (lo <= hi, so at
    // most one side is positive), squared once: q - hi is exactly -(hi - q),
    // so the square is that side's. One three-way max for floats.
    const D below = b.lo[i] - q[i];
    const D above = q[i] - b.hi[i];
    const D e = fmax(fmax(below, above), D(0));
    s = s + e * e;
  }
  return s;
}

// Upper bound of dist2 over a box, same rounding as pointkd::MaxDist2
// (box.h:171-198): >= dist2(q,p) for every p inside the box.
template <typename D, int DIM>
__device__ __forceinline__ D box_max_dist2(const BoxRec<D, DIM>& b, const D* q) {
  D s = D(0);
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    const D dmin = b.lo[i] - q[i];
    const D dmax = b.hi[i] - q[i];
    // the larger square (both are products, never NaN for finite boxes)
    s = s + fmax(dmin * dmin, dmax * dmax);
  }
  return s;
}

// box_min_dist2 for sgn = +1, box_max_dist2 for sgn = -1, bit for bit, in
// the same instructions: sq = sgn * q (exact), so fma(sgn, lo, -sq) and
// fma(-sgn, hi, sq) are sgn * (lo - q) and sgn * (q - hi), each one rounding
// of the exact difference as the plain subtractions are. For sgn = -1 the
// axis term is max(q - lo, hi - q)^2, which is max((lo - q)^2, (hi - q)^2)
// because rounding a square is monotone in the magnitude. Lets half a wave
// take the near bound and the other half the far bound of the same node.
template <typename D, int DIM>
__device__ __forceinline__ D box_near_far_dist2(const BoxRec<D, DIM>& b, const D* sq, D sgn) {
  D s = D(0);
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    const D x = fma(sgn, b.lo[i], -sq[i]);
    const D y = fma(-sgn, b.hi[i], sq[i]);
    const D e = fmax(fmax(x, y), D(0));
    s = s + e * e;
  }
  return s;
}

// Workgroups of a launch are dispatched to the chip's 8 XCDs in turn
// (workgroup i to XCD i mod 8), and each XCD has its own L2. xcd_block maps
// the dispatched workgroup b of a grid of g to the index of the work it
// serves so that, in every span of 8 * run consecutive workgroups, each XCD
// takes a run of `run` consecutive work indices: neighbouring work (queries
// that are neighbours on the Morton curve, reading the same leaves) then
// shares one L2. Launch order is kept from one span to the next; the last,
// partial span keeps the identity.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t g, int run) {
  const int64_t span = 8 * (int64_t)run;
  const int64_t s = b / span;
  if (run <= 1 || (s + 1) * span > g) return b;
  const int64_t r = b - s * span;
  return s * span + (r & 7) * run + (r >> 3);
}

// Run length of the self-query search kernels and the r-near count (lidar
// scene, 10M points: kNN + normals 4.50 -> 4.40 ms, count 1.65 -> 1.60 ms
// at 16; 2 to 64 measure alike).
constexpr int kXcdRun = 16;

template <typename D> inline D qnan_host() { return std::numeric_limits<D>::quiet_NaN(); }

template <typename D> __device__ __forceinline__ bool finite_val(D x) {
  return x - x == D(0);  // false for NaN and +-inf
}

// Tree node: points [a, b) in tree order; `child` is the first of the two
// adjacent children (child, child + 1), or -1 for a leaf; `parent` is -1 for
// the root. Node 0 is the root; children are allocated in pairs at odd ids,
// so a node's sibling is id + 1 (odd id) or id - 1 (even id).
struct NodeRec {
  int32_t a, b, child, parent;
};

// Device view of a built tree, passed by value to the query kernels.
template <typename D, int DIM>
struct TreeView {
  const PointRec<D, DIM>* pts;  // num_valid records, tree order
  const int32_t* orig;          // num_valid, tree position -> original index
  const int32_t* fwd;           // num_input, original index -> position or -1
  const BoxRec<D, DIM>* box;    // nn node boxes
  const NodeRec* node;          // nn nodes
  const int32_t* leafof;        // num_valid, tree position -> leaf node
  int32_t n;                    // num_valid
  int32_t leaf;                 // maximum points per leaf
  int32_t nn;                   // number of nodes
  int64_t n_input;
  const int32_t* n_dev;         // if set, num_valid is read from device memory
  int32_t nroots;               // roots are nodes [0, nroots) (one per cloud)
};

}  // namespace pk
