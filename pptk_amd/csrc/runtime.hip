// Library runtime: per-thread error text, device/stream set-up, the pooled
// device allocator and host<->device staging helpers.
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>

#include "common.hpp"
#include "pool.hpp"

namespace pk {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

void fail(pptk_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error{st, std::string(buf)};
}

// ---------------------------------------------------------------- device ----

namespace {
std::once_flag g_dev_once;
int g_device = -1;
int g_cu_count = 0;
hipStream_t g_stream = nullptr;
std::string g_dev_error;

void init_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    g_dev_error = "no HIP device available (hipGetDeviceCount: " +
                  std::string(hipGetErrorString(e)) + ", count " + std::to_string(n) + ")";
    return;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (hipSetDevice(dev) != hipSuccess) {
    g_dev_error = "hipSetDevice failed";
    return;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_cu_count = prop.multiProcessorCount;
  if (g_cu_count <= 0) g_cu_count = 256;
  // Default priority, non-blocking with respect to the null stream.
  if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
    g_dev_error = "hipStreamCreate failed";
    return;
  }
  g_device = dev;
}
}  // namespace

void ensure_device() {
  std::call_once(g_dev_once, init_device);
  if (g_device < 0) fail(PPTK_E_DEVICE, "%s", g_dev_error.c_str());
  PK_HIP(hipSetDevice(g_device));
}

hipStream_t lib_stream() {
  ensure_device();
  scope_note_stream(g_stream);
  return g_stream;
}

int device_cu_count() {
  ensure_device();
  return g_cu_count;
}

// ------------------------------------------------------------- allocator ----

namespace {

struct HipPoolDriver {
  using Stream = hipStream_t;
  using Event = hipEvent_t;
  bool malloc(void** p, size_t n) {
    if (hipMalloc(p, n) == hipSuccess) return true;
    (void)hipGetLastError();
    return false;
  }
  void free(void* p) { (void)hipFree(p); }
  bool record(Event* ev, Stream s) {
    if (hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if (hipEventRecord(*ev, s) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipEventDestroy(*ev);
      return false;
    }
    return true;
  }
  bool done(Event ev) {
    const hipError_t e = hipEventQuery(ev);
    if (e == hipSuccess) return true;
    (void)hipGetLastError();
    return false;  // not ready yet, or an error: never reuse early
  }
  void wait(Event ev) {
    if (hipEventSynchronize(ev) != hipSuccess) (void)hipGetLastError();
  }
  void destroy(Event ev) { (void)hipEventDestroy(ev); }
  bool sync(Stream s) {
    if (hipStreamSynchronize(s) == hipSuccess) return true;
    (void)hipGetLastError();
    return false;
  }
};

HipPoolDriver g_drv;
BlockPool<HipPoolDriver>& pool() {
  static BlockPool<HipPoolDriver>* p = new BlockPool<HipPoolDriver>(g_drv);  // lives to exit
  return *p;
}

thread_local CallScope* t_scope = nullptr;

}  // namespace

using Pool = BlockPool<HipPoolDriver>;

CallScope::CallScope() : outer_(t_scope) { t_scope = this; }

CallScope::~CallScope() {
  t_scope = outer_;
  // Blocks freed before any stream was picked were never used by this call.
  for (void* p : unqueued_) pool().release_now(p);
  if (!fence_) return;
  auto f = std::static_pointer_cast<Pool::Fence>(fence_);
  if (many_streams_) {
    // Work of this call went to more than one stream: wait for all of it.
    if (hipDeviceSynchronize() == hipSuccess) {
      pool().pass_fence(f);
      return;
    }
    (void)hipGetLastError();
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    // An event recorded now would join an unfinished capture: the fence stays
    // open and its blocks stay with this stream.
    (void)hipGetLastError();
    return;
  }
  pool().close_fence(f, stream_);
}

void scope_note_stream(hipStream_t s) {
  CallScope* c = t_scope;
  if (!c) return;
  if (c->has_stream_ && c->stream_ != s) c->many_streams_ = true;
  if (!c->has_stream_) {
    c->stream_ = s;
    c->has_stream_ = true;
  }
}

bool scope_defer_free(void* p) {
  CallScope* c = t_scope;
  if (!c) return false;
  if (!c->has_stream_) {
    c->unqueued_.push_back(p);
    return true;
  }
  if (!c->fence_) c->fence_ = pool().open_fence();
  pool().release_fenced(p, c->stream_, std::static_pointer_cast<Pool::Fence>(c->fence_));
  return true;
}

hipStream_t scope_stream() {
  CallScope* c = t_scope;
  return c && c->has_stream_ ? c->stream_ : nullptr;
}

void* dev_alloc(size_t bytes) {
  ensure_device();
  if (void* p = pool().take(bytes, scope_stream())) return p;
  const size_t want = Pool::round_size(bytes);
  void* p = nullptr;
  if (!g_drv.malloc(&p, want)) {
    // Return cached blocks to the driver and try once more.
    pool().trim();
    if (!g_drv.malloc(&p, want))
      fail(PPTK_E_MEMORY, "device allocation of %zu bytes failed", want);
  }
  pool().adopt(p, want);
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  if (!scope_defer_free(p)) pool().release_now(p);
}

// --------------------------------------------------------------- staging ----

void stage_input(DevInput& out, const void* src, size_t bytes, pptk_mem mem, hipStream_t s) {
  if (mem == PPTK_MEM_DEVICE || bytes == 0) {
    out.ptr = src;
    return;
  }
  out.owned.reset(bytes);
  PK_HIP(hipMemcpyAsync(out.owned.get(), src, bytes, hipMemcpyHostToDevice, s));
  out.ptr = out.owned.get();
}

void DevOutput::init(void* user, size_t nbytes, pptk_mem mem) {
  bytes = nbytes;
  if (!user) { dptr = nullptr; host = nullptr; return; }
  if (mem == PPTK_MEM_DEVICE) {
    dptr = user;
    host = nullptr;
  } else {
    owned.reset(nbytes);
    dptr = owned.get();
    host = user;
  }
}

void DevOutput::finish(hipStream_t s) {
  if (host && bytes) PK_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, s));
}

// ------------------------------------------------------------ dtype util ----

size_t dtype_size(pptk_dtype t) {
  switch (t) {
    case PPTK_F32: case PPTK_I32: case PPTK_U32: return 4;
    case PPTK_F64: case PPTK_I64: case PPTK_U64: return 8;
    case PPTK_I16: case PPTK_U16: return 2;
    case PPTK_I8: case PPTK_U8: return 1;
  }
  return 0;
}

bool dtype_valid(int t) { return t >= PPTK_F32 && t <= PPTK_U64; }

}  // namespace pk

// ------------------------------------------------------------------ C ABI ----

extern "C" const char* pptk_last_error(void) { return pk::g_last_error.c_str(); }

extern "C" int pptk_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" pptk_status pptk_trim(void) {
  return pk::guard([] { pk::pool().trim(); });
}
