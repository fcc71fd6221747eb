// Library runtime: per-thread error text, device/stream set-up, the pooled
// device allocator and host<->device staging helpers.
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>

#include "common.hpp"

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
  return g_stream;
}

int device_cu_count() {
  ensure_device();
  return g_cu_count;
}

// ------------------------------------------------------------- allocator ----

namespace {
std::mutex g_pool_mu;
std::multimap<size_t, void*> g_free;      // size -> block
std::map<void*, size_t> g_live;           // block -> size
size_t g_cached_bytes = 0;

size_t round_size(size_t b) {
  const size_t a = 256;
  return (b + a - 1) / a * a;
}
}  // namespace

void* dev_alloc(size_t bytes) {
  ensure_device();
  if (bytes == 0) bytes = 256;
  size_t want = round_size(bytes);
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_free.lower_bound(want);
    if (it != g_free.end() && it->first <= want + want / 4) {
      void* p = it->second;
      size_t sz = it->first;
      g_free.erase(it);
      g_cached_bytes -= sz;
      g_live[p] = sz;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    // Return cached blocks to the driver and try once more.
    (void)hipGetLastError();
    pptk_trim();
    e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(PPTK_E_MEMORY, "device allocation of %zu bytes failed: %s", want,
           hipGetErrorString(e));
    }
  }
  std::lock_guard<std::mutex> lk(g_pool_mu);
  g_live[p] = want;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto it = g_live.find(p);
  if (it == g_live.end()) return;
  g_free.emplace(it->second, p);
  g_cached_bytes += it->second;
  g_live.erase(it);
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
  return pk::guard([] {
    std::multimap<size_t, void*> blocks;
    {
      std::lock_guard<std::mutex> lk(pk::g_pool_mu);
      blocks.swap(pk::g_free);
      pk::g_cached_bytes = 0;
    }
    for (auto& kv : blocks) (void)hipFree(kv.second);
  });
}
