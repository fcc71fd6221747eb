// Host check of pk::BlockPool (pptk_amd/csrc/pool.hpp), the stream-ordered
// cache behind the library's device allocator, with a fake driver: streams
// and events are integers, an event completes when the check says so.
// Built and run by tests/test_pool_cpu.py (g++ with AddressSanitizer and
// UndefinedBehaviorSanitizer). Prints "pool_check: ok" on success.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>

#include "pool.hpp"

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "pool_check: %s:%d: CHECK(%s) failed\n",    \
                   __FILE__, __LINE__, #c);                           \
      std::exit(1);                                                   \
    }                                                                 \
  } while (0)

struct FakeDriver {
  using Stream = int;
  using Event = int;
  std::set<void*> allocated;
  std::map<int, bool> done_;
  std::map<int, int> ev_stream;
  std::set<int> destroyed;
  std::vector<int> waited;
  std::vector<int> synced;
  bool record_ok = true;
  bool sync_ok = true;
  int next_ev = 1;

  bool malloc(void** p, size_t n) {
    *p = std::malloc(n);
    allocated.insert(*p);
    return *p != nullptr;
  }
  void free(void* p) {
    CHECK(allocated.erase(p) == 1);
    std::free(p);
  }
  bool record(Event* ev, Stream s) {
    if (!record_ok) return false;
    *ev = next_ev++;
    done_[*ev] = false;
    ev_stream[*ev] = s;
    return true;
  }
  bool done(Event ev) { return done_.at(ev); }
  void wait(Event ev) {
    waited.push_back(ev);
    done_[ev] = true;
  }
  void destroy(Event ev) { CHECK(destroyed.insert(ev).second); }
  bool sync(Stream s) {
    synced.push_back(s);
    return sync_ok;
  }
  // Completes every event recorded on stream s (the stream went idle).
  void finish(Stream s) {
    for (auto& kv : ev_stream)
      if (kv.second == s) done_[kv.first] = true;
  }
};

using Pool = pk::BlockPool<FakeDriver>;

static void* fresh(FakeDriver& d, Pool& p, size_t n) {
  void* b = nullptr;
  CHECK(d.malloc(&b, Pool::round_size(n)));
  p.adopt(b, n);
  return b;
}

int main() {
  // 1. A block released with no fence goes to any stream; sizes within
  //    1.25x are served from the cache, larger requests are not.
  {
    FakeDriver d;
    Pool p(d);
    void* a = fresh(d, p, 4096);
    p.release_now(a);
    CHECK(p.take(8192, 7) == nullptr);
    CHECK(p.take(3300, 7) == a);  // 4096 <= 1.25 * 3328
    p.release_now(a);
    CHECK(p.take(2048, 7) == nullptr);  // 4096 > 1.25 * 2048
    CHECK(p.take(4096, 9) == a);
    p.release_now(a);
    p.trim();
    CHECK(d.allocated.empty());
  }
  // 2. A block released behind an open fence (the call that used it is
  //    still queuing work on stream 1): stream 1 may take it, stream 2 not,
  //    and trim keeps it.
  {
    FakeDriver d;
    Pool p(d);
    void* a = fresh(d, p, 1 << 20);
    auto f = p.open_fence();
    p.release_fenced(a, 1, f);
    CHECK(p.take(1 << 20, 2) == nullptr);
    p.trim();
    CHECK(d.allocated.count(a) == 1 && p.cached_blocks() == 1);
    CHECK(p.take(1 << 20, 1) == a);
    // released again under the same fence, then the call ends: the fence
    // is recorded on stream 1, not yet complete
    p.release_fenced(a, 1, f);
    p.close_fence(f, 1);
    f.reset();
    CHECK(p.take(1 << 20, 2) == nullptr);
    d.finish(1);  // the work queued on stream 1 has run
    CHECK(p.take(1 << 20, 2) == a);
    CHECK(d.destroyed.size() == 1);  // the last block of the fence left
    p.release_now(a);
    p.trim();
    CHECK(d.allocated.empty());
  }
  // 3. An error unwinds a call between its launch and its synchronisation:
  //    its temporaries are fenced on its stream, so a call on another
  //    stream (another serving slot) gets a fresh block instead.
  {
    FakeDriver d;
    Pool p(d);
    void* tmp = fresh(d, p, 65536);
    auto f = p.open_fence();
    p.release_fenced(tmp, 3, f);  // unwinding frees the temporary
    p.close_fence(f, 3);          // the scope ends: fence recorded
    f.reset();
    void* other = p.take(65536, 4);
    CHECK(other == nullptr);
    other = fresh(d, p, 65536);
    CHECK(other != tmp);
    // trim waits for the recorded fence before returning the block
    p.trim();
    CHECK(d.waited.size() == 1 && d.allocated.count(tmp) == 0);
    p.release_now(other);
    p.trim();
    CHECK(d.allocated.empty());
  }
  // 4. No event can be recorded: the stream is synchronised instead and the
  //    blocks are free for all; if that fails too they stay with their
  //    stream and are never freed.
  {
    FakeDriver d;
    Pool p(d);
    d.record_ok = false;
    void* a = fresh(d, p, 1024);
    auto f = p.open_fence();
    p.release_fenced(a, 5, f);
    p.close_fence(f, 5);
    CHECK(d.synced.size() == 1 && d.synced[0] == 5);
    CHECK(p.take(1024, 6) == a);
    p.release_now(a);
    d.sync_ok = false;
    void* b = fresh(d, p, 2048);
    auto g = p.open_fence();
    p.release_fenced(b, 5, g);
    p.close_fence(g, 5);
    CHECK(p.take(2048, 6) == nullptr);
    p.trim();
    CHECK(d.allocated.count(b) == 1 && d.allocated.count(a) == 0);
    CHECK(p.take(2048, 5) == b);  // still usable in stream order
    p.release_now(b);
    p.trim();
    CHECK(d.allocated.empty());
  }
  // 5. A fence passed by a full device synchronisation frees its blocks for
  //    every stream without an event.
  {
    FakeDriver d;
    Pool p(d);
    void* a = fresh(d, p, 512);
    auto f = p.open_fence();
    p.release_fenced(a, 1, f);
    p.pass_fence(f);
    CHECK(p.take(512, 2) == a);
    CHECK(d.next_ev == 1);
    p.release_now(a);
    p.trim();
    CHECK(d.allocated.empty());
  }
  // 6. Unknown pointers and double releases are ignored.
  {
    FakeDriver d;
    Pool p(d);
    void* a = fresh(d, p, 256);
    int x = 0;
    p.release_now(&x);
    p.release_now(a);
    p.release_now(a);
    CHECK(p.cached_blocks() == 1 && p.live_blocks() == 0);
    p.trim();
    CHECK(d.allocated.empty());
  }
  std::printf("pool_check: ok\n");
  return 0;
}
