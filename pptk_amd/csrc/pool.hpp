// Stream-ordered cache of device blocks behind dev_alloc / dev_free.
//
// A block freed while a library call may still have work in flight on its
// stream must not be handed to another stream before that work has finished.
// Each cached block therefore keeps the stream it was released on and a
// fence, an event recorded on that stream once the releasing call has queued
// all of its work:
//   * the same stream may take the block again at once, even before the fence
//     is recorded (its later work runs after the work that used the block);
//   * another stream takes it only once the fence's event has completed;
//   * blocks released with no fence (nothing can still use them) are free for
//     every stream.
// `trim` waits for recorded fences before it returns blocks to the driver and
// keeps blocks whose fence is not recorded yet.
//
// The class is written against a driver interface (allocation, event
// record/query/wait, stream synchronisation) so that these rules are checked
// on the host with a fake driver; runtime.hip instantiates it with the HIP
// runtime.
//
// Driver interface:
//   using Stream = ...; using Event = ...;   (Stream must be ==-comparable)
//   bool malloc(void** p, size_t bytes);
//   void free(void* p);
//   bool record(Event* ev, Stream s);   // create an event and record it on s
//   bool done(Event ev);                // the event has completed
//   void wait(Event ev);                // block until it completes
//   void destroy(Event ev);
//   bool sync(Stream s);                // block until s is idle
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace pk {

template <class Drv>
class BlockPool {
 public:
  using Stream = typename Drv::Stream;
  using Event = typename Drv::Event;

  // The point after which blocks released by one call may go to any stream.
  class Fence {
   public:
    explicit Fence(Drv& d) : drv_(d) {}
    ~Fence() {
      if (has_ev_) drv_.destroy(ev_);
    }
    Fence(const Fence&) = delete;
    Fence& operator=(const Fence&) = delete;

   private:
    friend class BlockPool;
    enum State { kOpen, kRecorded, kPassed, kDead };
    bool complete() {  // pool lock held
      if (state_ == kRecorded && drv_.done(ev_)) state_ = kPassed;
      return state_ == kPassed;
    }
    Drv& drv_;
    Event ev_{};
    bool has_ev_ = false;
    State state_ = kOpen;
  };
  using FencePtr = std::shared_ptr<Fence>;

  explicit BlockPool(Drv& drv) : drv_(drv) {}

  static size_t round_size(size_t b) {
    const size_t a = 256;
    if (b == 0) b = a;
    return (b + a - 1) / a * a;
  }

  // A cached block of at least `bytes` (at most 1.25x as large) that work on
  // `stream` may use, or nullptr if there is none.
  void* take(size_t bytes, Stream stream) {
    const size_t want = round_size(bytes);
    std::lock_guard<std::mutex> lk(mu_);
    for (auto it = free_.lower_bound(want); it != free_.end() && it->first <= want + want / 4;
         ++it) {
      Cached& c = it->second;
      if (c.fence && !(c.stream == stream) && !c.fence->complete()) continue;
      void* p = c.p;
      live_[p] = it->first;
      cached_bytes_ -= it->first;
      free_.erase(it);
      return p;
    }
    return nullptr;
  }

  // Registers a block fresh from the driver.
  void adopt(void* p, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    live_[p] = round_size(bytes);
  }

  // Returns `p` to the cache for every stream: nothing may use it any more.
  void release_now(void* p) { release(p, Stream(), nullptr); }

  // A fence for the blocks one call releases on `stream`; close it when the
  // call has queued all of its work.
  FencePtr open_fence() { return std::make_shared<Fence>(drv_); }

  // Returns `p` to the cache behind `fence`: `stream` may take it again at
  // once, other streams once the fence has passed.
  void release_fenced(void* p, Stream stream, const FencePtr& fence) { release(p, stream, fence); }

  // Records the fence on `stream`. If no event can be recorded the stream is
  // synchronised instead; if that fails too, the fence never passes and its
  // blocks stay with `stream` (never handed to another stream, never freed).
  void close_fence(const FencePtr& f, Stream stream) {
    Event ev{};
    const bool rec = drv_.record(&ev, stream);
    const bool synced = rec ? false : drv_.sync(stream);
    std::lock_guard<std::mutex> lk(mu_);
    if (rec) {
      f->ev_ = ev;
      f->has_ev_ = true;
      f->state_ = Fence::kRecorded;
    } else {
      f->state_ = synced ? Fence::kPassed : Fence::kDead;
    }
  }

  // Marks a fence passed: the caller has waited for all of its work.
  void pass_fence(const FencePtr& f) {
    std::lock_guard<std::mutex> lk(mu_);
    if (f->state_ != Fence::kDead) f->state_ = Fence::kPassed;
  }

  // Returns cached blocks to the driver: unfenced ones, and fenced ones after
  // waiting for their recorded event. Blocks behind an open or dead fence stay.
  void trim() {
    std::vector<Cached> out;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto it = free_.begin(); it != free_.end();) {
        const Cached& c = it->second;
        const bool keep = c.fence && (c.fence->state_ == Fence::kOpen ||
                                      c.fence->state_ == Fence::kDead);
        if (keep) { ++it; continue; }
        cached_bytes_ -= it->first;
        out.push_back(c);
        it = free_.erase(it);
      }
    }
    for (Cached& c : out) {
      if (c.fence && c.fence->state_ == Fence::kRecorded) drv_.wait(c.fence->ev_);
      drv_.free(c.p);
    }
  }

  size_t cached_bytes() const {
    std::lock_guard<std::mutex> lk(mu_);
    return cached_bytes_;
  }
  size_t cached_blocks() const {
    std::lock_guard<std::mutex> lk(mu_);
    return free_.size();
  }
  size_t live_blocks() const {
    std::lock_guard<std::mutex> lk(mu_);
    return live_.size();
  }

 private:
  struct Cached {
    void* p;
    Stream stream;
    FencePtr fence;
  };

  void release(void* p, Stream stream, const FencePtr& fence) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second, Cached{p, stream, fence});
    cached_bytes_ += it->second;
    live_.erase(it);
  }

  Drv& drv_;
  mutable std::mutex mu_;
  std::multimap<size_t, Cached> free_;
  std::map<void*, size_t> live_;
  size_t cached_bytes_ = 0;
};

}  // namespace pk
