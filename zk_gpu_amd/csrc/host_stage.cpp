// Host <-> device copies of caller (pageable) memory through pinned staging.
//
// The drop-in entry points take host buffers (reference src/api.rs:6-10,
// copied per call at cuda/fft.cu:284-285,388), so their time is mostly the
// copies. A copy here goes in chunks through two pinned buffers per stream:
// while the DMA engine moves chunk i, a small pool of host threads copies
// chunk i+1 between the caller's memory and the other pinned buffer. Pinned
// buffers are allocated (hipHostMalloc), never registered.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace zkgpu {

namespace {

constexpr size_t kChunk = 8u << 20;        // bytes per pinned buffer
constexpr size_t kDirectBelow = 1u << 20;  // smaller copies go straight through hipMemcpyAsync
constexpr size_t kSlice = 1u << 20;        // host-copy work unit

// Fixed pools of host threads for the copies into and out of pinned memory:
// one for uploads, one for downloads, so that a download's host copy does not
// queue behind an upload's (the server runs both at once).
class CopyPool {
 public:
  static CopyPool& get(int direction) {
    static CopyPool up, down;
    return direction ? down : up;
  }
  // dst[i] = src[i] for bytes, split over the pool; returns when done. One
  // copy at a time: concurrent callers queue here.
  void copy(void* dst, const void* src, size_t bytes) {
    const size_t n = (bytes + kSlice - 1) / kSlice;
    if (n <= 1 || workers_.empty()) {
      memcpy(dst, src, bytes);
      return;
    }
    std::lock_guard<std::mutex> one(call_mu_);
    std::unique_lock<std::mutex> lk(mu_);
    dst_ = (char*)dst;
    src_ = (const char*)src;
    bytes_ = bytes;
    next_ = 0;
    slices_ = n;
    done_ = 0;
    ++gen_;
    cv_.notify_all();
    lk.unlock();
    work();  // the caller takes slices too
    lk.lock();
    done_cv_.wait(lk, [&] { return done_ == slices_; });
  }

 private:
  CopyPool() {
    unsigned hc = std::thread::hardware_concurrency();
    const unsigned nw = std::min(6u, hc > 4 ? hc / 2 - 1 : 1u);
    for (unsigned i = 0; i < nw; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  // take slices until none are left
  void work() {
    for (;;) {
      size_t i;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (next_ >= slices_) return;
        i = next_++;
      }
      const size_t off = i * kSlice;
      memcpy(dst_ + off, src_ + off, std::min(kSlice, bytes_ - off));
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (++done_ == slices_) done_cv_.notify_all();
      }
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      work();
    }
  }
  std::mutex call_mu_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::thread> workers_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t bytes_ = 0, next_ = 0, slices_ = 0, done_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Two pinned buffers and their events per (device, stream); a host thread
// uses its own stream, so one staging set is never shared by two copies.
struct Staging {
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ok = false;
};

std::mutex g_stage_mu;
std::map<std::pair<int, hipStream_t>, Staging*> g_stage;

Staging* staging(int device, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  Staging*& st = g_stage[std::make_pair(device, s)];
  if (!st) {
    st = new Staging();
    st->ok = true;
    for (int i = 0; i < 2 && st->ok; ++i) {
      st->ok = hipHostMalloc(&st->pin[i], kChunk, hipHostMallocDefault) == hipSuccess &&
               hipEventCreateWithFlags(&st->ev[i], hipEventDisableTiming) == hipSuccess;
    }
    if (!st->ok) (void)hipGetLastError();
  }
  return st->ok ? st : nullptr;
}

}  // namespace

int stage_h2d(int device, hipStream_t s, void* dst, const void* src, size_t bytes) {
  Staging* st = bytes < kDirectBelow ? nullptr : staging(device, s);
  if (!st) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
  const size_t n = (bytes + kChunk - 1) / kChunk;
  for (size_t i = 0; i < n; ++i) {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    // the buffer's previous DMA (chunk i-2, or an earlier copy on s) has left
    if (hipEventSynchronize(st->ev[b]) != hipSuccess) return ZK_ERR_HIP;
    CopyPool::get(0).copy(st->pin[b], (const char*)src + off, len);
    if (hipMemcpyAsync((char*)dst + off, st->pin[b], len, hipMemcpyHostToDevice, s) != hipSuccess) return ZK_ERR_HIP;
    if (hipEventRecord(st->ev[b], s) != hipSuccess) return ZK_ERR_HIP;
  }
  return ZK_OK;
}

int stage_d2h(int device, hipStream_t s, void* dst, const void* src, size_t bytes) {
  Staging* st = bytes < kDirectBelow ? nullptr : staging(device, s);
  if (!st) {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) return ZK_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
  }
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto enqueue = [&](size_t i) {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    return hipMemcpyAsync(st->pin[b], (const char*)src + off, len, hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipEventRecord(st->ev[b], s) == hipSuccess;
  };
  for (size_t i = 0; i < std::min<size_t>(n, 2); ++i)
    if (!enqueue(i)) return ZK_ERR_HIP;
  for (size_t i = 0; i < n; ++i) {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    if (hipEventSynchronize(st->ev[b]) != hipSuccess) return ZK_ERR_HIP;
    CopyPool::get(1).copy((char*)dst + off, st->pin[b], len);
    if (i + 2 < n && !enqueue(i + 2)) return ZK_ERR_HIP;
  }
  return ZK_OK;
}

}  // namespace zkgpu
