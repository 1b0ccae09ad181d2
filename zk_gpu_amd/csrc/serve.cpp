// Serving: continuous batching of NTT jobs on one device.
//
// The reference admits one FFT call per GPU at a time (GpuLock,
// src/gpulock.rs:133-226: a job waits until a GPU has enough free memory for
// it, runs alone, then releases the memory). Here one scheduler thread per
// server runs an iteration loop instead:
//
//   retire  - iterations whose completion event has fired finish the jobs
//             whose last pass they ran (workspace back to the pool),
//   admit   - pending jobs, in arrival order, while their footprint fits the
//             HBM budget and fewer than max_active jobs are admitted (tables
//             fetched or built, output and workspace taken from the pool),
//   issue   - one persistent-kernel launch runs the NEXT pass of every
//             admitted job (passes of one job go to successive iterations on
//             one stream, so stream order is their dependency).
//
// Jobs on host buffers (ZKGPU_JOB_HOST_BUFFERS) get a device buffer from the
// pool at admission; an uploader thread stages their input to it on its own
// stream, a downloader thread stages the result back once the job's last
// iteration has retired, so one job's upload, other jobs' passes and a third
// job's download overlap (the reference copies in, transforms and copies out
// one call at a time, cuda/fft.cu:284-285,388). A host job joins iterations
// once its upload is enqueued; the iteration's stream waits for the upload's
// event.
//
// Up to kDepth iterations are in flight, so the device never waits for the
// host between iterations; a job submitted meanwhile joins at the next
// iteration issued. Per-iteration descriptors go through pinned staging (one
// host-to-device copy of the used part, which also zeroes the tile counter),
// so an iteration is one copy plus one kernel; an iteration that holds a
// single job launches its pass as a pass kernel instead, with the descriptor
// as the kernel argument.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/zkgpu_ntt.h"
#include "engine.hpp"

namespace zkgpu {

namespace {

constexpr int kDepth = 2;                         // iterations in flight
constexpr uint64_t kDefaultBudget = 64ull << 30;  // 64 GiB
constexpr uint32_t kDefaultMaxActive = 256;
constexpr uint64_t kIdleSpinNs = 200000;  // idle scheduler spins this long before sleeping

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device buffers recycled by exact size (outputs are 32 * 2^k bytes,
// workspaces 36 * 2^k), so a steady stream of jobs allocates nothing once
// every size class has been seen.
class Pool {
 public:
  explicit Pool(int device) : device_(device) {}
  ~Pool() { trim(); }
  void* get(size_t bytes) {
    auto& fl = free_[bytes];
    if (!fl.empty()) {
      void* p = fl.back();
      fl.pop_back();
      cached_ -= bytes;
      return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      trim();
      if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
      }
    }
    total_ += bytes;
    return p;
  }
  void put(void* p, size_t bytes) {
    if (!p) return;
    free_[bytes].push_back(p);
    cached_ += bytes;
  }
  // free every cached buffer (the caller has synchronised the device)
  void trim() {
    for (auto& kv : free_) {
      for (void* p : kv.second) {
        (void)hipFree(p);
        total_ -= kv.first;
      }
      kv.second.clear();
    }
    cached_ = 0;
  }
  size_t total() const { return total_; }

 private:
  int device_;
  std::map<size_t, std::vector<void*>> free_;
  size_t cached_ = 0, total_ = 0;
};

struct SJob {
  uint64_t ticket = 0;
  zkgpu_job cjob{};
  JobSpec spec;
  JobTables tables;
  PassDesc descs[MAX_PASSES];
  uint32_t npass = 0, next = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* out_owned = nullptr;
  size_t out_bytes = 0;
  size_t footprint = 0;
  int status = ZK_OK;
  bool done = false;
  bool host = false;           // d_in / d_out are host buffers
  bool ready = true;           // input on the device (host jobs: upload enqueued)
  void* dev_io = nullptr;      // host jobs: device buffer (input, then output in place)
  size_t io_bytes = 0;
  hipEvent_t ev_up = nullptr;  // host jobs: the upload's completion on the upload stream
  uint32_t iters = 0, iter_jobs_max = 0;  // iterations run, most jobs in one of them
  uint64_t t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
};

struct Iter {
  hipEvent_t done = nullptr;
  void* host = nullptr;  // pinned staging: descriptors | tile starts | counter
  void* dev = nullptr;
  std::vector<SJob*> finishing;
};

}  // namespace

class Server {
 public:
  Server(int device, Engine* e, uint64_t budget, uint32_t max_active)
      : device_(device), eng_(e), budget_(budget), max_active_(max_active), pool_(device) {}

  int init() {
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) return ZK_ERR_HIP;
    if (hipStreamCreateWithFlags(&up_stream_, hipStreamNonBlocking) != hipSuccess) return ZK_ERR_HIP;
    if (hipStreamCreateWithFlags(&down_stream_, hipStreamNonBlocking) != hipSuccess) return ZK_ERR_HIP;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_);
    grid_cap_ = (uint32_t)cus * persistent_wgs_per_cu();  // resident workgroups of the persistent kernel
    // staging layout of an iteration: counter | tile starts | descriptors;
    // only the used part of it is copied
    ts_off_ = 256;
    stage_bytes_ = ts_off_ + round256((size_t)max_active_ * sizeof(uint32_t)) + (size_t)max_active_ * sizeof(PassDesc);
    for (auto& it : ring_) {
      if (hipEventCreateWithFlags(&it.done, hipEventDisableTiming) != hipSuccess) return ZK_ERR_HIP;
      if (hipHostMalloc(&it.host, stage_bytes_, hipHostMallocDefault) != hipSuccess) return ZK_ERR_NOMEM;
      if (hipMalloc(&it.dev, stage_bytes_) != hipSuccess) return ZK_ERR_NOMEM;
      memset(it.host, 0, stage_bytes_);
    }
    thread_ = std::thread([this] { loop(); });
    uploader_ = std::thread([this] { io_loop(true); });
    downloader_ = std::thread([this] { io_loop(false); });
    return ZK_OK;
  }

  ~Server() {
    if (thread_.joinable()) {
      {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        ++gen_;
      }
      cv_work_.notify_all();
      thread_.join();
    }
    // the scheduler has retired every iteration; the I/O threads finish
    // their queues (the downloads of the last host jobs) and stop
    {
      std::lock_guard<std::mutex> lk(mu_);
      io_stop_ = true;
    }
    cv_io_.notify_all();
    if (uploader_.joinable()) uploader_.join();
    if (downloader_.joinable()) downloader_.join();
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto& kv : jobs_) {
      pool_.put(kv.second->out_owned, kv.second->out_bytes);
      pool_.put(kv.second->ws, kv.second->ws_bytes);
      pool_.put(kv.second->dev_io, kv.second->io_bytes);
      if (kv.second->ev_up) (void)hipEventDestroy(kv.second->ev_up);
    }
    pool_.trim();
    for (auto& it : ring_) {
      if (it.done) (void)hipEventDestroy(it.done);
      if (it.host) (void)hipHostFree(it.host);
      if (it.dev) (void)hipFree(it.dev);
    }
    if (stream_) (void)hipStreamDestroy(stream_);
    if (up_stream_) (void)hipStreamDestroy(up_stream_);
    if (down_stream_) (void)hipStreamDestroy(down_stream_);
  }

  int submit(const zkgpu_job& cj, uint64_t* ticket) {
    std::unique_ptr<SJob> j(new SJob());
    int st = job_from_c(cj, j->spec, true);
    if (st) return st;
    j->cjob = cj;
    const uint32_t k = cj.log_n;
    const Plan p = plan_for(k, PLAN_BATCH);
    j->npass = p.npass;
    j->ws_bytes = p.npass > 1 ? round256((size_t)36 << k) : 0;
    j->host = (cj.flags & ZKGPU_JOB_HOST_BUFFERS) != 0;
    if (j->host) {
      if (!cj.d_out) return ZK_ERR_INVALID;  // a host job's result goes to the caller's buffer
      j->io_bytes = round256((size_t)32 << k);
      j->ready = false;
    }
    j->out_bytes = (cj.d_out || j->host) ? 0 : round256((size_t)32 << k);
    j->footprint = j->ws_bytes + j->out_bytes + j->io_bytes;
    if (j->footprint > budget_) return ZK_ERR_BUDGET;
    j->t_submit = now_ns();
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stop_) return ZK_ERR_SHUTDOWN;
      j->ticket = ++next_ticket_;
      *ticket = j->ticket;
      pending_.push_back(j.get());
      jobs_[j->ticket] = std::move(j);
      ++submitted_;
      ++gen_;
    }
    submissions_.fetch_add(1, std::memory_order_release);
    cv_work_.notify_one();
    return ZK_OK;
  }

  int poll(uint64_t ticket, int32_t wait_ms, zkgpu_server_result* out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto f = jobs_.find(ticket);
    if (f == jobs_.end()) return ZK_ERR_INVALID;
    SJob* j = f->second.get();
    if (!j->done && wait_ms != 0) {
      auto pred = [&] { return j->done; };
      if (wait_ms < 0) cv_done_.wait(lk, pred);
      else cv_done_.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
    }
    if (out) {
      out->status = j->status;
      out->done = j->done ? 1 : 0;
      out->d_out = j->spec.out;
      out->t_submit_ns = j->t_submit;
      out->t_admit_ns = j->t_admit;
      out->t_start_ns = j->t_start;
      out->t_done_ns = j->t_done;
      out->iterations = j->iters;
      out->iter_jobs_max = j->iter_jobs_max;
    }
    return ZK_OK;
  }

  int release(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto f = jobs_.find(ticket);
    if (f == jobs_.end()) return ZK_ERR_INVALID;
    SJob* j = f->second.get();
    if (!j->done) return ZK_ERR_INVALID;
    if (j->out_owned) {
      pool_.put(j->out_owned, j->out_bytes);
      in_use_ -= j->out_bytes;
    }
    jobs_.erase(f);
    ++gen_;
    cv_work_.notify_one();  // memory came back: pending jobs may fit now
    return ZK_OK;
  }

  void stats(zkgpu_server_stats* s) {
    std::lock_guard<std::mutex> lk(mu_);
    s->submitted = submitted_;
    s->completed = completed_;
    s->failed = failed_;
    s->pending = pending_.size();
    s->active = active_.size() + finishing_count_ + downloading_;
    s->iterations = iterations_;
    s->budget_bytes = budget_;
    s->in_use_bytes = in_use_;
    s->peak_in_use_bytes = peak_in_use_;
    s->pool_bytes = pool_.total();
    s->elements_done = elements_done_;
  }

 private:
  void fail(SJob* j, int st) {
    j->status = st;
    j->done = true;
    j->t_done = now_ns();
    ++failed_;
  }

  // Admit pending jobs in arrival order while they fit. Runs with mu_ held;
  // table preparation happens here (fresh tables are built on the
  // scheduler's own table stream and synchronised there, never on stream_).
  void admit() {
    while (!pending_.empty() && active_.size() + finishing_count_ + downloading_ < max_active_) {
      SJob* j = pending_.front();
      if (in_use_ + j->footprint > budget_) break;  // arrival order: no overtaking
      pending_.pop_front();
      int st = eng_->prepare(j->spec, j->tables, nullptr, PLAN_BATCH);
      if (st == ZK_OK && j->tables.plan.npass != j->npass) st = ZK_ERR_INVALID;  // plan changed since submit
      if (st == ZK_OK && j->ws_bytes) {
        j->ws = pool_.get(j->ws_bytes);
        if (!j->ws) st = ZK_ERR_NOMEM;
      }
      if (st == ZK_OK && j->out_bytes) {
        j->out_owned = pool_.get(j->out_bytes);
        if (!j->out_owned) st = ZK_ERR_NOMEM;
        j->spec.out = j->out_owned;
      }
      if (st == ZK_OK && j->host) {
        j->dev_io = pool_.get(j->io_bytes);
        if (!j->dev_io) st = ZK_ERR_NOMEM;
        if (st == ZK_OK && hipEventCreateWithFlags(&j->ev_up, hipEventDisableTiming) != hipSuccess) {
          (void)hipGetLastError();
          j->ev_up = nullptr;
          st = ZK_ERR_HIP;
        }
        j->spec.in = j->dev_io;  // transformed in place on the device
        j->spec.out = j->dev_io;
      }
      uint32_t nd = 0;
      if (st == ZK_OK) st = eng_->build_descs(j->spec, j->tables, j->ws, j->descs, &nd);
      if (st != ZK_OK) {
        pool_.put(j->ws, j->ws_bytes);
        j->ws = nullptr;
        pool_.put(j->out_owned, j->out_bytes);
        j->out_owned = nullptr;
        pool_.put(j->dev_io, j->io_bytes);
        j->dev_io = nullptr;
        fail(j, st);
        cv_done_.notify_all();
        continue;
      }
      j->t_admit = now_ns();
      in_use_ += j->footprint;
      peak_in_use_ = std::max(peak_in_use_, in_use_);
      active_.push_back(j);
      if (j->host) {
        up_q_.push_back(j);
        cv_io_.notify_all();
      }
    }
  }

  // One persistent-kernel launch over the next pass of every admitted job.
  int issue() {
    Iter& it = ring_[(head_ + inflight_) % kDepth];
    it.finishing.clear();
    std::vector<SJob*> order;
    for (SJob* j : active_)
      if (j->ready) order.push_back(j);
    if (order.empty()) return ZK_OK;
    hipError_t e = hipSuccess;
    // host jobs starting now: this iteration waits for their uploads
    for (SJob* j : order)
      if (j->host && j->next == 0 && e == hipSuccess) e = hipStreamWaitEvent(stream_, j->ev_up, 0);
    // biggest transforms first so that the tile queue ends with short tiles
    std::stable_sort(order.begin(), order.end(),
                     [](const SJob* a, const SJob* b) { return a->cjob.log_n > b->cjob.log_n; });
    const size_t desc_off = ts_off_ + round256(order.size() * sizeof(uint32_t));
    PassDesc* hd = (PassDesc*)((char*)it.host + desc_off);
    uint32_t* hts = (uint32_t*)((char*)it.host + ts_off_);
    uint32_t* hcnt = (uint32_t*)it.host;
    uint32_t nd = 0, total = 0;
    const uint64_t t = now_ns();
    for (SJob* j : order) {
      const PassDesc& d = j->descs[j->next];
      hd[nd] = d;
      hts[nd] = total;
      total += batch_items(d);
      ++nd;
      if (j->next == 0) j->t_start = t;
      if (++j->next == j->npass) it.finishing.push_back(j);
    }
    hcnt[0] = 0;
    for (SJob* j : order) {
      ++j->iters;
      j->iter_jobs_max = std::max(j->iter_jobs_max, nd);
    }
    // jobs in their last pass leave the admitted list now; the iteration's
    // completion finishes them
    active_.erase(std::remove_if(active_.begin(), active_.end(), [](SJob* j) { return j->next == j->npass; }),
                  active_.end());
    finishing_count_ += it.finishing.size();
    if (e == hipSuccess && nd == 1) {
      // one job in the iteration: its pass as a pass-kernel launch (the
      // descriptor travels as the kernel argument; no copy, no counter)
      e = launch_pass(hd[0], stream_);
    } else if (e == hipSuccess) {
      e = hipMemcpyAsync(it.dev, it.host, desc_off + nd * sizeof(PassDesc), hipMemcpyHostToDevice, stream_);
      if (e == hipSuccess)
        e = launch_persistent((const PassDesc*)((char*)it.dev + desc_off), (const uint32_t*)((char*)it.dev + ts_off_),
                              nd, total, (uint32_t*)it.dev, std::min(total, grid_cap_), stream_);
    }
    if (e == hipSuccess) e = hipEventRecord(it.done, stream_);
    if (e != hipSuccess) {
      // nothing of this iteration is known to run: its finishing jobs fail
      // here (an unrecorded event would read as complete)
      (void)hipGetLastError();
      for (SJob* j : it.finishing) {
        pool_.put(j->ws, j->ws_bytes);
        j->ws = nullptr;
        in_use_ -= j->ws_bytes;
        release_io(j);
        fail(j, ZK_ERR_HIP);
      }
      finishing_count_ -= it.finishing.size();
      it.finishing.clear();
      cv_done_.notify_all();
      return ZK_ERR_HIP;
    }
    ++inflight_;
    ++iterations_;
    return ZK_OK;
  }

  // Finish the iterations whose event has fired, oldest first.
  void retire() {
    while (inflight_ > 0) {
      Iter& it = ring_[head_];
      const hipError_t q = hipEventQuery(it.done);
      if (q == hipErrorNotReady) break;
      const uint64_t t = now_ns();
      for (SJob* j : it.finishing) {
        pool_.put(j->ws, j->ws_bytes);
        j->ws = nullptr;
        in_use_ -= j->ws_bytes;
        if (q == hipSuccess && j->host) {
          // the result goes back to the caller's buffer first
          down_q_.push_back(j);
          ++downloading_;
          cv_io_.notify_all();
          continue;
        }
        release_io(j);
        j->done = true;
        j->t_done = t;
        if (q == hipSuccess) {
          ++completed_;
          elements_done_ += 1ull << j->cjob.log_n;
        } else {
          (void)hipGetLastError();
          j->status = ZK_ERR_HIP;
          ++failed_;
        }
      }
      finishing_count_ -= it.finishing.size();
      it.finishing.clear();
      head_ = (head_ + 1) % kDepth;
      --inflight_;
      cv_done_.notify_all();
    }
  }

  // A host job's device buffer back to the pool (mu_ held).
  void release_io(SJob* j) {
    if (!j->host) return;
    if (j->dev_io) {
      pool_.put(j->dev_io, j->io_bytes);
      in_use_ -= j->io_bytes;
      j->dev_io = nullptr;
    }
    if (j->ev_up) {
      (void)hipEventDestroy(j->ev_up);
      j->ev_up = nullptr;
    }
  }

  void fail_active(int st) {
    std::vector<SJob*> keep;
    for (SJob* j : active_) {
      if (j->host && !j->ready) {
        keep.push_back(j);  // the uploader still owns it; it fails it when done
        continue;
      }
      pool_.put(j->ws, j->ws_bytes);
      j->ws = nullptr;
      in_use_ -= j->ws_bytes;
      release_io(j);
      fail(j, st);
    }
    active_.swap(keep);
    broken_ = true;
    cv_done_.notify_all();
  }

  // Uploads (up = true) or downloads of host jobs, one job at a time, on the
  // server's upload or download stream.
  void io_loop(bool up) {
    (void)hipSetDevice(device_);
    hipStream_t s = up ? up_stream_ : down_stream_;
    std::deque<SJob*>& q = up ? up_q_ : down_q_;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_io_.wait(lk, [&] { return io_stop_ || !q.empty(); });
      if (q.empty()) {
        if (io_stop_) return;
        continue;
      }
      SJob* j = q.front();
      q.pop_front();
      const size_t bytes = (size_t)32 << j->cjob.log_n;
      lk.unlock();
      int st;
      if (up) {
        st = stage_h2d(device_, s, j->dev_io, j->cjob.d_in, bytes);
        if (st == ZK_OK && hipEventRecord(j->ev_up, s) != hipSuccess) st = ZK_ERR_HIP;
        if (st) (void)hipStreamSynchronize(s);
      } else {
        st = stage_d2h(device_, s, j->cjob.d_out, j->dev_io, bytes);
        if (st) (void)hipStreamSynchronize(s);
      }
      lk.lock();
      if (up) {
        if (st == ZK_OK && !broken_) {
          j->ready = true;
        } else {
          active_.erase(std::remove(active_.begin(), active_.end(), j), active_.end());
          pool_.put(j->ws, j->ws_bytes);
          j->ws = nullptr;
          in_use_ -= j->ws_bytes;
          release_io(j);
          fail(j, st ? st : ZK_ERR_HIP);
          cv_done_.notify_all();
        }
        ++gen_;
        submissions_.fetch_add(1, std::memory_order_release);  // wakes a waiting scheduler
        cv_work_.notify_one();
      } else {
        release_io(j);
        --downloading_;
        j->done = true;
        j->t_done = now_ns();
        if (st == ZK_OK) {
          ++completed_;
          elements_done_ += 1ull << j->cjob.log_n;
        } else {
          j->status = st;
          ++failed_;
        }
        ++gen_;
        cv_done_.notify_all();
        cv_work_.notify_one();  // memory came back
      }
    }
  }

  void loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      retire();
      if (stop_) {
        for (SJob* j : pending_) fail(j, ZK_ERR_SHUTDOWN);
        pending_.clear();
        cv_done_.notify_all();
      } else {
        admit();
      }
      const bool any_ready =
          std::any_of(active_.begin(), active_.end(), [](const SJob* j) { return j->ready; });
      if (any_ready && inflight_ < kDepth) {
        if (issue() != ZK_OK) {
          // the stream is unusable: wait for what was issued, then fail the rest
          lk.unlock();
          (void)hipStreamSynchronize(stream_);
          lk.lock();
          retire();
          fail_active(ZK_ERR_HIP);
        }
        continue;
      }
      if (stop_ && active_.empty() && inflight_ == 0) break;
      if (inflight_ > 0) {
        // wait for the oldest iteration, but come back at once when a job
        // arrives and an iteration slot is free
        const uint64_t seen = submissions_.load(std::memory_order_acquire);
        hipEvent_t ev = ring_[head_].done;
        const bool slot_free = inflight_ < kDepth;
        lk.unlock();
        while (hipEventQuery(ev) == hipErrorNotReady) {
          if (slot_free && submissions_.load(std::memory_order_acquire) != seen) break;
          std::this_thread::yield();
        }
        lk.lock();
      } else {
        // idle, or pending jobs that do not fit yet: wait until a
        // submission, a release or shutdown changes that -- spinning for a
        // short while first, so that a job arriving soon is not delayed by a
        // thread wake-up, then sleeping
        const uint64_t g = gen_.load(std::memory_order_relaxed);
        lk.unlock();
        const uint64_t t_end = now_ns() + kIdleSpinNs;
        while (gen_.load(std::memory_order_acquire) == g && now_ns() < t_end) std::this_thread::yield();
        lk.lock();
        cv_work_.wait(lk, [&] { return stop_ || gen_.load(std::memory_order_relaxed) != g; });
      }
    }
  }

  const int device_;
  Engine* eng_;
  const uint64_t budget_;
  const uint32_t max_active_;
  Pool pool_;
  hipStream_t stream_ = nullptr;
  hipStream_t up_stream_ = nullptr, down_stream_ = nullptr;
  std::thread uploader_, downloader_;
  std::condition_variable cv_io_;
  std::deque<SJob*> up_q_, down_q_;  // host jobs to upload / to download
  size_t downloading_ = 0;
  bool io_stop_ = false;
  bool broken_ = false;  // the compute stream failed
  uint32_t grid_cap_ = 512;
  size_t ts_off_ = 0, stage_bytes_ = 0;
  Iter ring_[kDepth];
  int head_ = 0, inflight_ = 0;

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::atomic<uint64_t> submissions_{0};
  std::thread thread_;
  bool stop_ = false;
  std::atomic<uint64_t> gen_{0};  // bumped (under mu_) by submit, release and shutdown
  uint64_t next_ticket_ = 0;
  std::deque<SJob*> pending_;
  std::vector<SJob*> active_;  // admitted, passes left to issue
  size_t finishing_count_ = 0; // last pass issued, iteration not retired
  std::unordered_map<uint64_t, std::unique_ptr<SJob>> jobs_;
  uint64_t in_use_ = 0, peak_in_use_ = 0;
  uint64_t submitted_ = 0, completed_ = 0, failed_ = 0, iterations_ = 0, elements_done_ = 0;
};

}  // namespace zkgpu

using namespace zkgpu;

extern "C" {

int zkgpu_server_create(int device_id, uint64_t budget_bytes, uint32_t max_active, uint32_t flags,
                        uint64_t* handle) {
  if (!handle || flags != 0) return ZK_ERR_INVALID;
  *handle = 0;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  if (budget_bytes == 0) budget_bytes = kDefaultBudget;
  if (max_active == 0) max_active = kDefaultMaxActive;
  std::unique_ptr<Server> s(new Server(device_id, e, budget_bytes, max_active));
  st = s->init();
  if (st) return st;
  *handle = (uint64_t)(uintptr_t)s.release();
  return ZK_OK;
}

int zkgpu_server_submit(uint64_t handle, const zkgpu_job* job, uint64_t* ticket) {
  Server* s = (Server*)(uintptr_t)handle;
  if (!s || !job || !ticket) return ZK_ERR_INVALID;
  return s->submit(*job, ticket);
}

int zkgpu_server_poll(uint64_t handle, uint64_t ticket, int32_t wait_ms, zkgpu_server_result* out) {
  Server* s = (Server*)(uintptr_t)handle;
  if (!s) return ZK_ERR_INVALID;
  return s->poll(ticket, wait_ms, out);
}

int zkgpu_server_release(uint64_t handle, uint64_t ticket) {
  Server* s = (Server*)(uintptr_t)handle;
  if (!s) return ZK_ERR_INVALID;
  return s->release(ticket);
}

int zkgpu_server_stats_get(uint64_t handle, zkgpu_server_stats* out) {
  Server* s = (Server*)(uintptr_t)handle;
  if (!s || !out) return ZK_ERR_INVALID;
  s->stats(out);
  return ZK_OK;
}

int zkgpu_server_destroy(uint64_t handle) {
  Server* s = (Server*)(uintptr_t)handle;
  if (!s) return ZK_ERR_INVALID;
  delete s;
  return ZK_OK;
}

}  // extern "C"
