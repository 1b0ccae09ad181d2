// Per-device NTT engine (see engine.hpp).
#include "engine.hpp"

#include "../../include/zkgpu_ntt.h"

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>

namespace zkgpu {

// ---------------------------------------------------------------- plans

static std::mutex g_plan_mu;
static std::map<uint32_t, Plan> g_plan_override[2];  // set at run time (autotuner, tests)
static std::map<uint32_t, Plan> g_plan_file[2];      // tuned plan files

uint32_t plan_tile_log(int kind) { return kind == PLAN_BATCH ? batch_tile_log() : TILE_LOG_MAX; }

bool plan_valid(const Plan& p, uint32_t max_tile_log) {
  if (p.log_n < 1 || p.log_n > MAX_LOG_N || p.npass < 1 || p.npass > MAX_PASSES) return false;
  uint32_t sum = 0;
  for (uint32_t i = 0; i < p.npass; ++i) {
    if (p.digit[i] < 1 || p.digit[i] > max_tile_log) return false;
    if (p.digit[i] + p.log_c[i] > max_tile_log) return false;
    // a tile's threads (8 elements each, fewer for 1- and 2-bit digits) must
    // fit one workgroup of the kernel that runs it (512 threads for 2^12
    // tiles, 256 in batches)
    if (tile_threads(p.digit[i], p.log_c[i], 3) > (1u << (max_tile_log - 3))) return false;
    sum += p.digit[i];
  }
  if (sum != p.log_n) return false;
  if (p.npass == 1) return p.log_c[0] == 0;
  uint32_t log_m = p.log_n;
  for (uint32_t i = 0; i + 1 < p.npass; ++i) {
    const uint32_t log_L = log_m - p.digit[i];  // columns of pass i
    if (p.log_c[i] > log_L) return false;
    log_m = log_L;
  }
  // last pass: its columns run over the first output digit
  return p.log_c[p.npass - 1] <= p.digit[0];
}

bool default_plan(uint32_t log_n, Plan& p, uint32_t max_tile_log) {
  p = Plan();
  if (log_n < 1 || log_n > MAX_LOG_N) return false;
  p.log_n = log_n;
  if (log_n <= max_tile_log) {
    p.npass = 1;
    p.digit[0] = log_n;
    return true;
  }
  p.npass = log_n <= 2 * max_tile_log ? 2 : 3;
  uint32_t rem = log_n;
  for (uint32_t i = 0; i < p.npass; ++i) {
    const uint32_t left = p.npass - i;
    p.digit[i] = (rem + left - 1) / left;
    rem -= p.digit[i];
  }
  // columns per tile: tiles of at most 2^max_tile_log elements, at least
  // ~1024 tiles per pass when the size allows (256 CUs, several tiles each).
  uint32_t log_m = log_n;
  for (uint32_t i = 0; i < p.npass; ++i) {
    const uint32_t avail = (i + 1 < p.npass) ? log_m - p.digit[i] : p.digit[0];
    const uint32_t log_dfts = log_n - p.digit[i];
    uint32_t c = log_dfts > 10 ? log_dfts - 10 : 0;
    c = std::min(c, std::min(avail, max_tile_log - p.digit[i]));
    p.log_c[i] = c;
    log_m -= p.digit[i];
  }
  return plan_valid(p, max_tile_log);
}

// Tuned splits: lines "log_n npass d0 d1 d2 c0 c1 c2 [ms]" in the files named
// by ZKGPU_PLAN_FILE / ZKGPU_BATCH_PLAN_FILE, else plans_gfx950.txt and
// plans_batch_gfx950.txt next to this library.
static void load_plan_file(int kind) {
  const char* env = getenv(kind == PLAN_BATCH ? "ZKGPU_BATCH_PLAN_FILE" : "ZKGPU_PLAN_FILE");
  const char* name = kind == PLAN_BATCH ? "/plans_batch_gfx950.txt" : "/plans_gfx950.txt";
  std::string path;
  if (env) {
    path = env;
  } else {
    Dl_info info;
    if (dladdr((void*)&load_plan_file, &info) && info.dli_fname) {
      path = info.dli_fname;
      const size_t slash = path.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + name;
    }
  }
  if (path.empty()) return;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  char line[256];
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#') continue;
    unsigned v[8];
    if (sscanf(line, "%u %u %u %u %u %u %u %u", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]) != 8)
      continue;
    Plan p;
    p.log_n = v[0];
    p.npass = v[1];
    for (int i = 0; i < 3; ++i) {
      p.digit[i] = v[2 + i];
      p.log_c[i] = v[5 + i];
    }
    if (plan_valid(p, plan_tile_log(kind))) g_plan_file[kind][p.log_n] = p;
  }
  fclose(f);
}

static void load_plan_files_once() {
  static std::once_flag once;
  std::call_once(once, [] {
    std::lock_guard<std::mutex> g(g_plan_mu);
    load_plan_file(PLAN_SINGLE);
    load_plan_file(PLAN_BATCH);
  });
}

Plan plan_for(uint32_t log_n, int kind) {
  kind = kind == PLAN_BATCH ? PLAN_BATCH : PLAN_SINGLE;
  load_plan_files_once();
  {
    std::lock_guard<std::mutex> g(g_plan_mu);
    auto it = g_plan_override[kind].find(log_n);
    if (it != g_plan_override[kind].end()) return it->second;
    auto f = g_plan_file[kind].find(log_n);
    if (f != g_plan_file[kind].end()) return f->second;
  }
  Plan p;
  default_plan(log_n, p, plan_tile_log(kind));
  return p;
}
void set_plan_override(const Plan& p, int kind) {
  kind = kind == PLAN_BATCH ? PLAN_BATCH : PLAN_SINGLE;
  std::lock_guard<std::mutex> g(g_plan_mu);
  g_plan_override[kind][p.log_n] = p;
}
void erase_plan_override(uint32_t log_n, int kind) {
  kind = kind == PLAN_BATCH ? PLAN_BATCH : PLAN_SINGLE;
  std::lock_guard<std::mutex> g(g_plan_mu);
  g_plan_override[kind].erase(log_n);
}
void clear_plan_overrides() {
  std::lock_guard<std::mutex> g(g_plan_mu);
  g_plan_override[0].clear();
  g_plan_override[1].clear();
}

// ---------------------------------------------------------------- engine

DevBuf::~DevBuf() {
  if (ptr) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    (void)hipFree(ptr);
    if (cur != device) (void)hipSetDevice(cur);
  }
}

static std::mutex g_engines_mu;
static std::map<int, std::unique_ptr<Engine>> g_engines;

Engine* Engine::get(int device, int* status) {
  static const int count = [] {
    int c = 0;
    return hipGetDeviceCount(&c) == hipSuccess ? c : 0;
  }();
  if (device < 0 || device >= count) {
    if (status) *status = ZK_ERR_INVALID;
    return nullptr;
  }
  Engine* e = nullptr;
  {
    std::lock_guard<std::mutex> g(g_engines_mu);
    auto it = g_engines.find(device);
    if (it == g_engines.end()) {
      if (hipSetDevice(device) != hipSuccess || configure_kernels() != hipSuccess) {
        if (status) *status = ZK_ERR_HIP;
        return nullptr;
      }
      noalloc_init_device(device);
      it = g_engines.emplace(device, std::unique_ptr<Engine>(new Engine(device))).first;
    }
    e = it->second.get();
  }
  // the device is per host thread; switch only when needed
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) {
    if (hipSetDevice(device) != hipSuccess) {
      if (status) *status = ZK_ERR_HIP;
      return nullptr;
    }
  }
  if (status) *status = ZK_OK;
  return e;
}

Engine::Engine(int device) : device_(device) {
  size_t mb = 32768;  // resident table budget
  if (const char* e = getenv("ZKGPU_TABLE_CACHE_MB")) mb = (size_t)strtoull(e, nullptr, 10);
  table_cap_ = mb << 20;
  size_t ws_mb = 65536;  // device-API workspaces: idle ones are freed above this
  if (const char* e = getenv("ZKGPU_WORKSPACE_CAP_MB")) ws_mb = (size_t)strtoull(e, nullptr, 10);
  ws_cap_ = ws_mb << 20;
  size_t host_mb = 32768;  // HBM budget of the host-buffer entry points
  if (const char* e = getenv("ZKGPU_HOST_BUDGET_MB")) host_mb = (size_t)strtoull(e, nullptr, 10);
  host_budget_ = host_budget_default_ = host_mb << 20;
  unsigned nctx = 4;  // host-call contexts (streams) shared by every caller thread
  if (const char* e = getenv("ZKGPU_HOST_STREAMS")) nctx = (unsigned)strtoul(e, nullptr, 10);
  nctx = std::max(1u, std::min(nctx, 64u));
  for (unsigned i = 0; i < nctx; ++i) host_ctx_.emplace_back(new HostCtx());
  (void)hipStreamCreateWithFlags(&build_stream_, hipStreamNonBlocking);
}

std::string Engine::last_error() {
  std::lock_guard<std::mutex> g(mu_);
  return last_error_;
}
void Engine::set_error(const std::string& e) {
  std::lock_guard<std::mutex> g(mu_);
  last_error_ = e;
}

static std::atomic<uint32_t> g_fail_skip{0}, g_fail_count{0};
void Engine::debug_fail_allocs(uint32_t skip, uint32_t count) {
  g_fail_count.store(0);
  g_fail_skip.store(skip);
  g_fail_count.store(count);
}
// true when the test hook asks this allocation to fail
static bool injected_alloc_failure() {
  if (g_fail_count.load(std::memory_order_relaxed) == 0) return false;
  uint32_t sk = g_fail_skip.load();
  while (sk > 0) {
    if (g_fail_skip.compare_exchange_weak(sk, sk - 1)) return false;
  }
  uint32_t c = g_fail_count.load();
  while (c > 0) {
    if (g_fail_count.compare_exchange_weak(c, c - 1)) return true;
  }
  return false;
}

std::shared_ptr<DevBuf> Engine::alloc(size_t bytes) {
  if (injected_alloc_failure()) return nullptr;
  auto b = std::make_shared<DevBuf>();
  b->device = device_;
  b->bytes = bytes;
  if (hipMalloc(&b->ptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    b->ptr = nullptr;
    return nullptr;
  }
  return b;
}

Marker* Engine::marker(hipStream_t s) {
  // mu_ held
  auto& m = markers_[s];
  if (!m) {
    m.reset(new Marker());
    m->stream = s;
    // completion only: nothing reads device memory through the marker, so no
    // system-scope release (an L2 write-back) is attached to the dispatch
    if (hipEventCreateWithFlags(&m->ev, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
      (void)hipGetLastError();
      m->ev = nullptr;
    }
  }
  return m.get();
}

Marker* Engine::stream_marker(hipStream_t s) {
  std::lock_guard<std::mutex> g(mu_);
  return marker(s);
}

void Engine::note_users(const JobTables& t, Marker* m) {
  std::lock_guard<std::mutex> g(mu_);
  auto add = [&](const std::shared_ptr<DevBuf>& b) {
    if (b && std::find(b->users.begin(), b->users.end(), m) == b->users.end()) b->users.push_back(m);
  };
  add(t.roots);
  for (const auto& w : t.w) add(w);
  add(t.pre);
  add(t.post);
}

void Engine::set_pass_timing(bool on) {
  timing_.store(on);
}

void Engine::timed_events(hipEvent_t* a, hipEvent_t* b) {
  std::lock_guard<std::mutex> g(timing_mu_);
  for (hipEvent_t* e : {a, b}) {
    if (!timing_pool_.empty()) {
      *e = timing_pool_.back();
      timing_pool_.pop_back();
    } else if (hipEventCreateWithFlags(e, hipEventDisableSystemFence) != hipSuccess) {
      (void)hipGetLastError();
      *e = nullptr;
    }
  }
}

void Engine::finish_timed(hipEvent_t a, hipEvent_t b, const char* name) {
  std::lock_guard<std::mutex> g(timing_mu_);
  if (!a || !b) {
    if (a) timing_pool_.push_back(a);
    if (b) timing_pool_.push_back(b);
    return;
  }
  TimedPass tp;
  tp.a = a;
  tp.b = b;
  tp.name = name;
  timed_.push_back(tp);
}

int Engine::pass_times(std::vector<KernelTime>& out) {
  out.clear();
  std::lock_guard<std::mutex> g(timing_mu_);
  int st = ZK_OK;
  for (auto& tp : timed_) {
    float ms = 0;
    if (hipEventSynchronize(tp.b) != hipSuccess || hipEventElapsedTime(&ms, tp.a, tp.b) != hipSuccess) {
      (void)hipGetLastError();
      st = ZK_ERR_HIP;
    }
    auto it = std::find_if(out.begin(), out.end(), [&](const KernelTime& k) { return k.name == tp.name; });
    if (it == out.end()) {
      out.push_back(KernelTime());
      it = out.end() - 1;
      it->name = tp.name;
    }
    ++it->launches;
    it->total_ms += ms;
    timing_pool_.push_back(tp.a);
    timing_pool_.push_back(tp.b);
  }
  timed_.clear();
  return st;
}

WsLease Engine::lease_workspace(hipStream_t s, size_t bytes, int slot, int* status) {
  WsLease l;
  *status = ZK_OK;
  Workspace* w = nullptr;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto& p = ws_[std::make_pair(s, slot)];
    if (!p) {
      p.reset(new Workspace());
      p->stream = s;
    }
    w = p.get();
  }
  l.lk_ = std::unique_lock<std::mutex>(w->mu);
  if (w->buf && w->buf->bytes >= bytes) {
    l.buf_ = w->buf;
    return l;
  }
  if (w->buf) {
    // the old buffer may still be read by work queued on s (only on s: the
    // lease is held while such work is enqueued)
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> g(mu_);
    ws_bytes_ -= w->buf->bytes;
    w->buf.reset();
  }
  {
    // over the cap: free workspaces of other streams that are neither leased
    // nor still in use by their stream's queued work
    std::lock_guard<std::mutex> g(mu_);
    for (auto it = ws_.begin(); ws_bytes_ + bytes > ws_cap_ && it != ws_.end(); ++it) {
      Workspace* o = it->second.get();
      if (o == w || !o->buf) continue;
      std::unique_lock<std::mutex> ol(o->mu, std::try_to_lock);
      if (!ol.owns_lock()) continue;
      auto mk = markers_.find(o->stream);
      const bool idle = mk == markers_.end() || !mk->second->ev || hipEventQuery(mk->second->ev) == hipSuccess;
      if (!idle) continue;
      ws_bytes_ -= o->buf->bytes;
      o->buf.reset();
    }
  }
  auto b = alloc(bytes);
  if (!b) {
    l.lk_.unlock();
    *status = ZK_ERR_NOMEM;
    return WsLease();
  }
  {
    std::lock_guard<std::mutex> g(mu_);
    ws_bytes_ += bytes;
  }
  w->buf = b;
  l.buf_ = b;
  return l;
}

int Engine::acquire_host(size_t io_bytes, size_t planar_bytes, size_t aux_bytes, HostCtx** out) {
  *out = nullptr;
  const size_t need[3] = {io_bytes, planar_bytes, aux_bytes};
  auto have = [](const std::shared_ptr<DevBuf>& b) { return b ? b->bytes : (size_t)0; };
  auto slots = [](HostCtx* h) {
    std::array<std::shared_ptr<DevBuf>*, 3> a = {&h->io, &h->planar, &h->aux};
    return a;
  };
  std::unique_lock<std::mutex> lk(host_mu_);
  HostCtx* c = nullptr;
  bool grows[3] = {false, false, false};
  for (;;) {
    c = nullptr;
    size_t best = SIZE_MAX;
    bool any_busy = false;
    for (auto& h : host_ctx_) {
      if (h->busy) {
        any_busy = true;
        continue;
      }
      // prefer the idle context whose buffers need the least growth
      size_t g = 0;
      auto sl = slots(h.get());
      for (int i = 0; i < 3; ++i) g += need[i] > have(*sl[i]) ? need[i] : 0;
      if (g < best) {
        best = g;
        c = h.get();
      }
    }
    if (c) {
      // buffers that grow are replaced: the old ones stop counting
      auto sl = slots(c);
      size_t freed = 0;
      for (int i = 0; i < 3; ++i)
        if (need[i] > have(*sl[i])) freed += have(*sl[i]);
      if (host_held_ - freed + best > host_budget_) {
        // release the buffers of the other idle contexts (their streams are idle)
        for (auto& h : host_ctx_) {
          if (h->busy || h.get() == c) continue;
          host_held_ -= h->held();
          h->io.reset();
          h->planar.reset();
          h->aux.reset();
        }
      }
      // over budget with other calls running: wait for them (a call larger
      // than the whole budget still runs, alone)
      if (host_held_ - freed + best <= host_budget_ || !any_busy) {
        // reserve the growth now, so that concurrent acquirers see it
        for (int i = 0; i < 3; ++i) {
          grows[i] = need[i] > have(*sl[i]);
          if (!grows[i]) continue;
          host_held_ -= have(*sl[i]);
          sl[i]->reset();  // the context's stream is idle
          host_held_ += need[i];
        }
        host_peak_ = std::max(host_peak_, host_held_);
        break;
      }
    }
    host_cv_.wait(lk);
  }
  c->busy = true;
  lk.unlock();
  int st = ZK_OK;
  if (!c->stream && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    c->stream = nullptr;
    st = ZK_ERR_HIP;
  }
  auto sl = slots(c);
  size_t unallocated = 0;
  for (int i = 0; i < 3; ++i) {
    if (!grows[i]) continue;
    if (st == ZK_OK) *sl[i] = alloc(need[i]);
    if (!*sl[i]) {
      unallocated += need[i];
      if (st == ZK_OK) st = ZK_ERR_NOMEM;
    }
  }
  if (st) {
    {
      std::lock_guard<std::mutex> g(host_mu_);
      host_held_ -= unallocated;
    }
    release_host(c);
    return st;
  }
  *out = c;
  return ZK_OK;
}

void Engine::release_host(HostCtx* c) {
  {
    std::lock_guard<std::mutex> g(host_mu_);
    c->busy = false;
  }
  host_cv_.notify_all();
}

size_t Engine::host_held_bytes() {
  std::lock_guard<std::mutex> g(host_mu_);
  return host_held_;
}
size_t Engine::host_peak_bytes() {
  std::lock_guard<std::mutex> g(host_mu_);
  return host_peak_;
}
size_t Engine::set_host_budget(size_t bytes) {
  std::lock_guard<std::mutex> g(host_mu_);
  const size_t prev = host_budget_;
  host_budget_ = bytes ? bytes : host_budget_default_;
  for (auto& h : host_ctx_) {
    if (host_held_ <= host_budget_) break;
    if (h->busy) continue;
    host_held_ -= h->held();
    h->io.reset();
    h->planar.reset();
    h->aux.reset();
  }
  host_peak_ = host_held_;
  return prev;
}

std::shared_ptr<DevBuf> Engine::cached_table(const TableKey& key) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = tables_.find(key);
  if (it == tables_.end()) return nullptr;
  lru_.remove(key);
  lru_.push_front(key);
  return it->second;
}

// A new (unpublished) table buffer, after evicting least recently used tables
// that no job holds. An evicted table is freed once every stream that ran
// work reading it has passed that work (its users' markers).
std::shared_ptr<DevBuf> Engine::new_table(size_t bytes) {
  std::vector<std::shared_ptr<DevBuf>> victims;
  {
    std::lock_guard<std::mutex> g(mu_);
    while (table_bytes_ + bytes > table_cap_ && !lru_.empty()) {
      bool evicted = false;
      for (auto r = lru_.rbegin(); r != lru_.rend(); ++r) {
        auto t = tables_.find(*r);
        if (t != tables_.end() && t->second.use_count() == 1) {
          table_bytes_ -= t->second->bytes;
          victims.push_back(t->second);
          tables_.erase(t);
          lru_.erase(std::next(r).base());
          evicted = true;
          break;
        }
      }
      if (!evicted) break;
    }
  }
  for (auto& v : victims) {
    for (Marker* m : v->users)
      if (m->ev) (void)hipEventSynchronize(m->ev);
  }
  victims.clear();
  return alloc(bytes);
}

void Engine::publish_tables(std::vector<std::pair<TableKey, std::shared_ptr<DevBuf>>>& fresh) {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& kv : fresh) {
    if (tables_.count(kv.first)) continue;
    tables_[kv.first] = kv.second;
    lru_.push_front(kv.first);
    table_bytes_ += kv.second->bytes;
  }
  fresh.clear();
}

size_t Engine::table_bytes() {
  std::lock_guard<std::mutex> g(mu_);
  return table_bytes_;
}
void Engine::drop_tables() {
  std::lock_guard<std::mutex> b(build_mu_);
  std::vector<std::shared_ptr<DevBuf>> all;
  {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& kv : tables_) all.push_back(kv.second);
    tables_.clear();
    lru_.clear();
    table_bytes_ = 0;
  }
  // tables still held by jobs (batches, served jobs) are freed when those
  // release them; the rest once their users' work has finished
  for (auto& v : all) {
    if (v.use_count() > 1) continue;
    for (Marker* m : v->users)
      if (m->ev) (void)hipEventSynchronize(m->ev);
  }
}

static TableKey make_key(char kind, uint32_t a, uint32_t b, uint32_t c, const host::Fr& x,
                         const host::Fr* y = nullptr) {
  TableKey k;
  k.kind = kind;
  k.a = a;
  k.b = b;
  k.c = c;
  for (int i = 0; i < 4; ++i) {
    k.e[i] = x.w[i];
    k.e[4 + i] = y ? y->w[i] : 0;
  }
  return k;
}

// lo[x] = base^(x*step), x < 2^h; hi[y] = base^(y*step*2^h), y < 2^(log_count-h); 48-byte limbs.
int Engine::build_pow_lohi(const host::Fr& base, uint64_t step, uint32_t log_count, uint32_t h,
                           std::shared_ptr<DevBuf>& lo, std::shared_ptr<DevBuf>& hi, hipStream_t s) {
  uint32_t w[8];
  host::to_words32(w, base);
  const uint32_t nlo = 1u << h, nhi = 1u << (log_count - h);
  lo = alloc((size_t)nlo * 48);
  hi = alloc((size_t)nhi * 48);
  if (!lo || !hi) return ZK_ERR_NOMEM;
  if (launch_pow_table(lo->ptr, w, step, nlo, true, s) != hipSuccess) return ZK_ERR_HIP;
  if (launch_pow_table(hi->ptr, w, step << h, nhi, true, s) != hipSuccess) return ZK_ERR_HIP;
  return ZK_OK;
}

int Engine::prepare(const JobSpec& job, JobTables& t, hipStream_t s, int plan_kind) {
  std::lock_guard<std::mutex> build_lock(build_mu_);
  if (!s) s = build_stream_;
  t = JobTables();
  t.plan = plan_for(job.log_n, plan_kind);
  if (!plan_valid(t.plan, plan_tile_log(plan_kind))) return ZK_ERR_INVALID;
  const Plan& p = t.plan;
  const uint32_t k = job.log_n;
  std::vector<std::shared_ptr<DevBuf>> temps;  // kept alive until the build is synchronised
  // tables built by this call: published only once their build has finished,
  // so an error on the way leaves no unbuilt table in the cache
  std::vector<std::pair<TableKey, std::shared_ptr<DevBuf>>> fresh;
  // on an error after builds were enqueued, let them finish before their
  // buffers are dropped
  struct BuildGuard {
    hipStream_t s;
    const std::vector<std::pair<TableKey, std::shared_ptr<DevBuf>>>& fresh;
    ~BuildGuard() {
      if (!fresh.empty() && hipStreamSynchronize(s) != hipSuccess) (void)hipGetLastError();
    }
  } guard{s, fresh};
  auto fetch = [&](const TableKey& key, size_t bytes, bool* is_new) -> std::shared_ptr<DevBuf> {
    auto b = cached_table(key);
    *is_new = !b;
    if (!b) {
      b = new_table(bytes);
      if (b) fresh.emplace_back(key, b);
    }
    return b;
  };

  // roots of the tile DFTs: omega_Nt^e, Nt = 2^min(k, 12)
  {
    const uint32_t lnt = std::min(k, TILE_LOG_MAX);
    bool is_new = false;
    t.roots = fetch(make_key('R', k, lnt, 0, job.omega), (size_t)48 << lnt, &is_new);
    if (!t.roots) return ZK_ERR_NOMEM;
    if (is_new) {
      uint32_t w[8];
      host::to_words32(w, job.omega);
      if (launch_pow_table(t.roots->ptr, w, 1ull << (k - lnt), 1u << lnt, true, s) != hipSuccess) return ZK_ERR_HIP;
    }
  }
  // inter-pass twiddles W_p[t * L + i] = omega_m^(t * i); for the inverse,
  // the last boundary's also carry the divisor
  const bool fold = fold_divisor(job.op, p);
  uint32_t divw[8];
  host::to_words32(divw, job.divisor);
  uint32_t log_m = k;
  for (uint32_t i = 0; i + 1 < p.npass; ++i) {
    const uint32_t log_L = log_m - p.digit[i];
    const bool scaled = fold && i + 2 == p.npass;
    bool is_new = false;
    const TableKey key = scaled ? make_key('V', k, log_m, log_L, job.omega, &job.divisor)
                                : make_key('W', k, log_m, log_L, job.omega);
    t.w[i] = fetch(key, (size_t)32 << log_m, &is_new);
    if (!t.w[i]) return ZK_ERR_NOMEM;
    if (is_new) {
      const uint32_t h = (log_m + 1) / 2;
      std::shared_ptr<DevBuf> lo, hi;
      int st = build_pow_lohi(job.omega, 1ull << (k - log_m), log_m, h, lo, hi, s);
      if (st) return st;
      if (launch_twiddle_table(t.w[i]->ptr, lo->ptr, hi->ptr, h, log_m, log_L, scaled ? divw : nullptr, s) !=
          hipSuccess)
        return ZK_ERR_HIP;
      temps.push_back(lo);
      temps.push_back(hi);
    }
    log_m = log_L;
  }
  // coset powers
  if (job.op == OP_COSET_FWD || job.op == OP_COSET_INV) {
    const bool post = job.op == OP_COSET_INV;
    const TableKey key = make_key(post ? 'Q' : 'P', k, 0, 0, job.coset, post ? &job.divisor : nullptr);
    bool is_new = false;
    auto tab = fetch(key, (size_t)32 << k, &is_new);
    if (!tab) return ZK_ERR_NOMEM;
    if (is_new) {
      const uint32_t h = (k + 1) / 2;
      std::shared_ptr<DevBuf> lo, hi;
      int st = build_pow_lohi(job.coset, 1, k, h, lo, hi, s);
      if (st) return st;
      uint32_t sc[8];
      host::to_words32(sc, job.divisor);
      if (launch_pow_table_2level(tab->ptr, lo->ptr, hi->ptr, h, 1ull << k, post ? sc : nullptr, s) != hipSuccess)
        return ZK_ERR_HIP;
      temps.push_back(lo);
      temps.push_back(hi);
    }
    (post ? t.post : t.pre) = tab;
  }
  if (job.op == OP_INV && !fold) {
    t.uscale = true;
    host::to_limbs261(t.scale261, job.divisor);
  }
  if (!fresh.empty()) {
    if (hipStreamSynchronize(s) != hipSuccess) {
      (void)hipGetLastError();
      return ZK_ERR_HIP;
    }
    publish_tables(fresh);
  }
  return ZK_OK;
}

TablePtrs table_ptrs(const JobTables& t) {
  TablePtrs p;
  p.plan = t.plan;
  p.roots = t.roots ? t.roots->ptr : nullptr;
  for (uint32_t i = 0; i < MAX_PASSES; ++i) p.w[i] = t.w[i] ? t.w[i]->ptr : nullptr;
  p.pre = t.pre ? t.pre->ptr : nullptr;
  p.post = t.post ? t.post->ptr : nullptr;
  p.uscale = t.uscale;
  memcpy(p.scale261, t.scale261, sizeof(p.scale261));
  return p;
}

int build_pass_descs(const JobSpec& job, const TablePtrs& t, void* ws, PassDesc* descs, uint32_t* ndesc) {
  const Plan& p = t.plan;
  const uint32_t k = job.log_n;
  const uint32_t lnt = std::min(k, TILE_LOG_MAX);
  if (p.npass > 1 && !ws) return ZK_ERR_INVALID;
  const uint32_t n = 1u << k;
  uint32_t log_m = k, prev_log_m = k;
  for (uint32_t i = 0; i < p.npass; ++i) {
    PassDesc d;
    memset(&d, 0, sizeof(d));
    const uint32_t N = p.digit[i], c = p.log_c[i];
    const bool last = i + 1 == p.npass;
    d.roots = t.roots;
    d.root_log = lnt;
    d.log_n_tile = N;
    d.log_c = c;
    // input: the caller's packed data in the first pass, the planar
    // workspace after it; the previous pass's twiddles are applied on load
    if (i == 0) {
      d.in = job.in;
      d.pre = t.pre;
      d.pre_mask = n - 1;
    } else {
      d.in = ws;
      d.in_plane = n;
      d.flags |= PASS_IN_PLANAR;
      d.pre = t.w[i - 1];
      d.pre_mask = (1u << prev_log_m) - 1u;
    }
    if (!last) {
      const uint32_t log_L = log_m - N;
      d.out = ws;
      d.out_plane = n;
      d.flags |= PASS_OUT_PLANAR;
      d.log_n_lo = log_L - c;
      d.num_tiles = 1u << (k - N - c);
      d.in_lo = d.out_lo = 1u << c;
      d.in_hi = d.out_hi = 1u << log_m;
      d.in_ts = d.out_ts = 1u << log_L;
      d.in_cs = d.out_cs = 1;
      prev_log_m = log_m;
      log_m = log_L;
    } else if (p.npass == 1) {
      d.out = job.out;
      d.log_n_lo = 0;
      d.num_tiles = 1;
      d.in_ts = d.out_ts = 1;
    } else {
      // columns over the first output digit k_1 (C consecutive values), one
      // tile row per value of the middle digits
      const uint32_t n1 = p.digit[0];
      const uint32_t log_hi = k - n1 - N;  // middle digits (0 or the single middle digit)
      d.out = job.out;
      d.log_n_lo = n1 - c;
      d.num_tiles = 1u << (n1 - c + log_hi);
      d.in_cs = 1u << (k - n1);
      d.in_lo = 1u << (k - n1 + c);
      d.in_hi = 1u << N;
      d.in_ts = 1;
      d.out_cs = 1;
      d.out_lo = 1u << c;
      d.out_hi = 1u << n1;
      d.out_ts = 1u << (k - N);
    }
    if (last) {
      d.post = t.post;
      d.flags |= PASS_CANON | (t.uscale ? PASS_USCALE : 0u);
      if (t.uscale) memcpy(d.scale, t.scale261, sizeof(t.scale261));
    }
    descs[i] = d;
  }
  *ndesc = p.npass;
  return ZK_OK;
}

int job_from_c(const zkgpu_job& j, JobSpec& out, bool out_optional) {
  if (!j.d_in || (!j.d_out && !out_optional) || j.log_n < 1 || j.log_n > MAX_LOG_N) return ZK_ERR_INVALID;
  if (j.op < OP_FWD || j.op > OP_COSET_INV) return ZK_ERR_INVALID;
  out = JobSpec();
  out.in = j.d_in;
  out.out = j.d_out;
  out.log_n = j.log_n;
  out.op = j.op;
  if (j.flags & ZKGPU_JOB_STD_DOMAIN) {
    const bool inverse = j.op == OP_INV || j.op == OP_COSET_INV;
    uint64_t w[4];
    zkgpu_domain_constant(j.log_n, inverse ? 1 : 0, w);
    out.omega = host::from_ptr(w);
    zkgpu_domain_constant(j.log_n, 2, w);
    out.divisor = host::from_ptr(w);
    zkgpu_domain_constant(j.log_n, j.op == OP_COSET_INV ? 4 : 3, w);
    out.coset = host::from_ptr(w);
  } else {
    out.omega = host::from_ptr(j.omega);
    out.divisor = host::from_ptr(j.divisor);
    out.coset = host::from_ptr(j.coset);
  }
  return ZK_OK;
}

int Engine::run(const JobSpec& job, hipStream_t s, void* planar_ws) {
  if (job.log_n < 1 || job.log_n > MAX_LOG_N || !job.in || !job.out) return ZK_ERR_INVALID;
  if (job.op < OP_FWD || job.op > OP_COSET_INV) return ZK_ERR_INVALID;
  JobTables t;
  int st = prepare(job, t, s);
  if (st) return st;
  WsLease lease;  // held until every pass is enqueued
  void* ws = planar_ws;
  if (t.plan.npass > 1 && !ws) {
    lease = lease_workspace(s, (size_t)36 << job.log_n, WS_PLANAR, &st);  // planar limb form
    if (st) return st;
    ws = lease.ptr();
  }
  PassDesc descs[MAX_PASSES];
  uint32_t nd = 0;
  st = build_descs(job, t, ws, descs, &nd);
  if (st) return st;
  // the stream's marker is tied to the last pass's dispatch (no extra packet)
  Marker* m = stream_marker(s);
  const bool timing = timing_.load(std::memory_order_relaxed);
  for (uint32_t i = 0; i < nd; ++i) {
    hipEvent_t a = nullptr, b = nullptr;
    if (timing) {
      std::lock_guard<std::mutex> g(timing_mu_);
      for (hipEvent_t* e : {&a, &b}) {
        if (!timing_pool_.empty()) {
          *e = timing_pool_.back();
          timing_pool_.pop_back();
        } else if (hipEventCreateWithFlags(e, hipEventDisableSystemFence) != hipSuccess) {
          (void)hipGetLastError();
          *e = nullptr;
        }
      }
    } else if (i + 1 == nd) {
      b = m->ev;
    }
    if (launch_pass(descs[i], s, a, b) != hipSuccess) return ZK_ERR_HIP;
    if (timing && a && b) {
      std::lock_guard<std::mutex> g(timing_mu_);
      TimedPass tp;
      tp.a = a;
      tp.b = b;
      tp.name = std::string(pass_kernel_narrow(descs[i]) ? "ntt_pass_kernel_narrow<"
                            : pass_kernel_quad(descs[i]) ? "ntt_pass_kernel_quad<"
                                                         : "ntt_pass_kernel<") +
                std::to_string(descs[i].log_n_tile) + ">";
      timed_.push_back(tp);
    }
  }
  if (timing && m->ev && hipEventRecord(m->ev, s) != hipSuccess) return ZK_ERR_HIP;
  note_users(t, m);
  return ZK_OK;
}

}  // namespace zkgpu
