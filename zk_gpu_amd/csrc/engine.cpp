// Per-device NTT engine (see engine.hpp).
#include "engine.hpp"

#include "../../include/zkgpu_ntt.h"

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

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
}

std::string Engine::last_error() {
  std::lock_guard<std::mutex> g(mu_);
  return last_error_;
}
void Engine::set_error(const std::string& e) {
  std::lock_guard<std::mutex> g(mu_);
  last_error_ = e;
}

std::shared_ptr<DevBuf> Engine::alloc(size_t bytes) {
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

hipStream_t Engine::thread_stream() {
  thread_local std::map<int, hipStream_t> streams;
  auto it = streams.find(device_);
  if (it != streams.end()) return it->second;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  streams[device_] = s;
  return s;
}

void* Engine::workspace(hipStream_t s, size_t bytes, int slot) {
  std::lock_guard<std::mutex> g(mu_);
  auto key = std::make_pair(s, slot);
  auto it = ws_.find(key);
  if (it != ws_.end() && it->second->bytes >= bytes) return it->second->ptr;
  if (it != ws_.end()) {
    // the old buffer may still be in use by work queued on s
    (void)hipStreamSynchronize(s);
    ws_.erase(it);
  }
  auto b = alloc(bytes);
  if (!b) return nullptr;
  ws_[key] = b;
  return b->ptr;
}

std::shared_ptr<DevBuf> Engine::table(const TableKey& key, size_t bytes, bool* fresh) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = tables_.find(key);
  if (it != tables_.end()) {
    lru_.remove(key);
    lru_.push_front(key);
    *fresh = false;
    return it->second;
  }
  // evict least recently used tables nobody else holds
  while (table_bytes_ + bytes > table_cap_ && !lru_.empty()) {
    bool evicted = false;
    for (auto r = lru_.rbegin(); r != lru_.rend(); ++r) {
      auto t = tables_.find(*r);
      if (t != tables_.end() && t->second.use_count() == 1) {
        (void)hipDeviceSynchronize();  // work on other streams may still read it
        table_bytes_ -= t->second->bytes;
        tables_.erase(t);
        lru_.erase(std::next(r).base());
        evicted = true;
        break;
      }
    }
    if (!evicted) break;
  }
  auto b = alloc(bytes);
  if (!b) return nullptr;
  tables_[key] = b;
  lru_.push_front(key);
  table_bytes_ += bytes;
  *fresh = true;
  return b;
}

size_t Engine::table_bytes() {
  std::lock_guard<std::mutex> g(mu_);
  return table_bytes_;
}
void Engine::drop_tables() {
  (void)hipDeviceSynchronize();
  std::lock_guard<std::mutex> g(mu_);
  tables_.clear();
  lru_.clear();
  table_bytes_ = 0;
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
  t = JobTables();
  t.plan = plan_for(job.log_n, plan_kind);
  if (!plan_valid(t.plan, plan_tile_log(plan_kind))) return ZK_ERR_INVALID;
  const Plan& p = t.plan;
  const uint32_t k = job.log_n;
  std::vector<std::shared_ptr<DevBuf>> temps;  // kept alive until the build is synchronised
  bool built = false;

  // roots of the tile DFTs: omega_Nt^e, Nt = 2^min(k, 12)
  {
    const uint32_t lnt = std::min(k, TILE_LOG_MAX);
    bool fresh = false;
    t.roots = table(make_key('R', k, lnt, 0, job.omega), (size_t)48 << lnt, &fresh);
    if (!t.roots) return ZK_ERR_NOMEM;
    if (fresh) {
      uint32_t w[8];
      host::to_words32(w, job.omega);
      if (launch_pow_table(t.roots->ptr, w, 1ull << (k - lnt), 1u << lnt, true, s) != hipSuccess) return ZK_ERR_HIP;
      built = true;
    }
  }
  // inter-pass twiddles W_p[t * L + i] = omega_m^(t * i)
  uint32_t log_m = k;
  for (uint32_t i = 0; i + 1 < p.npass; ++i) {
    const uint32_t log_L = log_m - p.digit[i];
    bool fresh = false;
    t.w[i] = table(make_key('W', k, log_m, log_L, job.omega), (size_t)32 << log_m, &fresh);
    if (!t.w[i]) return ZK_ERR_NOMEM;
    if (fresh) {
      const uint32_t h = (log_m + 1) / 2;
      std::shared_ptr<DevBuf> lo, hi;
      int st = build_pow_lohi(job.omega, 1ull << (k - log_m), log_m, h, lo, hi, s);
      if (st) return st;
      if (launch_twiddle_table(t.w[i]->ptr, lo->ptr, hi->ptr, h, log_m, log_L, s) != hipSuccess) return ZK_ERR_HIP;
      temps.push_back(lo);
      temps.push_back(hi);
      built = true;
    }
    log_m = log_L;
  }
  // coset powers
  if (job.op == OP_COSET_FWD || job.op == OP_COSET_INV) {
    const bool post = job.op == OP_COSET_INV;
    const TableKey key = make_key(post ? 'Q' : 'P', k, 0, 0, job.coset, post ? &job.divisor : nullptr);
    bool fresh = false;
    auto tab = table(key, (size_t)32 << k, &fresh);
    if (!tab) return ZK_ERR_NOMEM;
    if (fresh) {
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
      built = true;
    }
    (post ? t.post : t.pre) = tab;
  }
  if (job.op == OP_INV) {
    t.uscale = true;
    host::to_limbs261(t.scale261, job.divisor);
  }
  if (built) {
    // publish only finished tables: other streams may use them right away
    if (hipStreamSynchronize(s) != hipSuccess) return ZK_ERR_HIP;
  }
  return ZK_OK;
}

int Engine::build_descs(const JobSpec& job, const JobTables& t, void* ws, PassDesc* descs, uint32_t* ndesc) const {
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
    d.roots = t.roots->ptr;
    d.root_log = lnt;
    d.log_n_tile = N;
    d.log_c = c;
    // input: the caller's packed data in the first pass, the planar
    // workspace after it; the previous pass's twiddles are applied on load
    if (i == 0) {
      d.in = job.in;
      d.pre = t.pre ? t.pre->ptr : nullptr;
      d.pre_mask = n - 1;
    } else {
      d.in = ws;
      d.in_plane = n;
      d.flags |= PASS_IN_PLANAR;
      d.pre = t.w[i - 1]->ptr;
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
      d.post = t.post ? t.post->ptr : nullptr;
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

int Engine::run(const JobSpec& job, hipStream_t s) {
  if (job.log_n < 1 || job.log_n > MAX_LOG_N || !job.in || !job.out) return ZK_ERR_INVALID;
  if (job.op < OP_FWD || job.op > OP_COSET_INV) return ZK_ERR_INVALID;
  JobTables t;
  int st = prepare(job, t, s);
  if (st) return st;
  void* ws = nullptr;
  if (t.plan.npass > 1) {
    ws = workspace(s, (size_t)36 << job.log_n, WS_PLANAR);  // planar limb form
    if (!ws) return ZK_ERR_NOMEM;
  }
  PassDesc descs[MAX_PASSES];
  uint32_t nd = 0;
  st = build_descs(job, t, ws, descs, &nd);
  if (st) return st;
  for (uint32_t i = 0; i < nd; ++i) {
    if (launch_pass(descs[i], s) != hipSuccess) return ZK_ERR_HIP;
  }
  return ZK_OK;
}

}  // namespace zkgpu
