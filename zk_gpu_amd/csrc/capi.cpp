// extern "C" entry points of libzkgpu_ntt.so (declared in include/zkgpu_ntt.h).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>

#include "../../include/zkgpu_ntt.h"
#include "engine.hpp"

using namespace zkgpu;

namespace {

bool size_ok(uint64_t n, uint64_t log_n) {
  return log_n >= 1 && log_n <= MAX_LOG_N && n == (1ull << log_n);
}

std::atomic<int> g_validate{-1};  // -1: not read yet
bool validation_on() {
  int v = g_validate.load(std::memory_order_relaxed);
  if (v < 0) {
    const char* e = getenv("ZKGPU_VALIDATE");
    v = (e && *e && *e != '0') ? 1 : 0;
    int expect = -1;
    g_validate.compare_exchange_strong(expect, v);
    v = g_validate.load();
  }
  return v == 1;
}

// With validation on: ZK_ERR_NONCANONICAL when some of the count elements at
// d (device) are >= r. Synchronises s.
int validate_device(void* d, uint64_t count, uint32_t* d_flag, hipStream_t s) {
  uint32_t flag = 0;
  if (launch_reset_counters(d_flag, 1, s) != hipSuccess) return ZK_ERR_HIP;
  if (launch_check_canonical(d, count, d_flag, s) != hipSuccess) return ZK_ERR_HIP;
  if (hipMemcpyAsync(&flag, d_flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess) return ZK_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return ZK_ERR_HIP;
  return flag ? ZK_ERR_NONCANONICAL : ZK_OK;
}

// One synchronous host-buffer transform on a host-call context: upload,
// transform in place, download. The reference does the same per call with
// fresh allocations and device-wide syncs (cuda/fft.cu:275-405).
int host_call(size_t device_id, uint64_t* req, uint64_t log_n, int op, const uint64_t* omega,
              const uint64_t* divisor, const uint64_t* c) {
  if (!req || !omega) return ZK_ERR_INVALID;
  if ((op == OP_INV || op == OP_COSET_INV) && !divisor) return ZK_ERR_INVALID;
  if ((op == OP_COSET_FWD || op == OP_COSET_INV) && !c) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get((int)device_id, &st);
  if (!e) return st;
  const size_t bytes = (size_t)32 << log_n;
  const bool multi = plan_for((uint32_t)log_n, PLAN_SINGLE).npass > 1;
  const bool check = validation_on();
  HostLease ctx(e, bytes, multi ? (size_t)36 << log_n : 0, check ? 256 : 0);
  if (ctx.status()) return ctx.status();
  hipStream_t s = ctx->stream;
  void* d = ctx->io->ptr;
  JobSpec job;
  job.in = d;
  job.out = d;
  job.log_n = (uint32_t)log_n;
  job.op = op;
  job.omega = host::from_ptr(omega);
  if (divisor) job.divisor = host::from_ptr(divisor);
  if (c) job.coset = host::from_ptr(c);
  st = stage_h2d((int)device_id, s, d, req, bytes);
  if (st == ZK_OK && check) st = validate_device(d, 1ull << log_n, (uint32_t*)ctx->aux->ptr, s);
  if (st == ZK_OK) st = e->run(job, s, multi ? ctx->planar->ptr : nullptr);
  if (st) {
    (void)hipStreamSynchronize(s);  // the context goes back with an idle stream
    return st;
  }
  st = stage_d2h((int)device_id, s, req, d, bytes);
  if (st) (void)hipStreamSynchronize(s);
  return st;
}

// Standard-domain constants (omega, omega^-1, n^-1, g, g^-1) for every
// size, derived once: each needs field inversions on the host.
struct StdDomain {
  host::Fr c[MAX_LOG_N + 1][5];
  StdDomain() {
    const host::Fr g = host::from_u64(7), ginv = host::inv(g);
    for (uint32_t k = 1; k <= MAX_LOG_N; ++k) {
      c[k][0] = host::root_of_unity(k);
      c[k][1] = host::inv(c[k][0]);
      c[k][2] = host::inv(host::from_u64(1ull << k));
      c[k][3] = g;
      c[k][4] = ginv;
    }
  }
};

const host::Fr& std_constant(uint32_t log_n, int which) {
  static const StdDomain dom;
  return dom.c[log_n][which];
}

}  // namespace

extern "C" {

uint64_t do_fft_ffi(size_t device_id) {
  // round trip of a 2^10 synthetic vector through forward and inverse
  int st = ZK_OK;
  Engine* e = Engine::get((int)device_id, &st);
  if (!e) return (uint64_t)st;
  const uint32_t k = 10;
  const size_t bytes = (size_t)32 << k;
  HostLease ctx(e, 2 * bytes, 0, 0);
  if (ctx.status()) return (uint64_t)ctx.status();
  hipStream_t s = ctx->stream;
  void* a = ctx->io->ptr;
  void* b = (char*)a + bytes;
  std::vector<uint64_t> ha(4u << k), hb(4u << k);
  st = launch_fill_random(a, 1, 0, 1ull << k, s) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
  if (st == ZK_OK) st = zkgpu_ntt_std_device((int)device_id, a, b, k, ZKGPU_OP_FWD, s);
  if (st == ZK_OK) st = zkgpu_ntt_std_device((int)device_id, b, b, k, ZKGPU_OP_INV, s);
  if (st == ZK_OK && hipMemcpyAsync(ha.data(), a, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) st = ZK_ERR_HIP;
  if (st == ZK_OK && hipMemcpyAsync(hb.data(), b, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) st = ZK_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && st == ZK_OK) st = ZK_ERR_HIP;
  if (st) return (uint64_t)st;
  return ha == hb ? ZK_OK : ZK_ERR_HIP;
}

uint64_t parallel_fft_api(size_t device_id, uint64_t* req, const uint64_t* omega, const uint64_t* twiddle,
                          uint64_t n, uint64_t log_n, uint64_t twiddle_size) {
  (void)twiddle;
  (void)twiddle_size;
  if (!size_ok(n, log_n)) return ZK_ERR_INVALID;
  return (uint64_t)host_call(device_id, req, log_n, OP_FWD, omega, nullptr, nullptr);
}

uint64_t parallel_ifft_api(size_t device_id, uint64_t* req, const uint64_t* omega, const uint64_t* twiddle,
                           uint64_t n, uint64_t log_n, uint64_t twiddle_size, const uint64_t* divisor) {
  (void)twiddle;
  (void)twiddle_size;
  if (!size_ok(n, log_n)) return ZK_ERR_INVALID;
  return (uint64_t)host_call(device_id, req, log_n, OP_INV, omega, divisor, nullptr);
}

uint64_t parallel_coeff_to_extended_part_api(size_t device_id, uint64_t* req, const uint64_t* omega,
                                             const uint64_t* twiddle, uint64_t n, uint64_t log_n,
                                             uint64_t twiddle_size, const uint64_t* c) {
  (void)twiddle;
  (void)twiddle_size;
  if (!size_ok(n, log_n)) return ZK_ERR_INVALID;
  return (uint64_t)host_call(device_id, req, log_n, OP_COSET_FWD, omega, nullptr, c);
}

uint64_t zkgpu_coset_ifft_api(size_t device_id, uint64_t* req, const uint64_t* omega, uint64_t n, uint64_t log_n,
                              const uint64_t* divisor, const uint64_t* c) {
  if (!size_ok(n, log_n)) return ZK_ERR_INVALID;
  return (uint64_t)host_call(device_id, req, log_n, OP_COSET_INV, omega, divisor, c);
}

int zkgpu_ntt_device(int device_id, const void* d_in, void* d_out, uint32_t log_n, int op, const uint64_t* omega,
                     const uint64_t* divisor, const uint64_t* c, void* stream) {
  if (!d_in || !d_out || !omega || log_n < 1 || log_n > MAX_LOG_N) return ZK_ERR_INVALID;
  if (op < OP_FWD || op > OP_COSET_INV) return ZK_ERR_INVALID;
  if ((op == OP_INV || op == OP_COSET_INV) && !divisor) return ZK_ERR_INVALID;
  if ((op == OP_COSET_FWD || op == OP_COSET_INV) && !c) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  JobSpec job;
  job.in = d_in;
  job.out = d_out;
  job.log_n = log_n;
  job.op = op;
  job.omega = host::from_ptr(omega);
  if (divisor) job.divisor = host::from_ptr(divisor);
  if (c) job.coset = host::from_ptr(c);
  return e->run(job, (hipStream_t)stream);
}

int zkgpu_domain_constant(uint32_t log_n, int which, uint64_t* out4) {
  if (!out4 || log_n < 1 || log_n > MAX_LOG_N || which < 0 || which > 4) return ZK_ERR_INVALID;
  memcpy(out4, std_constant(log_n, which).w, 32);
  return ZK_OK;
}

int zkgpu_ntt_std_device(int device_id, const void* d_in, void* d_out, uint32_t log_n, int op, void* stream) {
  if (log_n < 1 || log_n > MAX_LOG_N) return ZK_ERR_INVALID;
  const bool inverse = op == OP_INV || op == OP_COSET_INV;
  const host::Fr& omega = std_constant(log_n, inverse ? 1 : 0);
  const host::Fr& div = std_constant(log_n, 2);
  const host::Fr& c = std_constant(log_n, op == OP_COSET_INV ? 4 : 3);
  return zkgpu_ntt_device(device_id, d_in, d_out, log_n, op, omega.w, div.w, c.w, stream);
}

int zkgpu_prepare(int device_id, uint32_t log_n, int op, const uint64_t* omega, const uint64_t* divisor,
                  const uint64_t* c) {
  if (!omega || log_n < 1 || log_n > MAX_LOG_N || op < OP_FWD || op > OP_COSET_INV) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  JobSpec job;
  job.log_n = log_n;
  job.op = op;
  job.omega = host::from_ptr(omega);
  if (divisor) job.divisor = host::from_ptr(divisor);
  if (c) job.coset = host::from_ptr(c);
  JobTables t;
  return e->prepare(job, t, nullptr);  // built on the engine's build stream and synchronised there
}

int zkgpu_fill_random(int device_id, void* d_out, uint64_t count, uint64_t seed, uint64_t start, void* stream) {
  if (!d_out) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  if (count == 0) return ZK_OK;
  return launch_fill_random(d_out, seed, start, count, (hipStream_t)stream) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
}

int zkgpu_check_canonical(int device_id, const void* d_in, uint64_t count, uint32_t* d_flag, void* stream) {
  if (!d_in || !d_flag) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  if (count == 0) return ZK_OK;
  return launch_check_canonical(d_in, count, d_flag, (hipStream_t)stream) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
}

int zkgpu_plan_get(uint32_t log_n, uint32_t* digits, uint32_t* log_cols) {
  return zkgpu_plan_get_kind(log_n, 0, digits, log_cols);
}

int zkgpu_plan_get_kind(uint32_t log_n, int kind, uint32_t* digits, uint32_t* log_cols) {
  if (log_n < 1 || log_n > MAX_LOG_N || kind < 0 || kind > 1) return 0;
  const Plan p = plan_for(log_n, kind);
  for (uint32_t i = 0; i < MAX_PASSES; ++i) {
    if (digits) digits[i] = p.digit[i];
    if (log_cols) log_cols[i] = p.log_c[i];
  }
  return (int)p.npass;
}

int zkgpu_plan_set(uint32_t log_n, uint32_t npass, const uint32_t* digits, const uint32_t* log_cols) {
  return zkgpu_plan_set_kind(log_n, 0, npass, digits, log_cols);
}

int zkgpu_plan_set_kind(uint32_t log_n, int kind, uint32_t npass, const uint32_t* digits, const uint32_t* log_cols) {
  if (log_n < 1 || log_n > MAX_LOG_N || kind < 0 || kind > 1) return ZK_ERR_INVALID;
  if (npass == 0) {
    erase_plan_override(log_n, kind);  // back to the tuned plan file, else the default split
    return ZK_OK;
  }
  if (npass > MAX_PASSES || !digits || !log_cols) return ZK_ERR_INVALID;
  Plan p;
  p.log_n = log_n;
  p.npass = npass;
  for (uint32_t i = 0; i < npass; ++i) {
    p.digit[i] = digits[i];
    p.log_c[i] = log_cols[i];
  }
  if (!plan_valid(p, plan_tile_log(kind))) return ZK_ERR_INVALID;
  set_plan_override(p, kind);
  return ZK_OK;
}

uint64_t zkgpu_table_bytes(int device_id) {
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  return e ? (uint64_t)e->table_bytes() : 0;
}

int zkgpu_drop_tables(int device_id) {
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  e->drop_tables();
  return ZK_OK;
}

const char* zkgpu_error_string(int code) {
  switch (code) {
    case ZK_OK: return "ok";
    case ZK_ERR_INVALID: return "invalid argument";
    case ZK_ERR_UNSUPPORTED: return "unsupported entry point";
    case ZK_ERR_HIP: return "HIP runtime error";
    case ZK_ERR_NOMEM: return "device allocation failed";
    case ZK_ERR_NONCANONICAL: return "input element not below r";
    case ZK_ERR_BUDGET: return "job exceeds the HBM budget";
    case ZK_ERR_SHUTDOWN: return "server stopping";
    case ZK_ERR_SCRATCH: return "size exceeds the no-allocation scratch limit";
    default: return "unknown error";
  }
}

int zkgpu_probe_mad_rate(int device_id, double* mads_per_s) {
  if (!mads_per_s) return ZK_ERR_INVALID;
  *mads_per_s = 0;
  int st = ZK_OK;
  if (!Engine::get(device_id, &st)) return st;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_id);
  const uint32_t blocks = (uint32_t)cus * 8;  // 8 waves per SIMD
  uint64_t* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint64_t)) != hipSuccess) {
    (void)hipGetLastError();
    return ZK_ERR_NOMEM;
  }
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_mad_probe(out, blocks, 1, s);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  if (e == hipSuccess) e = launch_mad_probe(out, blocks, 2, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess && ms > 0)
    *mads_per_s = (double)blocks * 256.0 * (double)mad_probe_ops_per_thread() / (ms * 1e-3);
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(out);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return ZK_ERR_HIP;
  }
  return ZK_OK;
}

int zkgpu_set_validation(int enable) {
  const int prev = validation_on() ? 1 : 0;
  g_validate.store(enable ? 1 : 0);
  return prev;
}

int zkgpu_debug_fail_allocs(uint32_t skip, uint32_t count) {
  Engine::debug_fail_allocs(skip, count);
  return ZK_OK;
}

uint64_t zkgpu_host_held_bytes(int device_id, uint64_t* peak_bytes) {
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (peak_bytes) *peak_bytes = e ? (uint64_t)e->host_peak_bytes() : 0;
  return e ? (uint64_t)e->host_held_bytes() : 0;
}

uint64_t zkgpu_set_host_budget(int device_id, uint64_t bytes) {
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  return e ? (uint64_t)e->set_host_budget((size_t)bytes) : 0;
}

int zkgpu_pass_timing(int device_id, int enable) {
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  e->set_pass_timing(enable != 0);
  return ZK_OK;
}

int zkgpu_pass_times(int device_id, zkgpu_kernel_time* out, uint32_t max_out, uint32_t* count) {
  if (!count || (max_out && !out)) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  std::vector<Engine::KernelTime> kt;
  st = e->pass_times(kt);
  *count = (uint32_t)kt.size();
  for (uint32_t i = 0; i < kt.size() && i < max_out; ++i) {
    memset(&out[i], 0, sizeof(out[i]));
    strncpy(out[i].name, kt[i].name.c_str(), sizeof(out[i].name) - 1);
    out[i].launches = kt[i].launches;
    out[i].total_ms = kt[i].total_ms;
  }
  return st;
}

int zkgpu_version(void) { return 200; }

}  // extern "C"
