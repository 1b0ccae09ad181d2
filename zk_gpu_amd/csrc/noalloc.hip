// Transform entry point that allocates nothing (zkgpu_ntt_device_noalloc).
//
// For callers that cannot call hipMalloc (work under HIP graph capture, tight
// serving loops, harnesses that forbid allocation). zkgpu_ntt_device keeps
// its twiddle tables in a resident cache and its inter-pass workspace in a
// per-stream buffer, both allocated on first use. This variant runs the same
// pass kernels, but everything beyond the caller's input and output -- the
// planar inter-pass workspace and the tables (roots, inter-pass twiddles,
// coset powers, their build temporaries) -- lives in `scratch`, a file-scope
// device array of NOALLOC_SCRATCH_BYTES that each GPU holds from the time
// the library's code object is loaded until the process ends.
//
// The host never uses the array's name as a device address: its address on
// each device is looked up once, with that device current, when the library
// sets the device up (noalloc_init_device, called from Engine::get), and
// kernels receive that pointer.
//
// Every call on a device shares that device's scratch, which keeps the last
// call's contents. The tables of the last call stay there and are reused when
// the next call has the same size, plan, op, omega, coset and divisor; no
// part of the scratch needs to start at zero (the pass kernels use no
// counters). Calls on one device must therefore be ordered (issued on one
// stream, or synchronised with each other): two calls that overlap on one
// device race on its scratch.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "../../include/zkgpu_ntt.h"
#include "engine.hpp"

namespace zkgpu {

// 2^20 elements at most: planar workspace 36 MiB, roots 192 KiB, inter-pass
// twiddles at most 32 + 16 MiB, coset powers 32 MiB, build temporaries 96 KiB.
constexpr uint32_t NOALLOC_MAX_LOG = 20;
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)120 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// per device: the scratch array's device address, and which tables it holds
struct ScratchState {
  uint8_t* ptr = nullptr;
  hipError_t lookup = hipErrorNotInitialized;
  bool tables_valid = false;
  uint32_t log_n = 0;
  int op = -1;
  Plan plan;
  host::Fr omega{}, coset{}, divisor{};
};
std::mutex g_mu;
ScratchState g_state[NOALLOC_MAX_DEVICES];

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

bool same_fr(const host::Fr& a, const host::Fr& b) { return memcmp(a.w, b.w, sizeof(a.w)) == 0; }
bool same_plan(const Plan& a, const Plan& b) {
  if (a.log_n != b.log_n || a.npass != b.npass) return false;
  for (uint32_t i = 0; i < MAX_PASSES; ++i)
    if (a.digit[i] != b.digit[i] || a.log_c[i] != b.log_c[i]) return false;
  return true;
}

// Offsets of the scratch regions for one job.
struct Layout {
  size_t planar = 0, roots = 0, w[MAX_PASSES] = {0, 0, 0}, coset = 0, temp = 0, end = 0;
};

Layout layout_for(const JobSpec& job, const Plan& p) {
  Layout L;
  const uint32_t k = job.log_n;
  size_t off = 0;
  if (p.npass > 1) {
    L.planar = off;
    off += up256((size_t)36 << k);
  }
  L.roots = off;
  off += up256((size_t)48 << std::min(k, TILE_LOG_MAX));
  uint32_t log_m = k;
  size_t temp = 0;
  for (uint32_t i = 0; i + 1 < p.npass; ++i) {
    L.w[i] = off;
    off += up256((size_t)32 << log_m);
    const uint32_t h = (log_m + 1) / 2;
    temp = std::max(temp, (size_t)(48 * ((1ull << h) + (1ull << (log_m - h)))));
    log_m -= p.digit[i];
  }
  if (job.op == OP_COSET_FWD || job.op == OP_COSET_INV) {
    L.coset = off;
    off += up256((size_t)32 << k);
    const uint32_t h = (k + 1) / 2;
    temp = std::max(temp, (size_t)(48 * ((1ull << h) + (1ull << (k - h)))));
  }
  L.temp = off;
  off += up256(temp);
  L.end = off;
  return L;
}

// lo[x] = base^(x*step), x < 2^h; hi[y] = base^(y*step*2^h) at temp (48-byte limb entries)
hipError_t pow_lohi(const host::Fr& base, uint64_t step, uint32_t log_count, uint32_t h, uint8_t* temp,
                    hipStream_t s) {
  uint32_t w[8];
  host::to_words32(w, base);
  uint8_t* lo = temp;
  uint8_t* hi = temp + 48 * ((size_t)1 << h);
  hipError_t e = launch_pow_table(lo, w, step, 1u << h, true, s);
  if (e == hipSuccess) e = launch_pow_table(hi, w, step << h, 1u << (log_count - h), true, s);
  return e;
}

// Build the job's tables at their scratch offsets (the same kernels and
// formats as Engine::prepare), stream-ordered on s.
hipError_t build_tables(const JobSpec& job, const Plan& p, const Layout& L, uint8_t* base, hipStream_t s) {
  const uint32_t k = job.log_n;
  const uint32_t lnt = std::min(k, TILE_LOG_MAX);
  uint32_t w[8];
  host::to_words32(w, job.omega);
  hipError_t e = launch_pow_table(base + L.roots, w, 1ull << (k - lnt), 1u << lnt, true, s);
  uint32_t divw[8];
  host::to_words32(divw, job.divisor);
  const bool fold = fold_divisor(job.op, p);
  uint32_t log_m = k;
  for (uint32_t i = 0; e == hipSuccess && i + 1 < p.npass; ++i) {
    const uint32_t log_L = log_m - p.digit[i];
    const uint32_t h = (log_m + 1) / 2;
    const bool scaled = fold && i + 2 == p.npass;
    e = pow_lohi(job.omega, 1ull << (k - log_m), log_m, h, base + L.temp, s);
    if (e == hipSuccess)
      e = launch_twiddle_table(base + L.w[i], base + L.temp, base + L.temp + 48 * ((size_t)1 << h), h, log_m, log_L,
                               scaled ? divw : nullptr, s);
    log_m = log_L;
  }
  if (e == hipSuccess && (job.op == OP_COSET_FWD || job.op == OP_COSET_INV)) {
    const bool post = job.op == OP_COSET_INV;
    const uint32_t h = (k + 1) / 2;
    uint32_t sc[8];
    host::to_words32(sc, job.divisor);
    e = pow_lohi(job.coset, 1, k, h, base + L.temp, s);
    if (e == hipSuccess)
      e = launch_pow_table_2level(base + L.coset, base + L.temp, base + L.temp + 48 * ((size_t)1 << h), h, 1ull << k,
                                  post ? sc : nullptr, s);
  }
  return e;
}

}  // namespace

// Once per device, with the device current: the device address of `scratch`.
void noalloc_init_device(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  std::lock_guard<std::mutex> g(g_mu);
  g_state[device].lookup = e;
  g_state[device].ptr = e == hipSuccess ? (uint8_t*)p : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
}

}  // namespace zkgpu

using namespace zkgpu;

extern "C" {

int zkgpu_ntt_device_noalloc(int device_id, const void* d_in, void* d_out, uint32_t log_n, int op,
                             const uint64_t* omega, const uint64_t* divisor, const uint64_t* c, void* stream) {
  if (!d_in || !d_out || !omega || log_n < 1 || log_n > MAX_LOG_N) return ZK_ERR_INVALID;
  if (op < OP_FWD || op > OP_COSET_INV) return ZK_ERR_INVALID;
  if ((op == OP_INV || op == OP_COSET_INV) && !divisor) return ZK_ERR_INVALID;
  if ((op == OP_COSET_FWD || op == OP_COSET_INV) && !c) return ZK_ERR_INVALID;
  if (log_n > NOALLOC_MAX_LOG) return ZK_ERR_SCRATCH;
  if (device_id < 0 || device_id >= NOALLOC_MAX_DEVICES) return ZK_ERR_INVALID;
  int st = ZK_OK;
  if (!Engine::get(device_id, &st)) return st;  // sets the device up (scratch address included)
  JobSpec job;
  job.in = d_in;
  job.out = d_out;
  job.log_n = log_n;
  job.op = op;
  job.omega = host::from_ptr(omega);
  if (divisor) job.divisor = host::from_ptr(divisor);
  if (c) job.coset = host::from_ptr(c);
  const Plan p = plan_for(log_n, PLAN_SINGLE);
  if (!plan_valid(p, TILE_LOG_MAX)) return ZK_ERR_INVALID;
  const Layout L = layout_for(job, p);
  if (L.end > NOALLOC_SCRATCH_BYTES) return ZK_ERR_SCRATCH;
  hipStream_t s = (hipStream_t)stream;

  std::lock_guard<std::mutex> g(g_mu);
  ScratchState& S = g_state[device_id];
  if (!S.ptr) return ZK_ERR_HIP;  // the symbol lookup failed at set-up
  const bool coset_op = op == OP_COSET_FWD || op == OP_COSET_INV;
  const bool reuse = S.tables_valid && S.log_n == log_n && S.op == op && same_plan(S.plan, p) &&
                     same_fr(S.omega, job.omega) && (!coset_op || same_fr(S.coset, job.coset)) &&
                     ((op != OP_COSET_INV && op != OP_INV) || same_fr(S.divisor, job.divisor));
  if (!reuse) {
    S.tables_valid = false;
    if (build_tables(job, p, L, S.ptr, s) != hipSuccess) {
      (void)hipGetLastError();
      return ZK_ERR_HIP;
    }
    S.tables_valid = true;
    S.log_n = log_n;
    S.op = op;
    S.plan = p;
    S.omega = job.omega;
    S.coset = job.coset;
    S.divisor = job.divisor;
  }
  TablePtrs t;
  t.plan = p;
  t.roots = S.ptr + L.roots;
  for (uint32_t i = 0; i + 1 < p.npass; ++i) t.w[i] = S.ptr + L.w[i];
  if (op == OP_COSET_FWD) t.pre = S.ptr + L.coset;
  if (op == OP_COSET_INV) t.post = S.ptr + L.coset;
  if (op == OP_INV && !fold_divisor(op, p)) {
    t.uscale = true;
    host::to_limbs261(t.scale261, job.divisor);
  }
  PassDesc descs[MAX_PASSES];
  uint32_t nd = 0;
  st = build_pass_descs(job, t, p.npass > 1 ? S.ptr + L.planar : nullptr, descs, &nd);
  if (st) return st;
  for (uint32_t i = 0; i < nd; ++i) {
    if (launch_pass(descs[i], s) != hipSuccess) {
      (void)hipGetLastError();
      S.tables_valid = false;
      return ZK_ERR_HIP;
    }
  }
  return ZK_OK;
}

uint64_t zkgpu_noalloc_scratch_bytes(uint32_t* max_log_n) {
  if (max_log_n) *max_log_n = NOALLOC_MAX_LOG;
  return (uint64_t)NOALLOC_SCRATCH_BYTES;
}

}  // extern "C"
