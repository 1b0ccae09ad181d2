// Per-device NTT engine: plans, device-resident tables, workspaces, streams.
//
// The reference allocates, uploads and frees everything on every call and
// synchronises the whole device after each of its six launches (reference
// cuda/fft.cu:275-405). Here each device keeps:
//   * twiddle / coset-power tables, built on the GPU once per (omega, size)
//     and kept resident in HBM (LRU-bounded); a table enters the cache only
//     once its build has finished, and an evicted table is freed only after
//     the streams that used it have passed their last use of it;
//   * device workspaces per (stream, slot), leased exclusively while the
//     passes that use one are enqueued, so concurrent callers sharing a
//     stream never interleave their passes on one workspace;
//   * a fixed pool of host-call contexts (stream, device buffers, pinned
//     staging) for the drop-in host-buffer entry points, checked out per
//     call under an HBM budget, so any number of caller threads shares a
//     bounded set of resources (the reference budgets each call by size,
//     src/gpulock.rs:37-44,133-203).
// All work of one transform is stream-ordered; nothing synchronises the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "host_field.hpp"
#include "ntt_desc.hpp"

struct zkgpu_job;  // include/zkgpu_ntt.h

namespace zkgpu {

enum Status : int {
  ZK_OK = 0,
  ZK_ERR_INVALID = 1,
  ZK_ERR_UNSUPPORTED = 2,
  ZK_ERR_HIP = 3,
  ZK_ERR_NOMEM = 4,
  ZK_ERR_NONCANONICAL = 5,
  ZK_ERR_BUDGET = 6,
  ZK_ERR_SHUTDOWN = 7,
  ZK_ERR_SCRATCH = 8,
};

enum Op : int { OP_FWD = 0, OP_INV = 1, OP_COSET_FWD = 2, OP_COSET_INV = 3 };

constexpr uint32_t MAX_LOG_N = host::TWO_ADICITY;  // 2^28: the largest power-of-two subgroup of Fr*
constexpr uint32_t MAX_PASSES = 3;

// Pass split of one transform: digit[p] = log2 of the pass-p DFT length,
// log_c[p] = log2 of the columns per tile.
struct Plan {
  uint32_t log_n = 0;
  uint32_t npass = 0;
  uint32_t digit[MAX_PASSES] = {0, 0, 0};
  uint32_t log_c[MAX_PASSES] = {0, 0, 0};
};

// Plans are kept per use: single transforms may use tiles of up to 2^12
// elements, batched passes (one persistent kernel for every size) up to 2^11;
// the two tables are tuned separately.
enum PlanKind : int { PLAN_SINGLE = 0, PLAN_BATCH = 1 };
uint32_t plan_tile_log(int kind);
bool default_plan(uint32_t log_n, Plan& out, uint32_t max_tile_log = TILE_LOG_MAX);
bool plan_valid(const Plan& p, uint32_t max_tile_log = TILE_LOG_MAX);
// The plan used for log_n: an installed override (autotuner) or the default.
Plan plan_for(uint32_t log_n, int kind = PLAN_SINGLE);
// Plan lookup order: override (set_plan_override), then the tuned plan file,
// then default_plan.
void set_plan_override(const Plan& p, int kind = PLAN_SINGLE);
void erase_plan_override(uint32_t log_n, int kind = PLAN_SINGLE);
void clear_plan_overrides();

// One transform, device-resident operands.
struct JobSpec {
  const void* in = nullptr;
  void* out = nullptr;
  uint32_t log_n = 0;
  int op = OP_FWD;
  host::Fr omega{};    // primitive 2^log_n-th root (its inverse for the inverse ops)
  host::Fr divisor{};  // INV / COSET_INV: output scale (n^-1 for halo2's ifft)
  host::Fr coset{};    // COSET_FWD: a[i] *= coset^i first; COSET_INV: out[i] *= coset^i last
};

// Resident-table key: kind ('R' roots, 'W' twiddles, 'V' twiddles scaled by
// the inverse's divisor, 'P' coset pre-scale, 'Q' coset post-scale), sizes,
// and the field elements the table derives from.
struct TableKey {
  char kind = 0;
  uint32_t a = 0, b = 0, c = 0;
  uint64_t e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool operator<(const TableKey& o) const {
    if (kind != o.kind) return kind < o.kind;
    if (a != o.a) return a < o.a;
    if (b != o.b) return b < o.b;
    if (c != o.c) return c < o.c;
    for (int i = 0; i < 8; ++i)
      if (e[i] != o.e[i]) return e[i] < o.e[i];
    return false;
  }
  bool operator==(const TableKey& o) const { return !(*this < o) && !(o < *this); }
};

// Per-stream completion marker: an event recorded on the stream after every
// transform Engine::run enqueues there. A later recording on the same stream
// covers every earlier one, so waiting for the marker's latest recording
// waits for all work that stream ran through the engine so far.
struct Marker {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
};

// Owning device allocation.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int device = 0;
  std::vector<Marker*> users;  // resident tables: streams that ran work reading it (Engine::mu_)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
};

struct JobTables {
  Plan plan;
  std::shared_ptr<DevBuf> roots;
  std::shared_ptr<DevBuf> w[MAX_PASSES];
  std::shared_ptr<DevBuf> pre, post;
  bool uscale = false;
  uint32_t scale261[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// The device tables of a job as plain pointers (build_descs).
struct TablePtrs {
  Plan plan;
  const void* roots = nullptr;
  const void* w[MAX_PASSES] = {nullptr, nullptr, nullptr};
  const void* pre = nullptr;
  const void* post = nullptr;
  bool uscale = false;
  uint32_t scale261[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};
TablePtrs table_ptrs(const JobTables& t);
// Pass descriptors of a job; ws must hold 36 * 2^log_n bytes (planar limb
// form) when the plan has more than one pass.
int build_pass_descs(const JobSpec& job, const TablePtrs& t, void* ws, PassDesc* descs, uint32_t* ndesc);

// Per-stream workspace slots (Engine::lease_workspace).
enum WsSlot : int {
  WS_PLANAR = 0,       // inter-pass workspace of a single transform
  WS_BATCH_ARENA = 1,  // workspaces of an uncaptured batch
  WS_BATCH_META = 2,   // descriptors and counters of an uncaptured batch
  WS_MSM = 3,          // workspace of a device-resident MSM
};

// One (stream, slot) workspace. Its mutex is held by a WsLease for as long as
// work that uses the buffer is being enqueued.
struct Workspace {
  std::mutex mu;
  std::shared_ptr<DevBuf> buf;
  hipStream_t stream = nullptr;
};

// Exclusive use of a workspace of at least the requested size while it is
// held. The buffer is grown (after the stream's earlier work that used the old
// one has finished) only under the lease.
class WsLease {
 public:
  WsLease() = default;
  WsLease(WsLease&&) = default;
  WsLease& operator=(WsLease&&) = default;
  void* ptr() const { return buf_ ? buf_->ptr : nullptr; }
  explicit operator bool() const { return ptr() != nullptr; }

 private:
  friend class Engine;
  std::unique_lock<std::mutex> lk_;
  std::shared_ptr<DevBuf> buf_;
};

// A host-call context: one stream with the device buffers of a host-buffer
// call (the caller's data, the inter-pass workspace, extra outputs). Checked
// out exclusively; its stream is idle whenever it is not checked out.
struct HostCtx {
  hipStream_t stream = nullptr;
  std::shared_ptr<DevBuf> io, planar, aux;
  bool busy = false;
  size_t held() const { return (io ? io->bytes : 0) + (planar ? planar->bytes : 0) + (aux ? aux->bytes : 0); }
};
class HostLease;

class Engine {
 public:
  static Engine* get(int device, int* status);
  int device() const { return device_; }

  // Fetch or build every table a job needs. Fresh tables are built on `s`
  // (the engine's build stream when s is null) and synchronised there before
  // they are published, so a published table is safe for every stream.
  int prepare(const JobSpec& job, JobTables& t, hipStream_t s, int plan_kind = PLAN_SINGLE);
  // Pass descriptors of a prepared job (build_pass_descs).
  int build_descs(const JobSpec& job, const JobTables& t, void* ws, PassDesc* descs, uint32_t* ndesc) const {
    return build_pass_descs(job, table_ptrs(t), ws, descs, ndesc);
  }
  // Enqueue one transform on `s` (asynchronous). planar_ws: the caller's
  // inter-pass workspace (36 * 2^log_n bytes), else the stream's leased one.
  int run(const JobSpec& job, hipStream_t s, void* planar_ws = nullptr);

  // Exclusive lease of the (s, slot) workspace, grown to `bytes` if needed.
  // *status is ZK_ERR_NOMEM when it cannot be allocated.
  WsLease lease_workspace(hipStream_t s, size_t bytes, int slot, int* status);
  // Stream for table builds of callers that have none (batches, servers).
  hipStream_t build_stream() const { return build_stream_; }
  // The stream's completion marker (created on first use).
  Marker* stream_marker(hipStream_t s);
  // Note that work just enqueued on m's stream (and recorded on m) read t.
  void note_users(const JobTables& t, Marker* m);

  // Per-pass kernel timing, for measurement: while on, every pass run()
  // launches carries start/stop events tied to its dispatch; pass_times()
  // waits for them and returns, per kernel, launches and total milliseconds
  // (then forgets them).
  struct KernelTime {
    std::string name;
    uint32_t launches = 0;
    double total_ms = 0;
  };
  void set_pass_timing(bool on);
  bool pass_timing() const { return timing_.load(std::memory_order_relaxed); }
  int pass_times(std::vector<KernelTime>& out);
  // Timing of other launches (the MSM's): a start and a stop event recorded
  // around one launch on s; finish_timed pairs them with the kernel's name.
  void timed_events(hipEvent_t* a, hipEvent_t* b);
  void finish_timed(hipEvent_t a, hipEvent_t b, const char* name);

  // Host-call contexts (host-buffer entry points).
  int acquire_host(size_t io_bytes, size_t planar_bytes, size_t aux_bytes, HostCtx** out);
  void release_host(HostCtx* c);
  size_t host_held_bytes();
  size_t host_peak_bytes();
  // New HBM budget of the host-call contexts (0: the default); idle
  // contexts' buffers are released down to it. Returns the previous budget.
  size_t set_host_budget(size_t bytes);

  size_t table_bytes();
  void drop_tables();
  std::string last_error();
  void set_error(const std::string& e);

  // Test hook: after `skip` more successful device allocations, the next
  // `count` fail (as if the device were out of memory).
  static void debug_fail_allocs(uint32_t skip, uint32_t count);

 private:
  explicit Engine(int device);
  std::shared_ptr<DevBuf> cached_table(const TableKey& key);
  std::shared_ptr<DevBuf> new_table(size_t bytes);
  void publish_tables(std::vector<std::pair<TableKey, std::shared_ptr<DevBuf>>>& fresh);
  std::shared_ptr<DevBuf> alloc(size_t bytes);
  Marker* marker(hipStream_t s);
  int build_pow_lohi(const host::Fr& base, uint64_t step, uint32_t log_count, uint32_t h,
                     std::shared_ptr<DevBuf>& lo, std::shared_ptr<DevBuf>& hi, hipStream_t s);

  int device_;
  hipStream_t build_stream_ = nullptr;
  std::mutex build_mu_;  // table builds, one at a time
  std::mutex mu_;        // tables_, lru_, markers_, ws_, buffer users
  std::map<TableKey, std::shared_ptr<DevBuf>> tables_;
  std::list<TableKey> lru_;
  size_t table_bytes_ = 0;
  size_t table_cap_ = 0;
  std::map<hipStream_t, std::unique_ptr<Marker>> markers_;
  std::map<std::pair<hipStream_t, int>, std::unique_ptr<Workspace>> ws_;
  size_t ws_bytes_ = 0, ws_cap_ = 0;
  // host-call contexts
  std::mutex host_mu_;
  std::condition_variable host_cv_;
  std::vector<std::unique_ptr<HostCtx>> host_ctx_;
  size_t host_held_ = 0, host_peak_ = 0, host_budget_ = 0, host_budget_default_ = 0;
  // pass timing
  struct TimedPass {
    hipEvent_t a = nullptr, b = nullptr;
    std::string name;
  };
  std::atomic<bool> timing_{false};
  std::mutex timing_mu_;
  std::vector<TimedPass> timed_;
  std::vector<hipEvent_t> timing_pool_;
  std::string last_error_;
};

// RAII checkout of a host-call context.
class HostLease {
 public:
  HostLease(Engine* e, size_t io_bytes, size_t planar_bytes, size_t aux_bytes) : e_(e) {
    status_ = e->acquire_host(io_bytes, planar_bytes, aux_bytes, &c_);
  }
  ~HostLease() {
    if (c_) e_->release_host(c_);
  }
  HostLease(const HostLease&) = delete;
  HostLease& operator=(const HostLease&) = delete;
  int status() const { return status_; }
  HostCtx* operator->() const { return c_; }

 private:
  Engine* e_;
  HostCtx* c_ = nullptr;
  int status_ = 0;
};

// A C-ABI job as a JobSpec (standard-domain constants filled in when the job
// asks for them). out_optional: d_out may be NULL (the caller fills it later).
int job_from_c(const ::zkgpu_job& j, JobSpec& out, bool out_optional);

// Copies between caller (pageable) host memory and the device through pinned
// staging, chunked and overlapped with host-side copies (host_stage.cpp).
// stage_h2d enqueues on s; stage_d2h returns once the data is in dst.
int stage_h2d(int device, hipStream_t s, void* dst, const void* src, size_t bytes);
int stage_d2h(int device, hipStream_t s, void* dst, const void* src, size_t bytes);

// HIP launchers (ntt_kernels.hip)
hipError_t configure_kernels();
// start/stop (optional): events tied to the dispatch (hipExtLaunchKernel)
hipError_t launch_pass(const PassDesc& d, hipStream_t s, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
bool pass_kernel_narrow(const PassDesc& d);  // the narrow pass-kernel family runs d
bool pass_kernel_quad(const PassDesc& d);    // the quad (four elements per thread) family runs d
hipError_t launch_pow_table(void* out, const uint32_t base[8], uint64_t step, uint32_t count, bool limbs48,
                            hipStream_t s);
hipError_t launch_pow_table_2level(void* out, const void* lo, const void* hi, uint32_t h, uint64_t count,
                                   const uint32_t* scale, hipStream_t s);
// scale (optional, halo2 words): a factor folded into every twiddle
hipError_t launch_twiddle_table(void* out, const void* lo, const void* hi, uint32_t h, uint32_t log_m,
                                uint32_t log_L, const uint32_t* scale, hipStream_t s);
// The inverse's divisor is folded into the twiddles of the last pass boundary
// (which the last pass applies to every element on load) when the plan has
// more than one pass; a one-pass inverse scales its outputs instead.
inline bool fold_divisor(int op, const Plan& p) { return op == OP_INV && p.npass > 1; }
hipError_t launch_fill_random(void* out, uint64_t seed, uint64_t start, uint64_t count, hipStream_t s);
hipError_t launch_check_canonical(const void* in, uint64_t count, uint32_t* flag, hipStream_t s);
hipError_t launch_persistent(const PassDesc* d_descs, const uint32_t* d_tile_start, uint32_t ndesc, uint32_t total,
                             uint32_t* d_counter, uint32_t grid, hipStream_t s);
hipError_t launch_reset_counters(uint32_t* d_counter, uint32_t n, hipStream_t s);
uint32_t batch_tile_log();
hipError_t launch_mad_probe(uint64_t* out, uint32_t blocks, uint32_t seed, hipStream_t s);
uint64_t mad_probe_ops_per_thread();
uint32_t tile_threads(uint32_t log_n_tile, uint32_t log_c, uint32_t le_max);  // threads of a tile
uint32_t batch_items(const PassDesc& d);  // queue items of a persistent-kernel descriptor
uint32_t persistent_wgs_per_cu();          // resident persistent-kernel workgroups per CU
// noalloc.hip: once per device, with it current, look up the device address
// of the no-allocation entry point's scratch array
void noalloc_init_device(int device);

}  // namespace zkgpu
