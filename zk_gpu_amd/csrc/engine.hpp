// Per-device NTT engine: plans, device-resident tables, workspaces, streams.
//
// The reference allocates, uploads and frees everything on every call and
// synchronises the whole device after each of its six launches (reference
// cuda/fft.cu:275-405). Here each device keeps:
//   * twiddle / coset-power tables, built on the GPU once per (omega, size)
//     and kept resident in HBM (LRU-bounded), never rebuilt by the caller;
//   * one grow-only workspace per stream (the out-of-place pass buffer);
//   * one stream per host thread, so concurrent callers overlap.
// All work of one transform is stream-ordered; nothing synchronises the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Resident-table key: kind ('R' roots, 'W' twiddles, 'P' coset pre-scale,
// 'Q' coset post-scale), sizes, and the field elements the table derives from.
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

// Owning device allocation.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int device = 0;
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

// Per-stream workspace slots (Engine::workspace).
enum WsSlot : int {
  WS_PLANAR = 0,       // inter-pass workspace of a single transform
  WS_HOST_IO = 1,      // device copy of a host-buffer call
  WS_SMOKE = 2,        // do_fft_ffi
  WS_BATCH_ARENA = 3,  // workspaces of an uncaptured batch
  WS_BATCH_META = 4,   // descriptors and counters of an uncaptured batch
  WS_PARTS_IN = 5,     // host-buffer extended parts: input
  WS_PARTS_OUT = 6,    // host-buffer extended parts: outputs
};

class Engine {
 public:
  static Engine* get(int device, int* status);
  int device() const { return device_; }

  // Fetch or build every table a job needs (stream-ordered on `s`, then
  // synchronised once so that the tables are safe for every stream).
  int prepare(const JobSpec& job, JobTables& t, hipStream_t s, int plan_kind = PLAN_SINGLE);
  // Pass descriptors of a prepared job; ws must hold 36 * 2^log_n bytes (planar
  // limb form) when the plan has more than one pass.
  int build_descs(const JobSpec& job, const JobTables& t, void* ws, PassDesc* descs, uint32_t* ndesc) const;
  // Enqueue one transform on `s` (asynchronous).
  int run(const JobSpec& job, hipStream_t s);

  // Grow-only buffers owned per stream (reused by every call on that
  // stream), one per WsSlot.
  void* workspace(hipStream_t s, size_t bytes, int slot);
  hipStream_t thread_stream();

  size_t table_bytes();
  void drop_tables();
  std::string last_error();
  void set_error(const std::string& e);

 private:
  explicit Engine(int device);
  std::shared_ptr<DevBuf> table(const TableKey& key, size_t bytes, bool* fresh);
  std::shared_ptr<DevBuf> alloc(size_t bytes);
  int build_pow_lohi(const host::Fr& base, uint64_t step, uint32_t log_count, uint32_t h,
                     std::shared_ptr<DevBuf>& lo, std::shared_ptr<DevBuf>& hi, hipStream_t s);

  int device_;
  std::mutex build_mu_;  // table lookups and builds: a table is visible only once built
  std::mutex mu_;
  std::map<TableKey, std::shared_ptr<DevBuf>> tables_;
  std::list<TableKey> lru_;
  size_t table_bytes_ = 0;
  size_t table_cap_ = 0;
  std::map<std::pair<hipStream_t, int>, std::shared_ptr<DevBuf>> ws_;
  std::string last_error_;
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
hipError_t launch_pass(const PassDesc& d, hipStream_t s);
hipError_t launch_pow_table(void* out, const uint32_t base[8], uint64_t step, uint32_t count, bool limbs48,
                            hipStream_t s);
hipError_t launch_pow_table_2level(void* out, const void* lo, const void* hi, uint32_t h, uint64_t count,
                                   const uint32_t* scale, hipStream_t s);
hipError_t launch_twiddle_table(void* out, const void* lo, const void* hi, uint32_t h, uint32_t log_m,
                                uint32_t log_L, hipStream_t s);
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

}  // namespace zkgpu
