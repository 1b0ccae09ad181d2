// Batched transforms: many polynomials of uneven sizes and mixed ops.
//
// All jobs' passes are flattened into pass levels (level p = pass p of every
// job that has one). Each level is ONE launch of the persistent pass kernel,
// which drains a queue holding every tile of that level, whatever job or tile
// length it belongs to. The launches (plus a counter reset) are captured once
// into a hipGraph and replayed on each launch of the batch.
//
// Levels run as separate launches because pass p+1 of a job reads what pass p
// wrote in every tile of that job. Inside one launch this would need an
// agent-scope release/acquire per tile (an L2 write-back across the eight
// XCD L2s of the MI355X); a kernel boundary orders them for the whole level.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/zkgpu_ntt.h"
#include "engine.hpp"

namespace zkgpu {

struct Batch {
  Engine* eng = nullptr;
  int device = 0;
  std::vector<JobSpec> jobs;
  std::vector<JobTables> tables;  // keeps the resident tables referenced
  std::shared_ptr<DevBuf> arena;  // planar workspaces of the multi-pass jobs
  std::shared_ptr<DevBuf> meta;   // descriptors, tile prefix sums, counters
  void* arena_ptr = nullptr;      // arena->ptr, or an engine workspace
  void* meta_ptr = nullptr;
  std::vector<char> host_meta;    // source of the (asynchronous) meta upload
  struct Level {
    uint32_t desc_off = 0, ndesc = 0, total = 0, grid = 0;  // total: queue items
    uint32_t tiles = 0;
  };
  std::vector<Level> levels;
  uint32_t* d_counters = nullptr;
  PassDesc* d_descs = nullptr;
  uint32_t* d_tile_start = nullptr;
  WsLease arena_lease, meta_lease;  // engine workspaces of an uncaptured batch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  uint64_t elements = 0;
  // an event per stream the batch was launched on, recorded after each launch
  std::mutex launch_mu;
  std::vector<std::pair<hipStream_t, hipEvent_t>> launched;
  // wait for every launch of the batch (only the streams it ran on)
  void wait_launches() {
    std::lock_guard<std::mutex> g(launch_mu);
    for (auto& kv : launched) (void)hipEventSynchronize(kv.second);
  }
  ~Batch() {
    for (auto& kv : launched) (void)hipEventDestroy(kv.second);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
};

// capture: record the launches in a hipGraph (zkgpu_batch_*); otherwise the
// batch is launched directly (batch_launch_direct). ws_stream: take the
// workspaces and the descriptor buffer from the engine's per-stream buffers
// of that stream instead of fresh allocations (the batch must then be run
// and finished on that stream before the stream's next such batch).
struct BatchOpts {
  bool capture = true;
  bool use_ws = false;
  hipStream_t ws_stream = nullptr;
};

int batch_build(int device, const zkgpu_job* cjobs, uint32_t count, Batch** out, const BatchOpts& opt) {
  *out = nullptr;
  if (!cjobs || count == 0) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get(device, &st);
  if (!e) return st;
  std::unique_ptr<Batch> b(new Batch());
  b->eng = e;
  b->device = device;
  b->jobs.resize(count);
  b->tables.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    if (cjobs[i].flags & ZKGPU_JOB_HOST_BUFFERS) return ZK_ERR_INVALID;  // batches take device buffers
    st = job_from_c(cjobs[i], b->jobs[i], false);
    if (st) return st;
    st = e->prepare(b->jobs[i], b->tables[i], nullptr, PLAN_BATCH);
    if (st) return st;
    b->elements += 1ull << b->jobs[i].log_n;
  }
  // workspaces (planar limb form, 36 bytes per element) for multi-pass jobs
  std::vector<size_t> ws_off(count, SIZE_MAX);
  size_t arena_bytes = 0;
  for (uint32_t i = 0; i < count; ++i) {
    if (b->tables[i].plan.npass > 1) {
      ws_off[i] = arena_bytes;
      arena_bytes += (((size_t)36 << b->jobs[i].log_n) + 255) & ~(size_t)255;
    }
  }
  if (arena_bytes && opt.use_ws) {
    b->arena_lease = e->lease_workspace(opt.ws_stream, arena_bytes, WS_BATCH_ARENA, &st);
    if (st) return st;
    b->arena_ptr = b->arena_lease.ptr();
  } else if (arena_bytes) {
    b->arena = std::make_shared<DevBuf>();
    b->arena->device = device;
    b->arena->bytes = arena_bytes;
    if (hipMalloc(&b->arena->ptr, arena_bytes) != hipSuccess) {
      (void)hipGetLastError();
      b->arena->ptr = nullptr;
      return ZK_ERR_NOMEM;
    }
    b->arena_ptr = b->arena->ptr;
  }
  // descriptors grouped by level; within a level the biggest jobs first so
  // that the queue ends with short tiles
  std::vector<std::vector<PassDesc>> per_job(count);
  uint32_t max_pass = 0;
  for (uint32_t i = 0; i < count; ++i) {
    void* ws = ws_off[i] == SIZE_MAX ? nullptr : (char*)b->arena_ptr + ws_off[i];
    PassDesc d[MAX_PASSES];
    uint32_t nd = 0;
    st = e->build_descs(b->jobs[i], b->tables[i], ws, d, &nd);
    if (st) return st;
    per_job[i].assign(d, d + nd);
    max_pass = std::max(max_pass, nd);
  }
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return b->jobs[x].log_n > b->jobs[y].log_n; });
  std::vector<PassDesc> all;
  std::vector<uint32_t> starts;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  for (uint32_t lvl = 0; lvl < max_pass; ++lvl) {
    Batch::Level L;
    L.desc_off = (uint32_t)all.size();
    uint32_t total = 0;
    for (uint32_t i : order) {
      if (lvl < per_job[i].size()) {
        starts.push_back(total);
        all.push_back(per_job[i][lvl]);
        total += batch_items(per_job[i][lvl]);
        L.tiles += per_job[i][lvl].num_tiles;
      }
    }
    L.ndesc = (uint32_t)all.size() - L.desc_off;
    L.total = total;
    // one resident workgroup per LDS/register slot of every CU
    L.grid = std::min<uint32_t>(total, (uint32_t)cus * persistent_wgs_per_cu());
    b->levels.push_back(L);
  }
  // one device buffer: descriptors | tile prefix sums | counters
  const size_t desc_bytes = all.size() * sizeof(PassDesc);
  const size_t ts_bytes = starts.size() * sizeof(uint32_t);
  const size_t cnt_bytes = b->levels.size() * sizeof(uint32_t);
  const size_t ts_off = (desc_bytes + 255) & ~(size_t)255;
  const size_t cnt_off = (ts_off + ts_bytes + 255) & ~(size_t)255;
  const size_t meta_bytes = cnt_off + cnt_bytes;
  b->host_meta.assign(meta_bytes, 0);
  memcpy(b->host_meta.data(), all.data(), desc_bytes);
  memcpy(b->host_meta.data() + ts_off, starts.data(), ts_bytes);
  if (opt.use_ws) {
    b->meta_lease = e->lease_workspace(opt.ws_stream, meta_bytes, WS_BATCH_META, &st);
    if (st) return st;
    b->meta_ptr = b->meta_lease.ptr();
    if (hipMemcpyAsync(b->meta_ptr, b->host_meta.data(), meta_bytes, hipMemcpyHostToDevice, opt.ws_stream) !=
        hipSuccess)
      return ZK_ERR_HIP;
  } else {
    b->meta = std::make_shared<DevBuf>();
    b->meta->device = device;
    b->meta->bytes = meta_bytes;
    if (hipMalloc(&b->meta->ptr, meta_bytes) != hipSuccess) {
      (void)hipGetLastError();
      b->meta->ptr = nullptr;
      return ZK_ERR_NOMEM;
    }
    b->meta_ptr = b->meta->ptr;
    if (hipMemcpy(b->meta_ptr, b->host_meta.data(), meta_bytes, hipMemcpyHostToDevice) != hipSuccess)
      return ZK_ERR_HIP;
  }
  b->d_descs = (PassDesc*)b->meta_ptr;
  b->d_tile_start = (uint32_t*)((char*)b->meta_ptr + ts_off);
  b->d_counters = (uint32_t*)((char*)b->meta_ptr + cnt_off);
  if (!opt.capture) {
    *out = b.release();
    return ZK_OK;
  }

  // capture: counter reset (a kernel, not a memset node) and one persistent
  // launch per level
  if (hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) return ZK_ERR_HIP;
  if (hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ZK_ERR_HIP;
  hipError_t le = launch_reset_counters(b->d_counters, (uint32_t)b->levels.size(), b->capture_stream);
  for (const auto& L : b->levels) {
    if (le != hipSuccess) break;
    le = launch_persistent(b->d_descs + L.desc_off, b->d_tile_start + L.desc_off, L.ndesc, L.total,
                           b->d_counters + (&L - b->levels.data()), L.grid, b->capture_stream);
  }
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(b->capture_stream, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return ZK_ERR_HIP;
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) return ZK_ERR_HIP;
  *out = b.release();
  return ZK_OK;
}

// Extended-domain parts of at least 2^19 elements run one after another as
// single transforms; smaller ones share persistent-kernel launches.
constexpr uint32_t kPartsAloneLog = 19;

// The batch's launches enqueued on s without a graph.
int batch_launch_direct(const Batch* b, hipStream_t s) {
  hipError_t le = launch_reset_counters(b->d_counters, (uint32_t)b->levels.size(), s);
  for (size_t i = 0; le == hipSuccess && i < b->levels.size(); ++i) {
    const auto& L = b->levels[i];
    le = launch_persistent(b->d_descs + L.desc_off, b->d_tile_start + L.desc_off, L.ndesc, L.total,
                           b->d_counters + i, L.grid, s);
  }
  return le == hipSuccess ? ZK_OK : ZK_ERR_HIP;
}

}  // namespace zkgpu

using namespace zkgpu;

extern "C" {

int zkgpu_batch_create(int device_id, const zkgpu_job* jobs, uint32_t count, uint64_t* handle) {
  if (!handle) return ZK_ERR_INVALID;
  *handle = 0;
  Batch* b = nullptr;
  const int st = batch_build(device_id, jobs, count, &b, BatchOpts());
  if (st) return st;
  *handle = (uint64_t)(uintptr_t)b;
  return ZK_OK;
}

int zkgpu_batch_launch(uint64_t handle, void* stream) {
  Batch* b = (Batch*)(uintptr_t)handle;
  if (!b) return ZK_ERR_INVALID;
  int st = ZK_OK;
  if (!Engine::get(b->device, &st)) return st;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(b->launch_mu);
  hipEvent_t ev = nullptr;
  for (auto& kv : b->launched)
    if (kv.first == s) ev = kv.second;
  if (!ev) {
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return ZK_ERR_HIP;
    b->launched.emplace_back(s, ev);
  }
  if (hipGraphLaunch(b->exec, s) != hipSuccess) return ZK_ERR_HIP;
  return hipEventRecord(ev, s) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
}

int zkgpu_batch_info(uint64_t handle, uint32_t* levels, uint32_t* tiles, uint64_t* elements) {
  Batch* b = (Batch*)(uintptr_t)handle;
  if (!b) return ZK_ERR_INVALID;
  if (levels) *levels = (uint32_t)b->levels.size();
  if (tiles) {
    uint32_t t = 0;
    for (const auto& L : b->levels) t += L.tiles;
    *tiles = t;
  }
  if (elements) *elements = b->elements;
  return ZK_OK;
}

int zkgpu_batch_destroy(uint64_t handle) {
  Batch* b = (Batch*)(uintptr_t)handle;
  if (!b) return ZK_ERR_INVALID;
  int st = ZK_OK;
  if (Engine::get(b->device, &st)) b->wait_launches();  // queued launches may still read its buffers
  delete b;
  return ZK_OK;
}

int zkgpu_coeff_to_extended_parts_device(int device_id, const void* d_in, void* const* d_outs, uint32_t log_n,
                                         const uint64_t* omega, const uint64_t* cosets, uint32_t count,
                                         void* stream) {
  if (!d_in || !d_outs || !omega || !cosets || count == 0 || count > ZKGPU_MAX_PARTS) return ZK_ERR_INVALID;
  if (log_n < 1 || log_n > MAX_LOG_N) return ZK_ERR_INVALID;
  std::vector<zkgpu_job> jobs(count);
  for (uint32_t j = 0; j < count; ++j) {
    if (!d_outs[j] || d_outs[j] == d_in) return ZK_ERR_INVALID;  // the input is read by every part
    zkgpu_job& jb = jobs[j];
    memset(&jb, 0, sizeof(jb));
    jb.d_in = d_in;
    jb.d_out = d_outs[j];
    jb.log_n = log_n;
    jb.op = ZKGPU_OP_COSET_FWD;
    memcpy(jb.omega, omega, 32);
    memcpy(jb.coset, cosets + 4 * (size_t)j, 32);
  }
  hipStream_t s = (hipStream_t)stream;
  if (log_n >= kPartsAloneLog) {
    // each part fills the device on its own: the single-transform plans
    // (tiles up to 2^12 elements, fewer passes) run them faster than a batch
    int st = ZK_OK;
    Engine* e = Engine::get(device_id, &st);
    if (!e) return st;
    for (uint32_t j = 0; j < count && st == ZK_OK; ++j) {
      JobSpec job;
      st = job_from_c(jobs[j], job, false);
      if (st == ZK_OK) st = e->run(job, s);
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == ZK_OK) st = ZK_ERR_HIP;
    return st;
  }
  BatchOpts opt;
  opt.capture = false;
  opt.use_ws = true;
  opt.ws_stream = s;
  Batch* b = nullptr;
  int st = batch_build(device_id, jobs.data(), count, &b, opt);
  if (st == ZK_OK) st = batch_launch_direct(b, s);
  // the batch's descriptors and workspaces are the stream's: finish here
  if (hipStreamSynchronize(s) != hipSuccess && st == ZK_OK) st = ZK_ERR_HIP;
  delete b;
  return st;
}

uint64_t zkgpu_coeff_to_extended_parts_api(size_t device_id, const uint64_t* coeffs, uint64_t* parts,
                                           const uint64_t* omega, uint64_t n, uint64_t log_n,
                                           const uint64_t* cosets, uint64_t count) {
  if (!coeffs || !parts || !omega || !cosets || count == 0 || count > ZKGPU_MAX_PARTS) return ZK_ERR_INVALID;
  if (log_n < 1 || log_n > MAX_LOG_N || n != (1ull << log_n)) return ZK_ERR_INVALID;
  int st = ZK_OK;
  Engine* e = Engine::get((int)device_id, &st);
  if (!e) return (uint64_t)st;
  const size_t bytes = (size_t)32 << log_n;
  HostLease ctx(e, bytes, 0, bytes * count);
  if (ctx.status()) return (uint64_t)ctx.status();
  hipStream_t s = ctx->stream;
  void* d_in = ctx->io->ptr;
  void* d_out = ctx->aux->ptr;
  std::vector<void*> outs(count);
  for (uint64_t j = 0; j < count; ++j) outs[j] = (char*)d_out + j * bytes;
  st = stage_h2d((int)device_id, s, d_in, coeffs, bytes);
  if (st == ZK_OK)
    st = zkgpu_coeff_to_extended_parts_device((int)device_id, d_in, outs.data(), (uint32_t)log_n, omega, cosets,
                                              (uint32_t)count, s);
  if (st) {
    (void)hipStreamSynchronize(s);
    return (uint64_t)st;
  }
  st = stage_d2h((int)device_id, s, parts, d_out, bytes * count);
  if (st) (void)hipStreamSynchronize(s);
  return (uint64_t)st;
}

int zkgpu_ntt_batch_device(int device_id, const zkgpu_job* jobs, uint32_t count, void* stream) {
  uint64_t h = 0;
  int st = zkgpu_batch_create(device_id, jobs, count, &h);
  if (st) return st;
  st = zkgpu_batch_launch(h, stream);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess && st == ZK_OK) st = ZK_ERR_HIP;
  zkgpu_batch_destroy(h);
  return st;
}

}  // extern "C"
