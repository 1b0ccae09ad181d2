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
  struct Level {
    uint32_t desc_off, ndesc, total, grid;
  };
  std::vector<Level> levels;
  uint32_t* d_counters = nullptr;
  PassDesc* d_descs = nullptr;
  uint32_t* d_tile_start = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  uint64_t elements = 0;
  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
};

int batch_build(int device, const zkgpu_job* cjobs, uint32_t count, Batch** out) {
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
  hipStream_t s = e->thread_stream();
  for (uint32_t i = 0; i < count; ++i) {
    st = job_from_c(cjobs[i], b->jobs[i], false);
    if (st) return st;
    st = e->prepare(b->jobs[i], b->tables[i], s, PLAN_BATCH);
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
  if (arena_bytes) {
    b->arena = std::make_shared<DevBuf>();
    b->arena->device = device;
    b->arena->bytes = arena_bytes;
    if (hipMalloc(&b->arena->ptr, arena_bytes) != hipSuccess) {
      (void)hipGetLastError();
      b->arena->ptr = nullptr;
      return ZK_ERR_NOMEM;
    }
  }
  // descriptors grouped by level; within a level the biggest jobs first so
  // that the queue ends with short tiles
  std::vector<std::vector<PassDesc>> per_job(count);
  uint32_t max_pass = 0;
  for (uint32_t i = 0; i < count; ++i) {
    void* ws = ws_off[i] == SIZE_MAX ? nullptr : (char*)b->arena->ptr + ws_off[i];
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
        total += per_job[i][lvl].num_tiles;
      }
    }
    L.ndesc = (uint32_t)all.size() - L.desc_off;
    L.total = total;
    // resident workgroups: two 2048-element tiles (72 KiB of LDS each) per CU
    L.grid = std::min<uint32_t>(total, (uint32_t)cus * 2);
    b->levels.push_back(L);
  }
  // one device buffer: descriptors | tile prefix sums | counters
  const size_t desc_bytes = all.size() * sizeof(PassDesc);
  const size_t ts_bytes = starts.size() * sizeof(uint32_t);
  const size_t cnt_bytes = b->levels.size() * sizeof(uint32_t);
  const size_t ts_off = (desc_bytes + 255) & ~(size_t)255;
  const size_t cnt_off = (ts_off + ts_bytes + 255) & ~(size_t)255;
  b->meta = std::make_shared<DevBuf>();
  b->meta->device = device;
  b->meta->bytes = cnt_off + cnt_bytes;
  if (hipMalloc(&b->meta->ptr, b->meta->bytes) != hipSuccess) {
    (void)hipGetLastError();
    b->meta->ptr = nullptr;
    return ZK_ERR_NOMEM;
  }
  std::vector<char> host(b->meta->bytes, 0);
  memcpy(host.data(), all.data(), desc_bytes);
  memcpy(host.data() + ts_off, starts.data(), ts_bytes);
  if (hipMemcpy(b->meta->ptr, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) return ZK_ERR_HIP;
  b->d_descs = (PassDesc*)b->meta->ptr;
  b->d_tile_start = (uint32_t*)((char*)b->meta->ptr + ts_off);
  b->d_counters = (uint32_t*)((char*)b->meta->ptr + cnt_off);

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

}  // namespace zkgpu

using namespace zkgpu;

extern "C" {

int zkgpu_batch_create(int device_id, const zkgpu_job* jobs, uint32_t count, uint64_t* handle) {
  if (!handle) return ZK_ERR_INVALID;
  *handle = 0;
  Batch* b = nullptr;
  const int st = batch_build(device_id, jobs, count, &b);
  if (st) return st;
  *handle = (uint64_t)(uintptr_t)b;
  return ZK_OK;
}

int zkgpu_batch_launch(uint64_t handle, void* stream) {
  Batch* b = (Batch*)(uintptr_t)handle;
  if (!b) return ZK_ERR_INVALID;
  int st = ZK_OK;
  if (!Engine::get(b->device, &st)) return st;
  return hipGraphLaunch(b->exec, (hipStream_t)stream) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
}

int zkgpu_batch_info(uint64_t handle, uint32_t* levels, uint32_t* tiles, uint64_t* elements) {
  Batch* b = (Batch*)(uintptr_t)handle;
  if (!b) return ZK_ERR_INVALID;
  if (levels) *levels = (uint32_t)b->levels.size();
  if (tiles) {
    uint32_t t = 0;
    for (const auto& L : b->levels) t += L.total;
    *tiles = t;
  }
  if (elements) *elements = b->elements;
  return ZK_OK;
}

int zkgpu_batch_destroy(uint64_t handle) {
  Batch* b = (Batch*)(uintptr_t)handle;
  if (!b) return ZK_ERR_INVALID;
  int st = ZK_OK;
  if (Engine::get(b->device, &st)) (void)hipDeviceSynchronize();  // queued launches may still read its buffers
  delete b;
  return ZK_OK;
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
