// api.cpp — host runtime behind include/rtx.h.
//
// A context owns one HIP stream on one device, the scene arena (copied or
// borrowed), the treelet partition and the scratch the kernels need.  Every
// entry point validates its arguments and returns an rtx_status; nothing
// aborts.  rtx_trace is synchronous: it returns once the rays are traced so
// that kernel-side errors (stack overflow, malformed nodes) reach the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "partition.h"
#include "rtx.h"

using namespace rtxk;

// bytes of the control block cleared before and read after a traversal
static constexpr uint32_t kSlots = 64;
// words of the mapped host block (err_host, 16 words): word 0 is the ray-error
// flag, words kReportWord.. receive launch_report_words
static constexpr uint32_t kReportWord = 8;
// most device memory the per-ray record slabs (8 bytes per record) may take
static constexpr uint64_t kSlabBytesMax = 16ull << 30;

struct rtx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evk0 = nullptr, evk1 = nullptr;  // around the last traversal kernel
  bool kernel_timed = false;
  int n_cus = 0;
  int lds_per_cu = 0;
  char name[128] = {0};
  // scene
  const uint8_t* arena = nullptr;
  uint8_t* arena_owned = nullptr;
  uint64_t arena_bytes = 0;
  uint64_t tlas = 0;
  bool ordered = false;  // every child box the TLAS reaches has lo <= hi (scene_check)
  // every reachable node has one tree level (scene_check); otherwise hits take
  // their max_level from the level pass
  bool levels_consistent = false;
  // partition
  Partition part;
  // scratch
  uint64_t* spill = nullptr;
  size_t spill_entries = 0;
  // One 256-byte control block: [0] ray counter, [1] error word, stats from
  // word 16 (16 x u64).  Cleared with one memset, read back with one copy into
  // pinned host memory.
  uint32_t* counters = nullptr;
  unsigned long long* stats = nullptr;
  // rtx_trace's rotating control slots (launch k uses slot k % kSlots; all are
  // cleared together when the index wraps) and the mapped host word its
  // kernels set on an error
  uint32_t* slots = nullptr;
  uint32_t slot = 0;
  bool slots_clean = true;
  // rtx_trace_async: the error word its launches share (read and cleared by
  // rtx_synchronize) and whether any was queued since
  uint32_t* async_err = nullptr;
  bool async_pending = false;
  uint32_t* err_host = nullptr;
  uint32_t* err_host_dev = nullptr;
  // workgroups per CU of each trace kernel, from the occupancy API on first use
  std::unordered_map<uint64_t, uint32_t> occupancy;
  uint64_t* scan_tmp = nullptr;
  size_t scan_tmp_entries = 0;
  // slab records (one traversal with records): per-ray slabs and the list of
  // rays whose records did not fit
  uint64_t* slabs = nullptr;
  size_t slab_entries = 0;
  uint32_t* overflow_list = nullptr;
  size_t overflow_entries = 0;
  uint32_t* rec_counts = nullptr;  // per-ray record counts of a slab traversal
  size_t rec_count_entries = 0;
  // level pass: per-lane tree_level_maps, zeroed when allocated (the pass
  // leaves them zeroed)
  uint8_t* lm_mem = nullptr;
  rtx_prefetch_stats last_stats{};
  // learned prefetch table of the current partition
  uint4* pred_table = nullptr;
  uint32_t pred_treelets = 0;
  // multi-scene batches
  struct SceneSlot {
    uint8_t* arena = nullptr;
    uint64_t bytes = 0;
    uint64_t tlas = 0;
    bool ordered = false;            // scene_check() of the arena
    bool levels_consistent = false;  // scene_check() of the arena
    Partition part;
    uint64_t part_gen = 0;  // bumped whenever the scene's partition is rebuilt
  };
  std::vector<SceneSlot> scenes;
  BatchJob* jobs_dev = nullptr;
  // batch graphs captured on this context (detached by rtx_destroy)
  std::vector<rtx_batch_graph*> graphs;
  // rtx_trace_static: this device's copy of the static scratch array (from
  // hipGetSymbolAddress) and the occupancy of the kernels it may launch,
  // [mode][lds_stack 4/8/16] and the eight-lanes-per-ray kernels [mode]
  uint8_t* static_scratch = nullptr;
  uint32_t static_occ[2][3] = {{0, 0, 0}, {0, 0, 0}};
  uint32_t static_wide_occ[2] = {0, 0};
};

struct rtx_batch_graph {
  rtx_ctx* ctx = nullptr;  // null once the context is destroyed
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BatchJob* jobs = nullptr;
  uint64_t* spill = nullptr;
  uint32_t* counters = nullptr;
  // treelet-order graphs read their scenes' partitions: (scene, partition
  // generation) at capture time
  std::vector<std::pair<uint32_t, uint64_t>> deps;
};

namespace {

rtx_status hip_status(hipError_t e) { return e == hipSuccess ? RTX_OK : RTX_ERR_HIP; }

#define HIP_TRY(expr)                         \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) {                   \
      last_hip_error = _e;                    \
      return RTX_ERR_HIP;                     \
    }                                         \
  } while (0)

thread_local hipError_t last_hip_error = hipSuccess;

rtx_status set_device(rtx_ctx* c) {
  HIP_TRY(hipSetDevice(c->device));
  return RTX_OK;
}

rtx_status kernel_error_status(uint32_t bits) {
  if (bits == 0) return RTX_OK;
  if (bits & kErrRange) return RTX_ERR_OFFSET_OUT_OF_ARENA;
  if (bits & kErrType) return RTX_ERR_BAD_NODE_TYPE;
  if (bits & kErrStack) return RTX_ERR_STACK_OVERFLOW;
  if (bits & kErrRecords) return RTX_ERR_TRACE_OVERFLOW;
  if (bits & kErrBudget) return RTX_ERR_NODE_LARGER_THAN_BUDGET;
  if (bits & kErrInstSlot) return RTX_ERR_UNSUPPORTED;
  if (bits & kErrSteps) return RTX_ERR_TIMEOUT;
  return RTX_ERR_HIP;
}

void free_scene(rtx_ctx* c) {
  if (c->arena_owned) hipFree(c->arena_owned);
  c->arena_owned = nullptr;
  c->arena = nullptr;
  c->arena_bytes = 0;
  c->ordered = false;
  c->levels_consistent = false;
}

rtx_status ensure_spill(rtx_ctx* c, size_t entries) {
  if (entries <= c->spill_entries) return RTX_OK;
  if (c->spill) HIP_TRY(hipFree(c->spill));
  c->spill = nullptr;
  c->spill_entries = 0;
  HIP_TRY(hipMalloc(&c->spill, entries * sizeof(uint64_t)));
  c->spill_entries = entries;
  return RTX_OK;
}

// The level pass's memory: one 256-lane workgroup per CU, each lane a map of
// kLevelMapCap entries and its position list (2.5 GiB on 256 CUs), zeroed
// once; the pass leaves every map zeroed for the next.
rtx_status ensure_level_mem(rtx_ctx* c) {
  if (c->lm_mem) return RTX_OK;
  const uint64_t bytes = (uint64_t)c->n_cus * 256 * level_pass_lane_bytes(kLevelMapCap);
  if (hipMalloc(&c->lm_mem, bytes) != hipSuccess) {
    (void)hipGetLastError();
    c->lm_mem = nullptr;
    return RTX_ERR_OUT_OF_MEMORY;
  }
  HIP_TRY(hipMemsetAsync(c->lm_mem, 0, bytes, c->stream));
  return RTX_OK;
}

// Point p's level-pass fields at `mem`, laid out for `lanes` lanes.
void set_level_maps(TraceParams& p, uint8_t* mem, uint64_t lanes) {
  p.lm_table = reinterpret_cast<uint64_t*>(mem);
  p.lm_list = reinterpret_cast<uint32_t*>(mem + lanes * kLevelMapCap * sizeof(uint64_t));
  p.lm_cap = kLevelMapCap;
}

// occupancy-cache keys of the eight-lanes-per-ray kernels: this plus the
// traversal mode (no traversal mode reaches it)
constexpr uint32_t kWideOccupancyKey = 0xff00u;

uint32_t cached_blocks_per_cu(rtx_ctx* c, uint32_t mode, uint32_t lds_stack, uint32_t stage, int records = kRecNone) {
  const uint64_t key = ((uint64_t)mode << 48) | ((uint64_t)records << 40) | ((uint64_t)lds_stack << 32) | stage;
  auto it = c->occupancy.find(key);
  if (it != c->occupancy.end()) return it->second;
  const uint32_t b = mode >= kWideOccupancyKey ? trace_wide_blocks_per_cu(mode - kWideOccupancyKey)
                                               : trace_blocks_per_cu(mode, records, lds_stack, stage);
  c->occupancy[key] = b;
  return b;
}

template <typename T>
rtx_status ensure_buffer(T*& buf, size_t& have, size_t entries) {
  if (entries <= have) return RTX_OK;
  if (buf) HIP_TRY(hipFree(buf));
  buf = nullptr;
  have = 0;
  if (hipMalloc(&buf, entries * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
  have = entries;
  return RTX_OK;
}

rtx_status ensure_scan_tmp(rtx_ctx* c, size_t entries) {
  if (entries <= c->scan_tmp_entries) return RTX_OK;
  if (c->scan_tmp) HIP_TRY(hipFree(c->scan_tmp));
  c->scan_tmp = nullptr;
  c->scan_tmp_entries = 0;
  HIP_TRY(hipMalloc(&c->scan_tmp, entries * sizeof(uint64_t)));
  c->scan_tmp_entries = entries;
  return RTX_OK;
}

}  // namespace

// The prediction table belongs to one partition: dropped whenever it changes.
static void drop_table(rtx_ctx* c) {
  if (c->pred_table) {
    hipStreamSynchronize(c->stream);
    hipFree(c->pred_table);
  }
  c->pred_table = nullptr;
  c->pred_treelets = 0;
}

namespace rtx_internal {
rtx_status ctx_stream(rtx_ctx* c, hipStream_t* s, int* device) {
  if (!c || !s || !device) return RTX_ERR_INVALID_ARGUMENT;
  *s = c->stream;
  *device = c->device;
  return RTX_OK;
}
}  // namespace rtx_internal

extern "C" {

const char* rtx_status_string(rtx_status s) {
  switch (s) {
    case RTX_OK: return "ok";
    case RTX_ERR_INVALID_ARGUMENT: return "invalid argument";
    case RTX_ERR_BAD_NODE_TYPE: return "malformed node (bad child type, non-zero PrimitiveIndex1Delta or cyclic BVH)";
    case RTX_ERR_OFFSET_OUT_OF_ARENA: return "node offset outside the arena";
    case RTX_ERR_NODE_LARGER_THAN_BUDGET: return "treelet budget smaller than a node";
    case RTX_ERR_TRACE_OVERFLOW: return "record buffer too small";
    case RTX_ERR_STACK_OVERFLOW: return "traversal stack overflow";
    case RTX_ERR_HIP: return "HIP runtime error";
    case RTX_ERR_NO_DEVICE: return "no HIP device";
    case RTX_ERR_NO_SCENE: return "no scene set";
    case RTX_ERR_NO_PARTITION: return "no treelet partition built";
    case RTX_ERR_UNSUPPORTED: return "unsupported input";
    case RTX_ERR_OUT_OF_MEMORY: return "out of memory";
    case RTX_ERR_TIMEOUT: return "timeout";
    case RTX_ERR_STALE: return "batch graph is stale (partition rebuilt or context destroyed)";
    case RTX_ERR_CYCLE: return "the BVH has a cycle";
  }
  return "unknown status";
}

const char* rtx_version(void) { return "rtx-mi355x 0.1 (gfx950)"; }

rtx_status rtx_device_count(int* count) {
  if (!count) return RTX_ERR_INVALID_ARGUMENT;
  *count = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return RTX_OK;
  }
  *count = n;
  return RTX_OK;
}

rtx_status rtx_create(int device, rtx_ctx** out) {
  if (!out) return RTX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return RTX_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= n) return RTX_ERR_INVALID_ARGUMENT;
  rtx_ctx* c = new (std::nothrow) rtx_ctx();
  if (!c) return RTX_ERR_OUT_OF_MEMORY;
  c->device = device;
  auto fail = [&](rtx_status st) {
    rtx_destroy(c);
    return st;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(RTX_ERR_HIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(RTX_ERR_HIP);
  c->n_cus = prop.multiProcessorCount;
  c->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
  snprintf(c->name, sizeof(c->name), "%s", prop.gcnArchName);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipEventCreate(&c->evk0) != hipSuccess || hipEventCreate(&c->evk1) != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipMalloc(&c->counters, 64 * sizeof(uint32_t)) != hipSuccess) return fail(RTX_ERR_OUT_OF_MEMORY);
  c->stats = reinterpret_cast<unsigned long long*>(c->counters + 16);
  if (hipMemset(c->counters, 0, 64 * sizeof(uint32_t)) != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipMalloc(&c->slots, kSlots * kSlotWords * sizeof(uint32_t)) != hipSuccess) return fail(RTX_ERR_OUT_OF_MEMORY);
  if (hipMemset(c->slots, 0, kSlots * kSlotWords * sizeof(uint32_t)) != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipMalloc(&c->async_err, 256) != hipSuccess) return fail(RTX_ERR_OUT_OF_MEMORY);
  if (hipMemset(c->async_err, 0, 256) != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipHostMalloc(&c->err_host, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return fail(RTX_ERR_OUT_OF_MEMORY);
  if (hipHostGetDevicePointer(reinterpret_cast<void**>(&c->err_host_dev), c->err_host, 0) != hipSuccess)
    return fail(RTX_ERR_HIP);
  // set-up of rtx_trace_static, so that it allocates and queries nothing
  void* sp = nullptr;
  if (static_scratch_init(device, &sp) != hipSuccess || !sp) return fail(RTX_ERR_HIP);
  c->static_scratch = static_cast<uint8_t*>(sp);
  const uint32_t ks[3] = {4, 8, 16};
  for (uint32_t m = 0; m < 2; ++m) {
    for (int k = 0; k < 3; ++k) c->static_occ[m][k] = trace_blocks_per_cu(m, kRecNone, ks[k], 0);
    c->static_wide_occ[m] = trace_wide_blocks_per_cu(m);
  }
  *out = c;
  return RTX_OK;
}

rtx_status rtx_destroy(rtx_ctx* c) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_scene(c);
  partition_free(c->part);
  drop_table(c);
  if (c->spill) hipFree(c->spill);
  if (c->counters) hipFree(c->counters);
  if (c->slots) hipFree(c->slots);
  if (c->async_err) hipFree(c->async_err);
  if (c->err_host) hipHostFree(c->err_host);
  if (c->scan_tmp) hipFree(c->scan_tmp);
  if (c->slabs) hipFree(c->slabs);
  if (c->overflow_list) hipFree(c->overflow_list);
  if (c->rec_counts) hipFree(c->rec_counts);
  if (c->lm_mem) hipFree(c->lm_mem);
  for (auto& sc : c->scenes) {
    if (sc.arena) hipFree(sc.arena);
    partition_free(sc.part);
  }
  for (rtx_batch_graph* g : c->graphs) g->ctx = nullptr;  // launches now report RTX_ERR_STALE
  if (c->jobs_dev) hipFree(c->jobs_dev);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->evk0) hipEventDestroy(c->evk0);
  if (c->evk1) hipEventDestroy(c->evk1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return RTX_OK;
}

rtx_status rtx_synchronize(rtx_ctx* c) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (!c->async_pending) return RTX_OK;
  // the first error of the traces queued by rtx_trace_async since the last wait
  c->async_pending = false;
  uint32_t err = 0;
  HIP_TRY(hipMemcpy(&err, c->async_err, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (err) HIP_TRY(hipMemset(c->async_err, 0, sizeof(uint32_t)));
  return kernel_error_status(err);
}

rtx_status rtx_get_stream(rtx_ctx* c, void** stream) {
  if (!c || !stream) return RTX_ERR_INVALID_ARGUMENT;
  *stream = reinterpret_cast<void*>(c->stream);
  return RTX_OK;
}

rtx_status rtx_malloc(rtx_ctx* c, size_t bytes, void** out) {
  if (!c || !out || bytes == 0) return RTX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  if (hipMalloc(out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return RTX_ERR_OUT_OF_MEMORY;
  }
  return RTX_OK;
}

rtx_status rtx_free(rtx_ctx* c, void* p) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (!p) return RTX_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(p));
  return RTX_OK;
}

rtx_status rtx_memcpy_h2d(rtx_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return RTX_ERR_INVALID_ARGUMENT;
  if (!bytes) return RTX_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RTX_OK;
}

rtx_status rtx_memcpy_d2d(rtx_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return RTX_ERR_INVALID_ARGUMENT;
  if (!bytes) return RTX_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RTX_OK;
}

rtx_status rtx_memcpy_d2h(rtx_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return RTX_ERR_INVALID_ARGUMENT;
  if (!bytes) return RTX_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RTX_OK;
}

rtx_status rtx_memset(rtx_ctx* c, void* dst, int value, size_t bytes) {
  if (!c || (!dst && bytes)) return RTX_ERR_INVALID_ARGUMENT;
  if (!bytes) return RTX_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemsetAsync(dst, value, bytes, c->stream));
  return RTX_OK;
}

rtx_status rtx_timer_start(rtx_ctx* c) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  return RTX_OK;
}

rtx_status rtx_timer_stop(rtx_ctx* c, float* ms) {
  if (!c || !ms) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return RTX_OK;
}

rtx_status rtx_device_info(rtx_ctx* c, int* n_cus, int* lds, char* name, size_t len) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (n_cus) *n_cus = c->n_cus;
  if (lds) *lds = c->lds_per_cu;
  if (name && len) snprintf(name, len, "%s", c->name);
  return RTX_OK;
}

static rtx_status check_scene_args(size_t bytes, uint64_t tlas) {
  if (bytes == 0 || (bytes % 64) != 0 || (tlas % 64) != 0 || tlas + 64 > bytes) return RTX_ERR_INVALID_ARGUMENT;
  if (bytes / 64 >= 0xffffffffull) return RTX_ERR_UNSUPPORTED;  // node slots are 32-bit
  return RTX_OK;
}

rtx_status rtx_set_scene(rtx_ctx* c, const void* host, size_t bytes, uint64_t tlas) {
  if (!c || !host) return RTX_ERR_INVALID_ARGUMENT;
  rtx_status st = check_scene_args(bytes, tlas);
  if (st != RTX_OK) return st;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_scene(c);
  partition_free(c->part);
  drop_table(c);
  uint8_t* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
  HIP_TRY(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
  c->arena_owned = d;
  c->arena = d;
  c->arena_bytes = bytes;
  c->tlas = tlas;
  const SceneCheck ck = scene_check(d, bytes, tlas, c->stream);
  c->ordered = ck.valid && ck.ordered;
  c->levels_consistent = ck.valid && ck.level_consistent;
  return RTX_OK;
}

rtx_status rtx_set_scene_device(rtx_ctx* c, const void* dev, size_t bytes, uint64_t tlas) {
  if (!c || !dev) return RTX_ERR_INVALID_ARGUMENT;
  rtx_status st = check_scene_args(bytes, tlas);
  if (st != RTX_OK) return st;
  if (((uintptr_t)dev & 63) != 0) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_scene(c);
  partition_free(c->part);
  drop_table(c);
  c->arena = (const uint8_t*)dev;
  c->arena_bytes = bytes;
  c->tlas = tlas;
  const SceneCheck ck = scene_check(c->arena, bytes, tlas, c->stream);
  c->ordered = ck.valid && ck.ordered;
  c->levels_consistent = ck.valid && ck.level_consistent;
  return RTX_OK;
}

rtx_status rtx_validate_scene(rtx_ctx* c, uint64_t* n_nodes) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->arena) return RTX_ERR_NO_SCENE;
  HIP_TRY(hipSetDevice(c->device));
  return partition_validate(c->arena, c->arena_bytes, c->tlas, c->stream, n_nodes);
}

rtx_status rtx_build_treelets(rtx_ctx* c, uint32_t budget, rtx_treelet_summary* out) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->arena) return RTX_ERR_NO_SCENE;
  if (budget < 128 || budget % 64 != 0 || budget > (1u << 20)) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  drop_table(c);
  rtx_status st = partition_build(c->arena, c->arena_bytes, c->tlas, budget, c->stream, c->part);
  if (st != RTX_OK) return st;
  if (out) *out = c->part.summary;
  return RTX_OK;
}

rtx_status rtx_get_treelets(rtx_ctx* c, uint64_t* roots, uint64_t* list_offsets, uint64_t* list_nodes,
                            uint32_t* owner) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->part.valid) return RTX_ERR_NO_PARTITION;
  HIP_TRY(hipSetDevice(c->device));
  return partition_copy_out(c->part, c->arena_bytes, c->stream, roots, list_offsets, list_nodes, owner);
}

rtx_status rtx_build_treelet_layout(rtx_ctx* c, uint32_t stride, rtx_treelet_layout_summary* out) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->part.valid) return RTX_ERR_NO_PARTITION;
  if (stride > (1u << 24)) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  rtx_status st = partition_layout(c->arena, c->arena_bytes, c->tlas, stride, c->stream, c->part);
  if (st != RTX_OK) return st;
  if (out) {
    memset(out, 0, sizeof(*out));
    out->stride = stride;
    out->block_bytes = c->part.budget + stride;
    out->layout_bytes = (uint64_t)c->part.n * out->block_bytes;
    out->mapped_nodes = c->part.layout_mapped;
    out->metadata_entries = c->part.budget / 64;
  }
  return RTX_OK;
}

rtx_status rtx_get_treelet_layout(rtx_ctx* c, uint64_t* remap, int32_t* metadata) {
  if (!c) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->part.valid || !c->part.layout_remap) return RTX_ERR_NO_PARTITION;
  HIP_TRY(hipSetDevice(c->device));
  if (remap)
    HIP_TRY(hipMemcpyAsync(remap, c->part.layout_remap, (c->arena_bytes / 64) * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, c->stream));
  if (metadata)
    HIP_TRY(hipMemcpyAsync(metadata, c->part.layout_metadata, (size_t)c->part.n * (c->part.budget / 64) * 4,
                           hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RTX_OK;
}

rtx_status rtx_set_treelets(rtx_ctx* c, uint32_t budget, uint32_t n, const uint64_t* roots,
                            const uint64_t* list_offsets, const uint64_t* list_nodes, const uint32_t* owner) {
  if (!c || !roots || !list_offsets || !list_nodes || !owner || n == 0) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->arena) return RTX_ERR_NO_SCENE;
  HIP_TRY(hipSetDevice(c->device));
  drop_table(c);
  return partition_install(c->arena, c->arena_bytes, c->tlas, budget, n, roots, list_offsets, list_nodes, owner,
                           c->stream, c->part);
}

}  // extern "C"

namespace {

// rtx_trace, and with `async` rtx_trace_async: the launches are queued and
// the call returns without waiting (hits only, no staging).
rtx_status trace_impl(rtx_ctx* c, const rtx_ray* rays, uint32_t n, const rtx_trace_options* opts, rtx_hit* hits,
                      rtx_record* records, uint64_t capacity, uint64_t* offsets, bool async) {
  if (!c || !opts || (n > 0 && (!rays || !hits))) return RTX_ERR_INVALID_ARGUMENT;
  if (async && (opts->emit_records || opts->prefetch != RTX_PREFETCH_NONE)) return RTX_ERR_UNSUPPORTED;
  if (opts->mode != RTX_MODE_PLAIN && opts->mode != RTX_MODE_TREELET) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->prefetch > RTX_PREFETCH_LATTER) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->emit_records && !offsets) return RTX_ERR_INVALID_ARGUMENT;
  if (n > 0 ? opts->claim_start >= n : opts->claim_start != 0) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->claim_order && (opts->claim_start != 0 || ((uintptr_t)opts->claim_order & 3)))
    return RTX_ERR_INVALID_ARGUMENT;
  if (opts->record_slab > 1024 || opts->record_layout > 1) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->arena) return RTX_ERR_NO_SCENE;
  if (opts->record_layout && (!c->part.valid || !c->part.layout_remap)) return RTX_ERR_NO_PARTITION;
  if ((opts->mode == RTX_MODE_TREELET || opts->prefetch != RTX_PREFETCH_NONE) && !c->part.valid)
    return RTX_ERR_NO_PARTITION;
  if (((uintptr_t)rays & 15) || ((uintptr_t)hits & 15) || ((uintptr_t)records & 15)) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  memset(&c->last_stats, 0, sizeof(c->last_stats));
  if (n == 0) {
    if (opts->emit_records) HIP_TRY(hipMemsetAsync(offsets, 0, sizeof(uint64_t), c->stream));
    if (!async) HIP_TRY(hipStreamSynchronize(c->stream));
    return RTX_OK;
  }
  // 16 KiB of LDS stack per 256-lane block (8 entries per lane in plain order,
  // 4 per list in treelet order), so the register budget, not LDS, bounds how
  // many waves share a CU
  uint32_t lds_stack = opts->lds_stack ? opts->lds_stack : 8;
  if (lds_stack != 4 && lds_stack != 8 && lds_stack != 16) return RTX_ERR_INVALID_ARGUMENT;
  // treelet staging: one wave per workgroup with stage_bytes of LDS for the image
  uint32_t stage = 0;
  if (opts->prefetch != RTX_PREFETCH_NONE) {
    stage = opts->stage_bytes;
    if (stage == 0) {
      stage = 4096;
      while (stage < c->part.budget && stage < 16384) stage *= 2;
    }
    if (stage < 4096 || stage > 65536 || (stage & (stage - 1))) return RTX_ERR_INVALID_ARGUMENT;
    if (opts->slot_bytes && (opts->slot_bytes < 64 || opts->slot_bytes > stage ||
                             (opts->slot_bytes & (opts->slot_bytes - 1))))
      return RTX_ERR_INVALID_ARGUMENT;
    if (!(opts->min_score >= 0.f)) return RTX_ERR_INVALID_ARGUMENT;
    if (!(opts->pred_fraction >= 0.f && opts->pred_fraction <= 1.f)) return RTX_ERR_INVALID_ARGUMENT;
    lds_stack = 8;
  }
  // Eight lanes per ray (no records or staging) for batches that leave most of
  // the GPU idle, at most 128 rays per CU in plain order and 256 in treelet
  // order (where the two kernels meet on the 1M-triangle scene): their time is
  // the step latency of the longest rays, and a group tests a node's six
  // children at once.
  if (opts->lanes_per_ray != 0 && opts->lanes_per_ray != 1 && opts->lanes_per_ray != kWideRayLanes)
    return RTX_ERR_INVALID_ARGUMENT;
  const bool wide_ok = !opts->emit_records && stage == 0;
  // records in one traversal: per-ray slabs of `slab` records (default 256: a
  // 1M-triangle diffuse batch has at most about 210 per ray), compacted into the
  // caller's CSR buffer; only rays with more records are traced again.  The
  // slabs take at most 16 GiB: a larger batch is traced in consecutive ray
  // ranges that fit, each compacted before the next reuses the slabs.  With a
  // claim order or staging such a batch takes a counting pass and a writing
  // pass instead.
  const uint32_t slab = opts->record_slab ? opts->record_slab : 256;
  const uint64_t chunk_rays = std::max<uint64_t>(1, kSlabBytesMax / ((uint64_t)slab_stride(slab) * sizeof(uint64_t)));
  const bool use_slab = opts->emit_records && ((uint64_t)n <= chunk_rays || (!opts->claim_order && stage == 0));
  const uint32_t per_chunk = (uint32_t)std::min<uint64_t>(n, chunk_rays);
  const uint32_t n_chunks = use_slab ? (n + per_chunk - 1) / per_chunk : 1;
  const int main_rec = !opts->emit_records ? kRecNone : use_slab ? kRecSlab : kRecNone;
  if (opts->lanes_per_ray == kWideRayLanes && !wide_ok) return RTX_ERR_INVALID_ARGUMENT;
  const bool wide = wide_ok && (opts->lanes_per_ray == kWideRayLanes ||
                                (opts->lanes_per_ray == 0 &&
                                 (uint64_t)n <= (opts->mode == RTX_MODE_PLAIN ? 128ull : 256ull) * c->n_cus));
  const uint32_t bs = trace_block_size(stage);
  // persistent blocks: by default as many as fit on a CU at once
  uint32_t blocks_per_cu = opts->waves_per_cu ? std::max(1u, opts->waves_per_cu * 64 / bs)
                           : wide             ? cached_blocks_per_cu(c, kWideOccupancyKey + opts->mode, 0, 0)
                                              : cached_blocks_per_cu(c, opts->mode, lds_stack, stage, main_rec);
  if (blocks_per_cu == 0) blocks_per_cu = 16 * 64 / bs;
  uint32_t grid = (uint32_t)c->n_cus * blocks_per_cu;
  const uint64_t lanes = (uint64_t)n * (wide ? kWideRayLanes : 1u);
  grid = (uint32_t)std::min<uint64_t>(grid, (lanes + bs - 1) / bs);
  grid = std::max(grid, 1u);
  const uint32_t spill_cap = 128;
  const uint32_t ns = opts->mode == RTX_MODE_TREELET ? 2 : 1;
  rtx_status st = ensure_spill(c, (size_t)grid * bs * spill_cap * ns);
  if (st != RTX_OK) return st;

  TraceParams p{};
  p.arena = c->arena;
  p.arena_bytes = c->arena_bytes;
  p.tlas = c->tlas;
  p.rays = rays;
  p.n = n;
  if (opts->refill > 64) return RTX_ERR_INVALID_ARGUMENT;
  // default refill: 40 idle lanes (on the 1M-triangle diffuse batch 2.6 %
  // faster than 32 and 0.5 % slower than 48; on the 21M-triangle scene 0.2 %
  // faster than 32 and 1.2 % faster than 48)
  p.refill = opts->refill ? opts->refill : 40;
  p.ordered_boxes = c->ordered ? 1u : 0u;
  p.hits = hits;
  p.records = nullptr;
  p.record_capacity = 0;
  p.record_offsets = nullptr;
  p.claim_start = opts->claim_start;
  p.ray_list = opts->claim_order;
  p.n_rays = n;
  p.layout_remap = opts->record_layout ? c->part.layout_remap : nullptr;
  p.owner = c->part.valid ? c->part.owner : nullptr;
  p.first_treelet = c->part.valid ? c->part.first_treelet : 0;
  p.spill = c->spill;
  p.spill_cap = spill_cap;
  p.prefetch = opts->prefetch;
  p.threshold = opts->threshold;
  if (c->part.valid) {
    p.image_slots = c->part.image_slots;
    p.image_offset = c->part.image_offset;
    p.image_bytes = c->part.image_bytes;
    p.images = c->part.images;
  }
  p.slot_units = opts->slot_bytes / 64;
  // defaults: the best hit rate / wasted bytes pair of a sweep on held-out
  // diffuse rays of the 1M-triangle instanced-sphere scene (1 KiB treelets,
  // 8 KiB staging area in 1 KiB slots)
  p.min_score = opts->min_score > 0.f ? opts->min_score : 16.0f;
  p.pred_fraction = opts->pred_fraction > 0.f ? opts->pred_fraction : 0.25f;
  if (opts->prefetch == RTX_PREFETCH_LEARNED && c->pred_table && c->pred_treelets == c->part.n) {
    p.pred_table = c->pred_table;
    p.pred_treelets = c->pred_treelets;
  }

  // Rotating control slots: each launch uses a fresh slot, and one memset
  // clears all of them every kSlots launches; any ray error also sets a mapped
  // host word, so an error-free launch needs no read-back.
  volatile uint32_t* herr = c->err_host;
  auto use_slot = [&]() -> rtx_status {
    if (!c->slots_clean || c->slot == 0) {
      HIP_TRY(hipMemsetAsync(c->slots, 0, kSlots * kSlotWords * sizeof(uint32_t), c->stream));
      c->slot = 0;
    }
    c->slots_clean = false;  // until the launch is known to have run
    uint32_t* cur = c->slots + (size_t)c->slot * kSlotWords;
    c->slot = (c->slot + 1) % kSlots;
    p.ray_counter = cur;
    p.error = async ? c->async_err : cur + 1;
    p.stats = reinterpret_cast<unsigned long long*>(cur + 16);
    p.host_err = async ? nullptr : c->err_host_dev;
    if (!async) *herr = 0;
    return RTX_OK;
  };
  auto launch_error = [&]() -> rtx_status {  // after a synchronize
    if (*herr == 0) return RTX_OK;
    uint32_t err = 0;
    HIP_TRY(hipMemcpy(&err, p.error, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return kernel_error_status(err);
  };
  c->kernel_timed = false;
  if (use_slab) {
    st = ensure_buffer(c->slabs, c->slab_entries, (size_t)per_chunk * slab_stride(slab));
    if (st != RTX_OK) return st;
    st = ensure_buffer(c->rec_counts, c->rec_count_entries, (size_t)n);
    if (st != RTX_OK) return st;
    st = ensure_buffer(c->overflow_list, c->overflow_entries, (size_t)n);
    if (st != RTX_OK) return st;
    p.slabs = c->slabs;
    p.record_slab = slab;
    p.rec_counts = c->rec_counts;
  }
  // scan tiles of the largest scanned range, and one word for a range's base offset
  const uint32_t scan_n = use_slab ? per_chunk : n;
  if (opts->emit_records) {
    st = ensure_scan_tmp(c, (size_t)(scan_n + 1023) / 1024 + 2);
    if (st != RTX_OK) return st;
  }
  uint64_t* const base_word = opts->emit_records ? c->scan_tmp + (scan_n + 1023) / 1024 + 1 : nullptr;
  st = use_slot();
  if (st != RTX_OK) return st;
  uint32_t* main_slot = p.ray_counter;
  const TraceParams scene_params = p;  // for the level pass
  // One ray range of a slab traversal (all of it when it fits): traced, its
  // counts scanned into offsets[b..e] (plus the previous ranges' total), its
  // slabs compacted.  The ranges share the control slot: its ray counter is
  // cleared between them, its error word and word 2 (the rays whose records
  // overflowed their slab) collect all of them.
  auto slab_range = [&](uint32_t b, uint32_t e, bool last) -> rtx_status {
    TraceParams q = p;
    q.rays = rays + b;
    q.hits = hits + b;
    q.n = q.n_rays = e - b;
    q.claim_start = opts->claim_start >= b && opts->claim_start < e ? opts->claim_start - b : 0;
    q.rec_counts = c->rec_counts + b;
    if (b > 0) HIP_TRY(launch_fill_zero(main_slot, 1, c->stream));
    const uint32_t gq = std::max(1u, (uint32_t)std::min<uint64_t>(grid, ((uint64_t)(e - b) + bs - 1) / bs));
    HIP_TRY(launch_trace(q, opts->mode, kRecSlab, lds_stack, stage, gq, c->stream));
    if (last) HIP_TRY(hipEventRecord(c->evk1, c->stream));
    if (b == 0)
      HIP_TRY(launch_fill_zero(reinterpret_cast<uint32_t*>(base_word), 2, c->stream));
    else
      HIP_TRY(hipMemcpyAsync(base_word, offsets + b, sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(launch_scan_records(nullptr, c->rec_counts + b, e - b, offsets + b, c->scan_tmp, c->stream));
    if (b > 0) HIP_TRY(launch_add_offset(offsets + b, (uint64_t)(e - b) + 1, base_word, c->stream));
    // the kernel copies nothing when the running total exceeds the capacity
    HIP_TRY(launch_compact_records(c->rec_counts + b, e - b, c->slabs, slab, p.owner, p.layout_remap, offsets + b,
                                   records, records ? capacity : 0, b, c->overflow_list, main_slot + 2, c->stream));
    return RTX_OK;
  };
  // the timed span: the traversal launch (from the first range's to the end of the last range's)
  HIP_TRY(hipEventRecord(c->evk0, c->stream));
  if (wide) {
    HIP_TRY(launch_trace_wide(p, opts->mode, lds_stack + spill_cap, grid, c->stream));
    HIP_TRY(hipEventRecord(c->evk1, c->stream));
  } else if (use_slab) {
    for (uint32_t k = 0; k < n_chunks; ++k) {
      st = slab_range(k * per_chunk, (uint32_t)std::min<uint64_t>(n, (uint64_t)(k + 1) * per_chunk), k + 1 == n_chunks);
      if (st != RTX_OK) return st;
    }
  } else {
    HIP_TRY(launch_trace(p, opts->mode, main_rec, lds_stack, stage, grid, c->stream));
    HIP_TRY(hipEventRecord(c->evk1, c->stream));
  }
  c->kernel_timed = true;
  // staging statistics of the main traversal (a second pass traces some rays again)
  unsigned long long stats[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (stage) HIP_TRY(hipMemcpyAsync(stats, p.stats, sizeof(stats), hipMemcpyDeviceToHost, c->stream));
  if (opts->emit_records) {
    if (!use_slab) HIP_TRY(launch_scan_records(hits, nullptr, n, offsets, c->scan_tmp, c->stream));
    // the rays the writing pass traces (slab overflows) and the records'
    // total, through mapped host words
    HIP_TRY(launch_report_words(use_slab ? main_slot + 2 : nullptr, offsets + n, c->err_host_dev + kReportWord,
                                c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    st = launch_error();
    if (st != RTX_OK) return st;
    const volatile uint32_t* rep = c->err_host + kReportWord;
    const uint32_t second = use_slab ? rep[0] : n;
    const uint64_t total = (uint64_t)rep[2] | ((uint64_t)rep[3] << 32);
    if (total > capacity || (total > 0 && !records)) {
      c->slots_clean = true;
      return RTX_ERR_TRACE_OVERFLOW;
    }
    p.records = records;
    p.record_capacity = capacity;
    p.record_offsets = offsets;
    if (use_slab) {
      p.ray_list = c->overflow_list;
      p.n = second;
      p.claim_start = 0;
      p.record_slab = 0;
      p.rec_counts = nullptr;
    }
    c->slots_clean = true;
    if (second > 0) {
      // the rays again, every record to its CSR place (hits are rewritten
      // with the same values)
      const uint32_t g2 = std::max(1u, (uint32_t)std::min<uint64_t>(grid, ((uint64_t)second + bs - 1) / bs));
      st = use_slot();
      if (st != RTX_OK) return st;
      HIP_TRY(launch_trace(p, opts->mode, kRecCsr, lds_stack, stage, g2, c->stream));
    }
  }
  if (!c->levels_consistent) {
    // a node is reachable at two levels: the rays again, keeping the
    // reference's tree_level_map, for max_level
    if (!async) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      st = launch_error();
      if (st != RTX_OK) return st;
    }
    st = ensure_level_mem(c);
    if (st != RTX_OK) return st;
    TraceParams lp = scene_params;
    const uint32_t lm_grid = std::max(1u, std::min<uint32_t>(c->n_cus, (n + 255) / 256));
    st = ensure_spill(c, (size_t)lm_grid * 256 * spill_cap * ns);
    if (st != RTX_OK) return st;
    lp.spill = c->spill;
    lp.records = nullptr;
    lp.slabs = nullptr;
    lp.record_slab = 0;
    lp.rec_counts = nullptr;
    lp.prefetch = RTX_PREFETCH_NONE;
    lp.ray_list = nullptr;
    lp.claim_start = 0;
    lp.n = n;
    set_level_maps(lp, c->lm_mem, (uint64_t)c->n_cus * 256);
    st = use_slot();
    if (st != RTX_OK) return st;
    lp.ray_counter = p.ray_counter;
    lp.error = p.error;
    lp.stats = p.stats;
    HIP_TRY(launch_level_pass(lp, opts->mode, lm_grid, c->stream));
  }
  if (async) {
    c->async_pending = true;
    return RTX_OK;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->slots_clean = true;
  c->last_stats.node_reads = stats[0];
  c->last_stats.lds_hits = stats[1];
  c->last_stats.staged_treelets = stats[2];
  c->last_stats.staged_bytes = stats[3];
  c->last_stats.wasted_bytes = stats[4];
  c->last_stats.predictions = stats[5];
  c->last_stats.prediction_hits = stats[6];
  c->last_stats.fallbacks = stats[7];
  c->last_stats.timely_bytes = stats[8];
  c->last_stats.too_early_bytes = stats[9];
  c->last_stats.too_late_bytes = stats[10];
  c->last_stats.late_bytes = 0;
  return launch_error();
}

}  // namespace

extern "C" {

rtx_status rtx_trace(rtx_ctx* c, const rtx_ray* rays, uint32_t n, const rtx_trace_options* opts, rtx_hit* hits,
                     rtx_record* records, uint64_t capacity, uint64_t* offsets) {
  return trace_impl(c, rays, n, opts, hits, records, capacity, offsets, false);
}

rtx_status rtx_trace_async(rtx_ctx* c, const rtx_ray* rays, uint32_t n, const rtx_trace_options* opts,
                           rtx_hit* hits) {
  return trace_impl(c, rays, n, opts, hits, nullptr, 0, nullptr, true);
}

// rtx_trace without allocation: hits only, the library's static device
// scratch for control words and stack spill (see rtx.h).
rtx_status rtx_trace_static(rtx_ctx* c, const rtx_ray* rays, uint32_t n, const rtx_trace_options* opts,
                            rtx_hit* hits) {
  if (!c || !opts || (n > 0 && (!rays || !hits))) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->mode != RTX_MODE_PLAIN && opts->mode != RTX_MODE_TREELET) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->prefetch != RTX_PREFETCH_NONE || opts->emit_records) return RTX_ERR_UNSUPPORTED;
  if (n > 0 ? opts->claim_start >= n : opts->claim_start != 0) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->claim_order && (opts->claim_start != 0 || ((uintptr_t)opts->claim_order & 3)))
    return RTX_ERR_INVALID_ARGUMENT;
  if (opts->refill > 64) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->arena) return RTX_ERR_NO_SCENE;
  if (opts->mode == RTX_MODE_TREELET && !c->part.valid) return RTX_ERR_NO_PARTITION;
  if (((uintptr_t)rays & 15) || ((uintptr_t)hits & 15)) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->static_scratch) return RTX_ERR_HIP;
  const uint32_t lds_stack = opts->lds_stack ? opts->lds_stack : 8;
  const int ki = lds_stack == 4 ? 0 : lds_stack == 8 ? 1 : lds_stack == 16 ? 2 : -1;
  if (ki < 0) return RTX_ERR_INVALID_ARGUMENT;
  if (opts->lanes_per_ray != 0 && opts->lanes_per_ray != 1 && opts->lanes_per_ray != kWideRayLanes)
    return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  c->kernel_timed = false;
  memset(&c->last_stats, 0, sizeof(c->last_stats));
  if (n == 0) return RTX_OK;
  const bool wide = opts->lanes_per_ray == kWideRayLanes ||
                    (opts->lanes_per_ray == 0 &&
                     (uint64_t)n <= (opts->mode == RTX_MODE_PLAIN ? 128ull : 256ull) * c->n_cus);
  const uint32_t bs = 256;
  uint32_t per_cu = opts->waves_per_cu ? std::max(1u, opts->waves_per_cu * 64 / bs)
                    : wide             ? c->static_wide_occ[opts->mode]
                                       : c->static_occ[opts->mode][ki];
  if (per_cu == 0) per_cu = 4;
  const uint64_t lanes = (uint64_t)n * (wide ? kWideRayLanes : 1u);
  uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)c->n_cus * per_cu, (lanes + bs - 1) / bs);
  grid = std::max(grid, 1u);
  // The spill area is what the scratch holds beyond its control words, split
  // over every lane of the launch (one region per list in treelet order);
  // the grid shrinks when fewer than kStaticMinSpill entries per lane would fit.
  const uint32_t ns = opts->mode == RTX_MODE_TREELET ? 2 : 1;
  const uint64_t spill_bytes = kStaticScratchBytes - kStaticControlBytes;
  uint32_t spill_cap = 0;
  if (!wide) {
    const uint64_t max_grid = spill_bytes / ((uint64_t)bs * ns * 8 * kStaticMinSpill);
    grid = (uint32_t)std::min<uint64_t>(grid, max_grid);
    spill_cap = (uint32_t)std::min<uint64_t>(128, spill_bytes / ((uint64_t)grid * bs * ns * 8));
  }
  uint32_t* control = reinterpret_cast<uint32_t*>(c->static_scratch);  // [0] ray counter, [1] error word

  TraceParams p{};
  p.arena = c->arena;
  p.arena_bytes = c->arena_bytes;
  p.tlas = c->tlas;
  p.rays = rays;
  p.n = n;
  p.refill = opts->refill ? opts->refill : 40;
  p.ordered_boxes = c->ordered ? 1u : 0u;
  p.hits = hits;
  p.owner = c->part.valid ? c->part.owner : nullptr;
  p.first_treelet = c->part.valid ? c->part.first_treelet : 0;
  p.spill = reinterpret_cast<uint64_t*>(c->static_scratch + kStaticControlBytes);
  p.spill_cap = spill_cap;
  p.claim_start = opts->claim_start;
  p.ray_list = opts->claim_order;
  p.n_rays = n;
  p.ray_counter = control;
  p.error = control + 1;
  p.host_err = c->err_host_dev;
  volatile uint32_t* herr = c->err_host;
  *herr = 0;
  // every call starts from cleared control words (a fill kernel, no memset)
  HIP_TRY(launch_fill_zero(control, 2, c->stream));
  HIP_TRY(hipEventRecord(c->evk0, c->stream));
  if (wide)
    HIP_TRY(launch_trace_wide(p, opts->mode, lds_stack + 128, grid, c->stream));
  else
    HIP_TRY(launch_trace(p, opts->mode, kRecNone, lds_stack, 0, grid, c->stream));
  HIP_TRY(hipEventRecord(c->evk1, c->stream));
  c->kernel_timed = true;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (*herr == 0 && !c->levels_consistent) {
    // The level pass in the same array, once the traversal is done with it:
    // per lane a zeroed map, its position list and 128 spill entries per list.
    const uint64_t lane_bytes = level_pass_lane_bytes(kLevelMapCap) + 128ull * 8 * ns;
    const uint32_t lm_grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((spill_bytes / lane_bytes) / 256, ((uint64_t)n + 255) / 256));
    const uint64_t lanes = (uint64_t)lm_grid * 256;
    uint8_t* base = c->static_scratch + kStaticControlBytes;
    TraceParams lp = p;
    lp.ray_list = nullptr;
    lp.claim_start = 0;
    set_level_maps(lp, base, lanes);
    lp.spill = reinterpret_cast<uint64_t*>(base + lanes * level_pass_lane_bytes(kLevelMapCap));
    lp.spill_cap = 128;
    HIP_TRY(launch_fill_zero(control, 2, c->stream));
    HIP_TRY(launch_fill_zero(reinterpret_cast<uint32_t*>(base),
                             (uint32_t)(lanes * level_pass_lane_bytes(kLevelMapCap) / 4), c->stream));
    HIP_TRY(launch_level_pass(lp, opts->mode, lm_grid, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  if (*herr == 0) return RTX_OK;
  uint32_t err = 0;
  HIP_TRY(hipMemcpy(&err, control + 1, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return kernel_error_status(err);
}

rtx_status rtx_coherence_order(rtx_ctx* c, const rtx_ray* rays, uint32_t n, const rtx_coherence_config* cfg,
                               uint32_t* order, uint32_t* hashes) {
  if (!c || !cfg || (n > 0 && (!rays || !order))) return RTX_ERR_INVALID_ARGUMENT;
  const char h = cfg->hash ? cfg->hash : 'd';
  if (h != 'f' && h != 'g' && h != 't' && h != 'd') return RTX_ERR_INVALID_ARGUMENT;
  // every hash below 2^32 (and the spherical shifts non-negative)
  if (h == 'f' && (cfg->francois_bits < 1 || cfg->francois_bits > 4)) return RTX_ERR_INVALID_ARGUMENT;
  if ((h == 'g' || h == 't') && (cfg->grid_bits < 1 || cfg->grid_bits > 10)) return RTX_ERR_INVALID_ARGUMENT;
  if ((h == 'g' || h == 'd') && (cfg->sphere_bits < 1 || cfg->sphere_bits > 8)) return RTX_ERR_INVALID_ARGUMENT;
  if (cfg->world > RTX_COHERENCE_WORLD_GIVEN || ((uintptr_t)rays & 15)) return RTX_ERR_INVALID_ARGUMENT;
  float wmin[3] = {0.f, 0.f, 0.f}, wmax[3] = {0.f, 0.f, 0.f};
  if (cfg->world == RTX_COHERENCE_WORLD_GIVEN) {
    for (int a = 0; a < 3; ++a) {
      wmin[a] = cfg->world_min[a];
      wmax[a] = cfg->world_max[a];
    }
  } else {
    if (!c->arena) return RTX_ERR_NO_SCENE;
    HIP_TRY(hipSetDevice(c->device));
    uint8_t hdr[64], node[64];
    HIP_TRY(hipMemcpy(hdr, c->arena + c->tlas, 64, hipMemcpyDeviceToHost));
    if (cfg->world == RTX_COHERENCE_WORLD_TLAS) {
      memcpy(wmin, hdr + 8, 12);
      memcpy(wmax, hdr + 20, 12);
    } else {
      // vulkan_ray_tracing.cc:2452-2465: both corners take a running min from
      // (0,0,0) over the TLAS root's child boxes (set_child_bounds: origin +
      // ldexpf(q, exponent - 8), exact)
      uint64_t root_off = 0;
      memcpy(&root_off, hdr, 8);
      const uint64_t root = c->tlas + root_off;
      if ((root & 63) || root + 64 > c->arena_bytes) return RTX_ERR_OFFSET_OUT_OF_ARENA;
      HIP_TRY(hipMemcpy(node, c->arena + root, 64, hipMemcpyDeviceToHost));
      float origin[3];
      memcpy(origin, node, 12);
      for (int i = 0; i < 6; ++i) {
        if ((node[22 + i] & 3u) == 0) continue;
        for (int a = 0; a < 3; ++a) {
          const int e = (int)(int8_t)node[18 + a];
          const float lo = origin[a] + std::ldexp((float)node[28 + 12 * a + i], e - 8);
          const float hi = origin[a] + std::ldexp((float)node[34 + 12 * a + i], e - 8);
          wmin[a] = lo < wmin[a] ? lo : wmin[a];  // std::min(world, lo)
          wmax[a] = hi < wmax[a] ? hi : wmax[a];  // std::min(world, hi), as the reference writes it
        }
      }
    }
  }
  HIP_TRY(hipSetDevice(c->device));
  rtx_coherence_config k = *cfg;
  k.hash = h;
  if (n == 0) return RTX_OK;
  const hipError_t e = coherence_order(rays, n, k, wmin, wmax, order, hashes, c->stream);
  if (e == hipErrorOutOfMemory) return RTX_ERR_OUT_OF_MEMORY;
  return e == hipSuccess ? RTX_OK : RTX_ERR_HIP;
}

rtx_status rtx_trace_stats_of(rtx_ctx* c, const rtx_hit* hits, uint32_t n, const rtx_record* records,
                              uint64_t n_records, rtx_trace_stats* out) {
  if (!c || !out || (n > 0 && !hits) || (n_records > 0 && !records)) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  unsigned long long* acc = nullptr;
  if (hipMalloc(&acc, 16 * sizeof(unsigned long long)) != hipSuccess) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
  unsigned long long h[16];
  hipError_t e = hipMemsetAsync(acc, 0, sizeof(h), c->stream);
  if (e == hipSuccess) e = launch_trace_stats(hits, n, n_records ? records : nullptr, n_records, acc, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(acc);
  if (e != hipSuccess) return RTX_ERR_HIP;
  memset(out, 0, sizeof(*out));
  for (int k = 0; k < 7; ++k) out->access_type[k] = h[k];
  out->num_hits = h[7];
  out->num_any_hits = h[8];
  out->n_rays = n;
  out->tot_nodes = h[9];
  out->max_nodes_per_ray = (uint32_t)h[10];
  out->max_tree_depth = (uint32_t)h[11];
  out->fetch_bytes = h[12];
  out->accessed_bytes = h[13];
  out->from_records = n_records ? 1u : 0u;
  return RTX_OK;
}

rtx_status rtx_set_prefetch_table(rtx_ctx* c, const uint32_t* table, uint64_t n_rows) {
  if (!c || !table) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->part.valid) return RTX_ERR_NO_PARTITION;
  if (n_rows != (uint64_t)c->part.n * RTX_OCTANTS) return RTX_ERR_INVALID_ARGUMENT;
  for (uint64_t i = 0; i < n_rows * 4; ++i) {
    const uint32_t t = table[i] & 0xffffffu;
    if (t != 0 && t - 1 >= c->part.n) return RTX_ERR_INVALID_ARGUMENT;
  }
  HIP_TRY(hipSetDevice(c->device));
  drop_table(c);
  uint4* d = nullptr;
  if (hipMalloc(&d, n_rows * sizeof(uint4)) != hipSuccess) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
  if (hipMemcpy(d, table, n_rows * sizeof(uint4), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    hipFree(d);
    return RTX_ERR_HIP;
  }
  c->pred_table = d;
  c->pred_treelets = c->part.n;
  return RTX_OK;
}

rtx_status rtx_last_kernel_ms(rtx_ctx* c, float* ms) {
  if (!c || !ms) return RTX_ERR_INVALID_ARGUMENT;
  if (!c->kernel_timed) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipEventSynchronize(c->evk1));
  HIP_TRY(hipEventElapsedTime(ms, c->evk0, c->evk1));
  return RTX_OK;
}

rtx_status rtx_get_prefetch_stats(rtx_ctx* c, rtx_prefetch_stats* out) {
  if (!c || !out) return RTX_ERR_INVALID_ARGUMENT;
  *out = c->last_stats;
  return RTX_OK;
}

}  // extern "C"

// ---- multi-scene batches -----------------------------------------------------
namespace {

constexpr uint32_t kBatchSpillCap = 128;

rtx_status make_jobs(rtx_ctx* c, const rtx_batch_job* jobs, uint32_t n_jobs, uint32_t mode,
                     std::vector<BatchJob>& out, uint64_t* total) {
  if (!jobs || n_jobs == 0 || n_jobs > kMaxBatchJobs) return RTX_ERR_INVALID_ARGUMENT;
  out.assign(n_jobs, BatchJob{});
  uint64_t base = 0;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const rtx_batch_job& jb = jobs[j];
    if (jb.scene >= c->scenes.size()) return RTX_ERR_INVALID_ARGUMENT;
    const auto& sc = c->scenes[jb.scene];
    if (mode == RTX_MODE_TREELET && !sc.part.valid) return RTX_ERR_NO_PARTITION;
    if (jb.n_rays > 0 && (!jb.rays_dev || !jb.hits_dev)) return RTX_ERR_INVALID_ARGUMENT;
    if (((uintptr_t)jb.rays_dev & 15) || ((uintptr_t)jb.hits_dev & 15)) return RTX_ERR_INVALID_ARGUMENT;
    BatchJob& b = out[j];
    b.arena = sc.arena;
    b.arena_bytes = sc.bytes;
    b.tlas = sc.tlas;
    b.rays = jb.rays_dev;
    b.hits = jb.hits_dev;
    b.owner = sc.part.valid ? sc.part.owner : nullptr;
    b.first_treelet = sc.part.valid ? sc.part.first_treelet : 0;
    b.n = jb.n_rays;
    b.ray_base = (uint32_t)base;
    base += jb.n_rays;
    if (base >= 0xffffffffull) return RTX_ERR_UNSUPPORTED;
  }
  *total = base;
  return RTX_OK;
}

struct BatchShape {
  uint32_t grid, lds_stack, refill, ns;
};

rtx_status batch_shape(const rtx_ctx* c, const rtx_trace_options* o, uint64_t total, BatchShape* s) {
  if (!o || (o->mode != RTX_MODE_PLAIN && o->mode != RTX_MODE_TREELET)) return RTX_ERR_INVALID_ARGUMENT;
  if (o->prefetch != RTX_PREFETCH_NONE || o->emit_records) return RTX_ERR_UNSUPPORTED;
  s->lds_stack = o->lds_stack ? o->lds_stack : 8;
  if (s->lds_stack != 4 && s->lds_stack != 8 && s->lds_stack != 16) return RTX_ERR_INVALID_ARGUMENT;
  if (o->refill > 64) return RTX_ERR_INVALID_ARGUMENT;
  // batches mix camera, shadow and bounce rays; larger refills keep the
  // coherent ones together
  s->refill = o->refill ? o->refill : 48;
  uint32_t per_cu = o->waves_per_cu ? std::max(1u, o->waves_per_cu / 4) : trace_batch_blocks_per_cu(o->mode, s->lds_stack);
  if (per_cu == 0) per_cu = 4;
  uint32_t grid = (uint32_t)c->n_cus * per_cu;
  grid = (uint32_t)std::min<uint64_t>(grid, (total + kBlock - 1) / kBlock);
  s->grid = std::max(grid, 1u);
  s->ns = o->mode == RTX_MODE_TREELET ? 2 : 1;
  return RTX_OK;
}

TraceParams batch_params(rtx_ctx* c, const BatchJob* jobs_dev, uint32_t n_jobs, uint64_t total, const BatchShape& s,
                         uint64_t* spill, uint32_t* counters) {
  TraceParams p{};
  p.n = (uint32_t)total;
  p.refill = s.refill;
  p.spill = spill;
  p.spill_cap = kBatchSpillCap;
  p.ray_counter = counters;
  p.error = counters + 1;
  p.stats = c->stats;
  p.jobs = jobs_dev;
  p.n_jobs = n_jobs;
  // sign-ordered slab tests only when every scene of the context allows them
  p.ordered_boxes = 1;
  for (const auto& sc : c->scenes)
    if (!sc.ordered) p.ordered_boxes = 0;
  return p;
}

// Whether any job of the batch traces a scene that needs the level pass.
bool batch_needs_levels(const rtx_ctx* c, const rtx_batch_job* jobs, uint32_t n_jobs) {
  for (uint32_t j = 0; j < n_jobs; ++j)
    if (jobs[j].n_rays > 0 && !c->scenes[jobs[j].scene].levels_consistent) return true;
  return false;
}

// The level pass of every job whose scene needs it, queued after the batch
// launch (capturable: the ray counter is cleared by a kernel, and the error
// word, counters[1], collects the passes' errors with the batch's).  `spill`
// holds kBatchSpillCap entries per list for min(n_cus, rays / 256) workgroups.
hipError_t batch_level_passes(rtx_ctx* c, const std::vector<BatchJob>& hj, const rtx_batch_job* jobs,
                              uint32_t n_jobs, uint32_t mode, uint64_t* spill, uint32_t* counters, hipStream_t s) {
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const auto& sc = c->scenes[jobs[j].scene];
    if (hj[j].n == 0 || sc.levels_consistent) continue;
    TraceParams p{};
    p.arena = sc.arena;
    p.arena_bytes = sc.bytes;
    p.tlas = sc.tlas;
    p.rays = hj[j].rays;
    p.hits = hj[j].hits;
    p.n = hj[j].n;
    p.n_rays = hj[j].n;
    p.refill = 40;
    p.ordered_boxes = sc.ordered ? 1u : 0u;
    p.owner = hj[j].owner;
    p.first_treelet = hj[j].first_treelet;
    p.spill = spill;
    p.spill_cap = kBatchSpillCap;
    p.ray_counter = counters;
    p.error = counters + 1;
    p.stats = c->stats;
    set_level_maps(p, c->lm_mem, (uint64_t)c->n_cus * kBlock);
    const uint32_t grid = std::max(1u, std::min<uint32_t>(c->n_cus, (hj[j].n + kBlock - 1) / kBlock));
    hipError_t e = launch_fill_zero(counters, 1, s);
    if (e == hipSuccess) e = launch_level_pass(p, mode, grid, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

rtx_status rtx_add_scene(rtx_ctx* c, const void* host, size_t bytes, uint64_t tlas, uint32_t* id) {
  if (!c || !host || !id) return RTX_ERR_INVALID_ARGUMENT;
  rtx_status st = check_scene_args(bytes, tlas);
  if (st != RTX_OK) return st;
  HIP_TRY(hipSetDevice(c->device));
  rtx_ctx::SceneSlot sc;
  if (hipMalloc(&sc.arena, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
  if (hipMemcpy(sc.arena, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(sc.arena);
    return RTX_ERR_HIP;
  }
  sc.bytes = bytes;
  sc.tlas = tlas;
  const SceneCheck ck = scene_check(sc.arena, bytes, tlas, c->stream);
  sc.ordered = ck.valid && ck.ordered;
  sc.levels_consistent = ck.valid && ck.level_consistent;
  c->scenes.push_back(sc);
  *id = (uint32_t)(c->scenes.size() - 1);
  return RTX_OK;
}

rtx_status rtx_build_scene_treelets(rtx_ctx* c, uint32_t id, uint32_t budget, rtx_treelet_summary* out) {
  if (!c || id >= c->scenes.size()) return RTX_ERR_INVALID_ARGUMENT;
  if (budget < 128 || budget % 64 != 0 || budget > (1u << 20)) return RTX_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(c->device));
  auto& sc = c->scenes[id];
  sc.part_gen++;  // graphs captured against the old partition are stale, also if the build fails
  rtx_status st = partition_build(sc.arena, sc.bytes, sc.tlas, budget, c->stream, sc.part);
  if (st != RTX_OK) return st;
  if (out) *out = sc.part.summary;
  return RTX_OK;
}

rtx_status rtx_trace_batch(rtx_ctx* c, const rtx_batch_job* jobs, uint32_t n_jobs, const rtx_trace_options* o) {
  if (!c || !o) return RTX_ERR_INVALID_ARGUMENT;
  std::vector<BatchJob> hj;
  uint64_t total = 0;
  rtx_status st = make_jobs(c, jobs, n_jobs, o->mode, hj, &total);
  if (st != RTX_OK) return st;
  BatchShape s;
  st = batch_shape(c, o, total, &s);
  if (st != RTX_OK) return st;
  HIP_TRY(hipSetDevice(c->device));
  if (total == 0) return RTX_OK;
  if (!c->jobs_dev) HIP_TRY(hipMalloc(&c->jobs_dev, kMaxBatchJobs * sizeof(BatchJob)));
  st = ensure_spill(c, (size_t)s.grid * kBlock * kBatchSpillCap * s.ns);
  if (st != RTX_OK) return st;
  HIP_TRY(hipMemcpyAsync(c->jobs_dev, hj.data(), n_jobs * sizeof(BatchJob), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->counters, 0, 2 * sizeof(uint32_t), c->stream));
  const bool levels = batch_needs_levels(c, jobs, n_jobs);
  if (levels) {
    st = ensure_level_mem(c);
    if (st != RTX_OK) return st;
  }
  TraceParams p = batch_params(c, c->jobs_dev, n_jobs, total, s, c->spill, c->counters);
  c->kernel_timed = false;
  HIP_TRY(hipEventRecord(c->evk0, c->stream));
  HIP_TRY(launch_trace_batch(p, o->mode, s.lds_stack, s.grid, c->stream));
  HIP_TRY(hipEventRecord(c->evk1, c->stream));
  c->kernel_timed = true;
  if (levels) HIP_TRY(batch_level_passes(c, hj, jobs, n_jobs, o->mode, c->spill, c->counters, c->stream));
  uint32_t err = 0;
  HIP_TRY(hipMemcpyAsync(&err, c->counters + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return kernel_error_status(err);
}

rtx_status rtx_batch_graph_create(rtx_ctx* c, const rtx_batch_job* jobs, uint32_t n_jobs,
                                  const rtx_trace_options* o, rtx_batch_graph** out) {
  if (!c || !o || !out) return RTX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::vector<BatchJob> hj;
  uint64_t total = 0;
  rtx_status st = make_jobs(c, jobs, n_jobs, o->mode, hj, &total);
  if (st != RTX_OK) return st;
  if (total == 0) return RTX_ERR_INVALID_ARGUMENT;
  BatchShape s;
  st = batch_shape(c, o, total, &s);
  if (st != RTX_OK) return st;
  HIP_TRY(hipSetDevice(c->device));
  rtx_batch_graph* g = new (std::nothrow) rtx_batch_graph();
  if (!g) return RTX_ERR_OUT_OF_MEMORY;
  g->ctx = c;
  c->graphs.push_back(g);
  if (o->mode == RTX_MODE_TREELET)
    for (uint32_t j = 0; j < n_jobs; ++j) g->deps.emplace_back(jobs[j].scene, c->scenes[jobs[j].scene].part_gen);
  auto fail = [&](rtx_status e) {
    rtx_batch_graph_destroy(g);
    return e;
  };
  if (hipMalloc(&g->jobs, n_jobs * sizeof(BatchJob)) != hipSuccess ||
      hipMalloc(&g->spill, (size_t)s.grid * kBlock * kBatchSpillCap * s.ns * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&g->counters, 64 * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(RTX_ERR_OUT_OF_MEMORY);
  }
  if (hipMemcpy(g->jobs, hj.data(), n_jobs * sizeof(BatchJob), hipMemcpyHostToDevice) != hipSuccess)
    return fail(RTX_ERR_HIP);
  const bool levels = batch_needs_levels(c, jobs, n_jobs);
  if (levels) {
    st = ensure_level_mem(c);  // zeroed before the capture, never inside it
    if (st != RTX_OK) return fail(st);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(RTX_ERR_HIP);
  TraceParams p = batch_params(c, g->jobs, n_jobs, total, s, g->spill, g->counters);
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail(RTX_ERR_HIP);
  // counters are cleared by a kernel: no memset node inside the graph
  hipError_t e1 = launch_fill_zero(g->counters, 2, c->stream);
  hipError_t e2 = e1 == hipSuccess ? launch_trace_batch(p, o->mode, s.lds_stack, s.grid, c->stream) : e1;
  if (e2 == hipSuccess && levels)
    e2 = batch_level_passes(c, hj, jobs, n_jobs, o->mode, g->spill, g->counters, c->stream);
  hipError_t e3 = hipStreamEndCapture(c->stream, &g->graph);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(RTX_ERR_HIP);
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) return fail(RTX_ERR_HIP);
  *out = g;
  return RTX_OK;
}

rtx_status rtx_batch_graph_launch(rtx_batch_graph* g, uint32_t replays, float* ms) {
  if (!g || !g->exec || replays == 0) return RTX_ERR_INVALID_ARGUMENT;
  rtx_ctx* c = g->ctx;
  if (!c) return RTX_ERR_STALE;
  for (const auto& d : g->deps)
    if (c->scenes[d.first].part_gen != d.second) return RTX_ERR_STALE;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (uint32_t r = 0; r < replays; ++r) HIP_TRY(hipGraphLaunch(g->exec, c->stream));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  uint32_t err = 0;
  HIP_TRY(hipMemcpyAsync(&err, g->counters + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (ms) HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return kernel_error_status(err);
}

rtx_status rtx_batch_graph_destroy(rtx_batch_graph* g) {
  if (!g) return RTX_ERR_INVALID_ARGUMENT;
  if (g->ctx) {
    rtx_ctx* c = g->ctx;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    c->graphs.erase(std::remove(c->graphs.begin(), c->graphs.end(), g), c->graphs.end());
  }
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->jobs) hipFree(g->jobs);
  if (g->spill) hipFree(g->spill);
  if (g->counters) hipFree(g->counters);
  delete g;
  return RTX_OK;
}

rtx_status rtx_autotune_batch(rtx_ctx* c, const rtx_batch_job* jobs, uint32_t n_jobs, rtx_trace_options* best,
                              float* best_ms) {
  if (!c || !best) return RTX_ERR_INVALID_ARGUMENT;
  const uint32_t waves[] = {0, 8, 12, 16, 20};  // 0: as many as fit on a CU
  const uint32_t stacks[] = {4, 8, 16};
  const uint32_t refills[] = {24, 32, 48};
  rtx_trace_options o = *best;
  o.prefetch = RTX_PREFETCH_NONE;
  o.emit_records = 0;
  float best_t = 1e30f;
  rtx_trace_options winner = o;
  for (uint32_t w : waves)
    for (uint32_t k : stacks)
      for (uint32_t r : refills) {
        o.waves_per_cu = w;
        o.lds_stack = k;
        o.refill = r;
        rtx_status st = rtx_trace_batch(c, jobs, n_jobs, &o);  // warm-up
        if (st != RTX_OK) return st;
        float t = 0.f;
        for (int rep = 0; rep < 2; ++rep) {
          st = rtx_trace_batch(c, jobs, n_jobs, &o);
          if (st != RTX_OK) return st;
          float ms = 0.f;
          st = rtx_last_kernel_ms(c, &ms);
          if (st != RTX_OK) return st;
          t += ms;
        }
        if (t < best_t) {
          best_t = t;
          winner = o;
        }
      }
  *best = winner;
  if (best_ms) *best_ms = best_t / 2;
  return RTX_OK;
}

rtx_status rtx_autotune_batch_treelets(rtx_ctx* c, const rtx_batch_job* jobs, uint32_t n_jobs, const uint32_t* budgets,
                                       uint32_t n_budgets, rtx_trace_options* best, uint32_t* best_budget,
                                       float* best_ms) {
  if (!c || !jobs || !budgets || n_budgets == 0 || !best || !best_budget || best->mode != RTX_MODE_TREELET)
    return RTX_ERR_INVALID_ARGUMENT;
  std::vector<uint32_t> ids;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    if (jobs[j].scene >= c->scenes.size()) return RTX_ERR_INVALID_ARGUMENT;
    ids.push_back(jobs[j].scene);
  }
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  for (uint32_t k = 0; k < n_budgets; ++k)
    if (budgets[k] < 256 || budgets[k] % 64 != 0 || budgets[k] > (1u << 20)) return RTX_ERR_INVALID_ARGUMENT;
  float win_ms = 1e30f;
  uint32_t win_budget = budgets[0];
  rtx_trace_options win = *best;
  for (uint32_t k = 0; k < n_budgets; ++k) {
    for (uint32_t id : ids) {
      rtx_status st = rtx_build_scene_treelets(c, id, budgets[k], nullptr);
      if (st != RTX_OK) return st;
    }
    rtx_trace_options o = *best;
    float ms = 0.f;
    rtx_status st = rtx_autotune_batch(c, jobs, n_jobs, &o, &ms);
    if (st != RTX_OK) return st;
    if (ms < win_ms) {
      win_ms = ms;
      win_budget = budgets[k];
      win = o;
    }
  }
  if (win_budget != budgets[n_budgets - 1]) {
    for (uint32_t id : ids) {
      rtx_status st = rtx_build_scene_treelets(c, id, win_budget, nullptr);
      if (st != RTX_OK) return st;
    }
  }
  *best = win;
  *best_budget = win_budget;
  if (best_ms) *best_ms = win_ms;
  return RTX_OK;
}

}  // extern "C"
