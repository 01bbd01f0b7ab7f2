// adapter.cpp — the per-lane traceRay call of the reference on top of rtx_trace.
//
// A simulator that keeps the reference's call site (one call per lane,
// src/cuda-sim/instructions.cc:7235-7254) links this file's entry points in
// place of VulkanRayTracing::traceRay / traceRayWithTreelets.  Each call copies
// the lanes' rays to the device, traces them with records, and copies the
// hits and records back; the device buffers grow as needed and are reused.
#include <algorithm>
#include <cstring>
#include <new>

#include "rtx.h"
#include "rtx_adapter.h"

struct rtx_adapter {
  rtx_ctx* ctx = nullptr;
  bool has_partition = false;
  uint32_t cap_rays = 0;     // rays the ray / hit / offset buffers hold
  uint64_t cap_records = 0;  // records the record buffer holds
  rtx_ray* rays = nullptr;
  rtx_hit* hits = nullptr;
  uint64_t* offsets = nullptr;
  rtx_record* records = nullptr;
};

namespace {

void release(rtx_adapter* a) {
  if (!a->ctx) return;
  rtx_free(a->ctx, a->rays);
  rtx_free(a->ctx, a->hits);
  rtx_free(a->ctx, a->offsets);
  rtx_free(a->ctx, a->records);
  a->rays = nullptr;
  a->hits = nullptr;
  a->offsets = nullptr;
  a->records = nullptr;
  a->cap_rays = 0;
  a->cap_records = 0;
}

rtx_status reserve_rays(rtx_adapter* a, uint32_t n) {
  if (n <= a->cap_rays) return RTX_OK;
  const uint32_t cap = std::max<uint32_t>(n, 64);
  rtx_free(a->ctx, a->rays);
  rtx_free(a->ctx, a->hits);
  rtx_free(a->ctx, a->offsets);
  a->rays = nullptr;
  a->hits = nullptr;
  a->offsets = nullptr;
  a->cap_rays = 0;
  rtx_status st = rtx_malloc(a->ctx, (size_t)cap * sizeof(rtx_ray), reinterpret_cast<void**>(&a->rays));
  if (st == RTX_OK) st = rtx_malloc(a->ctx, (size_t)cap * sizeof(rtx_hit), reinterpret_cast<void**>(&a->hits));
  if (st == RTX_OK)
    st = rtx_malloc(a->ctx, ((size_t)cap + 1) * sizeof(uint64_t), reinterpret_cast<void**>(&a->offsets));
  if (st == RTX_OK) a->cap_rays = cap;
  return st;
}

rtx_status reserve_records(rtx_adapter* a, uint64_t n) {
  if (n <= a->cap_records) return RTX_OK;
  const uint64_t cap = std::max<uint64_t>(n, 4096);
  rtx_free(a->ctx, a->records);
  a->records = nullptr;
  a->cap_records = 0;
  rtx_status st = rtx_malloc(a->ctx, (size_t)cap * sizeof(rtx_record), reinterpret_cast<void**>(&a->records));
  if (st == RTX_OK) a->cap_records = cap;
  return st;
}

// Trace n rays already in a->rays; hits and offsets stay on the device.
rtx_status trace_staged(rtx_adapter* a, int treelet_order, uint32_t n, bool want_records) {
  rtx_trace_options o;
  std::memset(&o, 0, sizeof o);
  o.mode = treelet_order ? RTX_MODE_TREELET : RTX_MODE_PLAIN;
  o.emit_records = want_records ? 1u : 0u;
  if (!want_records) return rtx_trace(a->ctx, a->rays, n, &o, a->hits, nullptr, 0, nullptr);
  rtx_status st = reserve_records(a, (uint64_t)n * 64);
  if (st != RTX_OK) return st;
  st = rtx_trace(a->ctx, a->rays, n, &o, a->hits, a->records, a->cap_records, a->offsets);
  if (st == RTX_ERR_TRACE_OVERFLOW) {
    // the offsets hold the total: grow the buffer and trace again
    uint64_t total = 0;
    st = rtx_memcpy_d2h(a->ctx, &total, a->offsets + n, sizeof total);
    if (st != RTX_OK) return st;
    st = reserve_records(a, total);
    if (st != RTX_OK) return st;
    st = rtx_trace(a->ctx, a->rays, n, &o, a->hits, a->records, a->cap_records, a->offsets);
  }
  return st;
}

}  // namespace

extern "C" {

rtx_status rtx_adapter_create(int device, const void* arena, size_t bytes, uint64_t tlas_offset,
                              uint32_t treelet_budget, rtx_adapter** out) {
  if (!out || !arena) return RTX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  rtx_adapter* a = new (std::nothrow) rtx_adapter();
  if (!a) return RTX_ERR_OUT_OF_MEMORY;
  rtx_status st = rtx_create(device, &a->ctx);
  if (st == RTX_OK) st = rtx_set_scene(a->ctx, arena, bytes, tlas_offset);
  if (st == RTX_OK && treelet_budget) {
    st = rtx_build_treelets(a->ctx, treelet_budget, nullptr);
    a->has_partition = st == RTX_OK;
  }
  if (st != RTX_OK) {
    rtx_adapter_destroy(a);
    return st;
  }
  *out = a;
  return RTX_OK;
}

rtx_status rtx_adapter_destroy(rtx_adapter* a) {
  if (!a) return RTX_ERR_INVALID_ARGUMENT;
  release(a);
  if (a->ctx) rtx_destroy(a->ctx);
  delete a;
  return RTX_OK;
}

rtx_status rtx_adapter_trace_ray(rtx_adapter* a, int treelet_order, uint32_t rayFlags, uint32_t cullMask,
                                 uint32_t sbtRecordOffset, uint32_t sbtRecordStride, uint32_t missIndex,
                                 const float origin[3], float Tmin, const float direction[3], float Tmax,
                                 rtx_hit* hit_out, rtx_record* records_out, uint64_t records_capacity,
                                 uint64_t* n_records) {
  if (!a || !origin || !direction || !hit_out || (records_out && !n_records)) return RTX_ERR_INVALID_ARGUMENT;
  // trace_ray_impl's operands as one rtx_ray (instructions.cc:7158-7227)
  rtx_ray r;
  std::memset(&r, 0, sizeof r);
  for (int k = 0; k < 3; ++k) {
    r.org[k] = origin[k];
    r.dir[k] = direction[k];
  }
  r.tmin = Tmin;
  r.tmax = Tmax;
  r.flags = rayFlags;
  r.cull_mask = cullMask;
  r.sbt_record_offset = sbtRecordOffset;
  r.sbt_record_stride = sbtRecordStride;
  r.miss_index = missIndex;
  uint64_t offs[2] = {0, 0};
  rtx_status st = rtx_adapter_trace_lanes(a, treelet_order, &r, 1, hit_out, records_out, records_capacity,
                                          records_out ? offs : nullptr);
  if (records_out && (st == RTX_OK || st == RTX_ERR_TRACE_OVERFLOW)) *n_records = offs[1];
  return st;
}

rtx_status rtx_adapter_trace_lanes(rtx_adapter* a, int treelet_order, const rtx_ray* rays, uint32_t n,
                                   rtx_hit* hits_out, rtx_record* records_out, uint64_t records_capacity,
                                   uint64_t* offsets_out) {
  if (!a || (n > 0 && (!rays || !hits_out)) || (records_out && !offsets_out)) return RTX_ERR_INVALID_ARGUMENT;
  if (treelet_order && !a->has_partition) return RTX_ERR_NO_PARTITION;
  if (n == 0) {
    if (offsets_out) offsets_out[0] = 0;
    return RTX_OK;
  }
  rtx_status st = reserve_rays(a, n);
  if (st != RTX_OK) return st;
  st = rtx_memcpy_h2d(a->ctx, a->rays, rays, (size_t)n * sizeof(rtx_ray));
  if (st != RTX_OK) return st;
  const bool want = records_out != nullptr;
  st = trace_staged(a, treelet_order, n, want);
  if (st != RTX_OK) return st;
  st = rtx_memcpy_d2h(a->ctx, hits_out, a->hits, (size_t)n * sizeof(rtx_hit));
  if (st != RTX_OK || !want) return st;
  st = rtx_memcpy_d2h(a->ctx, offsets_out, a->offsets, ((size_t)n + 1) * sizeof(uint64_t));
  if (st != RTX_OK) return st;
  const uint64_t total = offsets_out[n];
  if (total > records_capacity) return RTX_ERR_TRACE_OVERFLOW;
  return rtx_memcpy_d2h(a->ctx, records_out, a->records, (size_t)total * sizeof(rtx_record));
}

}  // extern "C"
