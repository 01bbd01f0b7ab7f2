/*
 * rtx_adapter.h — host-memory adapter with the reference's per-lane call.
 *
 * The reference calls, once per active lane of a trace_ray instruction
 * (src/cuda-sim/instructions.cc:7235-7254),
 *   VulkanRayTracing::traceRay(_topLevelAS, rayFlags, cullMask, sbtRecordOffset,
 *       sbtRecordStride, missIndex, origin, Tmin, direction, Tmax, payload, pI, thread)
 * or traceRayWithTreelets with the same operands
 * (src/cuda-sim/vulkan_ray_tracing.h:329-356), and returns the closest hit in
 * Traversal_data (src/cuda-sim/vulkan_rt_thread_data.h:43-58) and the node
 * transactions through thread->set_rt_transactions (vulkan_ray_tracing.cc:3040).
 *
 * rtx_adapter_trace_ray takes those operands (payload, pI and thread stay on
 * the simulator side) and returns the same two results in host memory: an
 * rtx_hit and the ray's rtx_records.  rtx_adapter_trace_lanes does the same
 * for every lane of a warp in one call, which is where a simulator gains:
 * one batched traversal instead of one per lane.  The adapter owns its
 * context, its device buffers, and the scene's treelet partition (built once,
 * replacing the createTreelets call of vulkan_ray_tracing.cc:1593-1596).
 */
#ifndef RTX_ADAPTER_H
#define RTX_ADAPTER_H

#include "rtx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtx_adapter rtx_adapter;

/* A context on `device` holding a copy of the arena; treelet_budget > 0 also
 * builds the partition (maxBytesPerTreelet, gpu-sim.cc:910-912), which the
 * treelet order needs and which labels every record with its treelet. */
rtx_status rtx_adapter_create(int device, const void* arena_host, size_t bytes, uint64_t tlas_offset,
                              uint32_t treelet_budget, rtx_adapter** out);
rtx_status rtx_adapter_destroy(rtx_adapter* a);

/* One lane: VulkanRayTracing::traceRay (treelet_order 0) or
 * traceRayWithTreelets (treelet_order 1).  Writes *hit_out and, when
 * records_out is not NULL, the ray's records (at most records_capacity; a
 * longer trace returns RTX_ERR_TRACE_OVERFLOW with *n_records set and
 * *hit_out written).  n_records may be NULL when records_out is. */
rtx_status rtx_adapter_trace_ray(rtx_adapter* a, int treelet_order, uint32_t rayFlags, uint32_t cullMask,
                                 uint32_t sbtRecordOffset, uint32_t sbtRecordStride, uint32_t missIndex,
                                 const float origin[3], float Tmin, const float direction[3], float Tmax,
                                 rtx_hit* hit_out, rtx_record* records_out, uint64_t records_capacity,
                                 uint64_t* n_records);

/* n lanes at once (a warp's trace_ray operands, already in rtx_ray form):
 * hits_out[n], and when records_out is not NULL the records of lane i at
 * [offsets_out[i], offsets_out[i + 1]) (offsets_out[n + 1] required then).
 * Results are those of n calls of rtx_adapter_trace_ray. */
rtx_status rtx_adapter_trace_lanes(rtx_adapter* a, int treelet_order, const rtx_ray* rays_host, uint32_t n,
                                   rtx_hit* hits_out, rtx_record* records_out, uint64_t records_capacity,
                                   uint64_t* offsets_out);

#ifdef __cplusplus
}
#endif
#endif /* RTX_ADAPTER_H */
