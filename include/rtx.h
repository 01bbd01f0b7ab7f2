/*
 * rtx.h — C ABI of the MI355X ray-traversal library (librtx.so).
 *
 * The library replaces the functional BVH traversal that Vulkan-Sim runs per
 * lane inside its simulated RT unit:
 *   VulkanRayTracing::traceRay              src/cuda-sim/vulkan_ray_tracing.cc:2309-3076
 *   VulkanRayTracing::traceRayWithTreelets  src/cuda-sim/vulkan_ray_tracing.cc:1522-2307
 *   VulkanRayTracing::createTreelets        src/cuda-sim/vulkan_ray_tracing.cc:823-1470
 *   VulkanRayTracing::buildNodeToRootMap    src/cuda-sim/vulkan_ray_tracing.cc:475-489
 * declared in src/cuda-sim/vulkan_ray_tracing.h:329-356.
 *
 * The reference is called once per ray with host pointers into a Mesa-owned
 * acceleration structure and returns its result through simulator globals.
 * This ABI is batched: the acceleration structure is one byte arena (TLAS and
 * every BLAS, node layout of src/cuda-sim/vulkan_acceleration_structure_util.h,
 * relative offsets unchanged) copied to or borrowed in device memory, and every
 * node is named by its byte offset in that arena.  That offset replaces the
 * simulated device address the reference stores in MemoryTransactionRecord
 * (src/abstract_hardware_model.h:315-321).
 *
 * All entry points return rtx_status and never abort (the reference asserts).
 * Pointers named *_dev are device pointers (from rtx_malloc or any HIP
 * allocation on the context's device); pointers named *_host are host memory.
 * Work is ordered on the context's stream; rtx_synchronize waits for it.
 * A context is not thread-safe; distinct contexts are independent.
 */
#ifndef RTX_H
#define RTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rtx_status {
  RTX_OK = 0,
  RTX_ERR_INVALID_ARGUMENT = 1,
  RTX_ERR_BAD_NODE_TYPE = 2,           /* ChildType not internal/instance/quad/procedural */
  RTX_ERR_OFFSET_OUT_OF_ARENA = 3,     /* a node or child offset points outside the arena */
  RTX_ERR_NODE_LARGER_THAN_BUDGET = 4, /* reference asserts at vulkan_ray_tracing.cc:893,962,976 */
  RTX_ERR_TRACE_OVERFLOW = 5,          /* record buffer smaller than the records produced */
  RTX_ERR_STACK_OVERFLOW = 6,          /* per-ray traversal stack exceeded its capacity */
  RTX_ERR_HIP = 7,
  RTX_ERR_NO_DEVICE = 8,
  RTX_ERR_NO_SCENE = 9,
  RTX_ERR_NO_PARTITION = 10,
  RTX_ERR_UNSUPPORTED = 11,
  RTX_ERR_OUT_OF_MEMORY = 12,
  RTX_ERR_TIMEOUT = 13,                /* a ray passed the per-ray limit of 2^24 steps: its traversal
                                          would not end (a cyclic BVH, where the reference loops) */
  RTX_ERR_STALE = 14,                  /* a batch graph whose scenes' partitions were rebuilt, or whose
                                          context was destroyed, after it was captured */
  RTX_ERR_CYCLE = 15                   /* rtx_validate_scene: a node the TLAS reaches reaches itself */
} rtx_status;

/* Traversal order. PLAIN follows traceRay, TREELET follows traceRayWithTreelets.
 * TREELET keeps each stack entry's instance leaf as a signed 22-bit slot distance
 * from the TLAS header: a ray that reaches an instance leaf 128 MiB or more away
 * from the header, on either side, reports RTX_ERR_UNSUPPORTED.  PLAIN has no
 * such limit. */
enum { RTX_MODE_PLAIN = 0, RTX_MODE_TREELET = 1 };

/* Treelet staging (prefetch into LDS).  Never changes hits or traces. */
enum {
  RTX_PREFETCH_NONE = 0,       /* every node read goes to global memory */
  RTX_PREFETCH_ALWAYS = 1,     /* heuristic 0: stage the wave's most popular treelet */
  RTX_PREFETCH_POPULARITY = 2, /* heuristic 1: stage it only if its share >= threshold */
  RTX_PREFETCH_PARTIAL = 3,    /* heuristic 2: stage the first round(n*share) nodes */
  RTX_PREFETCH_LEARNED = 4,    /* MLP predicts the next treelet; heuristic 0 is the fallback */
  RTX_PREFETCH_LATTER = 5      /* heuristic 3: stage the last round(n*share) units */
};
/* Heuristics 0-3 follow src/gpgpu-sim/shader.cc:3437-3590: the wave counts the
 * treelet of every active lane's next node and takes the most popular one,
 * ties to the lowest root offset (:3424-3446); share = its count / the lanes
 * counted.  A treelet already resident in the wave's staging area is not
 * staged again (:3564). */

/* SPIR-V ray flags used by the traversal (SpvRayFlags*KHRMask). */
enum {
  RTX_RAY_FLAG_OPAQUE = 0x1,
  RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT = 0x4,
  RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER = 0x8
};

/* TransactionType, src/abstract_hardware_model.h:201-211. */
enum {
  RTX_TX_BVH_STRUCTURE = 0,
  RTX_TX_BVH_INTERNAL_NODE = 1,
  RTX_TX_BVH_INSTANCE_LEAF = 2,
  RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR = 3,
  RTX_TX_BVH_QUAD_LEAF = 4,
  RTX_TX_BVH_QUAD_LEAF_HIT = 5,
  RTX_TX_BVH_PROCEDURAL_LEAF = 6
};

#define RTX_NO_TREELET 0xFFFFFFFFu
#define RTX_NO_NODE 0xFFFFFFFFFFFFFFFFull

/* One ray: the traceRay operands decoded by trace_ray_impl
 * (src/cuda-sim/instructions.cc:7158-7227).  cull_mask, sbt_* and miss_index
 * are carried through untouched, as in the reference (vulkan_ray_tracing.cc:2389-2395). */
typedef struct rtx_ray {
  float org[3];
  float tmin;
  float dir[3];
  float tmax;
  uint32_t flags;
  uint32_t cull_mask;
  uint32_t sbt_record_offset;
  uint32_t sbt_record_stride;
  uint32_t miss_index;
  uint32_t reserved[3];
} rtx_ray; /* 64 bytes */

/* Closest-hit record: the fields of Traversal_data / Hit_data
 * (src/cuda-sim/vulkan_rt_thread_data.h:29-58) plus per-ray counters that the
 * reference accumulates in its global statistics (gpu-sim.cc:2168-2182). */
typedef struct rtx_hit {
  uint32_t hit;             /* hit_geometry */
  uint32_t primitive_index; /* closest_leaf.PrimitiveIndex0 */
  uint32_t geometry_index;  /* closest_leaf.LeafDescriptor.GeometryIndex */
  uint32_t instance_index;  /* closest_instanceLeaf.InstanceID */
  float t;                  /* world_min_thit */
  float p[3];               /* intersection_point = origin + direction * t */
  float bary[3];            /* Barycentric(): (v, w, u) */
  uint32_t n_records;       /* memory transactions of this ray */
  uint32_t n_nodes;         /* total_nodes_accessed */
  uint32_t fetch_bytes;     /* sum of record sizes, leaf-descriptor records excluded */
  uint32_t n_any_hits;      /* accepted hits of a non-opaque ray (any-hit shader calls) */
  uint32_t max_level;       /* the ray's level: the maximum over its final tree_level_map
                               (vulkan_ray_tracing.cc:2274-2279, 3062-3067), where a node
                               assigned again keeps its last value; equal to the deepest level
                               assigned unless a node is reachable at two levels (a BLAS under
                               instance leaves of different depths), for which rtx_trace runs
                               a second traversal that keeps the map */
  uint64_t instance_leaf;   /* arena offset of closest_instanceLeaf (vulkan_ray_tracing.cc:2856), the
                               instance of the last accepted hit; Hit_data's worldToObjectMatrix and
                               objectToWorldMatrix (:3002-3003) are its rows at +16 and +80.
                               RTX_NO_NODE when the ray missed */
  uint32_t accessed_bytes;  /* sum of all record sizes, leaf descriptors included (accessedDataSize, :2260) */
  uint32_t reserved;
} rtx_hit; /* 80 bytes */

/* MemoryTransactionRecord (src/abstract_hardware_model.h:315-321) with the
 * node's treelet attached (addrToTreeletID, vulkan_ray_tracing.cc:468-472). */
typedef struct rtx_record {
  uint64_t offset;  /* arena byte offset of the node */
  uint32_t treelet; /* dense treelet index of the node's owner, or RTX_NO_TREELET */
  uint16_t size;    /* bytes: 64, 128 or 8 */
  uint8_t type;     /* RTX_TX_* */
  uint8_t level;    /* depth of the node along the path that reached it (the tree_level_map value
                       assigned with it, vulkan_ray_tracing.cc:2470,2569,2668): TLAS root 1, a child
                       its parent's + 1, a BLAS root and its header record their instance leaf's;
                       0 for the TLAS header; saturating at 255 */
} rtx_record; /* 16 bytes */

typedef struct rtx_trace_options {
  uint32_t mode;         /* RTX_MODE_* */
  uint32_t prefetch;     /* RTX_PREFETCH_* */
  float threshold;       /* POPULARITY share threshold (reference default 0.5, tested 0.75) */
  uint32_t emit_records; /* 1: fill the record buffer (see record_slab) */
  uint32_t waves_per_cu; /* persistent waves per CU; 0 = as many as fit on a CU at once */
  uint32_t lds_stack;    /* stack entries per lane (per list in TREELET order) held in LDS (4, 8, 16); 0 = 8 */
  uint32_t stage_bytes;  /* LDS bytes per wave for a staged treelet (4-64 KiB, power of two);
                            0 = the treelet budget, capped at 16 KiB */
  uint32_t refill;       /* a wave fetches new rays once this many lanes are idle (1-64); 0 = 40
                            (48 for batches) */
  uint32_t slot_bytes;   /* staging: LDS slot size; stage_bytes / slot_bytes treelets stay resident
                            at once, least recently used evicted first.  0 = one slot */
  float min_score;       /* LEARNED: summed lane probability a predicted treelet needs before it is
                            staged (0 = default 16.0) */
  float pred_fraction;   /* LEARNED: share of a predicted treelet's image staged, root first
                            (0 = default 0.25; 1 = the whole image) */
  uint32_t lanes_per_ray; /* 1: one lane per ray; 8: eight lanes per ray, testing a node's six
                             children at once (either order, no records or staging), for batches
                             too small to fill the GPU; 0 = 8 when those hold and the batch has
                             at most 128 rays per CU (PLAIN) or 256 (TREELET), else 1.  Results
                             are identical.
                             rtx_trace only; batches ignore it. */
  uint32_t claim_start;  /* persistent waves claim rays starting at this index and wrap around to 0
                            (outputs stay at input indices).  A batch that holds camera rays followed
                            by secondary rays traces faster when the longer, incoherent secondary
                            rays are claimed first: pass the index of the first secondary ray.
                            Must be < n (0 for an empty batch).  rtx_trace only. */
  uint32_t record_layout; /* 1: each record carries its node's offset in the treelet layout of
                             rtx_build_treelet_layout instead of its arena offset, as
                             -remap_to_treelet_layout makes the reference record remapped addresses
                             (vulkan_ray_tracing.cc:2446-2450, 2509-2512); RTX_NO_NODE for a node
                             placed nowhere.  Needs the layout (else RTX_ERR_NO_PARTITION) */
  uint32_t record_slab;  /* records, one traversal: each ray's first record_slab records (at most
                            1024; 0 = 256) go to a per-ray slab in library memory and are compacted
                            into the caller's CSR buffer; only rays with more records are traced a
                            second time.  The slabs (8 bytes per record) take at most 16 GiB: a
                            batch that needs more is traced in consecutive ray ranges that fit; with
                            a claim_order or staging it takes a counting pass and a writing pass
                            instead */
  const uint32_t* claim_order; /* device array of n ray indices or NULL: claim k traces ray
                                  claim_order[k] (outputs stay at input indices), e.g. the order of
                                  rtx_coherence_order; claim_start must then be 0.  An index >= n is
                                  reported as RTX_ERR_OFFSET_OUT_OF_ARENA.  rtx_trace only */
} rtx_trace_options;

/* Prefetch accounting of the last rtx_trace call.  The effectiveness classes
 * of the reference's prefetch tracker (prefetch_request_effectiveness,
 * src/gpgpu-sim/gpu-cache.h:99-131; classified at gpu-cache.cc:1960-2117,
 * printed at gpu-sim.cc:1750-1868), per 64-byte unit staged into a wave's LDS:
 *   TIMELY      read from LDS at least once before its eviction (a demand hit
 *               on a prefetched line);
 *   NEVER_USED  evicted unread (wasted_bytes; includes TOO_EARLY);
 *   TOO_EARLY   evicted unread, and its node then read from global memory by
 *               the same wave (a demand miss after the prefetch's eviction);
 *   TOO_LATE    staged after the same wave had read its node from global
 *               memory (a prefetch landing on data a demand load fetched);
 *   LATE        a demand read while the prefetch is in flight: always 0 here,
 *               because a wave finishes staging before its next node read.
 * TOO_EARLY and TOO_LATE remember node slots in 128-entry direct-mapped
 * tables per wave, so they are lower bounds; a node counts 64 bytes. */
typedef struct rtx_prefetch_stats {
  uint64_t node_reads;        /* node reads issued by the traversal */
  uint64_t lds_hits;          /* node reads served from a staged treelet */
  uint64_t staged_treelets;   /* treelet loads into LDS */
  uint64_t staged_bytes;      /* bytes copied from HBM into LDS */
  uint64_t wasted_bytes;      /* NEVER_USED: staged bytes never read before eviction */
  uint64_t predictions;       /* LEARNED: treelets staged because the table predicted them */
  uint64_t prediction_hits;   /* LEARNED: predicted stagings that served a node read before eviction */
  uint64_t fallbacks;         /* LEARNED: staging decisions where no lane had a table entry */
  uint64_t timely_bytes;      /* TIMELY: staged bytes read at least once; timely + wasted = staged */
  uint64_t too_early_bytes;   /* TOO_EARLY (part of wasted_bytes) */
  uint64_t too_late_bytes;    /* TOO_LATE */
  uint64_t late_bytes;        /* LATE: 0 */
} rtx_prefetch_stats;

typedef struct rtx_treelet_summary {
  uint32_t n_treelets;      /* "Treelet Count" (vulkan_ray_tracing.cc:1366) */
  uint32_t budget_bytes;    /* maxBytesPerTreelet */
  uint64_t n_list_entries;  /* total node-list entries after de-duplication */
  uint64_t total_bytes;     /* bytes placed in treelets, each distinct root counted once */
  uint32_t max_treelet_bytes;
  uint32_t max_treelet_nodes;
} rtx_treelet_summary;

typedef struct rtx_ctx rtx_ctx;

/* ---- context and memory ------------------------------------------------ */
const char* rtx_status_string(rtx_status s);
const char* rtx_version(void);
rtx_status rtx_device_count(int* count);
rtx_status rtx_create(int device, rtx_ctx** out);
rtx_status rtx_destroy(rtx_ctx* ctx);
/* Waits for the context's stream.  Returns the first error of the traces
 * queued with rtx_trace_async since the previous wait (and clears it). */
rtx_status rtx_synchronize(rtx_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work with it
 * (hipStreamWaitEvent, hipEventRecord). */
rtx_status rtx_get_stream(rtx_ctx* ctx, void** hip_stream);
rtx_status rtx_malloc(rtx_ctx* ctx, size_t bytes, void** out_dev);
rtx_status rtx_free(rtx_ctx* ctx, void* dev);
rtx_status rtx_memcpy_h2d(rtx_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
rtx_status rtx_memcpy_d2h(rtx_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
rtx_status rtx_memcpy_d2d(rtx_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);
rtx_status rtx_memset(rtx_ctx* ctx, void* dev, int value, size_t bytes);
/* Stream-ordered timer around work issued on the context's stream. */
rtx_status rtx_timer_start(rtx_ctx* ctx);
rtx_status rtx_timer_stop(rtx_ctx* ctx, float* ms);
rtx_status rtx_device_info(rtx_ctx* ctx, int* n_cus, int* lds_bytes_per_cu, char* name, size_t name_len);

/* ---- scene ---------------------------------------------------------------
 * Replaces the TLAS pointer handed to traceRay (instructions.cc:7161) and the
 * BLAS/TLAS host-to-device maps of allocTLAS/allocBLAS (vulkan_ray_tracing.cc:4891-4899).
 * tlas_offset is the arena offset of the TLAS GEN_RT_BVH header. */
rtx_status rtx_set_scene(rtx_ctx* ctx, const void* arena_host, size_t bytes, uint64_t tlas_offset);
rtx_status rtx_set_scene_device(rtx_ctx* ctx, const void* arena_dev, size_t bytes, uint64_t tlas_offset);
/* Walk the whole TLAS->BLAS DAG on the device and check node types and offsets. */
rtx_status rtx_validate_scene(rtx_ctx* ctx, uint64_t* n_nodes_reached);

/* ---- treelets ------------------------------------------------------------
 * Device replacement of createTreelets + buildNodeToRootMap with budget
 * maxBytesPerTreelet (gpu-sim.cc:910-912; tested 512 B, configs 8-64 KiB). */
rtx_status rtx_build_treelets(rtx_ctx* ctx, uint32_t budget_bytes, rtx_treelet_summary* out);
/* Copy the partition to host memory.  Any pointer may be NULL.
 *  roots[n_treelets]          treelet root offsets, ascending (treelet_addr_to_metadata_idx order)
 *  list_offsets[n_treelets+1] CSR offsets into list_nodes
 *  list_nodes[n_list_entries] de-duplicated node offsets of each treelet, in insertion order
 *  owner[arena_bytes/64]      dense treelet index owning the node at offset 64*i, or RTX_NO_TREELET */
rtx_status rtx_get_treelets(rtx_ctx* ctx, uint64_t* roots, uint64_t* list_offsets,
                            uint64_t* list_nodes, uint32_t* owner);
/* remapBVHToTreeletLayout (vulkan_ray_tracing.cc:1473-1509) of the current
 * partition, with -treelet_remap_stride `stride` (gpu-sim.cc:461-464): treelet
 * t occupies [t * (budget + stride), ...) of a treelet-contiguous address space,
 * its root first, then the other nodes of its list in order (instance leaf 128
 * bytes, every other node 64); a node several lists hold keeps the place of the
 * first (lowest root offset).  Also the per-treelet metadata table of
 * load_treelet_metadata (:1598-1605): budget/64 entries of 4 bytes per treelet,
 * entry j the signed byte distance from the treelet root's place to the place
 * of its list entry j, -1 past the list.  RTX_ERR_UNSUPPORTED where the
 * reference asserts (:1483): a treelet root that an earlier treelet's list
 * already placed.  Rebuilding the partition drops the layout. */
typedef struct rtx_treelet_layout_summary {
  uint32_t stride;           /* treelet_remap_stride */
  uint32_t block_bytes;      /* budget + stride: the distance between treelet roots */
  uint64_t layout_bytes;     /* n_treelets * block_bytes (the treelet_layout_bvh allocation, :1478) */
  uint64_t mapped_nodes;     /* nodes given a place */
  uint32_t metadata_entries; /* per treelet: budget / 64 */
  uint32_t reserved;
} rtx_treelet_layout_summary;
rtx_status rtx_build_treelet_layout(rtx_ctx* ctx, uint32_t stride, rtx_treelet_layout_summary* out);
/* Copy the layout to host memory (either pointer may be NULL):
 *  remap[arena_bytes/64]                layout offset of the node at arena offset 64*i, or RTX_NO_NODE
 *  metadata[n_treelets * budget/64]     the metadata table */
rtx_status rtx_get_treelet_layout(rtx_ctx* ctx, uint64_t* remap, int32_t* metadata);

/* Install a partition computed elsewhere (same arrays as rtx_get_treelets). */
rtx_status rtx_set_treelets(rtx_ctx* ctx, uint32_t budget_bytes, uint32_t n_treelets,
                            const uint64_t* roots, const uint64_t* list_offsets,
                            const uint64_t* list_nodes, const uint32_t* owner);

/* ---- traversal -----------------------------------------------------------
 * Trace n rays.  hits_dev[n] is always written.  When opts->emit_records is 1,
 * record_offsets_dev[n+1] receives the CSR offsets and records_dev the records
 * (record_capacity entries); a larger total returns RTX_ERR_TRACE_OVERFLOW with
 * record_offsets_dev still filled, so the caller can size the buffer. */
rtx_status rtx_trace(rtx_ctx* ctx, const rtx_ray* rays_dev, uint32_t n, const rtx_trace_options* opts,
                     rtx_hit* hits_dev, rtx_record* records_dev, uint64_t record_capacity,
                     uint64_t* record_offsets_dev);
rtx_status rtx_get_prefetch_stats(rtx_ctx* ctx, rtx_prefetch_stats* out);

/* rtx_trace without the wait, stream-ordered: the traversal, and the level
 * pass a scene may need, are queued on the context's stream and the call
 * returns.  Hits only, no staging (records or a prefetch policy:
 * RTX_ERR_UNSUPPORTED).  Argument errors are returned at once, errors of the
 * rays themselves by the next rtx_synchronize.  The rays and hits must stay
 * allocated until then. */
rtx_status rtx_trace_async(rtx_ctx* ctx, const rtx_ray* rays_dev, uint32_t n, const rtx_trace_options* opts,
                           rtx_hit* hits_dev);

/* Run-level statistics of a traced batch, the functional RT counters the
 * reference prints at the end of a kernel (src/gpgpu-sim/gpu-sim.cc:2168-2182;
 * kept in src/cuda-sim/cuda-sim.h:155-165), reduced on the device from the
 * hits and, when given, the records.  Without records the primitive-leaf
 * records are all counted as RTX_TX_BVH_QUAD_LEAF (the hit counters do not
 * tell quad leaves, accepted quad leaves and procedural leaves apart). */
typedef struct rtx_trace_stats {
  uint64_t access_type[7];     /* "Ray tracing memory access distribution": records per RTX_TX_* */
  uint64_t num_hits;           /* rt_num_hits: rays with hit_geometry */
  uint64_t num_any_hits;       /* rt_num_any_hits: accepted hits of non-opaque rays */
  uint64_t n_rays;             /* rt_n_total_rays */
  uint64_t tot_nodes;          /* rt_tot_nodes_per_ray (sum of total_nodes_accessed) */
  uint32_t max_nodes_per_ray;  /* rt_max_nodes_per_ray */
  uint32_t max_tree_depth;     /* rt_max_tree_depth */
  uint64_t fetch_bytes;        /* node bytes fetched, leaf descriptors excluded */
  uint64_t accessed_bytes;     /* total accessed data size: every record (gpu-sim.cc:2008) */
  uint32_t from_records;       /* 1 when the histogram comes from records */
  uint32_t reserved;
} rtx_trace_stats;
rtx_status rtx_trace_stats_of(rtx_ctx* ctx, const rtx_hit* hits_dev, uint32_t n, const rtx_record* records_dev,
                              uint64_t n_records, rtx_trace_stats* out);

/* rtx_trace that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory while they trace (tight serving
 * loops, harnesses that forbid hipMalloc).  Hits only, no staging: the options
 * are those of rtx_trace with prefetch RTX_PREFETCH_NONE and emit_records 0
 * (else RTX_ERR_UNSUPPORTED).  The kernels work in a static device array of
 * 128 MiB per GPU, held until the process ends: 256 bytes of control words and
 * the per-lane global stack spill.  The launch keeps at least 24 spill entries
 * per lane (per list in treelet order), shrinking its grid if needed, and at
 * most 128; a ray whose stack outgrows LDS plus that spill returns
 * RTX_ERR_STACK_OVERFLOW.  Calls on one GPU share the array, so calls that
 * overlap on one GPU (other threads, other contexts on the device) race;
 * each call clears the control words it uses and waits for its kernel. */
rtx_status rtx_trace_static(rtx_ctx* ctx, const rtx_ray* rays_dev, uint32_t n, const rtx_trace_options* opts,
                            rtx_hit* hits_dev);

/* ---- ray coherence engine ------------------------------------------------------
 * The RT unit's ray sorting (src/gpgpu-sim/ray_coherency_engine.cc; enabled by
 * -gpgpu_rt_coherence_engine, gpu-sim.cc:354-360) as a claim order for
 * rtx_trace.  Each ray is hashed (:393-474), rays of equal hash form a pool in
 * input order, and packets of `packet` rays are taken from the largest pool
 * first (ties: lowest hash; get_largest_packet :182-193).  order_dev[n]
 * receives the ray indices in packet order; hashes_dev[n] (may be NULL) each
 * ray's hash.  Hashes must fit 32 bits: francois_bits <= 4, grid_bits <= 10,
 * sphere_bits <= 8. */
enum {
  RTX_COHERENCE_WORLD_REFERENCE = 0, /* as the reference sets it (vulkan_ray_tracing.cc:2452-2465): running
                                        min from (0,0,0) over the TLAS root's child boxes, for the low
                                        corner and (as written there) the high corner alike */
  RTX_COHERENCE_WORLD_TLAS = 1,      /* the TLAS header bounds */
  RTX_COHERENCE_WORLD_GIVEN = 2      /* world_min / world_max below */
};
typedef struct rtx_coherence_config {
  char hash;               /* 'f' francois (:432-445), 'g' grid + spherical (:447-456),
                              't' two point (:462-469), 'd' direction only (:471-474; the default) */
  uint8_t world;           /* RTX_COHERENCE_WORLD_* */
  uint16_t reserved;
  uint32_t francois_bits;  /* hash_francois_bits */
  uint32_t grid_bits;      /* hash_grid_bits */
  uint32_t sphere_bits;    /* hash_sphere_bits */
  float two_point_ratio;   /* hash_two_point_est_length_ratio */
  uint32_t packet;         /* rays per packet (the reference's warp_size); 0 = 64, one wavefront */
  float world_min[3];
  float world_max[3];
} rtx_coherence_config;
rtx_status rtx_coherence_order(rtx_ctx* ctx, const rtx_ray* rays_dev, uint32_t n, const rtx_coherence_config* cfg,
                               uint32_t* order_dev, uint32_t* hashes_dev);

/* ---- learned next-treelet prefetch -----------------------------------------
 * No counterpart in the reference (its RT unit has popularity heuristics
 * only, src/gpgpu-sim/shader.cc:3419-3559).  RTX_PREFETCH_LEARNED reads a
 * prediction table of the current partition: row 8*t + o holds, for a ray in
 * treelet t whose direction lies in octant o (bit0 x<0, bit1 y<0, bit2 z<0),
 * up to 4 entries (treelet + 1) | (probability * 255) << 24, 0 = empty, most
 * likely first.  At a staging decision the wave stages the heuristic's choice
 * and then each predicted treelet whose summed lane probability reaches
 * min_score.  The table only moves node reads between LDS and global memory:
 * hits and records are the same whatever it holds. */
#define RTX_OCTANTS 8u
rtx_status rtx_set_prefetch_table(rtx_ctx* ctx, const uint32_t* table_host, uint64_t n_rows);

/* The network behind the table: a BF16 MLP (FP32 accumulation, MFMA
 *   16x16x32) that maps
 *   [E[cur] | octant one-hot | depth | E[h1..h4]] -> 128 -> 128 -> 32 -> C
 * classes, where E is a 32-wide embedding of n_treelets (+1 "none") treelets,
 * h1..h4 are the ray's last four treelets on its path from the root (nearest
 * first), and class c scores q . O[c] + o[c].  Samples live in device memory. */
typedef struct rtx_mlp_sample {
  uint32_t cur;     /* treelet the ray is in */
  uint32_t octant;  /* 0-7 */
  float depth;      /* tree level of the node / 32 */
  uint32_t hist[4]; /* treelets, RTX_NO_TREELET for none */
  uint32_t target;  /* class of the next treelet (training only) */
} rtx_mlp_sample;   /* 32 bytes */

typedef struct rtx_mlp rtx_mlp;
rtx_status rtx_mlp_create(rtx_ctx* ctx, uint32_t n_treelets, uint32_t n_classes, uint64_t seed, rtx_mlp** out);
rtx_status rtx_mlp_destroy(rtx_mlp* m);
/* Parameters are one FP32 vector: E, W1 (192x128, rows 169.. zero), b1, W2, b2,
 * W3 (128x32), b3, O (C x 32), o. */
rtx_status rtx_mlp_param_count(const rtx_mlp* m, uint64_t* count);
rtx_status rtx_mlp_get_params(rtx_mlp* m, float* params_host, float* grads_host);
rtx_status rtx_mlp_set_params(rtx_mlp* m, const float* params_host);
/* Mean softmax cross-entropy over n samples and its gradient (left in the
 * gradient vector, read with rtx_mlp_get_params). */
rtx_status rtx_mlp_loss_grad(rtx_mlp* m, const rtx_mlp_sample* samples_dev, uint32_t n, float* loss);
/* Adam (0.9, 0.999, 1e-8) over minibatches in order, `epochs` passes. */
rtx_status rtx_mlp_train(rtx_mlp* m, const rtx_mlp_sample* samples_dev, uint32_t n, uint32_t epochs,
                         uint32_t batch, float lr, float* last_epoch_loss);
/* Top-4 classes of each sample (top4_dev[4n]) and their softmax probabilities
 * (probs_dev[4n], may be NULL). */
rtx_status rtx_mlp_predict(rtx_mlp* m, const rtx_mlp_sample* samples_dev, uint32_t n, uint32_t* top4_dev,
                           float* probs_dev);
/* C = op(A) . op(B) on the MLP's MFMA GEMM (M x K by K x N, row-major FP32,
 * operands rounded to BF16): exposes the kernel for testing. */
rtx_status rtx_mlp_gemm(rtx_ctx* ctx, int M, int N, int K, int trans_a, int trans_b, const float* A_dev,
                        const float* B_dev, float* C_dev);

/* ---- multi-scene batches -------------------------------------------------
 * No counterpart in the reference, which traces one ray per call against the
 * one TLAS of the launch.  A context also holds a list of scenes; a batch
 * traces ray ranges of any of them in one persistent launch whose waves pull
 * rays from a single global counter, so batches of very different sizes share
 * the GPU without a launch per scene.  Batches write hits only (no records). */
typedef struct rtx_batch_job {
  uint32_t scene;          /* id returned by rtx_add_scene */
  uint32_t n_rays;
  const rtx_ray* rays_dev;
  rtx_hit* hits_dev;
} rtx_batch_job;

typedef struct rtx_batch_graph rtx_batch_graph;

rtx_status rtx_add_scene(rtx_ctx* ctx, const void* arena_host, size_t bytes, uint64_t tlas_offset,
                         uint32_t* scene_id);
rtx_status rtx_build_scene_treelets(rtx_ctx* ctx, uint32_t scene_id, uint32_t budget_bytes,
                                    rtx_treelet_summary* out);
/* opts->prefetch must be RTX_PREFETCH_NONE; opts->emit_records must be 0. At most 1024 jobs. */
rtx_status rtx_trace_batch(rtx_ctx* ctx, const rtx_batch_job* jobs, uint32_t n_jobs, const rtx_trace_options* opts);
/* Capture the batch launch (counter clear + persistent kernel) in a HIP graph. */
rtx_status rtx_batch_graph_create(rtx_ctx* ctx, const rtx_batch_job* jobs, uint32_t n_jobs,
                                  const rtx_trace_options* opts, rtx_batch_graph** out);
/* Replay the graph `replays` times on the context's stream, wait, report kernel errors.
 * A graph holds the device addresses of its scenes' arenas and, in treelet order, of
 * their partitions.  Rebuilding a partition it uses (rtx_build_scene_treelets,
 * rtx_autotune_batch_treelets) or destroying the context makes the graph stale:
 * launching it then returns RTX_ERR_STALE without running anything.  Destroy graphs
 * before their context; a graph left over after rtx_destroy can still be destroyed. */
rtx_status rtx_batch_graph_launch(rtx_batch_graph* g, uint32_t replays, float* ms);
rtx_status rtx_batch_graph_destroy(rtx_batch_graph* g);
/* Time candidate launch shapes (waves per CU, LDS stack depth, refill threshold)
 * on the batch and return the fastest in *best (mode taken from best->mode). */
rtx_status rtx_autotune_batch(rtx_ctx* ctx, const rtx_batch_job* jobs, uint32_t n_jobs, rtx_trace_options* best,
                              float* best_ms);
/* Treelet-order batches: for each candidate budget, rebuild the partition of
 * every scene the jobs name, then run rtx_autotune_batch; the fastest budget
 * and launch shape win.  The scenes keep the partitions of the winning budget.
 * best->mode must be RTX_MODE_TREELET; budgets are 64-byte multiples >= 256. */
rtx_status rtx_autotune_batch_treelets(rtx_ctx* ctx, const rtx_batch_job* jobs, uint32_t n_jobs,
                                       const uint32_t* budgets, uint32_t n_budgets, rtx_trace_options* best,
                                       uint32_t* best_budget, float* best_ms);
/* Duration of the traversal kernel of the last rtx_trace call (the hit pass),
 * measured with HIP events recorded on the context's stream right around it. */
rtx_status rtx_last_kernel_ms(rtx_ctx* ctx, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* RTX_H */
