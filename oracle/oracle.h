/*
 * oracle.h — CPU restatement of the reference ray-traversal path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing in the product path.  It restates, in plain C compiled with
 * -O3 -ffp-contract=off (the reference builds with g++ -O3 and no -march, so
 * SSE FP32 without contraction), these functions of the reference:
 *
 *   traceRay               src/cuda-sim/vulkan_ray_tracing.cc:2309-3076
 *   traceRayWithTreelets   src/cuda-sim/vulkan_ray_tracing.cc:1522-2307
 *   createTreelets         src/cuda-sim/vulkan_ray_tracing.cc:823-1470
 *   buildNodeToRootMap     src/cuda-sim/vulkan_ray_tracing.cc:475-489
 *   node decoders          src/cuda-sim/vulkan_acceleration_structure_util.h:66-508
 *   intersection math      src/cuda-sim/vulkan_ray_tracing.cc:148-257, 3089-3130
 *   float3 helpers         src/gpgpu-sim/vector-math.cc:28-108
 *
 * Node identity is the arena byte offset; the structs below have the same
 * layout as rtx_ray / rtx_hit / rtx_record of include/rtx.h.
 */
#ifndef RTX_ORACLE_H
#define RTX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ray {
  float org[3];
  float tmin;
  float dir[3];
  float tmax;
  uint32_t flags, cull_mask, sbt_record_offset, sbt_record_stride, miss_index;
  uint32_t reserved[3];
} orc_ray;

typedef struct orc_hit {
  uint32_t hit, primitive_index, geometry_index, instance_index;
  float t;
  float p[3];
  float bary[3];
  uint32_t n_records, n_nodes, fetch_bytes, n_any_hits, max_level;
  uint64_t instance_leaf;  /* closest_instanceLeaf (vulkan_ray_tracing.cc:2856), ~0 on a miss */
  uint32_t accessed_bytes; /* accessedDataSize of the ray (:2260): every record size */
  uint32_t path_level;     /* deepest level along the traversal paths (max_level is the map's maximum) */
} orc_hit;

typedef struct orc_record {
  uint64_t offset;
  uint32_t treelet;
  uint16_t size;
  uint8_t type;
  uint8_t level; /* depth of the node along the path that reached it (TLAS root 1, a BLAS root and
                    its header its instance leaf's, a child its parent's + 1; 0 for the TLAS
                    header), saturating at 255 */
} orc_record;

typedef struct orc_partition orc_partition;

/* Status codes share the numbering of rtx_status. */
enum {
  ORC_OK = 0,
  ORC_ERR_INVALID_ARGUMENT = 1,
  ORC_ERR_BAD_NODE_TYPE = 2,
  ORC_ERR_OFFSET_OUT_OF_ARENA = 3,
  ORC_ERR_NODE_LARGER_THAN_BUDGET = 4,
  ORC_ERR_TRACE_OVERFLOW = 5,
  ORC_ERR_OUT_OF_MEMORY = 12
};

/* createTreelets + buildNodeToRootMap.  expand_duplicates = 1 re-forms a
 * treelet every time its root is reached again through a shared BLAS, as the
 * reference does; 0 forms it once (identical maps, less work). */
int orc_create_treelets(const uint8_t* arena, uint64_t arena_bytes, uint64_t tlas_offset, uint32_t budget,
                        int expand_duplicates, orc_partition** out);
void orc_free_partition(orc_partition* p);
uint32_t orc_partition_count(const orc_partition* p);
uint64_t orc_partition_list_entries(const orc_partition* p);
/* Sum of node bytes over every treelet formation ("Total BVH Size", vulkan_ray_tracing.cc:1364). */
uint64_t orc_partition_total_bytes(const orc_partition* p);
/* Copies: roots[n], list_offsets[n+1], list_nodes[entries], owner[arena_bytes/64]. NULL skips. */
int orc_partition_copy(const orc_partition* p, uint64_t* roots, uint64_t* list_offsets, uint64_t* list_nodes,
                       uint32_t* owner);

/* remapBVHToTreeletLayout (vulkan_ray_tracing.cc:1473-1509): treelet t (roots
 * in ascending address order) occupies bytes [t*(budget+stride), ...) of the
 * treelet layout, its root first, then the other nodes of its list in order,
 * each advancing by its size (128 for an instance leaf, 64 otherwise); a node
 * already placed by an earlier treelet keeps that place.  remap[arena/64]
 * receives each node's layout offset (~0 for nodes in no list).  metadata
 * [n * budget/64] (the per-treelet table of :1598-1605) receives, for list
 * entry j of treelet t, the signed byte offset of the entry's layout place
 * from its root's, -1 past the end of the list.  Returns
 * ORC_ERR_INVALID_ARGUMENT where the reference asserts: a root already placed
 * by an earlier treelet (:1483). */
int orc_treelet_layout(const uint8_t* arena, uint64_t arena_bytes, uint64_t tlas_offset, const orc_partition* p,
                       uint32_t stride, uint64_t* remap, int32_t* metadata);

/* mode 0 = traceRay, 1 = traceRayWithTreelets (needs a partition).  The
 * partition, when given in mode 0, only labels records with treelets.
 * records may be NULL (then only record_offsets / counts are produced).
 * record_offsets[n+1] may be NULL. */
int orc_trace(const uint8_t* arena, uint64_t arena_bytes, uint64_t tlas_offset, const orc_partition* part,
              int mode, const orc_ray* rays, uint32_t n, orc_hit* hits, orc_record* records,
              uint64_t record_capacity, uint64_t* record_offsets);
/* Hits only, rays split over `threads` std threads (pthreads); for timing. */
int orc_trace_parallel(const uint8_t* arena, uint64_t arena_bytes, uint64_t tlas_offset, const orc_partition* part,
                       int mode, const orc_ray* rays, uint32_t n, orc_hit* hits, int threads);

/* Individual math routines, exposed so tests can pin them. */
int orc_ray_box_test(const float lo[3], const float hi[3], const float idir[3], const float org[3], float tmin,
                     float tmax, float* thit);
void orc_calculate_idir(const float dir[3], float idir[3]);
int orc_mt_ray_triangle_test(const float p0[3], const float p1[3], const float p2[3], const float org[3],
                             const float dir[3], float* thit);
void orc_barycentric(const float p[3], const float a[3], const float b[3], const float c[3], float out[3]);
void orc_set_child_bounds(const uint8_t* node, int child, float lo[3], float hi[3]);
/* make_transformed_ray: rows = 12 floats of an instance leaf matrix (rows 0..3, 3 columns). */
void orc_make_transformed_ray(const float rows[12], const float org[3], const float dir[3], float tmin,
                              float tmax, float out_org[3], float out_dir[3], float* out_tmin, float* out_tmax,
                              float* t_multiplier);
/* Structural walk of the TLAS->BLAS DAG (node types, offsets in range).
 * Returns ORC_OK and the number of distinct node offsets reached. */
int orc_validate(const uint8_t* arena, uint64_t arena_bytes, uint64_t tlas_offset, uint64_t* n_nodes);

#ifdef __cplusplus
}
#endif
#endif
