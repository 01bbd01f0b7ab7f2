/*
 * rtx_scene.h — host-side producer of the acceleration-structure arena and of
 * the synthetic workloads.
 *
 * In the reference this role is played by mesa-vulkan-sim + Embree 3.13.4
 * (README.md:36-45, 65), which build the BVH and hand the simulator a pointer
 * to it (gpgpusim_allocTLAS / gpgpusim_allocBLAS, gpgpusim_calls_from_mesa.cc:168-176),
 * and by the LumiBench / RayTracingInVulkan raygen shaders that generate rays.
 * Neither is part of the reference repository.  This builder emits the same
 * node layout (vulkan_acceleration_structure_util.h) with a binned-SAH binary
 * build collapsed to 6-wide nodes, one triangle per 64-byte quad leaf,
 * 128-byte instance leaves and conservative 8-bit child bounds.
 *
 * Arena order: TLAS header, TLAS nodes and instance leaves, then each BLAS
 * (header, nodes, leaves).  Every function here runs on the host CPU only.
 */
#ifndef RTX_SCENE_H
#define RTX_SCENE_H

#include <stddef.h>
#include <stdint.h>

#include "rtx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtx_builder rtx_builder;

rtx_status rtx_builder_create(rtx_builder** out);
rtx_status rtx_builder_destroy(rtx_builder* b);

/* Tree quality of the host build (the reference takes its trees from Embree, so
 * there is no counterpart in it).  DEFAULT: binned SAH on the axis of the
 * largest centroid extent, collapsed to 6-wide nodes by opening the child with
 * the largest surface area.  SAH_OPTIMAL: binned SAH over all three axes, and
 * the collapse that minimises the surface-area cost of the 6-wide tree (a
 * dynamic programme over the binary tree); fewer nodes per ray, a slower build.
 * Set before rtx_builder_finish (BLASes and TLAS) or rtx_builder_finish_device
 * (the TLAS only; the device builds its BLASes one way). */
enum { RTX_BUILD_DEFAULT = 0, RTX_BUILD_SAH_OPTIMAL = 1 };
rtx_status rtx_builder_set_quality(rtx_builder* b, uint32_t quality);
/* Add a triangle BLAS.  tris holds n_tris * 9 floats (v0 v1 v2, xyz each). */
rtx_status rtx_builder_add_blas(rtx_builder* b, const float* tris, uint32_t n_tris,
                                uint32_t geometry_index, uint32_t* blas_id);
/* Add an instance.  object_to_world is a 3x4 row-major matrix (p_world = M[:, :3] p + M[:, 3]).
 * The world-to-object matrix stored in the leaf is its inverse, rounded to float. */
rtx_status rtx_builder_add_instance(rtx_builder* b, uint32_t blas_id, const float object_to_world[12],
                                    uint32_t instance_id, uint32_t hit_group_offset);
/* Build every BLAS (threads = 0: all hardware threads) and the TLAS, and lay out the arena. */
rtx_status rtx_builder_finish(rtx_builder* b, uint32_t threads, uint64_t* arena_bytes, uint64_t* tlas_offset);
/* The same, with every BLAS built on the device of `ctx` in one set of launches
 * (Morton-ordered radix tree collapsed to 6-wide nodes, the same node layout);
 * the TLAS is built on the host.  Faster to build, more nodes per ray than the
 * SAH build. */
rtx_status rtx_builder_finish_device(rtx_builder* b, rtx_ctx* ctx, uint64_t* arena_bytes, uint64_t* tlas_offset);
rtx_status rtx_builder_copy_arena(const rtx_builder* b, void* dst, size_t capacity);
/* Size of the finished arena and offset of the TLAS header inside it. */
rtx_status rtx_builder_arena_info(const rtx_builder* b, uint64_t* arena_bytes, uint64_t* tlas_offset);
rtx_status rtx_builder_counts(const rtx_builder* b, uint64_t* n_triangles, uint64_t* n_blas,
                              uint64_t* n_instances, uint64_t* n_internal_nodes);
/* Arena offset of the header of BLAS blas_id (valid after finish). */
rtx_status rtx_builder_blas_offset(const rtx_builder* b, uint32_t blas_id, uint64_t* offset);

/* ---- procedural meshes ---------------------------------------------------- */
/* Displaced UV sphere: slices x stacks quads, 2 triangles each; every vertex is
 * moved radially by uniform[-disp, disp]; pole vertices are distinct per slice. */
rtx_status rtx_mesh_displaced_sphere(uint32_t slices, uint32_t stacks, float radius, float disp,
                                     uint64_t seed, float* tris_out /* slices*stacks*2*9 */);

/* ---- workloads ----------------------------------------------------------- */
typedef struct rtx_workload_desc {
  uint32_t kind;            /* RTX_WORKLOAD_* */
  uint32_t n_instances;     /* instanced spheres */
  uint32_t sphere_slices;   /* per-sphere grid; triangles = 2*slices*stacks */
  uint32_t sphere_stacks;
  uint32_t shared_blas;     /* 1: every instance references one BLAS */
  uint32_t width, height;   /* primary rays */
  uint64_t seed;
  float box_size;           /* instances placed uniformly in [0, box_size]^3 */
  float scale_min, scale_max;
  float displacement;
  uint32_t ray_flags;       /* flags of primary rays */
  uint32_t reserved;
} rtx_workload_desc;

enum {
  RTX_WORKLOAD_SPHERE_AND_QUAD = 0, /* one displaced sphere over a 2-triangle ground quad, 2 identity instances */
  RTX_WORKLOAD_INSTANCED_SPHERES = 1 /* n_instances displaced spheres scattered in a box */
};

/* Build the scene of a workload into a fresh builder (finished). */
rtx_status rtx_workload_scene(const rtx_workload_desc* d, uint32_t threads, rtx_builder** out);
/* The same scene with its BLASes built on the device of ctx (rtx_builder_finish_device). */
rtx_status rtx_workload_scene_device(const rtx_workload_desc* d, rtx_ctx* ctx, rtx_builder** out);
/* The workload's meshes and instances in a fresh builder, not yet built: finish
 * it with rtx_builder_finish or rtx_builder_finish_device (so the build can be
 * timed apart from mesh generation). */
rtx_status rtx_workload_builder(const rtx_workload_desc* d, rtx_builder** out);
/* Pinhole primary rays through pixel centres, one per pixel, ordered by 8x8
 * pixel tiles (tiles row-major, pixels row-major inside a tile). */
rtx_status rtx_workload_primary_rays(const rtx_workload_desc* d, rtx_ray* rays_out);
/* One secondary ray per hit: kind 0 = cosine-weighted diffuse bounce, 1 = shadow ray
 * toward light_pos (Opaque|TerminateOnFirstHit|SkipClosestHitShader, tmax = distance).
 * Writes n_out rays; src_index_out[k] names the ray that spawned ray k (may be NULL). */
rtx_status rtx_workload_secondary_rays(const rtx_builder* scene, const rtx_ray* rays, const rtx_hit* hits,
                                       uint32_t n, uint32_t kind, uint64_t seed, const float light_pos[3],
                                       rtx_ray* rays_out, uint32_t* src_index_out, uint32_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* RTX_SCENE_H */
