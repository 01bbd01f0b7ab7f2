// partition.h — device treelet partition (createTreelets + buildNodeToRootMap).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx.h"

namespace rtxk {

struct Partition {
  bool valid = false;
  uint32_t budget = 0;
  uint32_t n = 0;              // treelets
  uint32_t first_treelet = 0;  // dense index of the treelet rooted at the TLAS header
  uint64_t n_list = 0;
  uint32_t* owner = nullptr;         // [arena_bytes/64] dense treelet index or RTX_NO_TREELET, followed by
                                     // [arena_bytes/64] bytes of first-treelet child bits (first_bits_kernel)
  uint64_t* roots = nullptr;         // [n] root offsets, ascending
  uint64_t* list_offsets = nullptr;  // [n+1]
  uint64_t* list_nodes = nullptr;    // [n_list] node offsets
  // Treelet-contiguous images for LDS staging: image t holds the bytes of
  // the nodes treelet t owns, in list order; stage_slot[s] is the 64-byte
  // position of node slot s inside its owner's image.
  uint8_t* images = nullptr;
  uint64_t* image_offset = nullptr;  // [n]
  uint32_t* image_bytes = nullptr;   // [n]
  uint32_t* stage_slot = nullptr;    // [arena_bytes/64]
  uint32_t* image_slots = nullptr;   // [images_total/64] node slot at each 64-byte unit (instance tail: none)
  uint64_t images_total = 0;
  uint32_t max_image_bytes = 0;
  rtx_treelet_summary summary{};
  // treelet layout (remapBVHToTreeletLayout), built on request
  uint64_t* layout_remap = nullptr;    // [arena_bytes/64] layout offset of each node, ~0 for none
  int32_t* layout_metadata = nullptr;  // [n * budget/64] entry places relative to the root's, -1 past a list
  uint32_t layout_stride = 0;
  uint64_t layout_mapped = 0;          // nodes placed
};

void partition_free(Partition& p);
rtx_status partition_build(const uint8_t* arena, uint64_t bytes, uint64_t tlas, uint32_t budget, hipStream_t s,
                           Partition& out);
rtx_status partition_copy_out(const Partition& p, uint64_t arena_bytes, hipStream_t s, uint64_t* roots,
                              uint64_t* list_offsets, uint64_t* list_nodes, uint32_t* owner);
rtx_status partition_install(const uint8_t* arena, uint64_t bytes, uint64_t tlas, uint32_t budget, uint32_t n,
                             const uint64_t* roots, const uint64_t* list_offsets, const uint64_t* list_nodes,
                             const uint32_t* owner, hipStream_t s, Partition& out);
// What the structural walk of a scene found.  On a malformed arena (or no
// scratch memory) `valid` is false and the rest keeps its safe values: the
// compare-select slab tests, and the level pass.
struct SceneCheck {
  bool valid = false;
  bool ordered = false;           // every child box the TLAS reaches has lo <= hi bytes on all axes
  bool level_consistent = false;  // every reachable node has one tree level (every path agrees)
  bool cyclic = false;            // some node reaches itself
};
SceneCheck scene_check(const uint8_t* arena, uint64_t bytes, uint64_t tlas, hipStream_t s);
// remapBVHToTreeletLayout of the partition with -treelet_remap_stride `stride`
// (and the per-treelet metadata table); RTX_ERR_UNSUPPORTED where the
// reference asserts (a treelet root already placed by an earlier treelet).
rtx_status partition_layout(const uint8_t* arena, uint64_t bytes, uint64_t tlas, uint32_t stride, hipStream_t s,
                            Partition& p);
// The structural walk alone: RTX_ERR_CYCLE when a reachable node reaches itself.
rtx_status partition_validate(const uint8_t* arena, uint64_t bytes, uint64_t tlas, hipStream_t s,
                              uint64_t* n_nodes);

}  // namespace rtxk
