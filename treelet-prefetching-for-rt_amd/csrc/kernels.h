// kernels.h — launch parameters shared by the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx.h"

namespace rtxk {

// Error bits raised by kernels (mapped to rtx_status on the host).
enum : uint32_t {
  kErrStack = 1u,      // per-lane traversal stack overflow
  kErrType = 2u,       // malformed node type
  kErrRange = 4u,      // node outside the arena
  kErrRecords = 8u,    // record buffer overflow
  kErrInstSlot = 16u,  // treelet order: an instance leaf 128 MiB or more away from the TLAS header
  kErrSteps = 32u,     // per-ray step limit (cyclic or corrupt BVH)
  kErrBudget = 64u,    // treelet node larger than the budget
  kErrQueue = 128u     // partitioner work buffer overflow
};

// Record output of a trace kernel.
enum : int {
  kRecNone = 0,  // hits only
  kRecCsr = 1,   // records at record_offsets[ray]
  kRecSlab = 2   // a ray's first record_slab records at ray * slab_stride(record_slab)
};

constexpr uint32_t kBlock = 256;        // threads per block (4 waves of 64)
constexpr uint32_t kSlotWords = 64;     // control slot: [0] ray counter, [1] error, [2] overflow count, stats (u64) from word 16
constexpr uint32_t kMaxStepsPerRay = 1u << 24;
constexpr uint32_t kMaxBatchJobs = 1024;

// One entry of a multi-scene batch: a ray range of one scene (64 bytes).
struct alignas(64) BatchJob {
  const uint8_t* arena;
  uint64_t arena_bytes;
  uint64_t tlas;
  const rtx_ray* rays;
  rtx_hit* hits;
  const uint32_t* owner;
  uint32_t first_treelet;
  uint32_t n;
  uint32_t ray_base;  // rays of all earlier jobs
  uint32_t pad[3];
};

struct TraceParams {
  const uint8_t* arena;
  uint64_t arena_bytes;
  uint64_t tlas;
  const rtx_ray* rays;
  uint32_t n;
  uint32_t refill;  // refill the wave when at least this many lanes are idle
  rtx_hit* hits;
  rtx_record* records;
  uint64_t record_capacity;
  const uint64_t* record_offsets;
  // slab records: ray r's first record_slab records as 8-byte slab entries at
  // slabs[r * slab_stride(record_slab)] (`records` unused)
  uint64_t* slabs;
  uint32_t record_slab;
  uint32_t* rec_counts;    // slab records: ray r's record count at rec_counts[r]
  // level pass: per lane a tree_level_map of lm_cap u64 entries (zero between
  // rays) at lm_table + lane * lm_cap, and lm_cap / 2 positions at lm_list
  uint64_t* lm_table;
  uint32_t* lm_list;
  uint32_t lm_cap;
  const uint64_t* layout_remap;  // nullable: records carry these layout offsets (arena slot -> offset)
  uint32_t claim_start;    // rays are claimed from this index on, wrapping to 0
  const uint32_t* ray_list;  // nullable: claim k traces ray ray_list[k] (n = list length)
  uint32_t n_rays;           // rays in the ray / hit arrays (ray_list entries must lie below)
  const uint32_t* owner;  // node slot -> dense treelet index (nullable in plain mode)
  uint32_t first_treelet; // dense index of the treelet rooted at the TLAS header
  uint64_t* spill;        // global stack spill, spill_cap entries per stack per lane
  uint32_t spill_cap;
  uint32_t* ray_counter;
  uint32_t* error;
  unsigned long long* stats;  // rtx_prefetch_stats counters
  // single-scene launches: a mapped host word set to 1 when any ray reports
  // an error
  uint32_t* host_err;
  // treelet staging
  uint32_t prefetch;                 // RTX_PREFETCH_*
  float threshold;
  const uint32_t* image_slots;       // node slot held by each 64-byte unit of `images`
  const uint64_t* image_offset;      // dense treelet -> byte offset of its image in `images`
  const uint32_t* image_bytes;       // dense treelet -> image size
  const uint8_t* images;             // treelet-contiguous copies of the arena
  uint32_t slot_units;               // staging slot size in 64-byte units (0: one slot)
  float min_score;                   // learned policy: summed lane probability to stage a prediction
  float pred_fraction;               // learned policy: share of a predicted image staged (root first)
  const uint4* pred_table;           // learned policy: [pred_treelets * 8] rows of 4 packed predictions
  uint32_t pred_treelets;
  // multi-scene batch (batch kernels only): rays are numbered across jobs
  const BatchJob* jobs;
  uint32_t n_jobs;
  // every child box of the scene(s) has lower bound bytes <= upper ones, so a
  // finite ray may pick each axis's near and far bound by its direction sign
  uint32_t ordered_boxes;
};

// Most treelets resident at once in a wave's staging area.
constexpr uint32_t kMaxStageSlots = 16;

// Device scratch of rtx_trace_static (static_scratch.hip): 256 bytes of
// control words, then the per-lane global stack spill.
constexpr size_t kStaticScratchBytes = 128ull << 20;
constexpr size_t kStaticControlBytes = 256;
constexpr uint32_t kStaticMinSpill = 24;  // spill entries per lane and list it keeps at least
// Address of the calling thread's current device's copy of the scratch array
// (looked up once per device; `device` must be current).
hipError_t static_scratch_init(int device, void** out);

// Launchers (defined in the .hip files).
uint32_t trace_block_size(uint32_t stage_bytes);
hipError_t launch_trace(const TraceParams& p, uint32_t mode, int records, uint32_t lds_stack, uint32_t stage_bytes,
                        uint32_t grid, hipStream_t s);
hipError_t launch_trace_batch(const TraceParams& p, uint32_t mode, uint32_t lds_stack, uint32_t grid,
                              hipStream_t s);
// Workgroups of a trace kernel that fit on one CU at once (registers and LDS);
// 0 when the configuration has no kernel.
uint32_t trace_blocks_per_cu(uint32_t mode, int records, uint32_t lds_stack, uint32_t stage_bytes);
uint32_t trace_batch_blocks_per_cu(uint32_t mode, uint32_t lds_stack);
// Cooperative kernel for small batches (eight lanes per ray, no records or
// staging), plain or treelet order: `cap` entries per list and ray (at most
// 144), 256-lane blocks.
constexpr uint32_t kWideRayLanes = 8;
hipError_t launch_trace_wide(const TraceParams& p, uint32_t mode, uint32_t cap, uint32_t grid, hipStream_t s);
uint32_t trace_wide_blocks_per_cu(uint32_t mode);
// The level pass over p.n rays (p.ray_list / claim order as in launch_trace):
// each hit's max_level becomes the maximum of the reference's final
// tree_level_map.  p.lm_table must hold grid * 256 zeroed maps of p.lm_cap
// (a power of two) entries.
hipError_t launch_level_pass(const TraceParams& p, uint32_t mode, uint32_t grid, hipStream_t s);
// Bytes of level-pass memory per lane: the map and its position list.
constexpr uint64_t level_pass_lane_bytes(uint32_t lm_cap) { return (uint64_t)lm_cap * 8 + (uint64_t)lm_cap / 2 * 4; }
constexpr uint32_t kLevelMapCap = 4096;  // map entries per lane: at most 2,048 nodes assigned per ray
// Clears `words` 32-bit words with a kernel (usable inside graph capture).
hipError_t launch_fill_zero(uint32_t* dst, uint32_t words, hipStream_t s);
// host[0] = *word, host[2..3] = *dword (host: mapped host memory), so that the
// caller reads both after a stream synchronisation without a copy each.
hipError_t launch_report_words(const uint32_t* word, const uint64_t* dword, uint32_t* host, hipStream_t s);
// CSR offsets[0..n] of the rays' record counts: hits[i].n_records, or
// counts[i] when `counts` is given.
hipError_t launch_scan_records(const rtx_hit* hits, const uint32_t* counts, uint32_t n, uint64_t* offsets,
                               uint64_t* block_tmp, hipStream_t s);
// Statistics of a traced batch: acc[0..6] records per type, [7] hits, [8] any
// hits, [9] nodes, [10] max nodes, [11] max level, [12] fetch bytes, [13]
// accessed bytes (acc must start at 0; records may be null).
hipError_t launch_trace_stats(const rtx_hit* hits, uint32_t n, const rtx_record* records, uint64_t n_records,
                              unsigned long long* acc, hipStream_t s);
// The coherence engine's claim order of n rays (coherence.hip).
hipError_t coherence_order(const rtx_ray* rays, uint32_t n, const rtx_coherence_config& cfg, const float wmin[3],
                           const float wmax[3], uint32_t* order, uint32_t* hashes_out, hipStream_t s);
// Slab records: each lane holds its last kRecRing entries in LDS and writes
// them to the slab in aligned groups of four (32 bytes), so slab rows are
// padded to a multiple of four entries.
constexpr uint32_t kRecRing = 8;
__host__ __device__ inline uint32_t slab_stride(uint32_t record_slab) { return (record_slab + 3u) & ~3u; }
// A slab entry: the node's arena slot (offset / 64) in the low word, the
// record type and level in the high word.  The size follows from the type;
// the treelet and the layout offset are looked up when the entry is expanded.
__host__ __device__ inline uint64_t slab_entry(uint32_t slot, uint32_t type, uint32_t level) {
  return (uint64_t)slot | ((uint64_t)(type | ((level > 255u ? 255u : level) << 8)) << 32);
}
// Expand each ray's slab entries (at most `slab`) into rtx_records at its CSR
// position (owner and layout_remap nullable, as in TraceParams); rays with
// more than `slab` records are appended to overflow_list as ray_base + their
// index (count in overflow_count, which must start at 0).  Nothing is copied
// when the total (offsets[n]) exceeds `capacity`.
hipError_t launch_compact_records(const uint32_t* counts, uint32_t n, const uint64_t* slabs, uint32_t slab,
                                  const uint32_t* owner, const uint64_t* layout_remap, const uint64_t* offsets,
                                  rtx_record* records, uint64_t capacity, uint32_t ray_base, uint32_t* overflow_list,
                                  uint32_t* overflow_count, hipStream_t s);
// v[0..count) += *addend (a device word).
hipError_t launch_add_offset(uint64_t* v, uint64_t count, const uint64_t* addend, hipStream_t s);

}  // namespace rtxk
