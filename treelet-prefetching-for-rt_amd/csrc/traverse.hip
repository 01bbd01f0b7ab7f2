// traverse.hip — per-ray BVH traversal kernels for gfx950.
//
// One ray per lane.  Waves are persistent: whenever at least `refill` lanes of
// a wave are idle, the wave takes that many ray indices from a global counter
// with one atomic, so uneven ray costs do not leave lanes idle.
//
// Both traversal orders of the reference are flattened into "pop one entry,
// process it" steps:
//  * PLAIN (traceRay, vulkan_ray_tracing.cc:2309-3076): the nested top/bottom
//    loops visit the first hit internal child immediately and push the other
//    hit children in slot order; that is the same sequence as pushing all hit
//    children in slot order with the first internal child pushed last.  Only
//    one instance is open at a time, so the instance context lives in
//    registers.
//  * TREELET (traceRayWithTreelets, :1522-2307): two LIFO lists.  A child goes
//    to the "current" list while the current treelet is still the one rooted
//    at the TLAS header and the child's owner is that treelet; after the
//    first move from the "other" list the reference compares against a host
//    address (:1752), so every later child goes to "other".  Which children
//    the first treelet owns is precomputed with the partition, one byte per
//    node slot, so a node costs one extra byte load, not one per child.
//    Entries carry the instance leaf that opened them; the object-space ray
//    is recomputed from that leaf when it differs from the cached one (same
//    inputs, same floats).
//
// Stacks: a ring of K entries per lane in LDS, spilling the oldest entries to
// a per-lane region in global memory (exact LIFO order).
//
// One wave iteration runs three passes, each only for the lanes that need it:
// instance leaves, triangle leaves, then internal nodes.  A lane that finishes
// a leaf pops its next entry at once and joins the internal-node pass if that
// entry is an internal node; an instance leaf hands over its BLAS root the
// same way when that root is the next entry to be popped.  Completed rays
// keep their state until the wave's next refill, which writes their hit
// records together.
//
// Small batches (at most 128 rays per CU in plain order, 256 in treelet
// order) run trace_wide instead: eight lanes per ray, the six child boxes of
// a node tested at once (see the section before its definition).
//
// Treelet staging (STAGE > 0, one wave per workgroup): the wave keeps up to
// stage/slot treelet images in LDS slots, copied from the treelet-contiguous
// image buffer with coalesced 16-byte loads, and an LDS hash table from node
// slot to position in the staging area.  A node read first probes the table;
// hits read LDS, misses read global memory.  When most lanes miss, the wave
// gathers the owning treelet of each lane's next node and stages the most
// popular one into the least recently read slot, following the RT-unit policy
// of src/gpgpu-sim/shader.cc:3419-3559 (heuristic 0 always, 1 only above a
// popularity share, 2 a popularity-sized prefix).  The learned policy then
// also stages the treelets a prediction table names as the lanes' likely next
// ones.  Staging changes where node bytes are read from, never which nodes
// are visited.
#include <algorithm>
#include <type_traits>

#include "device_common.h"
#include "kernels.h"

using namespace rtxd;

namespace rtxk {
namespace {

constexpr uint32_t kFlTop = 1u, kFlLeaf = 2u;
constexpr uint32_t kNoSlot = 0xffffffffu;
constexpr uint32_t kNone = 0xffffffffu;

// A stack entry: node slot (32 bits), flags (2), tree level (8) and the
// instance leaf that opened the entry (22 bits, treelet order only).  The
// instance is stored as its slot's distance from the TLAS header slot, a
// signed 22-bit value, so instance leaves may lie anywhere within 128 MiB of
// the TLAS header whatever the arena's size.
constexpr int32_t kInstRelMax = 1 << 21;
__device__ __forceinline__ uint64_t make_entry(uint32_t slot, uint32_t flags, uint32_t level, uint32_t inst_rel) {
  uint32_t lv = level > 255u ? 255u : level;
  return (uint64_t)slot | ((uint64_t)(flags | (lv << 2) | ((inst_rel & 0x3fffffu) << 10)) << 32);
}
__device__ __forceinline__ uint32_t e_slot(uint64_t e) { return (uint32_t)e; }
__device__ __forceinline__ uint32_t e_flags(uint64_t e) { return (uint32_t)(e >> 32) & 3u; }
__device__ __forceinline__ uint32_t e_level(uint64_t e) { return ((uint32_t)(e >> 32) >> 2) & 255u; }
__device__ __forceinline__ uint32_t e_inst_rel(uint64_t e) { return (uint32_t)(e >> 42); }
// absolute instance-leaf slot of an entry, given the TLAS header slot
__device__ __forceinline__ uint32_t e_inst(uint64_t e, uint32_t tlas_slot) {
  return tlas_slot + (uint32_t)(((int32_t)(e_inst_rel(e) << 10)) >> 10);
}
// the relative field of instance-leaf slot `inst`, or false when it is out of range
__device__ __forceinline__ bool inst_rel_of(uint32_t inst, uint32_t tlas_slot, uint32_t& rel) {
  const int64_t d = (int64_t)inst - (int64_t)tlas_slot;
  rel = (uint32_t)d & 0x3fffffu;
  return d >= -kInstRelMax && d < kInstRelMax;
}

struct Stack {
  uint32_t bot, cnt, g;
};

// LDS and global stack words are reached through pointers typed by address
// space, so a pop compiles to ds_read or global_load and never to a flat load
// (which would also wait for every outstanding store of the lane).
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(1))) uint64_t glb_u64;

template <int K, uint32_t BS>
__device__ __forceinline__ bool st_push(Stack& s, uint64_t e, lds_u64* lds, glb_u64* g, uint32_t gcap) {
  if (s.cnt == (uint32_t)K) {
    if (s.g >= gcap) return false;
    g[s.g++] = lds[s.bot * BS];
    s.bot = (s.bot + 1) & (K - 1);
    s.cnt--;
  }
  lds[((s.bot + s.cnt) & (K - 1)) * BS] = e;
  s.cnt++;
  return true;
}

template <int K, uint32_t BS>
__device__ __forceinline__ uint64_t st_pop(Stack& s, lds_u64* lds, glb_u64* g) {
  if (s.cnt == 0) return g[--s.g];
  s.cnt--;
  return lds[((s.bot + s.cnt) & (K - 1)) * BS];
}

__device__ __forceinline__ bool st_empty(const Stack& s) { return s.cnt == 0 && s.g == 0; }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t slot_hash(uint32_t s) { return s * 2654435761u; }

// Ray traced by claim k of a single-scene launch: a listed ray, else rays in
// input order starting at claim_start and wrapping to 0.
__device__ __forceinline__ uint32_t claimed_ray(const TraceParams& p, uint32_t k) {
  if (p.ray_list) return p.ray_list[k];
  const uint32_t tail = p.n - p.claim_start;  // rays from claim_start to the end
  return k < tail ? k + p.claim_start : k - tail;
}

// The reference's accessedDataSize of one ray (vulkan_ray_tracing.cc:2260):
// every record size, the 8-byte leaf descriptors included.  It follows from
// the counters a ray keeps.  With a internal nodes, b instance leaves and c
// primitive leaves (the TLAS header record emitted):
//   fetch = 64 + 64a + 192b + 64c,  nodes = a + b + c,  records = 1 + a + 2b + 2c
// so b = (fetch/64 - 1 - nodes) / 2, c = records - 1 - nodes - b, and the
// descriptors add 8c bytes.
__device__ __forceinline__ uint32_t accessed_bytes_of(uint32_t n_rec, uint32_t n_nodes, uint32_t fetch) {
  if (n_rec == 0) return 0u;
  const uint32_t b = ((fetch >> 6) - 1u - n_nodes) >> 1;
  const uint32_t c = n_rec - 1u - n_nodes - b;
  return fetch + 8u * c;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Non-staged kernels are compiled for five waves per SIMD (at most 96 VGPRs);
// staged kernels are bounded by their LDS image instead.
// LM: the level pass.  The rays are traced again with the reference's
// tree_level_map kept per lane in global memory, and only each hit's max_level
// is written: the maximum over the final map (vulkan_ray_tracing.cc:2274-2279),
// where a node reassigned later keeps its last value.  Needed only for scenes
// in which a node is reachable at two levels; elsewhere every assignment of a
// node carries the same value and the traversal's running maximum is exact.
template <int MODE, int REC, int K, int STAGE, bool MULTI = false, bool LM = false>
__global__ __launch_bounds__(STAGE ? 64 : 256) __attribute__((amdgpu_waves_per_eu(STAGE ? 1 : (REC == kRecCsr || MODE == 1 || LM) ? 4 : 5)))
void trace_kernel(TraceParams p) {
  static_assert(!MULTI || (!REC && STAGE == 0), "batch kernels write hits only");
  static_assert(!LM || (!REC && STAGE == 0 && !MULTI), "the level pass writes max_level only");
  constexpr uint32_t BS = STAGE ? 64u : 256u;
  constexpr int NS = MODE == 1 ? 2 : 1;
  constexpr uint32_t UNITS = STAGE / 64;
  constexpr uint32_t HSIZE = STAGE ? 2 * UNITS : 1;
  __shared__ uint64_t lds_stack[NS * K * BS];
  __shared__ uint4 stage_img[STAGE ? STAGE / 16 : 1];
  __shared__ uint32_t hkey[HSIZE];
  __shared__ uint16_t hpos[HSIZE];
  __shared__ uint8_t used[STAGE ? UNITS : 1];
  // slot table of the staging area: treelet, units held, iteration of the last
  // read served, staged on a prediction
  constexpr uint32_t MAXSLOTS = STAGE ? kMaxStageSlots : 1;
  __shared__ uint32_t sl_t[MAXSLOTS], sl_units[MAXSLOTS], sl_last[MAXSLOTS], sl_pred[MAXSLOTS], sl_first[MAXSLOTS];
  // prefetch effectiveness (src/gpgpu-sim/gpu-cache.cc:1960-2117): node slots
  // of staged units evicted unread (a later global read of one is TOO_EARLY)
  // and of recent global reads (staging one afterwards is TOO_LATE), each in a
  // direct-mapped table (a collision forgets the older slot)
  constexpr uint32_t EV = STAGE ? 128 : 1;
  __shared__ uint32_t ev_key[EV], dm_key[EV];
  // slab records: each lane's last kRecRing slab entries (entry k at ring
  // position k % kRecRing), written to the slab four at a time
  __shared__ uint64_t rec_ring[REC == kRecSlab ? kRecRing * BS : 1];

  // Staging-only parameters are read from the kernel-argument segment where
  // they are used (a scalar load each time; the empty asm keeps the compiler
  // from hoisting them) instead of living in scalar registers for the whole
  // launch: the staged kernels ran short of SGPRs and spilled 81 of them to
  // VGPR lanes, now 48 (plain+stage 1.347 -> 1.302 ms on the 1M-triangle batch).
  typedef const __attribute__((address_space(4))) TraceParams* kparams_ptr;
  auto kp = []() {
    kparams_ptr q = (kparams_ptr)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(q));
    return q;
  };
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  lds_u64* ldsA = (lds_u64*)(lds_stack + tid);
  lds_u64* ldsB = (lds_u64*)(lds_stack + (NS - 1) * K * BS + tid);
  lds_u64* ring = (lds_u64*)(rec_ring + (REC == kRecSlab ? tid : 0u));
  const uint64_t lane_slot = (uint64_t)blockIdx.x * BS + tid;
  glb_u64* gA = (glb_u64*)(p.spill + lane_slot * (uint64_t)p.spill_cap * NS);
  glb_u64* gB = gA + (NS - 1) * (uint64_t)p.spill_cap;
  // Scene of the lane's current ray (a batch kernel switches it per ray).
  const uint8_t* __restrict__ A = p.arena;
  uint64_t bytes = p.arena_bytes;
  uint64_t tlas = p.tlas;
  const uint32_t* owner = p.owner;
  uint32_t first_treelet = p.first_treelet;
  rtx_hit* hits_out = p.hits;
  const rtx_ray* rays_in = p.rays;
  uint32_t local_id = 0;
  __shared__ uint32_t job_base[MULTI ? kMaxBatchJobs + 1 : 1];
  if constexpr (MULTI) {
    for (uint32_t j = threadIdx.x; j <= p.n_jobs; j += BS) job_base[j] = j < p.n_jobs ? p.jobs[j].ray_base : p.n;
    __syncthreads();
  }

  Stack sa{0, 0, 0}, sb{0, 0, 0};
  bool active = false, exhausted = false;
  bool pending = false;  // the lane's completed ray still has to write its hit record
  uint32_t err = 0;

  F3 ro{}, rd{}, ridir{};
  float rtmin = 0.f, rtmax = 0.f;
  uint32_t rflags = 0;
  float min_thit = 0.f, min_thit_obj = 0.f;
  uint32_t closest_leaf = kNoSlot, closest_inst = kNoSlot;
  ObjRay cur{};
  uint32_t cur_inst = kNoSlot;
  // the world ray / current object ray can take the v_min/v_max slab tests
  bool ray_ok = false, cur_ok = false;
  uint32_t n_rec = 0, n_nodes = 0, fetch_bytes = 0, n_any = 0, max_level = 0;
  uint64_t rec_base = 0;
  uint32_t rec_flushed = 0;  // slab records: entries of the ray already in its slab
  bool phase2 = false;

  // The entry a lane popped this iteration and processes below.
  bool has_e = false;
  uint64_t e = 0;

  // staging state (wave-uniform)
  uint32_t cooldown = 0, iter = 0, n_staged = 0;
  uint64_t st_treelets = 0, st_bytes = 0, st_wasted = 0, st_pred = 0, st_pred_hits = 0, st_fallback = 0;
  uint64_t st_timely = 0;
  uint32_t reads = 0, lds_hits = 0, too_early = 0, too_late = 0;
  uint32_t octant = 0;  // direction octant of the lane's ray (learned policy)
  // slots of `cap` units each (powers of two): nslots treelets stay resident
  uint32_t cap_shift = 0, nslots = 1;
  if constexpr (STAGE > 0) {
    uint32_t cap = UNITS;
    if (p.slot_units > 0 && p.slot_units < UNITS) cap = p.slot_units;
    while ((1u << cap_shift) < cap) cap_shift++;
    nslots = min(MAXSLOTS, UNITS >> cap_shift);
    for (uint32_t i = lane; i < HSIZE; i += 64) hkey[i] = 0;
    for (uint32_t i = lane; i < EV; i += 64) ev_key[i] = dm_key[i] = 0;
    for (uint32_t i = lane; i < MAXSLOTS; i += 64) {
      sl_t[i] = kNone;
      sl_units[i] = 0;
      sl_last[i] = 0;
      sl_pred[i] = 0;
      sl_first[i] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  auto emit = [&](uint64_t off, uint32_t size, uint32_t type, uint32_t lvl) {
    if (type != RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR) fetch_bytes += size;
    if constexpr (REC == kRecSlab) {
      // a slab keeps the ray's first record_slab records, the rest are written
      // by a second pass over the rays that had more.  The entry holds what
      // the traversal knows; the owner and layout lookups wait for the
      // compaction, off the traversal's dependent chain.  It goes to the
      // lane's LDS ring first (flush_slab_group writes it out).
      if (n_rec < p.record_slab) ring[(n_rec & (kRecRing - 1)) * BS] = slab_entry((uint32_t)(off >> 6), type, lvl);
    } else if constexpr (REC == kRecCsr) {
      const uint64_t idx = rec_base + n_rec;
      if (idx < p.record_capacity) {
        uint32_t tl = owner ? owner[off >> 6] : RTX_NO_TREELET;
        const uint64_t roff = p.layout_remap ? p.layout_remap[off >> 6] : off;
        uint4 v;
        v.x = (uint32_t)roff;
        v.y = (uint32_t)(roff >> 32);
        v.z = tl;
        v.w = size | (type << 16) | ((lvl > 255u ? 255u : lvl) << 24);
        reinterpret_cast<uint4*>(p.records)[idx] = v;
      } else {
        err |= kErrRecords;
      }
    }
    n_rec++;
  };

  // Slab records: entries [rec_flushed, rec_flushed + 4) of the lane's ray go
  // from the LDS ring to the slab once all four exist, as two 16-byte stores
  // that fill one aligned 32-byte sector (8-byte stores scattered over the
  // slabs cost partial-sector writes and fills).  Called at the top of every
  // iteration, which adds at most three entries, so at most six are ever
  // waiting and the ring of eight never overwrites one.
  auto flush_slab_group = [&]() {
    if constexpr (REC == kRecSlab) {
      if (rec_flushed + 4u <= min(n_rec, p.record_slab)) {
        const uint32_t q = rec_flushed & (kRecRing - 1);
        const uint64_t a0 = ring[q * BS], a1 = ring[(q + 1) * BS], a2 = ring[(q + 2) * BS], a3 = ring[(q + 3) * BS];
        uint4* dst = reinterpret_cast<uint4*>(p.slabs + (uint64_t)local_id * slab_stride(p.record_slab) + rec_flushed);
        dst[0] = make_uint4((uint32_t)a0, (uint32_t)(a0 >> 32), (uint32_t)a1, (uint32_t)(a1 >> 32));
        dst[1] = make_uint4((uint32_t)a2, (uint32_t)(a2 >> 32), (uint32_t)a3, (uint32_t)(a3 >> 32));
        rec_flushed += 4u;
      }
    }
  };
  // The entries still in the ring when the ray's hit is written, as whole
  // groups too: the row is padded to a multiple of four entries, and the
  // entries past the ray's count (stale ring words) are never read.
  auto flush_slab_rest = [&]() {
    if constexpr (REC == kRecSlab) {
      const uint32_t end = min(n_rec, p.record_slab);
      uint4* row = reinterpret_cast<uint4*>(p.slabs + (uint64_t)local_id * slab_stride(p.record_slab));
      for (; rec_flushed < end; rec_flushed += 4u) {
        const uint32_t q = rec_flushed & (kRecRing - 1);
        const uint64_t a0 = ring[q * BS], a1 = ring[(q + 1) * BS], a2 = ring[(q + 2) * BS], a3 = ring[(q + 3) * BS];
        row[rec_flushed / 2] = make_uint4((uint32_t)a0, (uint32_t)(a0 >> 32), (uint32_t)a1, (uint32_t)(a1 >> 32));
        row[rec_flushed / 2 + 1] = make_uint4((uint32_t)a2, (uint32_t)(a2 >> 32), (uint32_t)a3, (uint32_t)(a3 >> 32));
      }
      rec_flushed = end;
    }
  };

  auto push_entry = [&](uint64_t e, bool to_a) {
    bool ok;
    if (MODE == 0 || to_a)
      ok = st_push<K, BS>(sa, e, ldsA, gA, p.spill_cap);
    else
      ok = st_push<K, BS>(sb, e, ldsB, gB, p.spill_cap);
    if (!ok) err |= kErrStack;
  };

  // ---- the level pass's per-lane tree_level_map: open addressing over
  // lm_cap entries ((slot + 1) << 32 | level, 0 = empty), at most half full,
  // and the positions this ray wrote (for the final maximum and the clear)
  uint64_t* lm_t = nullptr;
  uint32_t* lm_l = nullptr;
  uint32_t lm_n = 0;
  if constexpr (LM) {
    lm_t = p.lm_table + lane_slot * (uint64_t)p.lm_cap;
    lm_l = p.lm_list + lane_slot * (uint64_t)(p.lm_cap / 2);
  }
  auto lm_pos = [&](uint32_t slot) -> uint32_t {
    uint32_t h = slot_hash(slot) & (p.lm_cap - 1);
    while (true) {
      const uint64_t v = lm_t[h];
      if (v == 0 || (uint32_t)(v >> 32) == slot + 1) return h;
      h = (h + 1) & (p.lm_cap - 1);
    }
  };
  auto lm_get = [&](uint32_t slot) -> uint32_t { return (uint32_t)lm_t[lm_pos(slot)]; };
  auto lm_set = [&](uint32_t slot, uint32_t level) {
    const uint32_t h = lm_pos(slot);
    if (lm_t[h] == 0) {
      if (lm_n >= p.lm_cap / 2) {  // the map is full: the ray reports a stack overflow
        err |= kErrStack;
        return;
      }
      lm_l[lm_n++] = h;
    }
    lm_t[h] = ((uint64_t)(slot + 1) << 32) | level;
  };

  // TLAS header slot of the lane's scene (instance leaves are stored relative to it)
  auto tlas_slot = [&]() -> uint32_t { return (uint32_t)(tlas >> 6); };
  // Treelet order: does `slot` belong to the current ("first") treelet?
  auto goes_current = [&](uint32_t slot) -> bool {
    if (MODE == 0) return true;
    if (phase2) return false;
    return owner[slot] == first_treelet;
  };
  // The first-treelet child bits of node slot `slot`: one byte per slot after
  // the owner words (bit i: child i of an internal node is owned by the first
  // treelet; bit 0 of an instance leaf: its BLAS root is).
  auto first_bits = [&](uint32_t slot) -> uint32_t {
    return reinterpret_cast<const uint8_t*>(owner + (bytes >> 6))[slot];
  };

  // Position of `slot` in the staged image, or kNone.
  auto staged_pos = [&](uint32_t slot) -> uint32_t {
    if constexpr (STAGE > 0) {
      if (n_staged == 0) return kNone;
      uint32_t h = slot_hash(slot) & (HSIZE - 1);
      for (uint32_t probe = 0; probe < HSIZE; ++probe) {
        const uint32_t k = hkey[h];
        if (k == slot + 1) return hpos[h];
        if (k == 0) return kNone;
        h = (h + 1) & (HSIZE - 1);
      }
    }
    return kNone;
  };

  // Read `count` 16-byte words of a node starting at byte `first` (a multiple
  // of 16): from the staged image when `pos` is its staged position, else
  // from global memory.  Explicit LDS and global paths keep both loads typed.
  auto read_words = [&](uint64_t off, uint32_t pos, uint32_t first, uint4* w, int count) {
    if constexpr (STAGE > 0) {
      reads++;
      if (pos != kNone) {
        lds_hits++;
        used[pos] = 1;
        sl_last[pos >> cap_shift] = iter;
        const uint4* q = stage_img + pos * 4 + first / 16;
        for (int k = 0; k < count; ++k) w[k] = q[k];
        return;
      }
      // a demand read from global memory: TOO_EARLY when the node's unit
      // was staged and evicted unread; remembered for TOO_LATE
      const uint32_t s = (uint32_t)(off >> 6), h = slot_hash(s) & (EV - 1);
      if (ev_key[h] == s + 1 && atomicCAS(&ev_key[h], s + 1, 0u) == s + 1) too_early++;
      dm_key[h] = s + 1;
    }
    const uint4* q = reinterpret_cast<const uint4*>(A + off + first);
    for (int k = 0; k < count; ++k) w[k] = q[k];
  };
  auto read_node = [&](uint64_t off, uint32_t pos) -> Words16 {
    uint4 w[4];
    read_words(off, pos, 0, w, 4);
    Words16 r;
    for (int k = 0; k < 4; ++k) {
      r.w[4 * k] = w[k].x;
      r.w[4 * k + 1] = w[k].y;
      r.w[4 * k + 2] = w[k].z;
      r.w[4 * k + 3] = w[k].w;
    }
    return r;
  };

  auto load_instance = [&](uint32_t inst) {
    const uint64_t ioff = (uint64_t)inst << 6;
    if (ioff + 128 > bytes) {
      err |= kErrRange;
      return;
    }
    const uint4* q = reinterpret_cast<const uint4*>(A + ioff + 16);
    uint4 a = q[0], b = q[1], c = q[2];
    float m[12] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w),
                   __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w),
                   __uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), __uint_as_float(c.w)};
    cur = transform_ray(m, ro, rd, rtmin, rtmax);
    cur_ok = slab_inputs_ok(cur.o, cur.d, cur.idir, cur.tmin, cur.tmax);
    cur_inst = inst;
  };

  auto wave_sync = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  // Account slot j's unused units, then empty it (wave-uniform).
  auto evict_slot = [&](uint32_t j) {
    if constexpr (STAGE > 0) {
      const uint32_t units = sl_units[j];
      if (sl_t[j] == kNone) return;
      const uint32_t base = j << cap_shift;
      uint32_t u = 0;
      for (uint32_t i = lane; i < units; i += 64) u += used[base + i] ? 1u : 0u;
      u = wave_sum(u);
      st_wasted += (uint64_t)(units - u) * 64;
      st_timely += (uint64_t)u * 64;
      if (sl_pred[j] && u > 0) st_pred_hits++;
      const uint64_t img = kp()->image_offset[sl_t[j]] / 64 + sl_first[j];
      for (uint32_t i = lane; i < units; i += 64) {
        if (used[base + i]) continue;
        const uint32_t s = kp()->image_slots[img + i];
        if (s != kNone) ev_key[slot_hash(s) & (EV - 1)] = s + 1;
      }
      wave_sync();
      if (lane == 0) {
        sl_t[j] = kNone;
        sl_units[j] = 0;
        sl_pred[j] = 0;
      }
      wave_sync();
      n_staged--;
    }
  };
  auto evict_all = [&]() {
    if constexpr (STAGE > 0)
      for (uint32_t j = 0; j < nslots; ++j) evict_slot(j);
  };

  // Copy treelet t (at most `max_units` 64-byte units of its image: the first
  // ones, or the last ones when `latter`) into a slot (wave-uniform): the slot
  // already holding it, else an empty one, else the least recently read one
  // that served no read in this iteration.  The hash table is rebuilt over
  // every resident slot.  Returns whether it copied.
  auto stage = [&](uint32_t t, uint32_t max_units, bool predicted, bool latter = false) -> bool {
    if constexpr (STAGE > 0) {
      const uint32_t tt = __builtin_amdgcn_readfirstlane(t);
      const uint32_t total = kp()->image_bytes[tt] / 64;
      uint32_t units = min(total, min(1u << cap_shift, max_units));
      units = __builtin_amdgcn_readfirstlane(units);
      if (units == 0) return false;
      const uint32_t first = latter ? total - units : 0u;
      // victim: the slot holding that part of t, else an empty slot, else the
      // least recently read slot that served no read in this iteration
      uint32_t victim = kNone, empty = kNone, lru = kNone, lru_it = 0;
      for (uint32_t j = 0; j < nslots; ++j) {
        const uint32_t sj = sl_t[j];
        if (sj == tt && sl_first[j] == first) victim = j;
        if (sj == kNone) {
          if (empty == kNone) empty = j;
        } else if (sl_last[j] < iter && (lru == kNone || sl_last[j] < lru_it)) {
          lru = j;
          lru_it = sl_last[j];
        }
      }
      if (victim != kNone) {
        if (sl_units[victim] >= units) {
          if (lane == 0) sl_last[victim] = iter;
          wave_sync();
          return false;
        }
      } else {
        victim = empty != kNone ? empty : lru;
        if (victim == kNone) return false;
      }
      victim = __builtin_amdgcn_readfirstlane(victim);
      evict_slot(victim);
      const uint32_t base = victim << cap_shift;
      const uint4* src = reinterpret_cast<const uint4*>(kp()->images + kp()->image_offset[tt]) + first * 4;
      for (uint32_t i = lane; i < units * 4; i += 64) stage_img[base * 4 + i] = src[i];
      for (uint32_t i = lane; i < units; i += 64) used[base + i] = 0;
      for (uint32_t i = lane; i < HSIZE; i += 64) hkey[i] = 0;
      wave_sync();
      if (lane == 0) {
        sl_t[victim] = tt;
        sl_units[victim] = units;
        sl_last[victim] = iter;
        sl_pred[victim] = predicted ? 1u : 0u;
        sl_first[victim] = first;
      }
      wave_sync();
      // insert every resident unit: unit i of the staging area is unit
      // i & (cap - 1) of slot i >> cap_shift
      for (uint32_t i = lane; i < (nslots << cap_shift); i += 64) {
        const uint32_t j = i >> cap_shift, u = i & ((1u << cap_shift) - 1);
        const uint32_t sj = sl_t[j];
        if (sj == kNone || u >= sl_units[j]) continue;
        const uint32_t s = kp()->image_slots[kp()->image_offset[sj] / 64 + sl_first[j] + u];
        if (s == kNone) continue;
        if (j == victim) {
          const uint32_t hd = slot_hash(s) & (EV - 1);
          if (dm_key[hd] == s + 1 && atomicCAS(&dm_key[hd], s + 1, 0u) == s + 1) too_late++;
        }
        uint32_t h = slot_hash(s) & (HSIZE - 1);
        for (uint32_t probe = 0; probe < HSIZE; ++probe) {
          if (atomicCAS(&hkey[h], 0u, s + 1) == 0u) {
            hpos[h] = (uint16_t)i;
            break;
          }
          h = (h + 1) & (HSIZE - 1);
        }
      }
      wave_sync();
      n_staged++;
      st_treelets++;
      st_bytes += (uint64_t)units * 64;
      if (predicted) st_pred++;
      return true;
    }
    return false;
  };

  // Pop the lane's next entry into `e`; false when its stacks are empty.
  auto pop_next = [&]() -> bool {
    if (MODE == 0 ? st_empty(sa) : (st_empty(sa) && st_empty(sb))) return false;
    if (MODE == 0) {
      e = st_pop<K, BS>(sa, ldsA, gA);
    } else if (!st_empty(sa)) {
      e = st_pop<K, BS>(sa, ldsA, gA);
    } else {
      phase2 = true;  // current_treelet_root := host address of the moved node (:1748-1754)
      e = st_pop<K, BS>(sb, ldsB, gB);
    }
    // every processed entry emits at least one record
    if (n_rec > kMaxStepsPerRay) err |= kErrSteps;
    return true;
  };

  // Write the hit record of the lane's completed ray (:2990-3038).  Deferred
  // to the wave's refill so the write path runs once per batch of completed
  // rays instead of in nearly every iteration.
  auto write_hit = [&]() {
    if constexpr (LM) {
      // the maximum over the final map; the entries are cleared for the next ray
      uint32_t mx = 0;
      for (uint32_t i = 0; i < lm_n; ++i) {
        const uint32_t h = lm_l[i];
        mx = max(mx, (uint32_t)lm_t[h]);
        lm_t[h] = 0;
      }
      lm_n = 0;
      hits_out[local_id].max_level = mx;
      if (err) {
        atomicOr(p.error, err);
        if (p.host_err) *reinterpret_cast<volatile uint32_t*>(p.host_err) = 1u;
      }
      return;
    }
    flush_slab_rest();
    uint4 o0 = make_uint4(0, 0, 0, 0), o1 = make_uint4(0, 0, 0, 0), o2 = make_uint4(0, 0, 0, 0);
    uint4 o4 = make_uint4(0xffffffffu, 0xffffffffu, 0, 0);
    if (err == 0 && min_thit < rtmax && closest_leaf != kNoSlot) {
      const Words16 q = load64B(A + ((uint64_t)closest_leaf << 6));
      const uint32_t inst_id = *reinterpret_cast<const uint32_t*>(A + ((uint64_t)closest_inst << 6) + 72);
      const F3 v0 = mk3(f_of(q, 4), f_of(q, 5), f_of(q, 6));
      const F3 v1 = mk3(f_of(q, 7), f_of(q, 8), f_of(q, 9));
      const F3 v2 = mk3(f_of(q, 10), f_of(q, 11), f_of(q, 12));
      const F3 ip = add3(ro, mk3(rd.x * min_thit, rd.y * min_thit, rd.z * min_thit));
      // the object-space ray of the closest hit's instance, recomputed from the
      // instance leaf (the same inputs as when the hit was accepted)
      const uint4* iq = reinterpret_cast<const uint4*>(A + ((uint64_t)closest_inst << 6) + 16);
      const uint4 ia = iq[0], ib = iq[1], ic = iq[2];
      const float im[12] = {__uint_as_float(ia.x), __uint_as_float(ia.y), __uint_as_float(ia.z), __uint_as_float(ia.w),
                            __uint_as_float(ib.x), __uint_as_float(ib.y), __uint_as_float(ib.z), __uint_as_float(ib.w),
                            __uint_as_float(ic.x), __uint_as_float(ic.y), __uint_as_float(ic.z), __uint_as_float(ic.w)};
      const ObjRay hr = transform_ray(im, ro, rd, rtmin, rtmax);
      const F3 op = add3(hr.o, mk3(hr.d.x * min_thit_obj, hr.d.y * min_thit_obj, hr.d.z * min_thit_obj));
      const F3 bc = barycentric(op, v0, v1, v2);
      o0 = make_uint4(1u, q.w[2], q.w[1] & 0x0fffffffu, inst_id);
      o1 = make_uint4(__float_as_uint(min_thit), __float_as_uint(ip.x), __float_as_uint(ip.y),
                      __float_as_uint(ip.z));
      o2 = make_uint4(__float_as_uint(bc.x), __float_as_uint(bc.y), __float_as_uint(bc.z), n_rec);
      o4.x = closest_inst << 6;
      o4.y = closest_inst >> 26;
    } else {
      o2.w = n_rec;
    }
    o4.z = accessed_bytes_of(n_rec, n_nodes, fetch_bytes);
    uint4* hq = reinterpret_cast<uint4*>(hits_out + local_id);
    hq[0] = o0;
    hq[1] = o1;
    hq[2] = o2;
    hq[3] = make_uint4(n_nodes, fetch_bytes, n_any, max_level);
    hq[4] = o4;
    if constexpr (REC == kRecSlab) p.rec_counts[local_id] = n_rec;
    if (err) {
      atomicOr(p.error, err);
      if (p.host_err) *reinterpret_cast<volatile uint32_t*>(p.host_err) = 1u;
    }
  };

  // the TLAS header of a single-scene launch, read once
  uint4 tl_h0 = make_uint4(0, 0, 0, 0), tl_h1 = make_uint4(0, 0, 0, 0);
  if (!MULTI && tlas + 64 <= bytes) {
    const uint4* hq = reinterpret_cast<const uint4*>(A + tlas);
    tl_h0 = hq[0];
    tl_h1 = hq[1];
  }
  while (true) {
    flush_slab_group();
    // ---- refill idle lanes from the global ray counter (wave converged) ----
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle > 0 && (nidle == 64u || nidle >= p.refill)) {
      if (pending) {
        write_hit();
        pending = false;
      }
      const uint32_t leader = (uint32_t)__ffsll((long long)idle) - 1u;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(p.ray_counter, nidle);
      base = __shfl(base, (int)leader, 64);
      if ((uint64_t)base + nidle >= p.n) exhausted = true;
      if (!active) {
        const uint32_t k = base + lanes_below(idle);
        const uint32_t r = k < p.n ? (MULTI ? k : claimed_ray(p, k)) : 0u;
        if (k < p.n && !MULTI && r >= p.n_rays) {
          // a listed ray outside the batch: reported, never read
          atomicOr(p.error, kErrRange);
          if (p.host_err) *reinterpret_cast<volatile uint32_t*>(p.host_err) = 1u;
        } else if (k < p.n) {
          active = true;
          local_id = r;
          if constexpr (MULTI) {
            // job holding ray r: job_base[j] <= r < job_base[j + 1]
            uint32_t lo = 0, hi = p.n_jobs;
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (job_base[mid] <= r)
                lo = mid;
              else
                hi = mid;
            }
            const BatchJob& jb = p.jobs[lo];
            A = jb.arena;
            bytes = jb.arena_bytes;
            tlas = jb.tlas;
            owner = jb.owner;
            first_treelet = jb.first_treelet;
            hits_out = jb.hits;
            rays_in = jb.rays;
            local_id = r - job_base[lo];
          }
          const uint4* rq = reinterpret_cast<const uint4*>(rays_in + local_id);
          uint4 r0 = rq[0], r1 = rq[1], r2 = rq[2];
          ro = mk3(__uint_as_float(r0.x), __uint_as_float(r0.y), __uint_as_float(r0.z));
          rtmin = __uint_as_float(r0.w);
          rd = mk3(__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z));
          rtmax = __uint_as_float(r1.w);
          rflags = r2.x;
          ridir = calc_idir(rd);
          ray_ok = slab_inputs_ok(ro, rd, ridir, rtmin, rtmax);
          octant = (rd.x < 0.f ? 1u : 0u) | (rd.y < 0.f ? 2u : 0u) | (rd.z < 0.f ? 4u : 0u);
          min_thit = rtmax;
          min_thit_obj = 0.f;
          closest_leaf = kNoSlot;
          closest_inst = kNoSlot;
          cur_inst = kNoSlot;
          n_rec = n_nodes = fetch_bytes = n_any = 0;
          rec_flushed = 0;
          max_level = 1;  // tree_level_map[topRootAddr] = 1
          if constexpr (REC == kRecCsr) rec_base = p.record_offsets[r];
          sa = Stack{0, 0, 0};
          sb = Stack{0, 0, 0};
          phase2 = false;
          err = 0;
          // TLAS header record and bounds test (vulkan_ray_tracing.cc:2444-2484)
          const uint64_t t = tlas;
          if (t + 64 > bytes) {
            err |= kErrRange;
          } else {
            emit(t, 64, RTX_TX_BVH_STRUCTURE, 0);
            uint4 h0, h1;
            if constexpr (MULTI) {
              const uint4* hq = reinterpret_cast<const uint4*>(A + t);
              h0 = hq[0];
              h1 = hq[1];
            } else {
              h0 = tl_h0;
              h1 = tl_h1;
            }
            const uint64_t root = t + ((uint64_t)h0.x | ((uint64_t)h0.y << 32));
            F3 lo = mk3(__uint_as_float(h0.z), __uint_as_float(h0.w), __uint_as_float(h1.x));
            F3 hi = mk3(__uint_as_float(h1.y), __uint_as_float(h1.z), __uint_as_float(h1.w));
            float th;
            // tree_level_map[topRootAddr] = 1, before the bounds test (:2470)
            if constexpr (LM) lm_set((uint32_t)(root >> 6), 1u);
            if (ray_box(lo, hi, ridir, ro, rtmin, rtmax, th)) {
              if ((root & 63) || root + 64 > bytes) {
                err |= kErrRange;
              } else {
                const uint32_t rs = (uint32_t)(root >> 6);
                push_entry(make_entry(rs, kFlTop, 1, 0), goes_current(rs));
              }
            }
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (exhausted) break;
      continue;
    }

    // ---- pop the lane's next entry, or complete the ray ---------------------
    if (active && has_e && err != 0) has_e = false;
    if (active && !has_e) {
      if (err == 0 && pop_next()) {
        has_e = true;
      } else {
        // ray complete; its hit record is written at the wave's next refill
        pending = true;
        active = false;
      }
    }

    const uint32_t slot = e_slot(e);
    const uint32_t flags = e_flags(e);
    const uint32_t level = e_level(e);
    const uint64_t off = (uint64_t)slot << 6;

    const bool step = has_e && err == 0;

    // ---- staging decision (wave converged) -----------------------------------
    uint32_t pos = kNone;
    if constexpr (STAGE > 0) {
      iter++;
      const bool want = has_e && err == 0;
      if (want) pos = staged_pos(slot);
      if (kp()->prefetch != RTX_PREFETCH_NONE) {
        const uint32_t nact = (uint32_t)__popcll(__ballot(want));
        const uint32_t nmiss = (uint32_t)__popcll(__ballot(want && pos == kNone));
        if (cooldown > 0) {
          cooldown--;
        } else if (nact > 0 && 2 * nmiss > nact) {
          // The RT unit's vote (src/gpgpu-sim/shader.cc:3424-3446): the
          // treelet of every lane's next node is counted, and the most
          // popular one wins, ties going to the lowest root offset (the
          // ordered map's first maximum; dense treelet indices follow root
          // offsets).  Each round takes one distinct treelet, so the loop
          // ends after at most 64 rounds.
          const uint32_t own = (want && off + 64 <= bytes) ? owner[slot] : kNone;
          uint64_t rem = __ballot(own != kNone);
          const uint32_t voters = (uint32_t)__popcll(rem);
          uint32_t best = kNone, bestc = 0;
          while (rem) {
            const uint32_t lead = (uint32_t)__ffsll((long long)rem) - 1u;
            const uint32_t cand = __shfl(own, (int)lead, 64);
            const uint64_t m = __ballot(own == cand);
            const uint32_t cnt = (uint32_t)__popcll(m);
            if (cnt > bestc || (cnt == bestc && cand < best)) {
              best = cand;
              bestc = cnt;
            }
            rem &= ~m;
          }
          // heuristics 0-3 (:3437-3559): always; share >= threshold; the
          // first round(n * share) units; the last round(n * share) units
          const float share = voters ? (float)bestc / (float)voters : 0.f;
          bool ok = best != kNone;
          if (kp()->prefetch == RTX_PREFETCH_POPULARITY && share < kp()->threshold) ok = false;
          bool staged = false;
          if (ok) {
            uint32_t maxu = UNITS;
            const bool part = kp()->prefetch == RTX_PREFETCH_PARTIAL || kp()->prefetch == RTX_PREFETCH_LATTER;
            if (part) maxu = (uint32_t)((float)(kp()->image_bytes[best] / 64) * share + 0.5f);
            // a treelet already resident is not staged again (the "same root as
            // the last prefetch" check of :3564)
            staged = stage(best, maxu, false, kp()->prefetch == RTX_PREFETCH_LATTER);
          }
          if (!staged) cooldown = 8;
          if (kp()->prefetch == RTX_PREFETCH_LEARNED) {
            // Each lane's predicted next treelets for (its treelet, its
            // octant).  A predicted treelet (its first pred_fraction of units)
            // is staged when the lanes' summed probability reaches min_score,
            // at most two per decision.
            uint4 row = make_uint4(0, 0, 0, 0);
            if (own != kNone && kp()->pred_table && own < kp()->pred_treelets)
              row = kp()->pred_table[(size_t)own * RTX_OCTANTS + octant];
            uint32_t ent[4] = {row.x, row.y, row.z, row.w};
            if (__ballot(row.x != 0) == 0) st_fallback++;
            uint32_t n_new = 0;
            for (int round = 0; round < 8 && n_new < 2; ++round) {
              uint32_t mine = 0;
              for (int k = 3; k >= 0; --k)
                if (ent[k] != 0) mine = ent[k];
              const uint64_t has = __ballot(mine != 0);
              if (has == 0) break;
              const uint32_t lead = (uint32_t)__ffsll((long long)has) - 1u;
              const uint32_t cand = __shfl(mine, (int)lead, 64) & 0xffffffu;
              float sc = 0.f;
              for (int k = 0; k < 4; ++k)
                if (ent[k] != 0 && (ent[k] & 0xffffffu) == cand) {
                  sc += (float)(ent[k] >> 24) * (1.f / 255.f);
                  ent[k] = 0;
                }
              for (int o = 32; o > 0; o >>= 1) sc += __shfl_xor(sc, o, 64);
              const uint32_t ct = cand - 1u;
              if (sc >= kp()->min_score && ct < kp()->pred_treelets) {
                const uint32_t units = kp()->image_bytes[ct] / 64;
                const uint32_t part = max(1u, (uint32_t)((float)units * kp()->pred_fraction + 0.5f));
                n_new += stage(ct, part, true) ? 1u : 0u;
              }
            }
          }
          if (want) pos = staged_pos(slot);
        }
      }
    }

    // ---- process the entry ----------------------------------------------------
    if (step) {
      has_e = false;
      if (MODE == 1 && !(flags & kFlTop) && e_inst(e, tlas_slot()) != cur_inst) load_instance(e_inst(e, tlas_slot()));
      // lanes on internal nodes (their own entry, or one reached through a
      // leaf in this iteration) share one internal-node pass at the end
      bool inner = err == 0 && !(flags & kFlLeaf);
      if (err != 0 || !(flags & kFlLeaf)) {
        // internal node: below; an error completes the ray next iteration
      } else if (flags & kFlTop) {
        // ---- instance leaf (:2602-2668 / :1872-1955) ----------------------
        if (off + 128 > bytes) {
          err |= kErrRange;
        } else {
          // a 128-byte leaf is served from LDS only when both of its units are staged
          uint32_t ipos = kNone;
          if constexpr (STAGE > 0)
            if (pos != kNone && (pos & ((1u << cap_shift) - 1)) + 1 < sl_units[pos >> cap_shift]) ipos = pos;
          uint4 w4[4];
          read_words(off, ipos, 16, w4, 4);
          // bytes 16..79 span both units of the leaf
          if constexpr (STAGE > 0)
            if (ipos != kNone) used[ipos + 1] = 1;
          const uint4 a = w4[0], b = w4[1], c = w4[2], d = w4[3];
          float m[12] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w),
                         __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w),
                         __uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), __uint_as_float(c.w)};
          const uint64_t bvh = (uint64_t)d.x | ((uint64_t)d.y << 32);
          emit(off, 128, RTX_TX_BVH_INSTANCE_LEAF, level);
          n_nodes++;
          const uint64_t blas = off + bvh;
          if ((blas & 63) || blas + 64 > bytes) {
            err |= kErrRange;
          } else {
            const uint2 ro2 = *reinterpret_cast<const uint2*>(A + blas);
            emit(blas, 64, RTX_TX_BVH_STRUCTURE, level);
            cur = transform_ray(m, ro, rd, rtmin, rtmax);
            cur_ok = slab_inputs_ok(cur.o, cur.d, cur.idir, cur.tmin, cur.tmax);
            cur_inst = slot;
            uint32_t irel = 0;
            if (MODE == 1 && !inst_rel_of(slot, tlas_slot(), irel)) err |= kErrInstSlot;
            const uint64_t root = blas + ((uint64_t)ro2.x | ((uint64_t)ro2.y << 32));
            if ((root & 63) || root + 64 > bytes) {
              err |= kErrRange;
            } else {
              const uint32_t rs = (uint32_t)(root >> 6);
              if (level > max_level) max_level = level;
              if constexpr (LM) lm_set(rs, lm_get(slot));  // :2668 / :1944
              const bool to_a = MODE == 0 || (!phase2 && (first_bits(slot) & 1u));
              // The BLAS root is popped next when it goes on top of the list
              // the next pop takes: always in plain order; in treelet order
              // when it joins the current list, or the other list while the
              // current one is empty (that pop starts phase 2).  Such a root is
              // tested in this iteration's internal-node pass instead.
              if (err == 0 && (MODE == 0 || to_a || st_empty(sa))) {
                if (MODE == 1 && !to_a) phase2 = true;  // current_treelet_root := host address (:1748-1754)
                e = make_entry(rs, 0u, level, irel);
                inner = true;
              } else {
                push_entry(make_entry(rs, 0u, level, irel), to_a);
              }
            }
          }
        }
      } else {
        // ---- primitive leaf (:2789-2985 / :2073-2203) ---------------------
        if (off + 64 > bytes) {
          err |= kErrRange;
        } else {
          emit(off, 8, RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR, level);
          const Words16 q = read_node(off, pos);
          if (((q.w[1] >> 29) & 1u) == 0u) {
            if ((q.w[3] & 0x1ffffu) != 0u) err |= kErrType;
            const F3 v0 = mk3(f_of(q, 4), f_of(q, 5), f_of(q, 6));
            const F3 v1 = mk3(f_of(q, 7), f_of(q, 8), f_of(q, 9));
            const F3 v2 = mk3(f_of(q, 10), f_of(q, 11), f_of(q, 12));
            float th = 0.f;
            const bool h = mt_test(v0, v1, v2, cur.o, cur.d, th);
            const float wt = th / cur.tmult;
            bool accept;
            if (MODE == 0) {
              accept = h && rtmin <= wt && wt <= rtmax;
              if (accept && (rflags & RTX_RAY_FLAG_OPAQUE) && wt < min_thit) min_thit = th / cur.tmult;
            } else {
              accept = h && rtmin <= wt && wt <= rtmax && wt < min_thit;
              if (accept) min_thit = th / cur.tmult;
            }
            if (accept) {
              min_thit_obj = th;
              closest_leaf = slot;
              closest_inst = cur_inst;
              emit(off, 64, RTX_TX_BVH_QUAD_LEAF_HIT, level);
              n_nodes++;
              if (MODE == 0 && !(rflags & RTX_RAY_FLAG_OPAQUE)) n_any++;
              if (rflags & RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT) {
                sa = Stack{0, 0, 0};
                sb = Stack{0, 0, 0};
              }
            } else {
              emit(off, 64, RTX_TX_BVH_QUAD_LEAF, level);
              n_nodes++;
            }
          } else {
            emit(off, 64, RTX_TX_BVH_PROCEDURAL_LEAF, level);
            n_nodes++;
          }
        }
        // the lane pops its next entry now; an internal node is tested in this
        // iteration's internal-node pass (in staged kernels its read still
        // probes the staged treelets; it only misses the iteration's vote)
        if (err == 0 && pop_next()) {
          if (err == 0 && !(e_flags(e) & kFlLeaf))
            inner = true;
          else
            has_e = true;
        }
      }
      if (inner) {
        const uint32_t slot = e_slot(e);
        const uint32_t flags = e_flags(e);
        const uint32_t level = e_level(e);
        const uint64_t off = (uint64_t)slot << 6;
        if (MODE == 1 && !(flags & kFlTop) && e_inst(e, tlas_slot()) != cur_inst) load_instance(e_inst(e, tlas_slot()));
        if (err == 0) {
          // ---- internal node (top :2499-2599 / bottom :2686-2786) ----------
          if (off + 64 > bytes) {
            err |= kErrRange;
          } else {
            const Words16 nd = read_node(off, STAGE > 0 ? staged_pos(slot) : pos);
            // Treelet order before the first switch: which children the first
            // treelet owns (read beside the node, used after its tests)
            const uint32_t fbits = (MODE == 1 && !phase2) ? first_bits(slot) : 0u;
            emit(off, 64, RTX_TX_BVH_INTERNAL_NODE, level);
            n_nodes++;
            const bool top = (flags & kFlTop) != 0;
            const int ex = (int)(int8_t)byte_of(nd, 18);
            const int ey = (int)(int8_t)byte_of(nd, 19);
            const int ez = (int)(int8_t)byte_of(nd, 20);
            F3 o, idir;
            float tmin, tmax, cull;
            if (top) {
              o = ro;
              idir = ridir;
              tmin = rtmin;
              tmax = rtmax;
              cull = min_thit;
            } else {
              o = cur.o;
              idir = cur.idir;
              tmin = cur.tmin;
              tmax = cur.tmax;
              cull = min_thit * cur.tmult;
            }
            const uint64_t first_child = off + (uint64_t)((int64_t)(int32_t)nd.w[3] * 64);
            const F3 sc = child_scales(ex, ey, ez);
            uint32_t hitmask = 0;
            // When no lane can produce a NaN (finite ray, non-NaN node origin)
            // and the arena's child boxes all have lo <= hi, each axis's near
            // and far bound is picked by the sign of the inverse direction and
            // the slab test needs only the max of the near and the min of the
            // far distances: for such boxes min(lo', hi') and max(lo', hi')
            // are exactly the picked products (rounding is monotonic), and the
            // v_min/v_max forms differ from the reference's compare-selects
            // only in the sign of a zero, which only enters comparisons.
            // Otherwise the reference's compare-selects run.
            const bool fast = (top ? ray_ok : cur_ok) && !__builtin_isnan(f_of(nd, 0) + f_of(nd, 1) + f_of(nd, 2)) &&
                              cull != -__builtin_inff();
            auto test_children = [&](auto fast_tag) {
              constexpr bool FAST = decltype(fast_tag)::value;
              if constexpr (FAST) {
                // near / far bound bytes per axis chosen by the ray's direction sign
                uint32_t nw[3][2], fw[3][2];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                  const int lw = 7 + 3 * a;
                  const float d = a == 0 ? idir.x : a == 1 ? idir.y : idir.z;
                  const bool neg = d < 0.0f;
                  const uint32_t l0 = nd.w[lw], l1 = nd.w[lw + 1];
                  const uint32_t u0 = __builtin_amdgcn_alignbyte(nd.w[lw + 2], nd.w[lw + 1], 2), u1 = nd.w[lw + 2] >> 16;
                  nw[a][0] = neg ? u0 : l0;
                  nw[a][1] = neg ? u1 : l1;
                  fw[a][0] = neg ? l0 : u0;
                  fw[a][1] = neg ? l1 : u1;
                }
                const float ox = f_of(nd, 0), oy = f_of(nd, 1), oz = f_of(nd, 2);
                // The cull test folded into the far bound: for a non-NaN tn
                // and a cull above -inf, !(tn >= cull) is tn < cull, which is
                // tn <= the largest float below cull; a NaN cull culls
                // nothing.  (A cull of -inf takes the compare-select path.)
                float tfc = tmax;
                if (!__builtin_isnan(cull)) {
                  const uint32_t cu = __float_as_uint(cull);
                  const uint32_t below = (int32_t)cu > 0 ? cu - 1u : ((cu << 1) == 0u ? 0x80000001u : cu + 1u);
                  tfc = __builtin_fminf(tmax, __uint_as_float(below));
                }
                // every child slot is tested (no branch per child) and the
                // empty slots are masked out: with a finite ray and a non-NaN
                // node origin an empty slot's bytes give finite or infinite
                // distances, never a NaN, and its result is discarded
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                  const uint32_t info = byte_of(nd, 22 + i) & 0x3fu;
                  const int wi = i >> 2, sh = 8 * (i & 3);
                  const float nx = __builtin_fmaf((float)((nw[0][wi] >> sh) & 0xffu), sc.x, ox);
                  const float ny = __builtin_fmaf((float)((nw[1][wi] >> sh) & 0xffu), sc.y, oy);
                  const float nz = __builtin_fmaf((float)((nw[2][wi] >> sh) & 0xffu), sc.z, oz);
                  const float fx = __builtin_fmaf((float)((fw[0][wi] >> sh) & 0xffu), sc.x, ox);
                  const float fy = __builtin_fmaf((float)((fw[1][wi] >> sh) & 0xffu), sc.y, oy);
                  const float fz = __builtin_fmaf((float)((fw[2][wi] >> sh) & 0xffu), sc.z, oz);
                  const float tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fmaxf((nx - o.x) * idir.x, (ny - o.y) * idir.y),
                                                                   (nz - o.z) * idir.z), tmin);
                  const float tf = __builtin_fminf(__builtin_fminf(__builtin_fminf((fx - o.x) * idir.x, (fy - o.y) * idir.y),
                                                                   (fz - o.z) * idir.z), tfc);
                  const bool h = (info & 3u) != 0u && tn <= tf;
                  hitmask |= h ? 1u << i : 0u;
                }
              } else {
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                  const uint32_t info = byte_of(nd, 22 + i) & 0x3fu;
                  const uint32_t size = info & 3u;
                  if (size > 0) {
                    F3 lo, hi;
                    child_box(nd, i, sc, lo, hi);
                    float th;
                    bool h = ray_box(lo, hi, idir, o, tmin, tmax, th);
                    if (h && th >= cull) h = false;
                    if (h) hitmask |= 1u << i;
                  }
                }
              }
            };
            if (__ballot(!(fast && p.ordered_boxes)) == 0)
              test_children(std::true_type{});
            else
              test_children(std::false_type{});
            if (hitmask) {
              const uint32_t lvl = level + 1;
              if (lvl > max_level) max_level = lvl;
              // the node's map value when it is processed (:2569-2581, :2756-2767,
              // :1837-1858, :2034-2055)
              const uint32_t lm_child = LM ? lm_get(slot) + 1 : 0u;
              const uint32_t inst = top ? 0u : e_inst_rel(e);
              const uint32_t base_flags = top ? kFlTop : 0u;
              bool pushed = false;
              if constexpr (MODE == 0 && !LM && K >= 8) {
                // Plain order without the level map, for a node whose children
                // lie in the arena and (under the TLAS) are internal nodes or
                // instance leaves: the pushes below give the same stack as one
                // push per child.  The ring's oldest entries are spilled first,
                // as many as pushing one by one would spill (up to six new
                // entries never spill each other when K >= 8), then each hit
                // child is stored at its place: children in slot order, the
                // first internal one last.
                const uint32_t iw0 = __builtin_amdgcn_alignbyte(nd.w[6], nd.w[5], 2);  // info bytes of children 0-3
                const uint32_t iw1 = nd.w[6] >> 16;                                    // children 4-5
                uint32_t pre[6], tot = 0, leafm = 0;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                  const uint32_t info = (i < 4 ? iw0 >> (8 * i) : iw1 >> (8 * (i - 4))) & 0x3fu;
                  pre[i] = tot;
                  tot += info & 3u;
                  leafm |= (info >> 2) != 0u ? 1u << i : 0u;
                }
                const uint32_t n_push = (uint32_t)__builtin_popcount(hitmask);
                const uint32_t need = sa.cnt + n_push;
                const uint32_t sp = need > (uint32_t)K ? need - (uint32_t)K : 0u;
                const bool ok = first_child < bytes && first_child + (uint64_t)tot * 64 <= bytes &&
                                !(top && ((iw0 & 0x38383838u) | (iw1 & 0x3838u)) != 0u) && sa.g + sp <= p.spill_cap;
                if (ok) {
                  pushed = true;
                  for (uint32_t j = 0; j < sp; ++j) {
                    gA[sa.g++] = ldsA[sa.bot * BS];
                    sa.bot = (sa.bot + 1) & (K - 1);
                    sa.cnt--;
                  }
                  const uint32_t im = hitmask & ~leafm;
                  const uint32_t dbit = im & (0u - im);  // the first internal hit child
                  const uint32_t pm = hitmask & ~dbit;
                  const uint32_t first_slot = (uint32_t)(first_child >> 6);
                  const uint32_t hi = base_flags | ((lvl > 255u ? 255u : lvl) << 2) | ((inst & 0x3fffffu) << 10);
                  const uint32_t top_pos = sa.bot + sa.cnt;
#pragma unroll
                  for (int i = 0; i < 6; ++i) {
                    if ((hitmask >> i) & 1u) {
                      // the deferred child takes the last place
                      const uint32_t r = ((dbit >> i) & 1u) ? n_push - 1u
                                                             : (uint32_t)__builtin_popcount(pm & ((1u << i) - 1u));
                      ldsA[((top_pos + r) & (K - 1)) * BS] =
                          (uint64_t)(first_slot + pre[i]) | ((uint64_t)(hi | (((leafm >> i) & 1u) ? kFlLeaf : 0u)) << 32);
                    }
                  }
                  sa.cnt += n_push;
                }
              } else if constexpr (MODE == 1 && !LM) {
                // Treelet order, same conditions: the children go in slot
                // order, a child owned by the first treelet to the current
                // list (A), the others to list B.  Each list first spills its
                // ring's oldest entries, as many as pushing one by one would
                // (at most K new entries per list, so new ones never spill
                // each other), then stores each child at its place.
                const uint32_t iw0 = __builtin_amdgcn_alignbyte(nd.w[6], nd.w[5], 2);
                const uint32_t iw1 = nd.w[6] >> 16;
                uint32_t pre[6], tot = 0, leafm = 0;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                  const uint32_t info = (i < 4 ? iw0 >> (8 * i) : iw1 >> (8 * (i - 4))) & 0x3fu;
                  pre[i] = tot;
                  tot += info & 3u;
                  leafm |= (info >> 2) != 0u ? 1u << i : 0u;
                }
                const uint32_t pmA = hitmask & fbits, pmB = hitmask & ~fbits;
                const uint32_t nA = (uint32_t)__builtin_popcount(pmA), nB = (uint32_t)__builtin_popcount(pmB);
                const uint32_t spA = sa.cnt + nA > (uint32_t)K ? sa.cnt + nA - (uint32_t)K : 0u;
                const uint32_t spB = sb.cnt + nB > (uint32_t)K ? sb.cnt + nB - (uint32_t)K : 0u;
                const bool ok = first_child < bytes && first_child + (uint64_t)tot * 64 <= bytes &&
                                !(top && ((iw0 & 0x38383838u) | (iw1 & 0x3838u)) != 0u) && nA <= (uint32_t)K &&
                                nB <= (uint32_t)K && sa.g + spA <= p.spill_cap && sb.g + spB <= p.spill_cap;
                if (ok) {
                  pushed = true;
                  const uint32_t first_slot = (uint32_t)(first_child >> 6);
                  const uint32_t hi = base_flags | ((lvl > 255u ? 255u : lvl) << 2) | ((inst & 0x3fffffu) << 10);
                  auto spill_list = [&](Stack& st, lds_u64* lds, glb_u64* g, uint32_t sp) {
                    for (uint32_t j = 0; j < sp; ++j) {
                      g[st.g++] = lds[st.bot * BS];
                      st.bot = (st.bot + 1) & (K - 1);
                      st.cnt--;
                    }
                  };
                  spill_list(sa, ldsA, gA, spA);
                  spill_list(sb, ldsB, gB, spB);
                  // one store per hit child, into the list it goes to
                  const uint32_t topA = sa.bot + sa.cnt, topB = sb.bot + sb.cnt;
#pragma unroll
                  for (int i = 0; i < 6; ++i) {
                    if ((hitmask >> i) & 1u) {
                      const bool a = (fbits >> i) & 1u;
                      const uint32_t r = (uint32_t)__builtin_popcount((a ? pmA : pmB) & ((1u << i) - 1u));
                      (a ? ldsA : ldsB)[(((a ? topA : topB) + r) & (K - 1)) * BS] =
                          (uint64_t)(first_slot + pre[i]) | ((uint64_t)(hi | (((leafm >> i) & 1u) ? kFlLeaf : 0u)) << 32);
                    }
                  }
                  sa.cnt += nA;
                  sb.cnt += nB;
                }
              }
              int deferred = -1;
              uint32_t deferred_cs = 0;
              const uint32_t cur_bits = fbits;
              uint64_t child = first_child;
#pragma unroll
              for (int i = 0; i < 6 && !pushed; ++i) {
                const uint32_t info = byte_of(nd, 22 + i) & 0x3fu;
                const uint64_t c = child;
                child += (uint64_t)(info & 3u) * 64;
                if (!(hitmask & (1u << i))) continue;
                const uint32_t ctype = info >> 2;
                const bool leaf = ctype != 0;
                if (top && leaf && ctype != 1u) err |= kErrType;
                if ((c & 63) || c + 64 > bytes) {
                  err |= kErrRange;
                  continue;
                }
                const uint32_t cs = (uint32_t)(c >> 6);
                if constexpr (LM) lm_set(cs, lm_child);
                if (MODE == 0 && !leaf && deferred < 0) {
                  deferred = i;
                  deferred_cs = cs;
                  continue;
                }
                push_entry(make_entry(cs, base_flags | (leaf ? kFlLeaf : 0u), lvl, inst), MODE == 0 || ((cur_bits >> i) & 1u));
              }
              if (MODE == 0 && deferred >= 0 && !pushed) {
                push_entry(make_entry(deferred_cs, base_flags, lvl, inst), true);
              }
            }
          }
        }
      }
    }
  }

  if (pending) write_hit();

  if constexpr (STAGE > 0) {
    evict_all();
    const uint32_t r = wave_sum(reads), hl = wave_sum(lds_hits);
    const uint32_t te = wave_sum(too_early), tl = wave_sum(too_late);
    if (lane == 0) {
      atomicAdd(&p.stats[8], (unsigned long long)st_timely);
      atomicAdd(&p.stats[9], (unsigned long long)te * 64ull);
      atomicAdd(&p.stats[10], (unsigned long long)tl * 64ull);
      atomicAdd(&p.stats[0], (unsigned long long)r);
      atomicAdd(&p.stats[1], (unsigned long long)hl);
      atomicAdd(&p.stats[2], (unsigned long long)st_treelets);
      atomicAdd(&p.stats[3], (unsigned long long)st_bytes);
      atomicAdd(&p.stats[4], (unsigned long long)st_wasted);
      atomicAdd(&p.stats[5], (unsigned long long)st_pred);
      atomicAdd(&p.stats[6], (unsigned long long)st_pred_hits);
      atomicAdd(&p.stats[7], (unsigned long long)st_fallback);
    }
  }
}

// ---- record offsets: exclusive scan of hits[i].n_records --------------------
constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanItems = 4;
constexpr uint32_t kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* smem, uint64_t& total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63) smem[wave] = x;
  __syncthreads();
  uint64_t wsum = 0;
  for (uint32_t w = 0; w < wave; ++w) wsum += smem[w];
  total = smem[0] + smem[1] + smem[2] + smem[3];
  __syncthreads();
  return wsum + x - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_tiles(const rtx_hit* hits, const uint32_t* counts, uint32_t n,
                                                         uint64_t* out, uint64_t* tile_sums) {
  __shared__ uint64_t smem[4];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t s = 0;
  if (counts && base + kScanItems <= n) {
    const uint4 c = *reinterpret_cast<const uint4*>(counts + base);  // base is a multiple of 4
    v[0] = c.x;
    v[1] = c.y;
    v[2] = c.z;
    v[3] = c.w;
    s = v[0] + v[1] + v[2] + v[3];
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
      v[k] = base + k < n ? (counts ? counts[base + k] : hits[base + k].n_records) : 0;
      s += v[k];
    }
  }
  uint64_t total;
  uint64_t ex = block_exclusive_scan(s, smem, total);
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_sums(uint64_t* tile_sums, uint32_t n_tiles, uint64_t* out,
                                                        uint32_t n) {
  __shared__ uint64_t smem[4];
  uint64_t carry = 0;
  for (uint32_t b = 0; b < n_tiles; b += kScanBlock) {
    const uint32_t i = b + threadIdx.x;
    uint64_t v = i < n_tiles ? tile_sums[i] : 0;
    uint64_t total;
    uint64_t ex = block_exclusive_scan(v, smem, total);
    if (i < n_tiles) tile_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_add(uint64_t* out, uint32_t n, const uint64_t* tile_sums) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  const uint64_t add = tile_sums[blockIdx.x];
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += add;
}

// ---- cooperative traversal for small batches (both orders) -----------------
//
// One ray per group of kWideLanes lanes.  Every lane of a group holds the same
// ray state and takes the same branches; an internal node's six child boxes
// are tested by lanes 0-5 of the group at once, so a traversal step issues a
// fraction of the instructions of trace_kernel, whose lanes each test all six.
// A batch too small to fill the GPU is bound by the step latency of its
// longest rays, which this shortens.  The per-ray sequence of popped entries,
// and so every hit record and counter, is that of trace_kernel<MODE, false, K,
// 0>: each list (one in plain order, two in treelet order) is an LDS array of
// `cap` = K + spill_cap entries per ray, and a push fails exactly when the
// per-lane kernel's LDS ring and spill area for that list are both full.
// Lane j pushes hit child j itself, at its place in that push sequence.
constexpr uint32_t kWideLanes = 8;
constexpr uint32_t kWideCap = 144;  // K + spill_cap for K <= 16
constexpr uint32_t kWideBlock = 256;

// Byte `base + j` (j in 0..5) of a node held in registers, for a lane-varying j:
// it lies in word base / 4 or the next one.
__device__ __forceinline__ uint32_t byte_at(const Words16& n, int base, uint32_t j) {
  const uint32_t b = (uint32_t)base + j;
  const uint32_t w = (b >> 2) == (uint32_t)(base >> 2) ? n.w[base >> 2] : n.w[(base >> 2) + 1];
  return (w >> ((b & 3u) * 8u)) & 0xffu;
}

template <int MODE>
__global__ __launch_bounds__(kWideBlock) void trace_wide(TraceParams p, uint32_t cap) {
  constexpr uint32_t NS = MODE == 1 ? 2 : 1;
  __shared__ uint64_t stk[kWideBlock / kWideLanes][NS * kWideCap];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t j = tid & (kWideLanes - 1);
  const uint32_t gbase = lane & ~(kWideLanes - 1);  // the group's first lane in the wave
  lds_u64* sta = (lds_u64*)&stk[tid / kWideLanes][0];
  lds_u64* stb = (lds_u64*)&stk[tid / kWideLanes][(NS - 1) * kWideCap];
  const uint8_t* __restrict__ A = p.arena;
  const uint64_t bytes = p.arena_bytes;
  const uint64_t tlas = p.tlas;
  const uint32_t* owner = p.owner;
  const uint32_t first_treelet = p.first_treelet;

  bool active = false, exhausted = false, pending = false;
  bool phase2 = false;
  uint32_t da = 0, db = 0, err = 0, local_id = 0;
  F3 ro{}, rd{}, ridir{};
  float rtmin = 0.f, rtmax = 0.f;
  uint32_t rflags = 0;
  float min_thit = 0.f, min_thit_obj = 0.f;
  uint32_t closest_leaf = kNoSlot, closest_inst = kNoSlot;
  ObjRay cur{};
  uint32_t cur_inst = kNoSlot;
  bool ray_ok = false, cur_ok = false;
  uint32_t n_rec = 0, n_nodes = 0, fetch_bytes = 0, n_any = 0, max_level = 0;
  bool has_e = false;
  uint64_t e = 0;

  auto emit = [&](uint32_t size, uint32_t type) {
    if (type != RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR) fetch_bytes += size;
    n_rec++;
  };
  // lane 0 of the group writes; the group's later pops read it back in order
  // (the list depths are updated through selects, not a reference to one of
  // them, which would put both in scratch memory)
  auto push = [&](uint64_t v, bool to_a) {
    const bool a = MODE == 0 || to_a;
    const uint32_t d = a ? da : db;
    if (d >= cap) {
      err |= kErrStack;
      return;
    }
    if (j == 0) *(a ? sta + d : stb + d) = v;
    da = a ? d + 1 : da;
    db = a ? db : d + 1;
  };
  auto pop_next = [&]() -> bool {
    if (da == 0 && db == 0) return false;
    const bool a = MODE == 0 || da > 0;
    // moving to the other list: current_treelet_root := host address of the moved node (:1748-1754)
    if (!a) phase2 = true;
    const uint32_t d = (a ? da : db) - 1;
    e = *(a ? sta + d : stb + d);
    da = a ? d : da;
    db = a ? db : d;
    if (n_rec > kMaxStepsPerRay) err |= kErrSteps;
    return true;
  };
  auto tlas_slot = [&]() -> uint32_t { return (uint32_t)(tlas >> 6); };
  auto goes_current = [&](uint32_t slot) -> bool {
    if (MODE == 0) return true;
    if (phase2) return false;
    return owner[slot] == first_treelet;
  };
  // first-treelet child bits of a node slot (one byte per slot after the owner words)
  auto first_bits = [&](uint32_t slot) -> uint32_t {
    return reinterpret_cast<const uint8_t*>(owner + (bytes >> 6))[slot];
  };
  auto load_instance = [&](uint32_t inst) {
    const uint64_t ioff = (uint64_t)inst << 6;
    if (ioff + 128 > bytes) {
      err |= kErrRange;
      return;
    }
    const uint4* q = reinterpret_cast<const uint4*>(A + ioff + 16);
    const uint4 a = q[0], b = q[1], c = q[2];
    const float m[12] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w),
                         __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w),
                         __uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), __uint_as_float(c.w)};
    cur = transform_ray(m, ro, rd, rtmin, rtmax);
    cur_ok = slab_inputs_ok(cur.o, cur.d, cur.idir, cur.tmin, cur.tmax);
    cur_inst = inst;
  };
  // the hit record (as trace_kernel writes it): lanes 0-3 store one 16-byte word each
  auto write_hit = [&]() {
    uint4 o0 = make_uint4(0, 0, 0, 0), o1 = make_uint4(0, 0, 0, 0), o2 = make_uint4(0, 0, 0, 0);
    uint4 o4 = make_uint4(0xffffffffu, 0xffffffffu, 0, 0);
    if (err == 0 && min_thit < rtmax && closest_leaf != kNoSlot) {
      const Words16 q = load64B(A + ((uint64_t)closest_leaf << 6));
      const uint32_t inst_id = *reinterpret_cast<const uint32_t*>(A + ((uint64_t)closest_inst << 6) + 72);
      const F3 v0 = mk3(f_of(q, 4), f_of(q, 5), f_of(q, 6));
      const F3 v1 = mk3(f_of(q, 7), f_of(q, 8), f_of(q, 9));
      const F3 v2 = mk3(f_of(q, 10), f_of(q, 11), f_of(q, 12));
      const F3 ip = add3(ro, mk3(rd.x * min_thit, rd.y * min_thit, rd.z * min_thit));
      const uint4* iq = reinterpret_cast<const uint4*>(A + ((uint64_t)closest_inst << 6) + 16);
      const uint4 ia = iq[0], ib = iq[1], ic = iq[2];
      const float im[12] = {__uint_as_float(ia.x), __uint_as_float(ia.y), __uint_as_float(ia.z), __uint_as_float(ia.w),
                            __uint_as_float(ib.x), __uint_as_float(ib.y), __uint_as_float(ib.z), __uint_as_float(ib.w),
                            __uint_as_float(ic.x), __uint_as_float(ic.y), __uint_as_float(ic.z), __uint_as_float(ic.w)};
      const ObjRay hr = transform_ray(im, ro, rd, rtmin, rtmax);
      const F3 op = add3(hr.o, mk3(hr.d.x * min_thit_obj, hr.d.y * min_thit_obj, hr.d.z * min_thit_obj));
      const F3 bc = barycentric(op, v0, v1, v2);
      o0 = make_uint4(1u, q.w[2], q.w[1] & 0x0fffffffu, inst_id);
      o1 = make_uint4(__float_as_uint(min_thit), __float_as_uint(ip.x), __float_as_uint(ip.y),
                      __float_as_uint(ip.z));
      o2 = make_uint4(__float_as_uint(bc.x), __float_as_uint(bc.y), __float_as_uint(bc.z), n_rec);
      o4.x = closest_inst << 6;
      o4.y = closest_inst >> 26;
    } else {
      o2.w = n_rec;
    }
    o4.z = accessed_bytes_of(n_rec, n_nodes, fetch_bytes);
    uint4* hq = reinterpret_cast<uint4*>(p.hits + local_id);
    if (j == 0)
      hq[0] = o0;
    else if (j == 1)
      hq[1] = o1;
    else if (j == 2)
      hq[2] = o2;
    else if (j == 3)
      hq[3] = make_uint4(n_nodes, fetch_bytes, n_any, max_level);
    else if (j == 4)
      hq[4] = o4;
    if (err && j == 0) {
      atomicOr(p.error, err);
      if (p.host_err) *reinterpret_cast<volatile uint32_t*>(p.host_err) = 1u;
    }
  };

  uint4 tl_h0 = make_uint4(0, 0, 0, 0), tl_h1 = make_uint4(0, 0, 0, 0);
  if (tlas + 64 <= bytes) {
    const uint4* hq = reinterpret_cast<const uint4*>(A + tlas);
    tl_h0 = hq[0];
    tl_h1 = hq[1];
  }
  const uint32_t groups = 64u / kWideLanes;
  while (true) {
    // ---- refill idle groups from the global ray counter (wave converged) ----
    const uint64_t idle = __ballot(j == 0 && !active);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle > 0 && (nidle == groups || nidle * kWideLanes >= p.refill)) {
      if (pending) {
        write_hit();
        pending = false;
      }
      const uint32_t leader = (uint32_t)__ffsll((long long)idle) - 1u;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(p.ray_counter, nidle);
      base = __shfl(base, (int)leader, 64);
      if ((uint64_t)base + nidle >= p.n) exhausted = true;
      if (!active) {
        const uint32_t k = base + (uint32_t)__popcll(idle & ((1ull << gbase) - 1ull));
        const uint32_t r = k < p.n ? claimed_ray(p, k) : 0u;
        if (k < p.n && r >= p.n_rays) {
          if (j == 0) {
            atomicOr(p.error, kErrRange);
            if (p.host_err) *reinterpret_cast<volatile uint32_t*>(p.host_err) = 1u;
          }
        } else if (k < p.n) {
          active = true;
          local_id = r;
          const uint4* rq = reinterpret_cast<const uint4*>(p.rays + r);
          const uint4 r0 = rq[0], r1 = rq[1], r2 = rq[2];
          ro = mk3(__uint_as_float(r0.x), __uint_as_float(r0.y), __uint_as_float(r0.z));
          rtmin = __uint_as_float(r0.w);
          rd = mk3(__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z));
          rtmax = __uint_as_float(r1.w);
          rflags = r2.x;
          ridir = calc_idir(rd);
          ray_ok = slab_inputs_ok(ro, rd, ridir, rtmin, rtmax);
          min_thit = rtmax;
          min_thit_obj = 0.f;
          closest_leaf = kNoSlot;
          closest_inst = kNoSlot;
          cur_inst = kNoSlot;
          n_rec = n_nodes = fetch_bytes = n_any = 0;
          max_level = 1;
          da = db = 0;
          phase2 = false;
          err = 0;
          has_e = false;
          if (tlas + 64 > bytes) {
            err |= kErrRange;
          } else {
            emit(64, RTX_TX_BVH_STRUCTURE);
            const uint64_t root = tlas + ((uint64_t)tl_h0.x | ((uint64_t)tl_h0.y << 32));
            const F3 lo = mk3(__uint_as_float(tl_h0.z), __uint_as_float(tl_h0.w), __uint_as_float(tl_h1.x));
            const F3 hi = mk3(__uint_as_float(tl_h1.y), __uint_as_float(tl_h1.z), __uint_as_float(tl_h1.w));
            float th;
            if (ray_box(lo, hi, ridir, ro, rtmin, rtmax, th)) {
              if ((root & 63) || root + 64 > bytes) {
                err |= kErrRange;
              } else {
                const uint32_t rs = (uint32_t)(root >> 6);
                push(make_entry(rs, kFlTop, 1, 0), goes_current(rs));
              }
            }
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (exhausted) break;
      continue;
    }

    // ---- pop the group's next entry, or complete the ray ---------------------
    if (active && has_e && err != 0) has_e = false;
    if (active && !has_e) {
      if (err == 0 && pop_next()) {
        has_e = true;
      } else {
        pending = true;
        active = false;
      }
    }
    const bool step = has_e && err == 0;
    const uint32_t slot = e_slot(e);
    const uint32_t flags = e_flags(e);
    const uint32_t level = e_level(e);
    const uint64_t off = (uint64_t)slot << 6;
    bool inner = false;
    if (step) {
      has_e = false;
      if (MODE == 1 && !(flags & kFlTop) && e_inst(e, tlas_slot()) != cur_inst) load_instance(e_inst(e, tlas_slot()));
      if (err != 0) {
        // an instance reload out of the arena completes the ray next iteration
      } else if (!(flags & kFlLeaf)) {
        inner = true;
      } else if (flags & kFlTop) {
        // ---- instance leaf: the BLAS root is the group's next entry ----------
        if (off + 128 > bytes) {
          err |= kErrRange;
        } else {
          const uint4* q4 = reinterpret_cast<const uint4*>(A + off + 16);
          const uint4 a = q4[0], b = q4[1], c = q4[2], d = q4[3];
          const float m[12] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w),
                               __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w),
                               __uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), __uint_as_float(c.w)};
          const uint64_t bvh = (uint64_t)d.x | ((uint64_t)d.y << 32);
          emit(128, RTX_TX_BVH_INSTANCE_LEAF);
          n_nodes++;
          const uint64_t blas = off + bvh;
          if ((blas & 63) || blas + 64 > bytes) {
            err |= kErrRange;
          } else {
            const uint2 ro2 = *reinterpret_cast<const uint2*>(A + blas);
            emit(64, RTX_TX_BVH_STRUCTURE);
            cur = transform_ray(m, ro, rd, rtmin, rtmax);
            cur_ok = slab_inputs_ok(cur.o, cur.d, cur.idir, cur.tmin, cur.tmax);
            cur_inst = slot;
            uint32_t irel = 0;
            if (MODE == 1 && !inst_rel_of(slot, tlas_slot(), irel)) err |= kErrInstSlot;
            const uint64_t root = blas + ((uint64_t)ro2.x | ((uint64_t)ro2.y << 32));
            if ((root & 63) || root + 64 > bytes) {
              err |= kErrRange;
            } else {
              if (level > max_level) max_level = level;
              const bool to_a = MODE == 0 || (!phase2 && (first_bits(slot) & 1u));
              // the root is the next entry when it goes on top of the list the
              // next pop takes (see trace_kernel); otherwise it is pushed
              if (err == 0 && (MODE == 0 || to_a || da == 0)) {
                if (MODE == 1 && !to_a) phase2 = true;
                e = make_entry((uint32_t)(root >> 6), 0u, level, irel);
                has_e = true;
              } else {
                push(make_entry((uint32_t)(root >> 6), 0u, level, irel), to_a);
              }
            }
          }
        }
      } else {
        // ---- primitive leaf ------------------------------------------------
        if (off + 64 > bytes) {
          err |= kErrRange;
        } else {
          emit(8, RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR);
          const Words16 q = load64B(A + off);
          if (((q.w[1] >> 29) & 1u) == 0u) {
            if ((q.w[3] & 0x1ffffu) != 0u) err |= kErrType;
            const F3 v0 = mk3(f_of(q, 4), f_of(q, 5), f_of(q, 6));
            const F3 v1 = mk3(f_of(q, 7), f_of(q, 8), f_of(q, 9));
            const F3 v2 = mk3(f_of(q, 10), f_of(q, 11), f_of(q, 12));
            float th = 0.f;
            const bool h = mt_test(v0, v1, v2, cur.o, cur.d, th);
            const float wt = th / cur.tmult;
            bool accept;
            if (MODE == 0) {
              accept = h && rtmin <= wt && wt <= rtmax;
              if (accept && (rflags & RTX_RAY_FLAG_OPAQUE) && wt < min_thit) min_thit = th / cur.tmult;
            } else {
              accept = h && rtmin <= wt && wt <= rtmax && wt < min_thit;
              if (accept) min_thit = th / cur.tmult;
            }
            if (accept) {
              min_thit_obj = th;
              closest_leaf = slot;
              closest_inst = cur_inst;
              emit(64, RTX_TX_BVH_QUAD_LEAF_HIT);
              n_nodes++;
              if (MODE == 0 && !(rflags & RTX_RAY_FLAG_OPAQUE)) n_any++;
              if (rflags & RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT) da = db = 0;
            } else {
              emit(64, RTX_TX_BVH_QUAD_LEAF);
              n_nodes++;
            }
          } else {
            emit(64, RTX_TX_BVH_PROCEDURAL_LEAF);
            n_nodes++;
          }
        }
      }
    }
    if (inner) {
      // ---- internal node: lane j of the group tests child j ------------------
      if (off + 64 > bytes) {
        err |= kErrRange;
      } else {
        const Words16 nd = load64B(A + off);
        // treelet order before the first switch: which children the first treelet owns
        const uint32_t cur_bits = (MODE == 1 && !phase2) ? first_bits(slot) : 0u;
        emit(64, RTX_TX_BVH_INTERNAL_NODE);
        n_nodes++;
        const bool top = (flags & kFlTop) != 0;
        const int ex = (int)(int8_t)byte_of(nd, 18);
        const int ey = (int)(int8_t)byte_of(nd, 19);
        const int ez = (int)(int8_t)byte_of(nd, 20);
        const F3 o = top ? ro : cur.o;
        const F3 idir = top ? ridir : cur.idir;
        const float tmin = top ? rtmin : cur.tmin;
        const float tmax = top ? rtmax : cur.tmax;
        const float cull = top ? min_thit : min_thit * cur.tmult;
        const bool fast = (top ? ray_ok : cur_ok) && !__builtin_isnan(f_of(nd, 0) + f_of(nd, 1) + f_of(nd, 2));
        bool h = false;
        if (j < 6u && (byte_at(nd, 22, j) & 3u) != 0u) {
          const F3 sc = child_scales(ex, ey, ez);
          const float ox = f_of(nd, 0), oy = f_of(nd, 1), oz = f_of(nd, 2);
          F3 lo, hi;
          lo.x = __builtin_fmaf((float)byte_at(nd, 28, j), sc.x, ox);
          lo.y = __builtin_fmaf((float)byte_at(nd, 40, j), sc.y, oy);
          lo.z = __builtin_fmaf((float)byte_at(nd, 52, j), sc.z, oz);
          hi.x = __builtin_fmaf((float)byte_at(nd, 34, j), sc.x, ox);
          hi.y = __builtin_fmaf((float)byte_at(nd, 46, j), sc.y, oy);
          hi.z = __builtin_fmaf((float)byte_at(nd, 58, j), sc.z, oz);
          float th;
          h = fast ? ray_box_fast(lo, hi, idir, o, tmin, tmax, th) : ray_box(lo, hi, idir, o, tmin, tmax, th);
          if (h && th >= cull) h = false;
        }
        const uint32_t hitmask = (uint32_t)(__ballot(h) >> gbase) & 0x3fu;
        if (hitmask) {
          const uint32_t lvl = level + 1;
          if (lvl > max_level) max_level = lvl;
          const uint32_t inst = top ? 0u : e_inst_rel(e);
          const uint32_t base_flags = top ? kFlTop : 0u;
          // Lane j pushes hit child j itself, at the position the per-lane
          // kernel's push sequence gives it: hit children in slot order (in
          // plain order the first internal one last), a list's pushes failing
          // from the one that would exceed its capacity on.
          uint64_t c = off + (uint64_t)((int64_t)(int32_t)nd.w[3] * 64);
#pragma unroll
          for (int i = 0; i < 5; ++i)
            if ((uint32_t)i < j) c += (uint64_t)(byte_of(nd, 22 + i) & 3u) * 64;
          const uint32_t ctype = j < 6u ? (byte_at(nd, 22, j) & 0x3fu) >> 2 : 0u;
          const bool leaf = ctype != 0;
          const bool mine = j < 6u && ((hitmask >> j) & 1u);
          const bool pushed = mine && !(c & 63) && c + 64 <= bytes;
          const bool to_a = MODE == 0 || ((cur_bits >> j) & 1u);
          const uint32_t m_push = (uint32_t)(__ballot(pushed) >> gbase) & 0x3fu;
          if ((__ballot(mine && top && leaf && ctype != 1u) >> gbase) & 0x3fu) err |= kErrType;
          if (hitmask & ~m_push) err |= kErrRange;  // hit children outside the arena
          const uint32_t below = (1u << j) - 1u;
          uint32_t pos;
          if (MODE == 0) {
            const uint32_t m_int = (uint32_t)(__ballot(pushed && !leaf) >> gbase) & 0x3fu;
            const uint32_t n = (uint32_t)__popc(m_push);
            const uint32_t d = m_int ? (uint32_t)__builtin_ctz(m_int) : 8u;  // deferred child, 8 = none
            const uint32_t rest = m_push & ~(d < 8u ? 1u << d : 0u);
            pos = j == d ? da + n - 1u : da + (uint32_t)__popc(rest & below);
            da = min(cap, da + n);
          } else {
            const uint32_t m_a = (uint32_t)(__ballot(pushed && to_a) >> gbase) & 0x3fu;
            const uint32_t m_b = m_push & ~m_a;
            pos = to_a ? da + (uint32_t)__popc(m_a & below) : db + (uint32_t)__popc(m_b & below);
            da = min(cap, da + (uint32_t)__popc(m_a));
            db = min(cap, db + (uint32_t)__popc(m_b));
          }
          const bool fits = pos < cap;
          if (pushed && fits)
            *(to_a ? sta + pos : stb + pos) =
                make_entry((uint32_t)(c >> 6), base_flags | (leaf ? kFlLeaf : 0u), lvl, inst);
          if ((__ballot(pushed && !fits) >> gbase) & 0x3fu) err |= kErrStack;
        }
      }
    }
  }
  if (pending) write_hit();
}

using TraceFn = void (*)(TraceParams);

template <int MODE, int REC>
TraceFn pick_k(uint32_t lds_stack, uint32_t stage) {
  if (stage == 0) {
    switch (lds_stack) {
      case 4: return trace_kernel<MODE, REC, 4, 0>;
      case 8: return trace_kernel<MODE, REC, 8, 0>;
      case 16: return trace_kernel<MODE, REC, 16, 0>;
      default: return nullptr;
    }
  }
  switch (stage) {
    case 4096: return trace_kernel<MODE, REC, 8, 4096>;
    case 8192: return trace_kernel<MODE, REC, 8, 8192>;
    case 16384: return trace_kernel<MODE, REC, 8, 16384>;
    case 32768: return trace_kernel<MODE, REC, 8, 32768>;
    case 65536: return trace_kernel<MODE, REC, 8, 65536>;
    default: return nullptr;
  }
}

template <int MODE>
TraceFn pick_rec(int records, uint32_t lds_stack, uint32_t stage) {
  switch (records) {
    case kRecNone: return pick_k<MODE, kRecNone>(lds_stack, stage);
    case kRecCsr: return pick_k<MODE, kRecCsr>(lds_stack, stage);
    case kRecSlab: return pick_k<MODE, kRecSlab>(lds_stack, stage);
    default: return nullptr;
  }
}

TraceFn pick_trace(uint32_t mode, int records, uint32_t lds_stack, uint32_t stage) {
  return mode == RTX_MODE_PLAIN ? pick_rec<0>(records, lds_stack, stage) : pick_rec<1>(records, lds_stack, stage);
}

// ---- slab records -> CSR ----------------------------------------------------
// A wave takes 64 consecutive rays, whose records fill one contiguous CSR
// range: lane i holds ray i's record count and the inclusive prefix of the
// counts, and output position f of the range (f = lane, lane + 64, ...)
// belongs to the ray whose prefix first exceeds f (a binary search over the
// wave's prefixes with shuffles).  Loads and stores are 16 bytes, the stores
// contiguous, and each lane has four loads in flight.  Records past a ray's
// slab are left to the second pass, which lists the ray.
__global__ void report_words_kernel(const uint32_t* word, const uint64_t* dword, uint32_t* host) {
  if (threadIdx.x == 0) {
    volatile uint32_t* h = host;
    const uint64_t d = *dword;
    h[0] = word ? *word : 0u;
    h[2] = (uint32_t)d;
    h[3] = (uint32_t)(d >> 32);
  }
}

__global__ __launch_bounds__(256) void compact_records(const uint32_t* counts, uint32_t n, const uint64_t* slabs,
                                                       uint32_t slab, const uint32_t* owner,
                                                       const uint64_t* layout_remap, const uint64_t* offsets,
                                                       uint4* records, uint64_t capacity, uint32_t ray_base,
                                                       uint32_t* overflow_list, uint32_t* overflow_count) {
  if (offsets[n] > capacity) return;  // the caller reports RTX_ERR_TRACE_OVERFLOW
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t waves = gridDim.x * 4;
  constexpr uint32_t kUnroll = 4;
  for (uint32_t base = wave * 64; base < n; base += waves * 64) {
    const uint32_t r = base + lane;
    uint32_t cnt = 0;
    if (r < n) {
      cnt = counts[r];
      if (cnt > slab) overflow_list[atomicAdd(overflow_count, 1u)] = ray_base + r;
    }
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(incl, d, 64);
      if (lane >= (uint32_t)d) incl += y;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    const uint64_t out = offsets[base];
    // wave-uniform loop: every lane takes part in the shuffles.  The kUnroll
    // searches run interleaved and every load is issued unconditionally (an
    // index past the end reads entry 0 and is not stored), so the slab loads
    // and then the owner loads of a pass are in flight together.
    const uint32_t stride = slab_stride(slab);
    for (uint32_t f0 = 0; f0 < total; f0 += 64 * kUnroll) {
      uint32_t lo[kUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kUnroll; ++u) lo[u] = 0;
      // first ray i with prefix(i) > f
#pragma unroll
      for (uint32_t step = 32; step > 0; step >>= 1) {
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
          const uint32_t pre = __shfl(incl, (int)(lo[u] + step - 1), 64);
          if (pre <= f0 + u * 64 + lane) lo[u] += step;
        }
      }
      uint64_t v[kUnroll];
      bool ok[kUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kUnroll; ++u) {
        const uint32_t f = f0 + u * 64 + lane;
        const uint32_t before = __shfl(incl, (int)(lo[u] ? lo[u] - 1 : 0), 64);
        const uint32_t k = f - (lo[u] ? before : 0u);
        ok[u] = f < total && k < slab;
        v[u] = slabs[ok[u] ? (uint64_t)(base + lo[u]) * stride + k : 0u];
      }
      uint32_t tl[kUnroll];
      uint64_t roff[kUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kUnroll; ++u) {
        const uint32_t slot = ok[u] ? (uint32_t)v[u] : 0u;
        tl[u] = owner ? owner[slot] : RTX_NO_TREELET;
        roff[u] = layout_remap ? layout_remap[slot] : (uint64_t)slot << 6;
      }
#pragma unroll
      for (uint32_t u = 0; u < kUnroll; ++u) {
        if (!ok[u]) continue;
        const uint32_t hw = (uint32_t)(v[u] >> 32);
        const uint32_t type = hw & 0xffu;
        const uint32_t size = type == RTX_TX_BVH_INSTANCE_LEAF ? 128u
                              : type == RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR ? 8u : 64u;
        records[out + f0 + u * 64 + lane] = make_uint4((uint32_t)roff[u], (uint32_t)(roff[u] >> 32), tl[u],
                                                       size | (type << 16) | ((hw >> 8) << 24));
      }
    }
  }
}

TraceFn pick_batch(uint32_t mode, uint32_t lds_stack) {
  const bool plain = mode == RTX_MODE_PLAIN;
  switch (lds_stack) {
    case 4: return plain ? trace_kernel<0, false, 4, 0, true> : trace_kernel<1, false, 4, 0, true>;
    case 8: return plain ? trace_kernel<0, false, 8, 0, true> : trace_kernel<1, false, 8, 0, true>;
    case 16: return plain ? trace_kernel<0, false, 16, 0, true> : trace_kernel<1, false, 16, 0, true>;
    default: return nullptr;
  }
}

hipError_t launch_fn(TraceFn f, const TraceParams& p, uint32_t grid, uint32_t bs, hipStream_t s) {
  if (!f) return hipErrorInvalidValue;
  TraceParams arg = p;
  void* args[] = {&arg};
  return hipLaunchKernel(reinterpret_cast<const void*>(f), dim3(grid), dim3(bs), args, 0, s);
}

uint32_t blocks_per_cu(TraceFn f, uint32_t bs) {
  int nb = 0;
  if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(f), (int)bs, 0) !=
                hipSuccess)
    return 0;
  return nb > 0 ? (uint32_t)nb : 0u;
}

__global__ void add_offset_kernel(uint64_t* v, uint64_t count, const uint64_t* addend) {
  const uint64_t a = *addend;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    v[i] += a;
}

__global__ void fill_zero_kernel(uint32_t* dst, uint32_t words) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) dst[i] = 0u;
}

}  // namespace

hipError_t launch_trace_batch(const TraceParams& p, uint32_t mode, uint32_t lds_stack, uint32_t grid,
                              hipStream_t s) {
  if (p.n_jobs == 0 || p.n_jobs > kMaxBatchJobs) return hipErrorInvalidValue;
  return launch_fn(pick_batch(mode, lds_stack), p, grid, 256, s);
}

uint32_t trace_batch_blocks_per_cu(uint32_t mode, uint32_t lds_stack) {
  return blocks_per_cu(pick_batch(mode, lds_stack), 256);
}

hipError_t launch_report_words(const uint32_t* word, const uint64_t* dword, uint32_t* host, hipStream_t s) {
  hipLaunchKernelGGL(report_words_kernel, dim3(1), dim3(64), 0, s, word, dword, host);
  return hipGetLastError();
}

hipError_t launch_fill_zero(uint32_t* dst, uint32_t words, hipStream_t s) {
  hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, s, dst, words);
  return hipGetLastError();
}

uint32_t trace_block_size(uint32_t stage_bytes) { return stage_bytes ? 64u : 256u; }

namespace {
// per-ray counters -> run totals; with no records the per-type counts follow
// from the counters: a internal nodes, b instance leaves, c primitive leaves
// (fetch = 64 + 64a + 192b + 64c, nodes = a + b + c, records = 1 + a + 2b + 2c)
__global__ __launch_bounds__(256) void k_stats_hits(const rtx_hit* hits, uint32_t n, bool derive_types,
                                                    unsigned long long* acc) {
  unsigned long long t[7] = {0, 0, 0, 0, 0, 0, 0}, nh = 0, na = 0, nodes = 0, fb = 0, ab = 0;
  uint32_t mx_nodes = 0, mx_level = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const rtx_hit& h = hits[i];
    nh += h.hit;
    na += h.n_any_hits;
    nodes += h.n_nodes;
    fb += h.fetch_bytes;
    ab += h.accessed_bytes;
    mx_nodes = max(mx_nodes, h.n_nodes);
    mx_level = max(mx_level, h.max_level);
    if (derive_types && h.n_records) {
      const uint32_t b = ((h.fetch_bytes >> 6) - 1u - h.n_nodes) >> 1;
      const uint32_t c = h.n_records - 1u - h.n_nodes - b;
      const uint32_t a = h.n_nodes - b - c;
      t[RTX_TX_BVH_STRUCTURE] += 1u + b;
      t[RTX_TX_BVH_INTERNAL_NODE] += a;
      t[RTX_TX_BVH_INSTANCE_LEAF] += b;
      t[RTX_TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR] += c;
      t[RTX_TX_BVH_QUAD_LEAF] += c;
    }
  }
  const unsigned long long v[14] = {t[0], t[1], t[2], t[3], t[4], t[5], t[6], nh, na, nodes, 0, 0, fb, ab};
  for (int k = 0; k < 14; ++k) {
    if (k == 10 || k == 11) continue;
    unsigned long long x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63u) == 0 && x) atomicAdd(&acc[k], x);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx_nodes = max(mx_nodes, (uint32_t)__shfl_xor(mx_nodes, o, 64));
    mx_level = max(mx_level, (uint32_t)__shfl_xor(mx_level, o, 64));
  }
  if ((threadIdx.x & 63u) == 0) {
    atomicMax(&acc[10], (unsigned long long)mx_nodes);
    atomicMax(&acc[11], (unsigned long long)mx_level);
  }
}

__global__ __launch_bounds__(256) void k_stats_records(const rtx_record* rec, uint64_t n, unsigned long long* acc) {
  unsigned long long t[7] = {0, 0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t ty = rec[i].type;
    if (ty < 7) t[ty]++;
  }
  for (int k = 0; k < 7; ++k) {
    unsigned long long x = t[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63u) == 0 && x) atomicAdd(&acc[k], x);
  }
}
}  // namespace

hipError_t launch_trace_stats(const rtx_hit* hits, uint32_t n, const rtx_record* records, uint64_t n_records,
                              unsigned long long* acc, hipStream_t s) {
  const uint32_t g = std::max(1u, std::min<uint32_t>(2048, (n + 255) / 256));
  if (n) hipLaunchKernelGGL(k_stats_hits, dim3(g), dim3(256), 0, s, hits, n, records == nullptr, acc);
  if (records && n_records) {
    const uint32_t gr = (uint32_t)std::min<uint64_t>(4096, (n_records + 255) / 256);
    hipLaunchKernelGGL(k_stats_records, dim3(gr), dim3(256), 0, s, records, n_records, acc);
  }
  return hipGetLastError();
}

hipError_t launch_compact_records(const uint32_t* counts, uint32_t n, const uint64_t* slabs, uint32_t slab,
                                  const uint32_t* owner, const uint64_t* layout_remap, const uint64_t* offsets,
                                  rtx_record* records, uint64_t capacity, uint32_t ray_base, uint32_t* overflow_list,
                                  uint32_t* overflow_count, hipStream_t s) {
  if (n == 0 || slab == 0) return hipSuccess;
  const uint32_t grid = std::max(1u, std::min<uint32_t>(4096, (n + 255) / 256));
  hipLaunchKernelGGL(compact_records, dim3(grid), dim3(256), 0, s, counts, n, slabs, slab, owner, layout_remap, offsets,
                     reinterpret_cast<uint4*>(records), capacity, ray_base, overflow_list, overflow_count);
  return hipGetLastError();
}

hipError_t launch_add_offset(uint64_t* v, uint64_t count, const uint64_t* addend, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (count + 255) / 256));
  hipLaunchKernelGGL(add_offset_kernel, dim3(grid), dim3(256), 0, s, v, count, addend);
  return hipGetLastError();
}

hipError_t launch_level_pass(const TraceParams& p, uint32_t mode, uint32_t grid, hipStream_t s) {
  if (!p.lm_table || !p.lm_list || p.lm_cap < 2 || (p.lm_cap & (p.lm_cap - 1))) return hipErrorInvalidValue;
  TraceFn f = mode == RTX_MODE_PLAIN ? trace_kernel<0, kRecNone, 8, 0, false, true>
                                     : trace_kernel<1, kRecNone, 8, 0, false, true>;
  return launch_fn(f, p, grid, 256, s);
}

hipError_t launch_trace_wide(const TraceParams& p, uint32_t mode, uint32_t cap, uint32_t grid, hipStream_t s) {
  if (cap > kWideCap) return hipErrorInvalidValue;
  if (mode == RTX_MODE_PLAIN)
    hipLaunchKernelGGL(trace_wide<0>, dim3(grid), dim3(kWideBlock), 0, s, p, cap);
  else
    hipLaunchKernelGGL(trace_wide<1>, dim3(grid), dim3(kWideBlock), 0, s, p, cap);
  return hipGetLastError();
}

uint32_t trace_wide_blocks_per_cu(uint32_t mode) {
  int nb = 0;
  const void* f = mode == RTX_MODE_PLAIN ? reinterpret_cast<const void*>(trace_wide<0>)
                                         : reinterpret_cast<const void*>(trace_wide<1>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, (int)kWideBlock, 0) != hipSuccess) return 0;
  return nb > 0 ? (uint32_t)nb : 0u;
}

hipError_t launch_trace(const TraceParams& p, uint32_t mode, int records, uint32_t lds_stack, uint32_t stage_bytes,
                        uint32_t grid, hipStream_t s) {
  return launch_fn(pick_trace(mode, records, lds_stack, stage_bytes), p, grid, trace_block_size(stage_bytes), s);
}

uint32_t trace_blocks_per_cu(uint32_t mode, int records, uint32_t lds_stack, uint32_t stage_bytes) {
  return blocks_per_cu(pick_trace(mode, records, lds_stack, stage_bytes), trace_block_size(stage_bytes));
}

hipError_t launch_scan_records(const rtx_hit* hits, const uint32_t* counts, uint32_t n, uint64_t* offsets,
                               uint64_t* tile_sums, hipStream_t s) {
  static_assert(kScanItems == 4, "scan_tiles reads four counts as one uint4");
  const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles > 0) {
    hipLaunchKernelGGL(scan_tiles, dim3(tiles), dim3(kScanBlock), 0, s, hits, counts, n, offsets, tile_sums);
  }
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, s, tile_sums, tiles, offsets, n);
  if (tiles > 0) {
    hipLaunchKernelGGL(scan_add, dim3(tiles), dim3(kScanBlock), 0, s, offsets, n, tile_sums);
  }
  return hipGetLastError();
}

}  // namespace rtxk
