// partition.hip — device treelet partitioner.
//
// Restates createTreelets (vulkan_ray_tracing.cc:823-1470) and
// buildNodeToRootMap (:475-489) for the GPU:
//  * A treelet is the longest breadth-first prefix of its root's subtree whose
//    byte sizes fit the budget (internal node 64, instance leaf 128 plus its
//    BLAS header 64, primitive leaf 64); the root always enters and must fit
//    (:893,962,976).  The first treelet is the TLAS header plus the prefix
//    that starts at the TLAS root with budget-64 bytes left (:846-866).
//  * When the next breadth-first entry does not fit, every entry still queued
//    becomes a pending root (:1132-1211).
//  * A treelet depends only on its root, so all pending roots of one round
//    are formed in parallel, one wave each, and a root reached twice (a BLAS
//    shared by several instances) is formed once; the reference forms it
//    again with an identical result.
//  * The wave walks the breadth-first queue in chunks of 64 entries: an
//    inclusive scan of the entry sizes finds how many fit, the accepted
//    entries append their children in order (exclusive scan of child counts),
//    which is exactly the sequential pop-front / push-back order.
//  * Node lists keep first occurrences (:1310-1336); roots are numbered in
//    ascending offset order (:1332); a node belongs to the largest root offset
//    whose list holds it, which is what iterating the ordered root map and
//    overwriting produces (:479-488).
#include <algorithm>
#include <vector>

#include "kernels.h"
#include "layout.h"
#include "partition.h"

namespace rtxk {
namespace {

// queue entry kinds
constexpr uint32_t kTopInternal = 0, kTopLeaf = 1, kBotInternal = 2, kBotLeaf = 3;
// node kinds recorded per slot by the validator
constexpr uint8_t kSlotNone = 0, kSlotHeader = 1, kSlotInternal = 2, kSlotInstance = 3, kSlotPrim = 4;

__device__ __forceinline__ uint64_t qentry(uint32_t slot, uint32_t kind) { return (uint64_t)slot | ((uint64_t)kind << 32); }
__device__ __forceinline__ uint32_t q_slot(uint64_t e) { return (uint32_t)e; }
__device__ __forceinline__ uint32_t q_kind(uint64_t e) { return (uint32_t)(e >> 32); }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += y;
  }
  return v;
}

struct FormParams {
  const uint8_t* A;
  uint64_t bytes;
  uint64_t tlas;
  uint32_t budget;
  uint32_t first_round;
  const uint64_t* frontier_in;
  uint32_t n_in;
  uint64_t* frontier_out;
  uint32_t* n_out;
  uint32_t* root_flag;
  uint32_t* owner_root1;
  uint32_t* raw_list;
  uint64_t raw_cap;
  unsigned long long* raw_count;
  uint4* desc;
  uint32_t* desc_count;
  uint64_t* scratch;
  uint32_t qcap;
  uint32_t lcap;
  uint32_t* work;
  uint32_t* error;
};

__global__ __launch_bounds__(256) void form_treelets(FormParams p) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  uint64_t* Q = p.scratch + (uint64_t)wave * (p.qcap + (p.lcap + 1) / 2);
  uint32_t* L = reinterpret_cast<uint32_t*>(Q + p.qcap);
  const uint8_t* A = p.A;
  while (true) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(p.work, 1u);
    idx = __shfl(idx, 0, 64);
    if (idx >= p.n_in) break;
    uint32_t err = 0;
    uint32_t head = 0, tail = 1, nl = 0;
    int64_t remaining = p.budget;
    uint64_t acc_bytes = 0;
    uint32_t root_slot;
    if (p.first_round) {
      root_slot = (uint32_t)(p.tlas >> 6);
      const uint64_t rootoff = *reinterpret_cast<const uint64_t*>(A + p.tlas);
      const uint64_t top = p.tlas + rootoff;
      if ((top & 63) || top + 64 > p.bytes) err |= kErrRange;
      if (lane == 0) {
        L[0] = root_slot;
        Q[0] = qentry((uint32_t)(top >> 6), kTopInternal);
      }
      nl = 1;
      remaining -= 64;
      acc_bytes = 64;
    } else {
      const uint64_t e = p.frontier_in[idx];
      root_slot = q_slot(e);
      if (lane == 0) Q[0] = e;
    }
    __threadfence();
    bool first_chunk = true;
    while (err == 0) {
      const uint32_t chunk = min(64u, tail - head);
      uint64_t e = 0;
      uint32_t sz = 0;
      if (lane < chunk) {
        e = Q[head + lane];
        sz = q_kind(e) == kTopLeaf ? 192u : 64u;
      }
      const uint32_t incl = wave_incl_scan<uint32_t>(sz, lane);
      const bool acc = lane < chunk && (int64_t)incl <= remaining;
      const uint64_t accmask = __ballot(acc);
      const uint32_t count = (uint32_t)__popcll(accmask);
      if (count == 0) {
        if (first_chunk) err |= kErrBudget;
        break;
      }
      // children and list additions of accepted entries
      uint32_t ladd = 0, nch = 0;
      uint64_t ch[6];
      uint32_t extra_slot = 0;
      if (acc) {
        const uint32_t slot = q_slot(e), kind = q_kind(e);
        const uint64_t off = (uint64_t)slot << 6;
        ladd = 1;
        if (kind == kTopInternal || kind == kBotInternal) {
          if (off + 64 > p.bytes) {
            err |= kErrRange;
          } else {
            const uint8_t* nd = A + off;
            const int32_t co = *reinterpret_cast<const int32_t*>(nd + 12);
            uint64_t c = off + (uint64_t)((int64_t)co * 64);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
              const uint32_t info = nd[22 + i] & 0x3fu;
              const uint32_t size = info & 3u, type = info >> 2;
              if (size > 0) {
                if ((c & 63) || c + 64 > p.bytes) err |= kErrRange;
                uint32_t k;
                if (kind == kTopInternal) {
                  if (type != 0 && type != 1) err |= kErrType;
                  k = type != 0 ? kTopLeaf : kTopInternal;
                } else {
                  k = type != 0 ? kBotLeaf : kBotInternal;
                }
                ch[nch++] = qentry((uint32_t)(c >> 6), k);
              }
              c += (uint64_t)size * 64;
            }
          }
        } else if (kind == kTopLeaf) {
          if (off + 128 > p.bytes) {
            err |= kErrRange;
          } else {
            const uint64_t bvh = *reinterpret_cast<const uint64_t*>(A + off + 64);
            const uint64_t h = off + bvh;
            if ((h & 63) || h + 64 > p.bytes) {
              err |= kErrRange;
            } else {
              const uint64_t r = h + *reinterpret_cast<const uint64_t*>(A + h);
              if ((r & 63) || r + 64 > p.bytes) err |= kErrRange;
              extra_slot = (uint32_t)(h >> 6);
              ladd = 2;
              ch[nch++] = qentry((uint32_t)(r >> 6), kBotInternal);
            }
          }
        } else {
          if (off + 64 > p.bytes) err |= kErrRange;
        }
      }
      if (__ballot(err != 0)) {
        for (int o = 32; o > 0; o >>= 1) err |= __shfl_xor(err, o, 64);
        break;
      }
      const uint32_t lincl = wave_incl_scan<uint32_t>(ladd, lane);
      const uint32_t cincl = wave_incl_scan<uint32_t>(nch, lane);
      const uint32_t ltot = __shfl(lincl, 63, 64);
      const uint32_t ctot = __shfl(cincl, 63, 64);
      if (nl + ltot > p.lcap || tail + ctot > p.qcap) {
        err |= kErrQueue;
        break;
      }
      if (acc) {
        uint32_t lp = nl + lincl - ladd;
        L[lp] = q_slot(e);
        if (ladd == 2) L[lp + 1] = extra_slot;
        uint32_t cp = tail + cincl - nch;
        for (uint32_t k = 0; k < nch; ++k) Q[cp + k] = ch[k];
      }
      __threadfence();
      const uint32_t used = __shfl(incl, (int)count - 1, 64);
      remaining -= used;
      acc_bytes += used;
      head += count;
      tail += ctot;
      nl += ltot;
      first_chunk = false;
      if (count < chunk || head == tail) break;
    }
    if (err) {
      if (lane == 0) atomicOr(p.error, err);
      continue;
    }
    // publish the node list, owner candidates, descriptor and pending roots
    unsigned long long pos = 0;
    if (lane == 0) pos = atomicAdd(p.raw_count, (unsigned long long)nl);
    pos = __shfl(pos, 0, 64);
    if (pos + nl > p.raw_cap) {
      if (lane == 0) atomicOr(p.error, kErrQueue);
      continue;
    }
    for (uint32_t k = lane; k < nl; k += 64) {
      const uint32_t s = L[k];
      p.raw_list[pos + k] = s;
      atomicMax(&p.owner_root1[s], root_slot + 1);
    }
    if (lane == 0) {
      const uint32_t d = atomicAdd(p.desc_count, 1u);
      p.desc[d] = make_uint4(root_slot, (uint32_t)pos, nl, (uint32_t)acc_bytes);
    }
    for (uint32_t k = head + lane; k < tail; k += 64) {
      const uint64_t e = Q[k];
      if (atomicCAS(&p.root_flag[q_slot(e)], 0u, 1u) == 0u) {
        const uint32_t o = atomicAdd(p.n_out, 1u);
        p.frontier_out[o] = e;
      }
    }
  }
}

// ---- generic exclusive scans ------------------------------------------------
__device__ __forceinline__ uint64_t block_excl(uint64_t v, uint64_t* sm, uint64_t& total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint64_t x = wave_incl_scan<uint64_t>(v, lane);
  if (lane == 63) sm[w] = x;
  __syncthreads();
  uint64_t off = 0;
  for (uint32_t k = 0; k < w; ++k) off += sm[k];
  total = sm[0] + sm[1] + sm[2] + sm[3];
  __syncthreads();
  return off + x - v;
}

constexpr uint32_t kTile = 1024;  // 256 threads x 4 items

template <typename In>
__global__ __launch_bounds__(256) void scan_tile_k(In in, uint64_t n, uint64_t* out, uint64_t* sums) {
  __shared__ uint64_t sm[4];
  const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x * 4;
  uint64_t v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? (uint64_t)in(base + k) : 0;
    s += v[k];
  }
  uint64_t tot;
  uint64_t ex = block_excl(s, sm, tot);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void scan_sums_k(uint64_t* sums, uint64_t tiles, uint64_t* total_out) {
  __shared__ uint64_t sm[4];
  uint64_t carry = 0;
  for (uint64_t b = 0; b < tiles; b += 256) {
    const uint64_t i = b + threadIdx.x;
    uint64_t v = i < tiles ? sums[i] : 0, tot;
    uint64_t ex = block_excl(v, sm, tot);
    if (i < tiles) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(256) void scan_add_k(uint64_t* out, uint64_t n, const uint64_t* sums) {
  const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x * 4;
  const uint64_t a = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) out[base + k] += a;
}

struct FlagIn {
  const uint32_t* f;
  __device__ uint64_t operator()(uint64_t i) const { return f[i] != 0 ? 1u : 0u; }
};
struct U32In {
  const uint32_t* v;
  __device__ uint64_t operator()(uint64_t i) const { return v[i]; }
};

template <typename In>
hipError_t exclusive_scan(In in, uint64_t n, uint64_t* out, uint64_t* sums, uint64_t* total, hipStream_t s) {
  const uint64_t tiles = (n + kTile - 1) / kTile;
  if (tiles) hipLaunchKernelGGL((scan_tile_k<In>), dim3((uint32_t)tiles), dim3(256), 0, s, in, n, out, sums);
  hipLaunchKernelGGL(scan_sums_k, dim3(1), dim3(256), 0, s, sums, tiles, total);
  if (tiles) hipLaunchKernelGGL(scan_add_k, dim3((uint32_t)tiles), dim3(256), 0, s, out, n, sums);
  return hipGetLastError();
}

// ---- finalisation --------------------------------------------------------
__global__ void owner_and_roots(uint64_t slots, const uint32_t* root_flag, const uint64_t* rank,
                                const uint32_t* owner_root1, uint32_t* owner, uint64_t* roots) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < slots; s += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t o1 = owner_root1[s];
    owner[s] = o1 ? (uint32_t)rank[o1 - 1] : RTX_NO_TREELET;
    if (root_flag[s]) roots[rank[s]] = s << 6;
  }
}

// One wave per descriptor: de-duplicated length (first occurrence kept).
__global__ __launch_bounds__(256) void dedupe_count(const uint4* desc, uint32_t n_desc, const uint32_t* raw,
                                                    const uint64_t* rank, uint32_t* dlen, uint32_t* desc_of_rank,
                                                    uint32_t check_dupes) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (uint32_t d = wave; d < n_desc; d += gridDim.x * 4) {
    const uint4 ds = desc[d];
    const uint32_t* l = raw + ds.y;
    uint32_t kept = 0;
    for (uint32_t j = lane; j < ds.z; j += 64) {
      bool keep = true;
      if (check_dupes) {
        const uint32_t v = l[j];
        for (uint32_t k = 0; k < j; ++k)
          if (l[k] == v) {
            keep = false;
            break;
          }
      }
      kept += keep ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) kept += __shfl_xor(kept, o, 64);
    if (lane == 0) {
      const uint32_t r = (uint32_t)rank[ds.x];
      dlen[r] = kept;
      desc_of_rank[r] = d;
    }
  }
}

__global__ __launch_bounds__(256) void dedupe_copy(const uint4* desc, uint32_t n, const uint32_t* raw,
                                                   const uint32_t* desc_of_rank, const uint64_t* list_offsets,
                                                   uint64_t* list_nodes, uint32_t check_dupes) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (uint32_t r = wave; r < n; r += gridDim.x * 4) {
    const uint4 ds = desc[desc_of_rank[r]];
    const uint32_t* l = raw + ds.y;
    uint64_t out = list_offsets[r];
    for (uint32_t j0 = 0; j0 < ds.z; j0 += 64) {
      const uint32_t j = j0 + lane;
      bool keep = j < ds.z;
      uint32_t v = 0;
      if (keep) {
        v = l[j];
        if (check_dupes)
          for (uint32_t k = 0; k < j; ++k)
            if (l[k] == v) {
              keep = false;
              break;
            }
      }
      const uint64_t m = __ballot(keep);
      if (keep) {
        const uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        list_nodes[out + pos] = (uint64_t)v << 6;
      }
      out += (uint64_t)__popcll(m);
    }
  }
}

// ---- structural validation: breadth-first walk with per-slot reference counts
struct ValParams {
  const uint8_t* A;
  uint64_t bytes;
  const uint64_t* in;
  uint32_t n_in;
  uint64_t* out;
  uint32_t* n_out;
  uint32_t* refs;
  uint8_t* kind;
  uint32_t* shared;
  uint32_t* error;
  uint32_t* unordered;  // set when a valid child box has a lower bound byte above its upper one
  // tree levels (tree_level_map values): nodes processed in this round are at
  // level round + 1 (the TLAS root at 1), their children and a BLAS root
  // reached through an instance leaf at round + 2; lvl[slot] keeps the first
  // level seen and `inconsistent` is set when a node is reached at another
  uint32_t* lvl;
  uint32_t round;
  uint32_t* inconsistent;
};

__device__ __forceinline__ void level_ref(const ValParams& p, uint64_t off, uint32_t level) {
  const uint32_t prev = atomicCAS(&p.lvl[off >> 6], 0u, level);
  if (prev != 0u && prev != level) atomicOr(p.inconsistent, 1u);
}

__device__ void val_ref(const ValParams& p, uint64_t off, uint8_t k, bool push, uint32_t qkind) {
  const uint32_t s = (uint32_t)(off >> 6);
  const uint32_t old = atomicAdd(&p.refs[s], 1u);
  if (old != 0) {
    atomicOr(p.shared, 1u);
    return;
  }
  p.kind[s] = k;
  if (push) {
    const uint32_t o = atomicAdd(p.n_out, 1u);
    p.out[o] = qentry(s, qkind);
  }
}

__global__ void validate_level(ValParams p) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n_in; i += gridDim.x * blockDim.x) {
    const uint64_t e = p.in[i];
    const uint64_t off = (uint64_t)q_slot(e) << 6;
    const uint32_t kind = q_kind(e);
    const uint8_t* nd = p.A + off;
    const int32_t co = *reinterpret_cast<const int32_t*>(nd + 12);
    uint64_t c = off + (uint64_t)((int64_t)co * 64);
    const uint32_t child_level = p.round + 2;
    for (int k = 0; k < 6; ++k) {
      const uint32_t info = nd[22 + k] & 0x3fu;
      const uint32_t size = info & 3u, type = info >> 2;
      if (size == 0) continue;
      // lower x/y/z at bytes 28/40/52 + k, upper at 34/46/58 + k
      if (nd[28 + k] > nd[34 + k] || nd[40 + k] > nd[46 + k] || nd[52 + k] > nd[58 + k]) atomicOr(p.unordered, 1u);
      const uint64_t len = (uint64_t)size * 64;
      if ((c & 63) || c + len > p.bytes) {
        atomicOr(p.error, kErrRange);
        return;
      }
      level_ref(p, c, child_level);
      if (type == 0) {
        val_ref(p, c, kSlotInternal, true, kind);
      } else if (kind == kTopInternal) {
        if (type != 1 || size != 2) {
          atomicOr(p.error, kErrType);
          return;
        }
        const uint64_t h = c + *reinterpret_cast<const uint64_t*>(p.A + c + 64);
        if ((h & 63) || h + 64 > p.bytes) {
          atomicOr(p.error, kErrRange);
          return;
        }
        const uint64_t r = h + *reinterpret_cast<const uint64_t*>(p.A + h);
        if ((r & 63) || r + 64 > p.bytes) {
          atomicOr(p.error, kErrRange);
          return;
        }
        val_ref(p, c, kSlotInstance, false, 0);
        level_ref(p, r, child_level);  // tree_level_map[botLevelRootAddr] = tree_level_map[leaf_addr]
        const uint32_t hs = (uint32_t)(h >> 6);
        if (atomicAdd(&p.refs[hs], 1u) == 0u) {
          p.kind[hs] = kSlotHeader;
          val_ref(p, r, kSlotInternal, true, kBotInternal);
        } else {
          atomicOr(p.shared, 1u);
        }
      } else {
        if (type == 1) {
          atomicOr(p.error, kErrType);
          return;
        }
        val_ref(p, c, kSlotPrim, false, 0);
      }
      c += len;
    }
  }
}

// ---- cycle check: Kahn's peeling over the graph the walk above counted ------
// refs[s] holds each reachable slot's in-degree from the walk (every reference
// an expanded node made; a BLAS header one per instance leaf, its root one from
// the header).  A round takes the internal nodes whose in-degree reached zero
// and removes their edges in the same order the walk added them; the graph has
// a cycle exactly when some count never reaches zero.
struct KahnParams {
  const uint8_t* A;
  const uint64_t* in;
  uint32_t n_in;
  uint64_t* out;
  uint32_t* n_out;
  uint32_t* refs;
};

__device__ __forceinline__ void kahn_push(const KahnParams& p, uint64_t off, uint32_t kind) {
  const uint32_t o = atomicAdd(p.n_out, 1u);
  p.out[o] = qentry((uint32_t)(off >> 6), kind);
}

__global__ void kahn_level(KahnParams p) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n_in; i += gridDim.x * blockDim.x) {
    const uint64_t e = p.in[i];
    const uint64_t off = (uint64_t)q_slot(e) << 6;
    const uint32_t kind = q_kind(e);
    const uint8_t* nd = p.A + off;
    const int32_t co = *reinterpret_cast<const int32_t*>(nd + 12);
    uint64_t c = off + (uint64_t)((int64_t)co * 64);
    for (int k = 0; k < 6; ++k) {
      const uint32_t info = nd[22 + k] & 0x3fu;
      const uint32_t size = info & 3u, type = info >> 2;
      if (size == 0) continue;
      if (type == 0) {
        if (atomicSub(&p.refs[c >> 6], 1u) == 1u) kahn_push(p, c, kind);
      } else if (kind == kTopInternal) {
        atomicSub(&p.refs[c >> 6], 1u);
        const uint64_t h = c + *reinterpret_cast<const uint64_t*>(p.A + c + 64);
        const uint64_t r = h + *reinterpret_cast<const uint64_t*>(p.A + h);
        if (atomicSub(&p.refs[h >> 6], 1u) == 1u && atomicSub(&p.refs[r >> 6], 1u) == 1u) kahn_push(p, r, kBotInternal);
      } else {
        atomicSub(&p.refs[c >> 6], 1u);
      }
      c += (uint64_t)size * 64;
    }
  }
}

__global__ void count_nonzero(const uint32_t* v, uint64_t n, uint32_t* count) {
  uint32_t local = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    local += v[i] != 0u ? 1u : 0u;
  if (local) atomicAdd(count, local);
}

template <typename T>
rtx_status dmalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return RTX_ERR_OUT_OF_MEMORY;
  }
  return RTX_OK;
}

struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* q : ptrs) hipFree(q);
  }
  template <typename T>
  rtx_status alloc(T** p, size_t count) {
    rtx_status st = dmalloc(p, count);
    if (st == RTX_OK) ptrs.push_back(*p);
    return st;
  }
};

rtx_status map_err(uint32_t bits) {
  if (bits & kErrRange) return RTX_ERR_OFFSET_OUT_OF_ARENA;
  if (bits & kErrType) return RTX_ERR_BAD_NODE_TYPE;
  if (bits & kErrBudget) return RTX_ERR_NODE_LARGER_THAN_BUDGET;
  if (bits & kErrQueue) return RTX_ERR_OUT_OF_MEMORY;
  return bits ? RTX_ERR_HIP : RTX_OK;
}

#define CK(expr)                          \
  do {                                    \
    if ((expr) != hipSuccess) {           \
      (void)hipGetLastError();            \
      return RTX_ERR_HIP;                 \
    }                                     \
  } while (0)
#define CS(expr)                  \
  do {                            \
    rtx_status _s = (expr);       \
    if (_s != RTX_OK) return _s;  \
  } while (0)

// Walks the DAG; fills kind[] and reports whether any node is referenced twice.
rtx_status run_validate(const uint8_t* A, uint64_t bytes, uint64_t tlas, hipStream_t s, Scratch& sc, uint8_t* kind,
                        uint32_t* shared_out, uint64_t* n_nodes, uint32_t* unordered_out = nullptr,
                        SceneCheck* check = nullptr) {
  const uint64_t slots = bytes / 64;
  uint32_t *refs, *lvl;
  uint64_t *qa, *qb;
  uint32_t* ctr;  // [0] n_out, [1] error, [2] shared, [3] unordered, [4] inconsistent, [5] nonzero
  CS(sc.alloc(&refs, slots));
  CS(sc.alloc(&lvl, slots));
  CS(sc.alloc(&qa, slots));
  CS(sc.alloc(&qb, slots));
  CS(sc.alloc(&ctr, 8));
  CK(hipMemsetAsync(refs, 0, slots * 4, s));
  CK(hipMemsetAsync(lvl, 0, slots * 4, s));
  CK(hipMemsetAsync(kind, 0, slots, s));
  CK(hipMemsetAsync(ctr, 0, 32, s));
  uint64_t hdr[1];
  CK(hipMemcpyAsync(hdr, A + tlas, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  const uint64_t root = tlas + hdr[0];
  if ((root & 63) || root + 64 > bytes) return RTX_ERR_OFFSET_OUT_OF_ARENA;
  const uint32_t one = 1;
  const uint8_t khdr = kSlotHeader, kint = kSlotInternal;
  CK(hipMemcpyAsync(refs + tlas / 64, &one, 4, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(kind + tlas / 64, &khdr, 1, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(refs + root / 64, &one, 4, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(lvl + root / 64, &one, 4, hipMemcpyHostToDevice, s));  // tree_level_map[topRootAddr] = 1
  CK(hipMemcpyAsync(kind + root / 64, &kint, 1, hipMemcpyHostToDevice, s));
  uint64_t e0 = (root / 64) | ((uint64_t)kTopInternal << 32);
  CK(hipMemcpyAsync(qa, &e0, 8, hipMemcpyHostToDevice, s));
  uint32_t n_in = 1;
  uint64_t total = 2;
  uint32_t round = 0;
  while (n_in > 0) {
    CK(hipMemsetAsync(ctr, 0, 4, s));
    ValParams vp{A, bytes, qa, n_in, qb, ctr, refs, kind, ctr + 2, ctr + 1, ctr + 3, lvl, round++, ctr + 4};
    uint32_t grid = std::min<uint32_t>(4096, (n_in + 255) / 256);
    hipLaunchKernelGGL(validate_level, dim3(grid), dim3(256), 0, s, vp);
    CK(hipGetLastError());
    uint32_t c[3];
    CK(hipMemcpyAsync(c, ctr, 12, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    if (c[1]) return map_err(c[1]);
    total += c[0];
    n_in = c[0];
    std::swap(qa, qb);
  }
  // count every referenced slot (leaves and headers included)
  uint32_t c[5];
  CK(hipMemcpyAsync(c, ctr, 20, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  if (shared_out) *shared_out = c[2];
  if (unordered_out) *unordered_out = c[3];
  if (check) {
    check->ordered = c[3] == 0;
    check->level_consistent = c[4] == 0;
    check->cyclic = false;
    // A cycle reaches a node at two levels, so only inconsistent scenes are peeled.
    if (!check->level_consistent) {
      uint32_t r0 = 0;
      CK(hipMemcpyAsync(&r0, refs + root / 64, 4, hipMemcpyDeviceToHost, s));
      CK(hipStreamSynchronize(s));
      r0 -= 1;  // the walk's seed reference
      CK(hipMemcpyAsync(refs + root / 64, &r0, 4, hipMemcpyHostToDevice, s));
      const uint32_t zero = 0;
      CK(hipMemcpyAsync(refs + tlas / 64, &zero, 4, hipMemcpyHostToDevice, s));
      n_in = r0 == 0 ? 1u : 0u;  // an edge into the TLAS root closes a cycle
      CK(hipMemcpyAsync(qa, &e0, 8, hipMemcpyHostToDevice, s));
      while (n_in > 0) {
        CK(hipMemsetAsync(ctr, 0, 4, s));
        KahnParams kp{A, qa, n_in, qb, ctr, refs};
        uint32_t grid = std::min<uint32_t>(4096, (n_in + 255) / 256);
        hipLaunchKernelGGL(kahn_level, dim3(grid), dim3(256), 0, s, kp);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(&n_in, ctr, 4, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        std::swap(qa, qb);
      }
      const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (slots + 255) / 256));
      hipLaunchKernelGGL(count_nonzero, dim3(g), dim3(256), 0, s, refs, slots, ctr + 5);
      CK(hipGetLastError());
      uint32_t left = 0;
      CK(hipMemcpyAsync(&left, ctr + 5, 4, hipMemcpyDeviceToHost, s));
      CK(hipStreamSynchronize(s));
      check->cyclic = left != 0;
    }
  }
  if (n_nodes) {
    std::vector<uint8_t> k(slots);
    CK(hipMemcpyAsync(k.data(), kind, slots, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    uint64_t cnt = 0;
    for (uint8_t v : k) cnt += v != kSlotNone;
    *n_nodes = cnt;
  }
  (void)total;
  return RTX_OK;
}

// ---- treelet-contiguous images (LDS staging source) -------------------------
// Image t = the nodes owned by treelet t, in list order; an instance leaf takes
// 128 bytes, every other node 64.  One wave per treelet.
__global__ __launch_bounds__(256) void image_layout(uint32_t n, const uint64_t* lo, const uint64_t* ln,
                                                    const uint32_t* owner, const uint8_t* kind, uint32_t* stage_slot,
                                                    uint32_t* image_bytes) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (uint32_t t = wave; t < n; t += gridDim.x * 4) {
    uint32_t pos = 0;
    for (uint64_t j0 = lo[t]; j0 < lo[t + 1]; j0 += 64) {
      const uint64_t j = j0 + lane;
      uint32_t units = 0, s = 0;
      if (j < lo[t + 1]) {
        s = (uint32_t)(ln[j] >> 6);
        if (owner[s] == t) units = kind[s] == kSlotInstance ? 2u : 1u;
      }
      const uint32_t incl = wave_incl_scan<uint32_t>(units, lane);
      if (units) stage_slot[s] = pos + incl - units;
      pos += __shfl(incl, 63, 64);
    }
    if (lane == 0) image_bytes[t] = pos * 64;
  }
}

__global__ __launch_bounds__(256) void image_copy(uint32_t n, const uint64_t* lo, const uint64_t* ln,
                                                  const uint32_t* owner, const uint8_t* kind,
                                                  const uint32_t* stage_slot, const uint64_t* image_offset,
                                                  const uint8_t* A, uint8_t* images, uint32_t* image_slots) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (uint32_t t = wave; t < n; t += gridDim.x * 4) {
    uint8_t* img = images + image_offset[t];
    for (uint64_t j = lo[t] + lane; j < lo[t + 1]; j += 64) {
      const uint32_t s = (uint32_t)(ln[j] >> 6);
      if (owner[s] != t) continue;
      const uint32_t units = kind[s] == kSlotInstance ? 2u : 1u;
      const uint4* src = reinterpret_cast<const uint4*>(A + ((uint64_t)s << 6));
      uint4* dst = reinterpret_cast<uint4*>(img + (uint64_t)stage_slot[s] * 64);
      for (uint32_t k = 0; k < 4 * units; ++k) dst[k] = src[k];
      uint32_t* idx = image_slots + image_offset[t] / 64 + stage_slot[s];
      idx[0] = s;
      if (units == 2) idx[1] = 0xffffffffu;
    }
  }
}

}  // namespace

// Children owned by the first treelet, one byte per node slot after the owner
// words: for an internal node bit i is set when child i is in the arena and
// owner[child i] == first; for an instance leaf bit 0 is set when its BLAS
// root is.  Treelet-order traversal tests exactly these ownerships before its
// first treelet switch, so it reads this byte instead of six owner words.
__global__ void first_bits_kernel(const uint8_t* A, uint64_t bytes, const uint8_t* kind, const uint32_t* owner,
                                  uint32_t first, uint64_t slots, uint8_t* out) {
  for (uint64_t sl = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; sl < slots;
       sl += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t off = sl << 6;
    uint32_t bits = 0;
    if (kind[sl] == kSlotInternal) {
      const uint8_t* nd = A + off;
      uint64_t c = off + (uint64_t)((int64_t)*reinterpret_cast<const int32_t*>(nd + 12) * 64);
      for (int i = 0; i < 6; ++i) {
        const uint32_t size = nd[22 + i] & 3u;
        if (size != 0 && !(c & 63) && c + 64 <= bytes && owner[c >> 6] == first) bits |= 1u << i;
        c += (uint64_t)size * 64;
      }
    } else if (kind[sl] == kSlotInstance && off + 128 <= bytes) {
      const uint64_t h = off + *reinterpret_cast<const uint64_t*>(A + off + 64);
      if (!(h & 63) && h + 64 <= bytes) {
        const uint64_t r = h + *reinterpret_cast<const uint64_t*>(A + h);
        if (!(r & 63) && r + 64 <= bytes && owner[r >> 6] == first) bits = 1u;
      }
    }
    out[sl] = (uint8_t)bits;
  }
}

rtx_status build_images(const uint8_t* A, uint64_t bytes, hipStream_t s, const uint8_t* kind, Partition& p) {
  const uint64_t slots = bytes / 64;
  hipLaunchKernelGGL(first_bits_kernel, dim3((uint32_t)std::min<uint64_t>(8192, (slots + 255) / 256)), dim3(256), 0, s, A,
                     bytes, kind, p.owner, p.first_treelet, slots, reinterpret_cast<uint8_t*>(p.owner + slots));
  CK(hipGetLastError());
  CS(dmalloc(&p.stage_slot, slots));
  CS(dmalloc(&p.image_bytes, p.n));
  CS(dmalloc(&p.image_offset, (size_t)p.n + 1));
  CK(hipMemsetAsync(p.stage_slot, 0xff, slots * 4, s));
  const uint32_t g = std::min<uint32_t>(4096, (p.n + 3) / 4);
  hipLaunchKernelGGL(image_layout, dim3(g), dim3(256), 0, s, p.n, p.list_offsets, p.list_nodes, p.owner, kind,
                     p.stage_slot, p.image_bytes);
  CK(hipGetLastError());
  Scratch sc;
  uint64_t* sums;
  CS(sc.alloc(&sums, p.n / kTile + 2));
  CK(exclusive_scan(U32In{p.image_bytes}, p.n, p.image_offset, sums, p.image_offset + p.n, s));
  uint64_t total = 0;
  CK(hipMemcpyAsync(&total, p.image_offset + p.n, 8, hipMemcpyDeviceToHost, s));
  std::vector<uint32_t> ib(p.n);
  CK(hipMemcpyAsync(ib.data(), p.image_bytes, (size_t)p.n * 4, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  p.images_total = total;
  p.max_image_bytes = 0;
  for (uint32_t v : ib) p.max_image_bytes = std::max(p.max_image_bytes, v);
  CS(dmalloc(&p.images, total));
  CS(dmalloc(&p.image_slots, total / 64));
  // units no node claims stay "none", so staging never inserts them
  if (total) CK(hipMemsetAsync(p.image_slots, 0xff, total / 64 * 4, s));
  hipLaunchKernelGGL(image_copy, dim3(g), dim3(256), 0, s, p.n, p.list_offsets, p.list_nodes, p.owner, kind,
                     p.stage_slot, p.image_offset, A, p.images, p.image_slots);
  CK(hipGetLastError());
  return RTX_OK;
}

// ---- treelet layout: remapBVHToTreeletLayout (vulkan_ray_tracing.cc:1473-1509) ----
// Treelet t's block starts at t * (budget + stride): its root first, then the
// other nodes of its list in order, each advancing by its size (instance leaf
// 128 bytes, every other node 64).  A node listed by several treelets keeps
// the place of the first (lowest index, i.e. lowest root address): each list
// entry proposes (treelet << 32 | place in block) and the slot keeps the
// minimum.  A wave scans one treelet's list.
__global__ __launch_bounds__(256) void layout_propose(uint32_t n, const uint64_t* roots, const uint64_t* lo,
                                                      const uint64_t* ln, const uint8_t* kind,
                                                      unsigned long long* key) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (uint32_t t = wave; t < n; t += gridDim.x * 4) {
    const uint32_t rs = (uint32_t)(roots[t] >> 6);
    // the root sits at the block start; the others follow it
    uint32_t place = kind[rs] == kSlotInstance ? 128u : 64u;
    if (lane == 0) atomicMin(&key[rs], (unsigned long long)t << 32);
    for (uint64_t j0 = lo[t]; j0 < lo[t + 1]; j0 += 64) {
      const uint64_t j = j0 + lane;
      uint32_t sz = 0, s = 0;
      if (j < lo[t + 1]) {
        s = (uint32_t)(ln[j] >> 6);
        if (s != rs) sz = kind[s] == kSlotInstance ? 128u : 64u;
      }
      const uint32_t incl = wave_incl_scan<uint32_t>(sz, lane);
      if (sz) atomicMin(&key[s], ((unsigned long long)t << 32) | (place + incl - sz));
      place += __shfl(incl, 63, 64);
    }
  }
}

__global__ __launch_bounds__(256) void layout_place(uint64_t slots, const unsigned long long* key, uint64_t block,
                                                    uint64_t* remap, unsigned long long* placed) {
  uint64_t count = 0;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = key[s];
    remap[s] = k == ~0ull ? ~0ull : (k >> 32) * block + (k & 0xffffffffull);
    count += k == ~0ull ? 0u : 1u;
  }
  for (int o = 32; o > 0; o >>= 1) count += __shfl_xor(count, o, 64);
  if ((threadIdx.x & 63u) == 0 && count) atomicAdd(placed, (unsigned long long)count);
}

// metadata[t][j]: place of list entry j relative to the root's place; also
// flags a root that an earlier treelet placed (the reference asserts there)
__global__ __launch_bounds__(256) void layout_metadata_kernel(uint32_t n, uint32_t per, const uint64_t* roots,
                                                              const uint64_t* lo, const uint64_t* ln,
                                                              const unsigned long long* key, const uint64_t* remap,
                                                              int32_t* meta, uint32_t* conflict) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (uint32_t t = wave; t < n; t += gridDim.x * 4) {
    const uint32_t rs = (uint32_t)(roots[t] >> 6);
    if (lane == 0 && (uint32_t)(key[rs] >> 32) != t) atomicOr(conflict, 1u);
    const int64_t r = (int64_t)remap[rs];
    const uint64_t len = lo[t + 1] - lo[t];
    for (uint32_t j = lane; j < per; j += 64) {
      int32_t v = -1;
      if (j < len) v = (int32_t)((int64_t)remap[ln[lo[t] + j] >> 6] - r);
      meta[(uint64_t)t * per + j] = v;
    }
  }
}

void partition_free(Partition& p) {
  hipFree(p.layout_remap);
  hipFree(p.layout_metadata);
  hipFree(p.owner);
  hipFree(p.roots);
  hipFree(p.list_offsets);
  hipFree(p.list_nodes);
  hipFree(p.images);
  hipFree(p.image_offset);
  hipFree(p.image_bytes);
  hipFree(p.stage_slot);
  hipFree(p.image_slots);
  p = Partition{};
}

rtx_status partition_layout(const uint8_t* A, uint64_t bytes, uint64_t tlas, uint32_t stride, hipStream_t s,
                            Partition& p) {
  if (!p.valid) return RTX_ERR_NO_PARTITION;
  const uint64_t slots = bytes / 64;
  const uint32_t per = p.budget / 64;
  hipFree(p.layout_remap);
  hipFree(p.layout_metadata);
  p.layout_remap = nullptr;
  p.layout_metadata = nullptr;
  Scratch sc;
  uint8_t* kind;
  unsigned long long* key;
  unsigned long long* placed;
  uint32_t* conflict;
  CS(sc.alloc(&kind, slots));
  CS(sc.alloc(&key, slots));
  CS(sc.alloc(&placed, 1));
  CS(sc.alloc(&conflict, 1));
  CS(run_validate(A, bytes, tlas, s, sc, kind, nullptr, nullptr));
  uint64_t* remap;
  int32_t* meta;
  CS(dmalloc(&remap, slots));
  rtx_status st = dmalloc(&meta, (size_t)p.n * per);
  if (st != RTX_OK) {
    hipFree(remap);
    return st;
  }
  auto fail = [&](rtx_status e) {
    hipFree(remap);
    hipFree(meta);
    return e;
  };
  if (hipMemsetAsync(key, 0xff, slots * 8, s) != hipSuccess || hipMemsetAsync(placed, 0, 8, s) != hipSuccess ||
      hipMemsetAsync(conflict, 0, 4, s) != hipSuccess)
    return fail(RTX_ERR_HIP);
  const uint32_t g = std::max(1u, std::min<uint32_t>(4096, (p.n + 3) / 4));
  hipLaunchKernelGGL(layout_propose, dim3(g), dim3(256), 0, s, p.n, p.roots, p.list_offsets, p.list_nodes, kind, key);
  const uint32_t gs = (uint32_t)std::min<uint64_t>(8192, (slots + 255) / 256);
  hipLaunchKernelGGL(layout_place, dim3(std::max(gs, 1u)), dim3(256), 0, s, slots, key,
                     (uint64_t)p.budget + stride, remap, placed);
  hipLaunchKernelGGL(layout_metadata_kernel, dim3(g), dim3(256), 0, s, p.n, per, p.roots, p.list_offsets,
                     p.list_nodes, key, remap, meta, conflict);
  if (hipGetLastError() != hipSuccess) return fail(RTX_ERR_HIP);
  uint32_t bad = 0;
  unsigned long long mapped = 0;
  if (hipMemcpyAsync(&bad, conflict, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&mapped, placed, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(RTX_ERR_HIP);
  if (bad) return fail(RTX_ERR_UNSUPPORTED);
  p.layout_remap = remap;
  p.layout_metadata = meta;
  p.layout_stride = stride;
  p.layout_mapped = mapped;
  return RTX_OK;
}

// Whether every internal node reachable from the TLAS has, for each child it
// uses, lower bound bytes no greater than the upper ones on all three axes (so
// the child's decoded box has lo <= hi).  False as well when the walk finds a
// malformed arena or cannot allocate its scratch: traversal then keeps the
// compare-select slab tests, and a malformed arena still reports at trace time.
SceneCheck scene_check(const uint8_t* A, uint64_t bytes, uint64_t tlas, hipStream_t s) {
  SceneCheck out;
  Scratch sc;
  uint8_t* kind;
  if (sc.alloc(&kind, bytes / 64) != RTX_OK) return out;
  SceneCheck ck;
  if (run_validate(A, bytes, tlas, s, sc, kind, nullptr, nullptr, nullptr, &ck) != RTX_OK) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s);
    return out;
  }
  ck.valid = true;
  return ck;
}

rtx_status partition_validate(const uint8_t* A, uint64_t bytes, uint64_t tlas, hipStream_t s, uint64_t* n_nodes) {
  Scratch sc;
  uint8_t* kind;
  CS(sc.alloc(&kind, bytes / 64));
  SceneCheck ck;
  CS(run_validate(A, bytes, tlas, s, sc, kind, nullptr, n_nodes, nullptr, &ck));
  return ck.cyclic ? RTX_ERR_CYCLE : RTX_OK;
}

rtx_status build_images(const uint8_t* A, uint64_t bytes, hipStream_t s, const uint8_t* kind, Partition& p);

rtx_status partition_build(const uint8_t* A, uint64_t bytes, uint64_t tlas, uint32_t budget, hipStream_t s,
                           Partition& out) {
  partition_free(out);
  const uint64_t slots = bytes / 64;
  Scratch sc;
  uint8_t* kind;
  CS(sc.alloc(&kind, slots));
  uint32_t shared = 0;
  CS(run_validate(A, bytes, tlas, s, sc, kind, &shared, nullptr));

  uint32_t *root_flag, *owner_root1, *ctr;  // ctr: [0] n_out [1] error [2] desc_count [3] work
  uint64_t *fa, *fb;
  unsigned long long* raw_count;
  CS(sc.alloc(&root_flag, slots));
  CS(sc.alloc(&owner_root1, slots));
  CS(sc.alloc(&fa, slots));
  CS(sc.alloc(&fb, slots));
  CS(sc.alloc(&ctr, 8));
  CS(sc.alloc(&raw_count, 1));
  uint4* desc;
  CS(sc.alloc(&desc, slots));
  uint64_t raw_cap = std::min<uint64_t>(0xffffffffull, 3 * slots + 1024);
  uint32_t* raw_list;
  CS(sc.alloc(&raw_list, raw_cap));
  const uint32_t qcap = 6 * (budget / 64) + 8;
  const uint32_t lcap = budget / 64 + 4;
  const uint64_t per_wave = qcap + (lcap + 1) / 2;
  uint32_t waves = 1024;
  while (waves > 64 && per_wave * waves * 8 > (256ull << 20)) waves /= 2;
  uint64_t* scratch;
  CS(sc.alloc(&scratch, per_wave * waves));

  CK(hipMemsetAsync(root_flag, 0, slots * 4, s));
  CK(hipMemsetAsync(owner_root1, 0, slots * 4, s));
  CK(hipMemsetAsync(ctr, 0, 32, s));
  CK(hipMemsetAsync(raw_count, 0, 8, s));
  const uint32_t one = 1;
  CK(hipMemcpyAsync(root_flag + tlas / 64, &one, 4, hipMemcpyHostToDevice, s));
  uint64_t e0 = tlas / 64;
  CK(hipMemcpyAsync(fa, &e0, 8, hipMemcpyHostToDevice, s));
  uint32_t n_in = 1;
  bool first = true;
  while (n_in > 0) {
    CK(hipMemsetAsync(ctr, 0, 4, s));      // n_out
    CK(hipMemsetAsync(ctr + 3, 0, 4, s));  // work counter
    FormParams fp{A, bytes, tlas, budget, first ? 1u : 0u, fa, n_in, fb, ctr, root_flag, owner_root1, raw_list,
                  raw_cap, raw_count, desc, ctr + 2, scratch, qcap, lcap, ctr + 3, ctr + 1};
    const uint32_t grid = std::min<uint32_t>(waves / 4, (n_in + 3) / 4);
    hipLaunchKernelGGL(form_treelets, dim3(std::max(grid, 1u)), dim3(256), 0, s, fp);
    CK(hipGetLastError());
    uint32_t c[2];
    CK(hipMemcpyAsync(c, ctr, 8, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    if (c[1]) return map_err(c[1]);
    n_in = c[0];
    std::swap(fa, fb);
    first = false;
  }
  uint32_t n_desc = 0;
  CK(hipMemcpyAsync(&n_desc, ctr + 2, 4, hipMemcpyDeviceToHost, s));

  // dense numbering of roots in ascending offset order
  uint64_t *rank, *sums, *tot;
  CS(sc.alloc(&rank, slots));
  CS(sc.alloc(&sums, slots / kTile + 2));
  CS(sc.alloc(&tot, 2));
  CK(exclusive_scan(FlagIn{root_flag}, slots, rank, sums, tot, s));
  uint64_t n_roots = 0;
  CK(hipMemcpyAsync(&n_roots, tot, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  if (n_roots != n_desc || n_roots == 0) return RTX_ERR_HIP;
  const uint32_t n = (uint32_t)n_roots;

  Partition p;
  p.budget = budget;
  p.n = n;
  CS(dmalloc(&p.owner, slots + (slots + 3) / 4));  // owner words, then the first-treelet bytes
  CS(dmalloc(&p.roots, n));
  CS(dmalloc(&p.list_offsets, (size_t)n + 1));
  hipLaunchKernelGGL(owner_and_roots, dim3(2048), dim3(256), 0, s, slots, root_flag, rank, owner_root1, p.owner,
                     p.roots);
  CK(hipGetLastError());
  uint32_t *dlen, *desc_of_rank;
  CS(sc.alloc(&dlen, n));
  CS(sc.alloc(&desc_of_rank, n));
  const uint32_t g = std::min<uint32_t>(4096, (n + 3) / 4);
  hipLaunchKernelGGL(dedupe_count, dim3(g), dim3(256), 0, s, desc, n, raw_list, rank, dlen, desc_of_rank, shared);
  CK(hipGetLastError());
  CK(exclusive_scan(U32In{dlen}, n, p.list_offsets, sums, p.list_offsets + n, s));
  uint64_t n_list = 0;
  CK(hipMemcpyAsync(&n_list, p.list_offsets + n, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  p.n_list = n_list;
  CS(dmalloc(&p.list_nodes, n_list));
  hipLaunchKernelGGL(dedupe_copy, dim3(g), dim3(256), 0, s, desc, n, raw_list, desc_of_rank, p.list_offsets,
                     p.list_nodes, shared);
  CK(hipGetLastError());

  // summary
  std::vector<uint4> hd(n_desc);
  std::vector<uint32_t> hl(n);
  CK(hipMemcpyAsync(hd.data(), desc, (size_t)n_desc * sizeof(uint4), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(hl.data(), dlen, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  uint32_t ft = 0;
  {
    uint64_t r = 0;
    CK(hipMemcpyAsync(&r, rank + tlas / 64, 8, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    ft = (uint32_t)r;
  }
  p.first_treelet = ft;
  rtx_treelet_summary sm{};
  sm.n_treelets = n;
  sm.budget_bytes = budget;
  sm.n_list_entries = n_list;
  for (const uint4& d : hd) {
    sm.total_bytes += d.w;
    sm.max_treelet_bytes = std::max(sm.max_treelet_bytes, d.w);
  }
  for (uint32_t v : hl) sm.max_treelet_nodes = std::max(sm.max_treelet_nodes, v);
  p.summary = sm;
  rtx_status st = build_images(A, bytes, s, kind, p);
  if (st != RTX_OK) {
    partition_free(p);
    return st;
  }
  CK(hipStreamSynchronize(s));
  p.valid = true;
  out = p;
  return RTX_OK;
}

rtx_status partition_copy_out(const Partition& p, uint64_t bytes, hipStream_t s, uint64_t* roots, uint64_t* lo,
                              uint64_t* ln, uint32_t* owner) {
  if (!p.valid) return RTX_ERR_NO_PARTITION;
  if (roots) CK(hipMemcpyAsync(roots, p.roots, (size_t)p.n * 8, hipMemcpyDeviceToHost, s));
  if (lo) CK(hipMemcpyAsync(lo, p.list_offsets, ((size_t)p.n + 1) * 8, hipMemcpyDeviceToHost, s));
  if (ln && p.n_list) CK(hipMemcpyAsync(ln, p.list_nodes, p.n_list * 8, hipMemcpyDeviceToHost, s));
  if (owner) CK(hipMemcpyAsync(owner, p.owner, (bytes / 64) * 4, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  return RTX_OK;
}

rtx_status partition_install(const uint8_t* A, uint64_t bytes, uint64_t tlas, uint32_t budget, uint32_t n,
                             const uint64_t* roots, const uint64_t* lo, const uint64_t* ln, const uint32_t* owner,
                             hipStream_t s, Partition& out) {
  const uint64_t slots = bytes / 64;
  // host-side consistency checks
  for (uint32_t i = 0; i < n; ++i) {
    if (roots[i] % 64 || roots[i] >= bytes || (i && roots[i] <= roots[i - 1])) return RTX_ERR_INVALID_ARGUMENT;
    if (lo[i + 1] < lo[i]) return RTX_ERR_INVALID_ARGUMENT;
  }
  if (lo[0] != 0) return RTX_ERR_INVALID_ARGUMENT;
  const uint64_t n_list = lo[n];
  for (uint64_t i = 0; i < n_list; ++i)
    if (ln[i] % 64 || ln[i] >= bytes) return RTX_ERR_INVALID_ARGUMENT;
  // Each list must be de-duplicated (VRT:1312-1356 removes repeats): a node
  // listed twice would get two image positions, and image_layout keeps only one.
  {
    std::vector<uint64_t> sorted;
    for (uint32_t i = 0; i < n; ++i) {
      sorted.assign(ln + lo[i], ln + lo[i + 1]);
      std::sort(sorted.begin(), sorted.end());
      if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return RTX_ERR_INVALID_ARGUMENT;
    }
  }
  uint32_t ft = RTX_NO_TREELET;
  for (uint32_t i = 0; i < n; ++i)
    if (roots[i] == tlas) ft = i;
  if (ft == RTX_NO_TREELET) return RTX_ERR_INVALID_ARGUMENT;
  for (uint64_t sidx = 0; sidx < slots; ++sidx)
    if (owner[sidx] != RTX_NO_TREELET && owner[sidx] >= n) return RTX_ERR_INVALID_ARGUMENT;
  partition_free(out);
  Scratch sc;
  uint8_t* kind;
  CS(sc.alloc(&kind, slots));
  CS(run_validate(A, bytes, tlas, s, sc, kind, nullptr, nullptr));
  Partition p;
  p.budget = budget;
  p.n = n;
  p.n_list = n_list;
  p.first_treelet = ft;
  CS(dmalloc(&p.owner, slots + (slots + 3) / 4));  // owner words, then the first-treelet bytes
  CS(dmalloc(&p.roots, n));
  CS(dmalloc(&p.list_offsets, (size_t)n + 1));
  CS(dmalloc(&p.list_nodes, n_list));
  CK(hipMemcpyAsync(p.owner, owner, slots * 4, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(p.roots, roots, (size_t)n * 8, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(p.list_offsets, lo, ((size_t)n + 1) * 8, hipMemcpyHostToDevice, s));
  if (n_list) CK(hipMemcpyAsync(p.list_nodes, ln, n_list * 8, hipMemcpyHostToDevice, s));
  rtx_treelet_summary sm{};
  sm.n_treelets = n;
  sm.budget_bytes = budget;
  sm.n_list_entries = n_list;
  for (uint32_t i = 0; i < n; ++i) sm.max_treelet_nodes = std::max<uint32_t>(sm.max_treelet_nodes, (uint32_t)(lo[i + 1] - lo[i]));
  p.summary = sm;
  rtx_status st = build_images(A, bytes, s, kind, p);
  if (st != RTX_OK) {
    partition_free(p);
    return st;
  }
  CK(hipStreamSynchronize(s));
  p.valid = true;
  out = p;
  return RTX_OK;
}

}  // namespace rtxk
