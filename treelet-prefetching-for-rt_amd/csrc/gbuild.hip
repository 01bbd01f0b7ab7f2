// gbuild.hip — device builder for the bottom-level structures.
//
// Every BLAS of a scene is built in the same launches (the reference leaves
// BVH construction to Embree inside mesa-vulkan-sim; this is the device
// counterpart of the host SAH builder in builder.cpp):
//   1. centroid bounds per BLAS, 30-bit Morton codes inside them, and one
//      radix sort of (BLAS << 30 | Morton, triangle) pairs;
//   2. a binary radix tree per BLAS (Karras, "Maximizing parallelism in the
//      construction of BVHs, octrees, and k-d trees", HPG 2012), built one
//      thread per internal node, ties broken by sorted position;
//   3. bounding boxes bottom-up, the second child to arrive forming the parent;
//   4. 6-wide nodes level by level: each node opens its children with the
//      largest surface area until it has six (as the host builder does), the
//      children of a level are placed contiguously after an exclusive scan,
//      and each node is written with the conservative 8-bit child bounds of the
//      node layout (vulkan_acceleration_structure_util.h).
// The BLAS of all instances share one pool: slot b holds the header of BLAS b,
// slot B + b its root, and every later level follows.  Offsets inside the pool
// are relative, so the pool drops into an arena after the TLAS unchanged.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <vector>

#include "layout.h"
#include "rtx.h"

using namespace rtxl;

namespace {

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kNoParent = 0xffffffffu;

__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

struct BuildParams {
  const float* tris;        // 9 floats per triangle
  const uint32_t* blas_of;  // triangle -> BLAS
  const uint32_t* pofs;     // [B + 1] first triangle of each BLAS
  const uint32_t* iofs;     // [B + 1] first internal node of each BLAS (pofs[b] - b)
  uint32_t n_tris, n_blas, n_int;
  uint32_t* cmin;           // [3B] ordered-int centroid bounds
  uint32_t* cmax;
  unsigned long long* keys;
  uint32_t* vals;
  const unsigned long long* skeys;  // sorted
  const uint32_t* svals;            // sorted triangle ids
  uint32_t* left;    // [n_int] child refs (kLeafFlag | sorted position for leaves)
  uint32_t* right;
  uint32_t* iparent;  // [n_int]
  uint32_t* lparent;  // [n_tris]
  float* ibox;        // [6 n_int]
  float* lbox;        // [6 n_tris]
  uint32_t* flags;    // [n_int]
};

// BLAS b owns triangles [pofs[b], pofs[b+1]): one workgroup per BLAS writes
// its id over that range.
__global__ void k_blas_of(BuildParams p, uint32_t* blas_of) {
  const uint32_t b = blockIdx.x;
  for (uint32_t t = p.pofs[b] + threadIdx.x; t < p.pofs[b + 1]; t += blockDim.x) blas_of[t] = b;
}

// Centroid bounds per BLAS as ordered-int atomics.  Triangles of a BLAS are
// contiguous, so most waves hold one BLAS: those reduce across the wave first
// and issue one atomic per bound instead of 64.
__global__ void k_centroid_bounds(BuildParams p) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = t < p.n_tris;
  const uint32_t b = ok ? p.blas_of[t] : 0xffffffffu;
  uint32_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = 0xffffffffu;
    hi[a] = 0u;
    if (ok) {
      const float* v = p.tris + 9 * (size_t)t;
      lo[a] = hi[a] = fkey((v[a] + v[3 + a] + v[6 + a]) * (1.0f / 3.0f));
    }
  }
  // lane 0 holds the lowest t of the wave, so it is valid whenever any lane is
  const uint32_t b0 = __shfl(b, 0);
  if (__all(!ok || b == b0)) {
    if (b0 == 0xffffffffu) return;
    for (int off = 32; off >= 1; off >>= 1)
      for (int a = 0; a < 3; ++a) {
        lo[a] = min(lo[a], (uint32_t)__shfl_xor((int)lo[a], off));
        hi[a] = max(hi[a], (uint32_t)__shfl_xor((int)hi[a], off));
      }
    if ((threadIdx.x & 63u) == 0u)
      for (int a = 0; a < 3; ++a) {
        atomicMin(&p.cmin[3 * b0 + a], lo[a]);
        atomicMax(&p.cmax[3 * b0 + a], hi[a]);
      }
  } else if (ok) {
    for (int a = 0; a < 3; ++a) {
      atomicMin(&p.cmin[3 * b + a], lo[a]);
      atomicMax(&p.cmax[3 * b + a], hi[a]);
    }
  }
}

__global__ void k_morton(BuildParams p) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.n_tris) return;
  const float* v = p.tris + 9 * (size_t)t;
  const uint32_t b = p.blas_of[t];
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    const float c = (v[a] + v[3 + a] + v[6 + a]) * (1.0f / 3.0f);
    const float lo = funkey(p.cmin[3 * b + a]), hi = funkey(p.cmax[3 * b + a]);
    float x = hi > lo ? (c - lo) / (hi - lo) * 1024.0f : 0.0f;
    x = fminf(fmaxf(x, 0.0f), 1023.0f);
    q[a] = (uint32_t)x;
  }
  const uint32_t m = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  p.keys[t] = ((unsigned long long)b << 30) | m;
  p.vals[t] = t;
}

// Common prefix length of the codes at local positions i and j of a segment
// of n codes (-1 outside), extended by the positions on equal codes.
__device__ __forceinline__ int delta(const unsigned long long* k, uint32_t base, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = (uint32_t)(k[base + i] & 0x3fffffffu), b = (uint32_t)(k[base + j] & 0x3fffffffu);
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

__global__ void k_karras(BuildParams p) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= p.n_int) return;
  // segment: the last b with iofs[b] <= gi (empty segments share iofs with the next)
  uint32_t lo = 0, hi = p.n_blas;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (p.iofs[mid] <= gi)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t b = lo;
  const uint32_t base = p.pofs[b];
  const int n = (int)(p.pofs[b + 1] - base);
  const uint32_t ibase = p.iofs[b];
  const int i = (int)(gi - ibase);
  const unsigned long long* k = p.skeys;
  const int d = (delta(k, base, n, i, i + 1) - delta(k, base, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(k, base, n, i, i - d);
  int lmax = 2;
  while (delta(k, base, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(k, base, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(k, base, n, i, j);
  int s = 0;
  for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
    if (delta(k, base, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  const int gamma = i + s * d + min(d, 0);
  const int mn = min(i, j), mx = max(i, j);
  const uint32_t lref = (mn == gamma) ? (kLeafFlag | (base + (uint32_t)gamma)) : (ibase + (uint32_t)gamma);
  const uint32_t rref = (mx == gamma + 1) ? (kLeafFlag | (base + (uint32_t)gamma + 1)) : (ibase + (uint32_t)gamma + 1);
  p.left[gi] = lref;
  p.right[gi] = rref;
  if (lref & kLeafFlag)
    p.lparent[lref & ~kLeafFlag] = gi;
  else
    p.iparent[lref] = gi;
  if (rref & kLeafFlag)
    p.lparent[rref & ~kLeafFlag] = gi;
  else
    p.iparent[rref] = gi;
}

__device__ __forceinline__ float ld_coherent(const float* a) {
  return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(a), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT));
}

// Box of the triangle at sorted position pos.
__device__ __forceinline__ void leaf_box(const BuildParams& p, uint32_t pos, float bx[6]) {
  const float* v = p.tris + 9 * (size_t)p.svals[pos];
  for (int a = 0; a < 3; ++a) {
    bx[a] = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
    bx[3 + a] = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
  }
}

// Largest BLAS (in triangles) whose boxes k_boxes_lds forms; k_boxes takes
// the larger ones.
constexpr uint32_t kLdsBoxLeaves = 2048;

// Boxes of one BLAS per workgroup, for BLAS of at most kLdsBoxLeaves
// triangles: the bottom-up climb of k_boxes with the arrival counters and the
// internal boxes in LDS (cap internal nodes: 6 floats each, then one counter
// each), so every hand-over stays inside the workgroup.  The second child to
// arrive forms the parent from both children, left then right, as k_boxes
// does, so both kernels give the same bits; a leaf child's box is recomputed
// from its triangle.
__global__ void k_boxes_lds(BuildParams p, uint32_t cap) {
  extern __shared__ uint32_t lds[];
  float* sbox = reinterpret_cast<float*>(lds);
  uint32_t* sflag = lds + 6 * cap;
  const uint32_t b = blockIdx.x;
  const uint32_t base = p.pofs[b], n = p.pofs[b + 1] - base, ibase = p.iofs[b];
  if (n > kLdsBoxLeaves || n - 1 > cap) return;
  for (uint32_t i = threadIdx.x; i + 1 < n; i += blockDim.x) sflag[i] = 0u;
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < n; l += blockDim.x) {
    const uint32_t pos = base + l;
    float u[6];
    leaf_box(p, pos, u);
    for (int c = 0; c < 6; ++c) p.lbox[6 * (size_t)pos + c] = u[c];
    uint32_t cur = kLeafFlag | pos, node = p.lparent[pos];
    while (node != kNoParent) {
      const uint32_t li = node - ibase;
      if (__hip_atomic_fetch_add(&sflag[li], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        break;  // the sibling forms the parent
      const uint32_t lref = p.left[node], rref = p.right[node];
      const uint32_t sib = lref == cur ? rref : lref;
      float s[6];
      if (sib & kLeafFlag)
        leaf_box(p, sib & ~kLeafFlag, s);
      else
        for (int c = 0; c < 6; ++c) s[c] = sbox[6 * (sib - ibase) + c];
      const bool own_left = lref == cur;
      for (int c = 0; c < 6; ++c) {
        const float l0 = own_left ? u[c] : s[c], r0 = own_left ? s[c] : u[c];
        u[c] = c < 3 ? fminf(fminf(INFINITY, l0), r0) : fmaxf(fmaxf(-INFINITY, l0), r0);
      }
      for (int c = 0; c < 6; ++c) {
        sbox[6 * li + c] = u[c];
        p.ibox[6 * (size_t)node + c] = u[c];
      }
      cur = node;
      node = p.iparent[node];
    }
  }
}

// Boxes bottom-up over the leaves of BLAS larger than kLdsBoxLeaves (all
// BLAS when lds_leaves is 0), through device-wide counters.
__global__ void k_boxes(BuildParams p, uint32_t lds_leaves) {
  const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= p.n_tris) return;
  if (lds_leaves) {
    const uint32_t b = p.blas_of[p.svals[pos]];
    if (p.pofs[b + 1] - p.pofs[b] <= lds_leaves) return;
  }
  float bx[6];
  leaf_box(p, pos, bx);
  for (int c = 0; c < 6; ++c) p.lbox[6 * (size_t)pos + c] = bx[c];
  uint32_t node = p.lparent[pos];
  while (node != kNoParent) {
    __threadfence();
    if (atomicAdd(&p.flags[node], 1u) == 0u) return;  // the sibling forms the parent
    __threadfence();
    float u[6];
    const uint32_t refs[2] = {p.left[node], p.right[node]};
    for (int c = 0; c < 6; ++c) u[c] = c < 3 ? INFINITY : -INFINITY;
    for (int r = 0; r < 2; ++r) {
      const float* src = (refs[r] & kLeafFlag) ? p.lbox + 6 * (size_t)(refs[r] & ~kLeafFlag) : p.ibox + 6 * (size_t)refs[r];
      for (int c = 0; c < 3; ++c) u[c] = fminf(u[c], ld_coherent(src + c));
      for (int c = 3; c < 6; ++c) u[c] = fmaxf(u[c], ld_coherent(src + c));
    }
    for (int c = 0; c < 6; ++c) p.ibox[6 * (size_t)node + c] = u[c];
    node = p.iparent[node];
  }
}

__device__ __forceinline__ const float* ref_box(const BuildParams& p, uint32_t r) {
  return (r & kLeafFlag) ? p.lbox + 6 * (size_t)(r & ~kLeafFlag) : p.ibox + 6 * (size_t)r;
}

__device__ __forceinline__ double box_area(const float* b) {
  const double dx = (double)b[3] - b[0], dy = (double)b[4] - b[1], dz = (double)b[5] - b[2];
  if (!(dx >= 0.0)) return 0.0;
  return 2.0 * (dx * dy + dy * dz + dz * dx);
}

// Up to six tree references that become the children of the 6-wide node made
// from reference r: open the internal child with the largest area until six.
__device__ int collapse(const BuildParams& p, uint32_t r, uint32_t out[6]) {
  if (r & kLeafFlag) {
    out[0] = r;
    return 1;
  }
  int n = 2;
  out[0] = p.left[r];
  out[1] = p.right[r];
  while (n < 6) {
    int best = -1;
    double best_area = -1.0;
    for (int k = 0; k < n; ++k) {
      if (out[k] & kLeafFlag) continue;
      const double a = box_area(ref_box(p, out[k]));
      if (a > best_area) {
        best_area = a;
        best = k;
      }
    }
    if (best < 0) break;
    const uint32_t open = out[best];
    for (int k = n; k > best + 1; --k) out[k] = out[k - 1];
    out[best] = p.left[open];
    out[best + 1] = p.right[open];
    ++n;
  }
  return n;
}

struct LevelParams {
  const uint32_t* front;      // tree refs of this level's nodes
  const uint32_t* front_slot; // pool slot of each node
  uint32_t n_front;
  uint32_t* n_children;       // out: children per node
  uint32_t* n_internal;       // out: internal children per node
  const uint32_t* child_base; // exclusive scan of n_children
  const uint32_t* int_base;   // exclusive scan of n_internal
  uint32_t slot_cursor;       // first slot of this level's child blocks
  uint32_t* next_front;
  uint32_t* next_slot;
  uint8_t* pool;
  const uint32_t* geom_of;    // BLAS -> geometry index (per sorted triangle via blas_of)
  uint32_t* error;
};

__global__ void k_level_count(BuildParams p, LevelParams L) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n_front) return;
  uint32_t ch[6];
  const int n = collapse(p, L.front[i], ch);
  uint32_t ni = 0;
  for (int k = 0; k < n; ++k) ni += (ch[k] & kLeafFlag) ? 0u : 1u;
  L.n_children[i] = (uint32_t)n;
  L.n_internal[i] = ni;
}

__device__ __forceinline__ float decode_bound(float origin, int q, int e) { return origin + ldexpf((float)q, e - 8); }

__device__ bool quantize_axis(float origin, const float* lo, const float* hi, int n, int e, uint8_t* qlo,
                              uint8_t* qhi) {
  const double s = ldexp(1.0, e - 8);
  for (int k = 0; k < n; ++k) {
    long q = (long)floor(((double)lo[k] - (double)origin) / s);
    q = max(0L, min(255L, q));
    while (q > 0 && decode_bound(origin, (int)q, e) > lo[k]) --q;
    if (decode_bound(origin, (int)q, e) > lo[k]) return false;
    long qh = (long)ceil(((double)hi[k] - (double)origin) / s);
    qh = max(0L, qh);
    if (qh > 255) return false;
    while (qh < 255 && decode_bound(origin, (int)qh, e) < hi[k]) ++qh;
    if (decode_bound(origin, (int)qh, e) < hi[k]) return false;
    qlo[k] = (uint8_t)q;
    qhi[k] = (uint8_t)qh;
  }
  return true;
}

__device__ void st32(uint8_t* d, uint32_t v) {
  d[0] = (uint8_t)v;
  d[1] = (uint8_t)(v >> 8);
  d[2] = (uint8_t)(v >> 16);
  d[3] = (uint8_t)(v >> 24);
}
__device__ void stf(uint8_t* d, float f) { st32(d, __float_as_uint(f)); }

__global__ void k_level_emit(BuildParams p, LevelParams L) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n_front) return;
  uint32_t ch[6];
  const int n = collapse(p, L.front[i], ch);
  const uint32_t node_slot = L.front_slot[i];
  const uint32_t block = L.slot_cursor + L.child_base[i];
  uint32_t next = L.int_base[i];
  float lo[3][6], hi[3][6];
  for (int k = 0; k < n; ++k) {
    const float* b = ref_box(p, ch[k]);
    for (int a = 0; a < 3; ++a) {
      lo[a][k] = b[a];
      hi[a][k] = b[3 + a];
    }
  }
  uint8_t* nd = L.pool + (size_t)node_slot * 64;
  for (int c = 0; c < 64; ++c) nd[c] = 0;
  float origin[3];
  int8_t exps[3];
  uint8_t q[6][6];
  for (int a = 0; a < 3; ++a) {
    origin[a] = INFINITY;
    double extent = 0.0;
    for (int k = 0; k < n; ++k) origin[a] = fminf(origin[a], lo[a][k]);
    for (int k = 0; k < n; ++k) extent = fmax(extent, (double)hi[a][k] - (double)origin[a]);
    int e = -128;
    if (extent > 0.0) e = max(-128, min(127, (int)ceil(log2(extent)) - 1));
    while (!quantize_axis(origin[a], lo[a], hi[a], n, e, q[2 * a], q[2 * a + 1])) {
      if (++e > 127) {
        atomicOr(L.error, 1u);
        return;
      }
    }
    exps[a] = (int8_t)e;
  }
  stf(nd + kInOrigin + 0, origin[0]);
  stf(nd + kInOrigin + 4, origin[1]);
  stf(nd + kInOrigin + 8, origin[2]);
  st32(nd + kInChildOffset, (uint32_t)(int32_t)(block - node_slot));
  nd[kInNodeType] = (uint8_t)kTypeInternal;
  for (int a = 0; a < 3; ++a) nd[kInExp + a] = (uint8_t)exps[a];
  nd[kInMask] = 0xff;
  const uint32_t field[6] = {kInLowerX, kInUpperX, kInLowerY, kInUpperY, kInLowerZ, kInUpperZ};
  for (int k = 0; k < 6; ++k) {
    uint8_t info = 0;
    if (k < n) info = (uint8_t)(1u | (((ch[k] & kLeafFlag) ? kTypeQuad : kTypeInternal) << 2));
    nd[kInChildInfo + k] = info;
    for (int f = 0; f < 6; ++f) nd[field[f] + k] = k < n ? q[f][k] : 0;
  }
  for (int k = 0; k < n; ++k) {
    const uint32_t slot = block + (uint32_t)k;
    if (ch[k] & kLeafFlag) {
      // quad leaf with one triangle (vertex 3 repeats vertex 2)
      const uint32_t pos = ch[k] & ~kLeafFlag;
      const uint32_t tri = p.svals[pos];
      const uint32_t b = p.blas_of[tri];
      const float* v = p.tris + 9 * (size_t)tri;
      uint8_t* lf = L.pool + (size_t)slot * 64;
      for (int c = 0; c < 64; ++c) lf[c] = 0;
      st32(lf + 0, 0xff000000u);
      st32(lf + kQlDescriptor1, (L.geom_of[b] & 0x0fffffffu) | (1u << 30));
      st32(lf + kQlPrim0, tri - p.pofs[b]);
      st32(lf + kQlPrim1Delta, (1u << 19) | (2u << 21));
      for (int c = 0; c < 9; ++c) stf(lf + kQlV0 + 4 * c, v[c]);
      for (int c = 0; c < 3; ++c) stf(lf + kQlV0 + 36 + 4 * c, v[6 + c]);
    } else {
      L.next_front[next] = ch[k];
      L.next_slot[next] = slot;
      ++next;
    }
  }
}

__global__ void k_headers(BuildParams p, uint8_t* pool, float* root_boxes, uint32_t* roots, uint32_t* slots) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.n_blas) return;
  const uint32_t n = p.pofs[b + 1] - p.pofs[b];
  const uint32_t r = n >= 2 ? p.iofs[b] : (kLeafFlag | p.pofs[b]);
  roots[b] = r;
  slots[b] = p.n_blas + b;
  const float* bx = ref_box(p, r);
  uint8_t* h = pool + (size_t)b * 64;
  for (int c = 0; c < 64; ++c) h[c] = 0;
  const uint64_t root_off = (uint64_t)p.n_blas * 64;  // header b -> root slot B + b
  st32(h + kHdrRootOffset, (uint32_t)root_off);
  st32(h + kHdrRootOffset + 4, (uint32_t)(root_off >> 32));
  for (int a = 0; a < 3; ++a) {
    stf(h + kHdrBoundsMin + 4 * a, bx[a]);
    stf(h + kHdrBoundsMax + 4 * a, bx[3 + a]);
    root_boxes[6 * b + a] = bx[a];
    root_boxes[6 * b + 3 + a] = bx[3 + a];
  }
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  bool alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess; }
};

}  // namespace

namespace rtx_internal {

// Builds the bottom levels of B meshes on the device.  tris: all triangles,
// 9 floats each, BLAS b owning [pofs[b], pofs[b+1]).  Returns the pool in
// device memory (*pool_dev, *pool_bytes; header of BLAS b at byte 64 b, freed
// by the caller with hipFree) and each BLAS's bounds.
rtx_status gpu_build_blas_pool(hipStream_t s, const float* tris_host, const uint32_t* pofs_host,
                               const uint32_t* geom_host, uint32_t n_blas, uint8_t** pool_dev, uint64_t* pool_bytes,
                               std::vector<float>& boxes_out) {
  if (n_blas == 0 || !pool_dev || !pool_bytes) return RTX_ERR_INVALID_ARGUMENT;
  *pool_dev = nullptr;
  *pool_bytes = 0;
  const uint32_t N = pofs_host[n_blas];
  for (uint32_t b = 0; b < n_blas; ++b)
    if (pofs_host[b + 1] <= pofs_host[b]) return RTX_ERR_INVALID_ARGUMENT;  // every BLAS holds a triangle
  const uint32_t n_int = N - n_blas;
  std::vector<uint32_t> iofs(n_blas + 1);
  // lds_cap: internal nodes of the largest BLAS that k_boxes_lds takes
  uint32_t lds_cap = 0, max_n = 0, min_n = 0xffffffffu;
  for (uint32_t b = 0; b < n_blas; ++b) {
    iofs[b] = pofs_host[b] - b;
    const uint32_t n = pofs_host[b + 1] - pofs_host[b];
    max_n = std::max(max_n, n);
    min_n = std::min(min_n, n);
    if (n <= kLdsBoxLeaves) lds_cap = std::max(lds_cap, n - 1);
  }
  iofs[n_blas] = n_int;
  const uint64_t max_slots = 2ull * n_blas + 2ull * N;
  DevBuf<float> tris, ibox, lbox, rboxes;
  DevBuf<uint32_t> d_blas_of, d_pofs, d_iofs, d_geom, cmin, cmax, vals, svals, left, right, ipar, lpar, flags, err;
  DevBuf<unsigned long long> keys, skeys;
  DevBuf<uint32_t> f0, f1, s0, s1, nch, nin, cb, ib, roots;
  DevBuf<uint8_t> pool, tmp;
  const uint32_t fcap = std::max(n_blas, N);
  if (!tris.alloc((size_t)N * 9) || !ibox.alloc((size_t)n_int * 6) || !lbox.alloc((size_t)N * 6) ||
      !rboxes.alloc((size_t)n_blas * 6) || !d_blas_of.alloc(N) || !d_pofs.alloc(n_blas + 1) ||
      !d_iofs.alloc(n_blas + 1) || !d_geom.alloc(n_blas) || !cmin.alloc(3ull * n_blas) || !cmax.alloc(3ull * n_blas) ||
      !vals.alloc(N) || !svals.alloc(N) || !left.alloc(n_int) || !right.alloc(n_int) || !ipar.alloc(n_int) ||
      !lpar.alloc(N) || !flags.alloc(n_int) || !err.alloc(1) || !keys.alloc(N) || !skeys.alloc(N) ||
      !f0.alloc(fcap) || !f1.alloc(fcap) || !s0.alloc(fcap) || !s1.alloc(fcap) || !nch.alloc(fcap) ||
      !nin.alloc(fcap) || !cb.alloc(fcap) || !ib.alloc(fcap) || !roots.alloc(n_blas) || !pool.alloc(max_slots * 64)) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
#define GB_TRY(x)                 \
  do {                            \
    if ((x) != hipSuccess) {      \
      (void)hipGetLastError();    \
      return RTX_ERR_HIP;         \
    }                             \
  } while (0)
  GB_TRY(hipMemcpyAsync(tris.p, tris_host, (size_t)N * 36, hipMemcpyHostToDevice, s));
  GB_TRY(hipMemcpyAsync(d_pofs.p, pofs_host, (size_t)(n_blas + 1) * 4, hipMemcpyHostToDevice, s));
  GB_TRY(hipMemcpyAsync(d_iofs.p, iofs.data(), (size_t)(n_blas + 1) * 4, hipMemcpyHostToDevice, s));
  GB_TRY(hipMemcpyAsync(d_geom.p, geom_host, (size_t)n_blas * 4, hipMemcpyHostToDevice, s));
  GB_TRY(hipMemsetAsync(cmin.p, 0xff, 12ull * n_blas, s));
  GB_TRY(hipMemsetAsync(cmax.p, 0, 12ull * n_blas, s));
  GB_TRY(hipMemsetAsync(ipar.p, 0xff, (size_t)n_int * 4, s));
  GB_TRY(hipMemsetAsync(lpar.p, 0xff, (size_t)N * 4, s));
  GB_TRY(hipMemsetAsync(flags.p, 0, (size_t)n_int * 4, s));
  GB_TRY(hipMemsetAsync(err.p, 0, 4, s));

  BuildParams p{};
  p.tris = tris.p;
  p.blas_of = d_blas_of.p;
  p.pofs = d_pofs.p;
  p.iofs = d_iofs.p;
  p.n_tris = N;
  p.n_blas = n_blas;
  p.n_int = n_int;
  p.cmin = cmin.p;
  p.cmax = cmax.p;
  p.keys = keys.p;
  p.vals = vals.p;
  p.skeys = skeys.p;
  p.svals = svals.p;
  p.left = left.p;
  p.right = right.p;
  p.iparent = ipar.p;
  p.lparent = lpar.p;
  p.ibox = ibox.p;
  p.lbox = lbox.p;
  p.flags = flags.p;
  const uint32_t bs = 256;
  hipLaunchKernelGGL(k_blas_of, dim3(n_blas), dim3(bs), 0, s, p, d_blas_of.p);
  hipLaunchKernelGGL(k_centroid_bounds, dim3((N + bs - 1) / bs), dim3(bs), 0, s, p);
  hipLaunchKernelGGL(k_morton, dim3((N + bs - 1) / bs), dim3(bs), 0, s, p);
  GB_TRY(hipGetLastError());
  int end_bit = 30;
  while ((1ull << (end_bit - 30)) < n_blas) ++end_bit;
  size_t tmp_bytes = 0;
  GB_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.p, skeys.p, vals.p, svals.p, (int)N, 0, end_bit, s));
  size_t scan_bytes = 0;
  GB_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, nch.p, cb.p, (int)fcap, s));
  if (!tmp.alloc(std::max(tmp_bytes, scan_bytes))) {
    (void)hipGetLastError();
    return RTX_ERR_OUT_OF_MEMORY;
  }
  GB_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys.p, skeys.p, vals.p, svals.p, (int)N, 0, end_bit, s));
  if (n_int > 0) hipLaunchKernelGGL(k_karras, dim3((n_int + bs - 1) / bs), dim3(bs), 0, s, p);
  if (min_n <= kLdsBoxLeaves)
    hipLaunchKernelGGL(k_boxes_lds, dim3(n_blas), dim3(bs), (size_t)lds_cap * 28, s, p, lds_cap);
  if (max_n > kLdsBoxLeaves)
    hipLaunchKernelGGL(k_boxes, dim3((N + bs - 1) / bs), dim3(bs), 0, s, p, kLdsBoxLeaves);
  hipLaunchKernelGGL(k_headers, dim3((n_blas + bs - 1) / bs), dim3(bs), 0, s, p, pool.p, rboxes.p, f0.p, s0.p);
  GB_TRY(hipGetLastError());

  uint32_t n_front = n_blas, cursor = 2 * n_blas;
  uint32_t *front = f0.p, *fslot = s0.p, *nfront = f1.p, *nslot = s1.p;
  while (n_front > 0) {
    LevelParams L{};
    L.front = front;
    L.front_slot = fslot;
    L.n_front = n_front;
    L.n_children = nch.p;
    L.n_internal = nin.p;
    L.child_base = cb.p;
    L.int_base = ib.p;
    L.slot_cursor = cursor;
    L.next_front = nfront;
    L.next_slot = nslot;
    L.pool = pool.p;
    L.geom_of = d_geom.p;
    L.error = err.p;
    hipLaunchKernelGGL(k_level_count, dim3((n_front + bs - 1) / bs), dim3(bs), 0, s, p, L);
    GB_TRY(hipGetLastError());
    size_t sb = scan_bytes;
    GB_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, sb, nch.p, cb.p, (int)n_front, s));
    sb = scan_bytes;
    GB_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, sb, nin.p, ib.p, (int)n_front, s));
    uint32_t last[4];
    GB_TRY(hipMemcpyAsync(&last[0], nch.p + n_front - 1, 4, hipMemcpyDeviceToHost, s));
    GB_TRY(hipMemcpyAsync(&last[1], cb.p + n_front - 1, 4, hipMemcpyDeviceToHost, s));
    GB_TRY(hipMemcpyAsync(&last[2], nin.p + n_front - 1, 4, hipMemcpyDeviceToHost, s));
    GB_TRY(hipMemcpyAsync(&last[3], ib.p + n_front - 1, 4, hipMemcpyDeviceToHost, s));
    GB_TRY(hipStreamSynchronize(s));
    const uint32_t total_children = last[0] + last[1], total_internal = last[2] + last[3];
    if ((uint64_t)cursor + total_children > max_slots || total_internal > fcap) return RTX_ERR_HIP;
    hipLaunchKernelGGL(k_level_emit, dim3((n_front + bs - 1) / bs), dim3(bs), 0, s, p, L);
    GB_TRY(hipGetLastError());
    cursor += total_children;
    n_front = total_internal;
    std::swap(front, nfront);
    std::swap(fslot, nslot);
  }
  uint32_t e = 0;
  GB_TRY(hipMemcpyAsync(&e, err.p, 4, hipMemcpyDeviceToHost, s));
  boxes_out.resize((size_t)n_blas * 6);
  GB_TRY(hipMemcpyAsync(boxes_out.data(), rboxes.p, boxes_out.size() * 4, hipMemcpyDeviceToHost, s));
  GB_TRY(hipStreamSynchronize(s));
#undef GB_TRY
  if (e) return RTX_ERR_INVALID_ARGUMENT;
  *pool_dev = pool.p;  // the caller frees it with hipFree
  pool.p = nullptr;
  *pool_bytes = (uint64_t)cursor * 64;
  return RTX_OK;
}

}  // namespace rtx_internal
