// coherence.hip — the RT unit's ray coherence engine as a claim order.
//
// The reference's engine (src/gpgpu-sim/ray_coherency_engine.cc) hashes each
// ray a warp brings in (:43), pools rays of equal hash in insertion order
// (:46-54), and whenever a packet slot is free moves up to warp_size rays
// from the largest pool into it (:155-175, get_largest_packet :182-193: the
// first maximum in ascending hash order).  For a batch whose rays are all
// present, that schedule is a sort: pool g of size S contributes chunks
// k = 0, 1, ... with S - k * W rays left before each, and always taking the
// largest remaining pool merges those decreasing sequences, so the chunks
// come out ordered by (rays left before the chunk, descending; hash,
// ascending), each chunk holding its pool's next W rays in input order.
//
// Here: a hash per ray (the four functions of :393-470), a stable radix sort
// by hash, pool starts and sizes, and a second stable radix sort by the chunk
// key (~rays left << 32 | hash).  The permutation is rtx_trace's claim order:
// rays of one packet start together in one wave.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "kernels.h"
#include "rtx.h"

namespace rtxk {
namespace {

struct CohParams {
  const rtx_ray* rays;
  uint32_t n;
  uint32_t hash;  // 'f', 'g', 't', 'd'
  uint32_t francois_bits, grid_bits, sphere_bits;
  float two_point_ratio;
  float wmin[3], wmax[3];
};

// hash_comp (:393-402): sign, the top `b` exponent bits and the top `b`
// mantissa bits of the float's pattern
__device__ __forceinline__ uint64_t hash_comp(float x, uint32_t b) {
  const uint32_t mask = 0xffffffffu >> (32 - b);
  const uint32_t o = __float_as_uint(x);
  const uint64_t sign = o >> 31;
  const uint64_t ex = (o >> (31 - b)) & mask;
  const uint64_t mant = (o >> (23 - b)) & mask;
  return (sign << (2 * b)) | (ex << b) | mant;
}

// float -> uint64 as the reference's implicit conversions; a NaN (a world
// axis of zero extent) becomes 0, where the reference's conversion is undefined
__device__ __forceinline__ uint64_t to_u64(float v) { return __builtin_isnan(v) ? 0ull : (uint64_t)v; }

// clamp (vector-math.cc:68-71): std::max(std::min(x, hi), lo)
__device__ __forceinline__ float vm_clamp(float x, float lo, float hi) {
  const float m = hi < x ? hi : x;
  return m < lo ? lo : m;
}

// hash_origin_grid (:423-430)
__device__ __forceinline__ uint64_t hash_grid(const CohParams& c, float ox, float oy, float oz, uint32_t b) {
  const uint32_t g = 1u << b;
  const float gs = (float)g, top = (float)g - 1.f;
  const uint64_t x = to_u64(vm_clamp((ox - c.wmin[0]) / (c.wmax[0] - c.wmin[0]) * gs, 0.f, top));
  const uint64_t y = to_u64(vm_clamp((oy - c.wmin[1]) / (c.wmax[1] - c.wmin[1]) * gs, 0.f, top));
  const uint64_t z = to_u64(vm_clamp((oz - c.wmin[2]) / (c.wmax[2] - c.wmin[2]) * gs, 0.f, top));
  return (x << (2 * b)) | (y << b) | z;
}

// hash_direction_spherical (:408-421)
__device__ __forceinline__ uint64_t hash_sphere(float dx, float dy, float dz, uint32_t b) {
  const float kPi = 3.14159265358979323846f;
  const uint64_t theta = to_u64(acosf(vm_clamp(dz, -1.f, 1.f)) / kPi * 180.f);
  const uint64_t phi = to_u64((atan2f(dy, dx) + kPi) / kPi * 180.f);
  const uint64_t q_theta = theta >> (8 - b);
  const uint64_t q_phi = phi >> (9 - (b + 1));
  return (q_phi << b) | q_theta;
}

__global__ __launch_bounds__(256) void k_ray_hash(CohParams c, uint32_t* keys, uint32_t* idx) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
    const rtx_ray r = c.rays[i];
    const float ox = r.org[0], oy = r.org[1], oz = r.org[2];
    const float dx = r.dir[0], dy = r.dir[1], dz = r.dir[2];
    uint64_t h = 0;
    if (c.hash == 'f') {  // hash_francois (:432-445)
      const uint32_t b = c.francois_bits, cb = 2 * b + 1;
      const uint64_t hd = (hash_comp(dz, b) << (2 * cb)) | (hash_comp(dy, b) << cb) | hash_comp(dx, b);
      const uint64_t ho = (hash_comp(ox, b) << (2 * cb)) | (hash_comp(oy, b) << cb) | hash_comp(oz, b);
      h = ho ^ hd;
    } else if (c.hash == 'g') {  // hash_grid_spherical (:447-456)
      h = hash_grid(c, ox, oy, oz, c.grid_bits) ^ hash_sphere(dx, dy, dz, c.sphere_bits);
    } else if (c.hash == 't') {  // hash_two_point (:462-469)
      const float ex = c.wmax[0] - c.wmin[0], ey = c.wmax[1] - c.wmin[1], ez = c.wmax[2] - c.wmin[2];
      const float mx = ex > ey ? ex : ey;  // std::max(std::max(x, y), z)
      const float ext = mx > ez ? mx : ez;
      const float s = c.two_point_ratio * ext;
      h = hash_grid(c, ox, oy, oz, c.grid_bits) ^ hash_grid(c, ox + s * dx, oy + s * dy, oz + s * dz, c.grid_bits);
    } else {  // hash_direction_only (:471-474)
      h = hash_sphere(dx, dy, dz, c.sphere_bits);
    }
    keys[i] = (uint32_t)h;  // the host keeps every configuration's hash below 2^32
    idx[i] = i;
  }
}

// pool start of each sorted position (positions where the hash changes)
__global__ __launch_bounds__(256) void k_pool_heads(const uint32_t* h, uint32_t n, uint32_t* head_pos) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    head_pos[p] = (p == 0 || h[p] != h[p - 1]) ? p : 0u;
}

// chunk key of each sorted position: rays left in its pool before its chunk
// (descending, so inverted) and the hash
__global__ __launch_bounds__(256) void k_chunk_keys(const uint32_t* h, const uint32_t* start, uint32_t n,
                                                    uint32_t packet, unsigned long long* key) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const uint32_t s = start[p];
    // the pool ends where the next pool starts (or at n)
    uint32_t lo = p, hi = n;  // first position > p whose start differs: binary search on h
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (h[mid] == h[p])
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t size = hi - s;
    const uint32_t left = size - ((p - s) / packet) * packet;
    key[p] = ((unsigned long long)(~left) << 32) | h[p];
  }
}

struct MaxOp {
  __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

}  // namespace

hipError_t coherence_order(const rtx_ray* rays, uint32_t n, const rtx_coherence_config& cfg, const float wmin[3],
                           const float wmax[3], uint32_t* order, uint32_t* hashes_out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  CohParams c{};
  c.rays = rays;
  c.n = n;
  c.hash = (uint32_t)(uint8_t)cfg.hash;
  c.francois_bits = cfg.francois_bits;
  c.grid_bits = cfg.grid_bits;
  c.sphere_bits = cfg.sphere_bits;
  c.two_point_ratio = cfg.two_point_ratio;
  for (int a = 0; a < 3; ++a) {
    c.wmin[a] = wmin[a];
    c.wmax[a] = wmax[a];
  }
  const uint32_t packet = cfg.packet ? cfg.packet : 64u;
  uint32_t *k0 = nullptr, *k1 = nullptr, *i0 = nullptr, *i1 = nullptr, *st = nullptr;
  unsigned long long* ck = nullptr;
  unsigned long long* ck1 = nullptr;
  void* tmp = nullptr;
  auto cleanup = [&]() {
    hipFree(k0);
    hipFree(k1);
    hipFree(i0);
    hipFree(i1);
    hipFree(st);
    hipFree(ck);
    hipFree(ck1);
    hipFree(tmp);
  };
  hipError_t e;
#define COH_TRY(expr)    \
  do {                   \
    e = (expr);          \
    if (e != hipSuccess) { \
      cleanup();         \
      return e;          \
    }                    \
  } while (0)
  COH_TRY(hipMalloc(&k0, n * 4ull));
  COH_TRY(hipMalloc(&k1, n * 4ull));
  COH_TRY(hipMalloc(&i0, n * 4ull));
  COH_TRY(hipMalloc(&i1, n * 4ull));
  COH_TRY(hipMalloc(&st, n * 4ull));
  COH_TRY(hipMalloc(&ck, n * 8ull));
  COH_TRY(hipMalloc(&ck1, n * 8ull));
  size_t b1 = 0, b2 = 0, b3 = 0;
  COH_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, k0, k1, i0, i1, (int)n, 0, 32, s));
  COH_TRY(hipcub::DeviceScan::InclusiveScan(nullptr, b2, st, st, MaxOp(), (int)n, s));
  COH_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b3, ck, ck1, i1, order, (int)n, 0, 64, s));
  COH_TRY(hipMalloc(&tmp, std::max(b1, std::max(b2, b3))));
  const uint32_t g = std::max(1u, std::min<uint32_t>(8192, (n + 255) / 256));
  hipLaunchKernelGGL(k_ray_hash, dim3(g), dim3(256), 0, s, c, k0, i0);
  COH_TRY(hipGetLastError());
  // pools: rays of one hash, in input order (the radix sort is stable)
  COH_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b1, k0, k1, i0, i1, (int)n, 0, 32, s));
  hipLaunchKernelGGL(k_pool_heads, dim3(g), dim3(256), 0, s, k1, n, st);
  COH_TRY(hipGetLastError());
  COH_TRY(hipcub::DeviceScan::InclusiveScan(tmp, b2, st, st, MaxOp(), (int)n, s));
  hipLaunchKernelGGL(k_chunk_keys, dim3(g), dim3(256), 0, s, k1, st, n, packet, ck);
  COH_TRY(hipGetLastError());
  // chunks: most rays left first, then ascending hash; a chunk's rays stay in order
  COH_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b3, ck, ck1, i1, order, (int)n, 0, 64, s));
  if (hashes_out) COH_TRY(hipMemcpyAsync(hashes_out, k0, n * 4ull, hipMemcpyDeviceToDevice, s));
  COH_TRY(hipStreamSynchronize(s));
#undef COH_TRY
  cleanup();
  return hipSuccess;
}

}  // namespace rtxk
