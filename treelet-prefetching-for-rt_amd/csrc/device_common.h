// device_common.h — FP32 arithmetic of the traversal, written so that every
// operation and its order match the reference expressions bit for bit.
//
// Compile with -ffp-contract=off (no FMA fusion), correctly rounded FP32
// division/sqrt and no denormal flushing.  Min/max are compare-selects with
// the operand order of the reference macros (vulkan_ray_tracing.h:53-56),
// never v_min/v_max, whose NaN and signed-zero behaviour differ.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace rtxd {

struct F3 {
  float x, y, z;
};

__device__ __forceinline__ F3 mk3(float a, float b, float c) { return F3{a, b, c}; }
// vector-math.cc:37-80
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 add3(F3 a, F3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }

// calculate_idir, vulkan_ray_tracing.cc:221-237 (exp2f(-80) == 2^-80 exactly)
__device__ __forceinline__ F3 calc_idir(F3 d) {
  const float ooeps = 0x1p-80f;
  F3 r;
  r.x = 1.0f / (fabsf(d.x) > ooeps ? d.x : copysignf(ooeps, d.x));
  r.y = 1.0f / (fabsf(d.y) > ooeps ? d.y : copysignf(ooeps, d.y));
  r.z = 1.0f / (fabsf(d.z) > ooeps ? d.z : copysignf(ooeps, d.z));
  return r;
}

// get_t_bound + magic_max7/min7 + ray_box_test, vulkan_ray_tracing.cc:183-257
__device__ __forceinline__ bool ray_box(F3 lo3, F3 hi3, F3 idir, F3 o, float tmin, float tmax, float& thit) {
  float lox = (lo3.x - o.x) * idir.x, loy = (lo3.y - o.y) * idir.y, loz = (lo3.z - o.z) * idir.z;
  float hix = (hi3.x - o.x) * idir.x, hiy = (hi3.y - o.y) * idir.y, hiz = (hi3.z - o.z) * idir.z;
  float t1 = sel_max(sel_min(lox, hix), tmin);
  float t2 = sel_max(sel_min(loy, hiy), t1);
  float mn = sel_max(sel_min(loz, hiz), t2);
  float u1 = sel_min(sel_max(lox, hix), tmax);
  float u2 = sel_min(sel_max(loy, hiy), u1);
  float mx = sel_min(sel_max(loz, hiz), u2);
  thit = mn;
  return mn <= mx;
}

// ray_box with the compare-selects done by v_min_f32 / v_max_f32.  The two
// agree whenever no operand is NaN: they can differ only in the sign of a
// zero result, and thit and the returned flag only ever enter comparisons, in
// which -0 == +0.  Callers take this form only when the ray (origin, inverse
// direction, t range) and the node origin hold no NaN or infinity, which keeps
// every slab product NaN-free.
__device__ __forceinline__ bool ray_box_fast(F3 lo3, F3 hi3, F3 idir, F3 o, float tmin, float tmax, float& thit) {
  float lox = (lo3.x - o.x) * idir.x, loy = (lo3.y - o.y) * idir.y, loz = (lo3.z - o.z) * idir.z;
  float hix = (hi3.x - o.x) * idir.x, hiy = (hi3.y - o.y) * idir.y, hiz = (hi3.z - o.z) * idir.z;
  float mn = __builtin_fmaxf(__builtin_fminf(loz, hiz),
                             __builtin_fmaxf(__builtin_fminf(loy, hiy), __builtin_fmaxf(__builtin_fminf(lox, hix), tmin)));
  float mx = __builtin_fminf(__builtin_fmaxf(loz, hiz),
                             __builtin_fminf(__builtin_fmaxf(loy, hiy), __builtin_fminf(__builtin_fmaxf(lox, hix), tmax)));
  thit = mn;
  return mn <= mx;
}

__device__ __forceinline__ bool finite3(F3 v) {
  return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

// A ray whose slab products can never be NaN (given a non-NaN node origin):
// a finite origin, and a finite direction, so that no inverse-direction
// component is zero (|idir| lies in [2^-128, 2^80]) and an infinite bound
// times idir stays infinite instead of becoming 0 * inf.
__device__ __forceinline__ bool slab_inputs_ok(F3 o, F3 d, F3 idir, float tmin, float tmax) {
  return finite3(o) && finite3(d) && finite3(idir) && !__builtin_isnan(tmin) && !__builtin_isnan(tmax);
}

// mt_ray_triangle_test, vulkan_ray_tracing.cc:3089-3111
__device__ __forceinline__ bool mt_test(F3 p0, F3 p1, F3 p2, F3 o, F3 d, float& thit) {
  F3 v0v1 = sub3(p1, p0);
  F3 v0v2 = sub3(p2, p0);
  F3 pvec = cross3(d, v0v2);
  float det = dot3(v0v1, pvec);
  float idet = 1.0f / det;
  F3 tvec = sub3(o, p0);
  float u = dot3(tvec, pvec) * idet;
  if (u < 0.0f || u > 1.0f) return false;
  F3 qvec = cross3(tvec, v0v1);
  float v = dot3(d, qvec) * idet;
  if (v < 0.0f || (u + v) > 1.0f) return false;
  thit = dot3(v0v2, qvec) * idet;
  return true;
}

// Barycentric, vulkan_ray_tracing.cc:3113-3130 -> (v, w, u)
__device__ __forceinline__ F3 barycentric(F3 p, F3 a, F3 b, F3 c) {
  F3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
  float d00 = dot3(v0, v0);
  float d01 = dot3(v0, v1);
  float d11 = dot3(v1, v1);
  float d20 = dot3(v2, v0);
  float d21 = dot3(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  float v = (d11 * d20 - d01 * d21) / denom;
  float w = (d00 * d21 - d01 * d20) / denom;
  float u = 1.0f - v - w;
  return mk3(v, w, u);
}

// make_transformed_ray, vulkan_ray_tracing.cc:169-181, with float4x4::operator*
// (util.h:47-55): res[i] = (((0 + m0i*v0) + m1i*v1) + m2i*v2) + m3i*v3, and
// the 4x4 matrix of instance_leaf_matrix_to_float4x4 (util.h:210-224): rows
// 0..3 hold the 12 stored floats in columns 0..2, column 3 is (0,0,0,1).
struct ObjRay {
  F3 o, d, idir;
  float tmin, tmax, tmult;
};

__device__ __forceinline__ ObjRay transform_ray(const float m[12], F3 o, F3 d, float tmin, float tmax) {
  // column i of the product (i = 0..2) and w (column 3)
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  ox += m[0] * o.x;
  ox += m[3] * o.y;
  ox += m[6] * o.z;
  ox += m[9] * 1.0f;
  oy += m[1] * o.x;
  oy += m[4] * o.y;
  oy += m[7] * o.z;
  oy += m[10] * 1.0f;
  oz += m[2] * o.x;
  oz += m[5] * o.y;
  oz += m[8] * o.z;
  oz += m[11] * 1.0f;
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  dx += m[0] * d.x;
  dx += m[3] * d.y;
  dx += m[6] * d.z;
  dx += m[9] * 0.0f;
  dy += m[1] * d.x;
  dy += m[4] * d.y;
  dy += m[7] * d.z;
  dy += m[10] * 0.0f;
  dz += m[2] * d.x;
  dz += m[5] * d.y;
  dz += m[8] * d.z;
  dz += m[11] * 0.0f;
  ObjRay r;
  // w = (((0 + 0*o.x) + 0*o.y) + 0*o.z) + 1 is exactly 1 for a finite origin,
  // and x / 1 = x: only a non-finite origin takes the divisions
  if (finite3(o)) {
    r.o = mk3(ox, oy, oz);
  } else {
    float ow = 0.0f;
    ow += 0.0f * o.x;
    ow += 0.0f * o.y;
    ow += 0.0f * o.z;
    ow += 1.0f * 1.0f;
    r.o = mk3(ox / ow, oy / ow, oz / ow);
  }
  float norm = sqrtf(dx * dx + dy * dy + dz * dz);
  r.tmult = norm;
  r.d = mk3(dx / norm, dy / norm, dz / norm);
  r.tmin = tmin * norm;
  r.tmax = tmax * norm;
  r.idir = calc_idir(r.d);
  return r;
}

// ---- node words ---------------------------------------------------------
struct Words16 {
  uint32_t w[16];
};

__device__ __forceinline__ Words16 load64B(const uint8_t* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1], c = q[2], e = q[3];
  Words16 r;
  r.w[0] = a.x; r.w[1] = a.y; r.w[2] = a.z; r.w[3] = a.w;
  r.w[4] = b.x; r.w[5] = b.y; r.w[6] = b.z; r.w[7] = b.w;
  r.w[8] = c.x; r.w[9] = c.y; r.w[10] = c.z; r.w[11] = c.w;
  r.w[12] = e.x; r.w[13] = e.y; r.w[14] = e.z; r.w[15] = e.w;
  return r;
}

__device__ __forceinline__ uint32_t byte_of(const Words16& n, int b) { return (n.w[b >> 2] >> (8 * (b & 3))) & 0xffu; }
__device__ __forceinline__ float f_of(const Words16& n, int word) { return __uint_as_float(n.w[word]); }

// set_child_bounds (util.h:499-508) on an internal node held in registers.
// Per-axis scales 2^(e-8) of a node's quantized child bounds.
__device__ __forceinline__ F3 child_scales(int ex, int ey, int ez) {
  return mk3(ldexpf(1.0f, ex - 8), ldexpf(1.0f, ey - 8), ldexpf(1.0f, ez - 8));
}

// origin + ldexpf(q, e - 8) per bound.  q * 2^(e-8) is exact for an 8-bit q
// and e in [-128, 127] (at most 255 * 2^119, at least one unit of 2^-136), so
// the single rounding of fma(q, 2^(e-8), origin) is the rounding of the sum.
__device__ __forceinline__ void child_box(const Words16& n, int i, F3 sc, F3& lo, F3& hi) {
  const float ox = f_of(n, 0), oy = f_of(n, 1), oz = f_of(n, 2);
  lo.x = __builtin_fmaf((float)byte_of(n, 28 + i), sc.x, ox);
  lo.y = __builtin_fmaf((float)byte_of(n, 40 + i), sc.y, oy);
  lo.z = __builtin_fmaf((float)byte_of(n, 52 + i), sc.z, oz);
  hi.x = __builtin_fmaf((float)byte_of(n, 34 + i), sc.x, ox);
  hi.y = __builtin_fmaf((float)byte_of(n, 46 + i), sc.y, oy);
  hi.z = __builtin_fmaf((float)byte_of(n, 58 + i), sc.z, oz);
}

}  // namespace rtxd
