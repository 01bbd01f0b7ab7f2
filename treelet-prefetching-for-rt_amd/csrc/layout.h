// layout.h — byte layout of the acceleration-structure arena.
//
// The arena holds the node formats that the reference decodes in
// src/cuda-sim/vulkan_acceleration_structure_util.h (Intel-style 6-wide
// quantized BVH).  Offsets below are byte offsets inside one node:
//
//   GEN_RT_BVH header (accounted 64 B, util.h:89-107)
//     +0  u64 RootNodeOffset (relative to the header)  +8 f32[3] BoundsMin  +20 f32[3] BoundsMax
//   internal node (64 B, util.h:109-207)
//     +0 f32[3] Origin  +12 i32 ChildOffset (x64 B, relative to the node)  +16 u8 NodeType
//     +18..20 i8 exponent x/y/z  +21 u8 NodeRayMask  +22..27 childInfo[6]
//       (bits 0-1 ChildSize in 64 B blocks, bits 2-5 ChildType)
//     +28 LowerX[6] +34 UpperX[6] +40 LowerY[6] +46 UpperY[6] +52 LowerZ[6] +58 UpperZ[6]
//   instance leaf (128 B, util.h:226-386)
//     +4 u24 InstanceContributionToHitGroupIndex  +16 f32[12] world->object rows (row 3 = translation)
//     +64 u64 BVHAddress (byte offset from this leaf to the BLAS header)  +72 u32 InstanceID
//     +80 f32[12] object->world rows
//   primitive leaf descriptor (first 8 B of a leaf, util.h:389-422)
//     dword1 bits 0-27 GeometryIndex, bit 29 LeafType (0 = quad)
//   quad leaf (64 B, util.h:426-466)
//     +8 u32 PrimitiveIndex0  +12 u32 PrimitiveIndex1Delta (bits 0-16)  +16 f32[3] v0 +28 v1 +40 v2 +52 v3
//
// Children of an internal node are contiguous: child0 at node + 64*ChildOffset,
// child i+1 at child i + 64*ChildSize[i] (vulkan_ray_tracing.cc:2555,2592).
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define RTX_HD __host__ __device__ __forceinline__
#else
#define RTX_HD inline
#endif

namespace rtxl {

constexpr uint32_t kNodeBytes = 64;
constexpr uint32_t kHeaderBytes = 64;
constexpr uint32_t kInternalBytes = 64;
constexpr uint32_t kInstanceLeafBytes = 128;
constexpr uint32_t kQuadLeafBytes = 64;
constexpr uint32_t kDescriptorBytes = 8;

constexpr uint32_t kTypeInternal = 0;
constexpr uint32_t kTypeInstance = 1;
constexpr uint32_t kTypeProcedural = 3;
constexpr uint32_t kTypeQuad = 4;

// header
constexpr uint32_t kHdrRootOffset = 0;
constexpr uint32_t kHdrBoundsMin = 8;
constexpr uint32_t kHdrBoundsMax = 20;
// internal node
constexpr uint32_t kInOrigin = 0;
constexpr uint32_t kInChildOffset = 12;
constexpr uint32_t kInNodeType = 16;
constexpr uint32_t kInExp = 18;
constexpr uint32_t kInMask = 21;
constexpr uint32_t kInChildInfo = 22;
constexpr uint32_t kInLowerX = 28;
constexpr uint32_t kInUpperX = 34;
constexpr uint32_t kInLowerY = 40;
constexpr uint32_t kInUpperY = 46;
constexpr uint32_t kInLowerZ = 52;
constexpr uint32_t kInUpperZ = 58;
// instance leaf
constexpr uint32_t kIlHitGroup = 4;
constexpr uint32_t kIlWorldToObject = 16;
constexpr uint32_t kIlBvhAddress = 64;
constexpr uint32_t kIlInstanceId = 72;
constexpr uint32_t kIlInstanceIndex = 76;
constexpr uint32_t kIlObjectToWorld = 80;
// quad leaf
constexpr uint32_t kQlDescriptor1 = 4;
constexpr uint32_t kQlPrim0 = 8;
constexpr uint32_t kQlPrim1Delta = 12;
constexpr uint32_t kQlV0 = 16;

RTX_HD uint32_t child_size(uint8_t info) { return info & 0x3u; }
RTX_HD uint32_t child_type(uint8_t info) { return (info & 0x3fu) >> 2; }

template <typename T>
RTX_HD T load_as(const uint8_t* p) {
  T v;
  memcpy(&v, p, sizeof(T));
  return v;
}
template <typename T>
inline void store_as(uint8_t* p, T v) {
  memcpy(p, &v, sizeof(T));
}

}  // namespace rtxl
