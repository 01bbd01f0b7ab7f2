// builder.cpp — host BVH builder emitting the 6-wide quantized arena layout.
//
// Role in the reference: the Embree-built, mesa-vulkan-sim-packed acceleration
// structure that traceRay reads (vulkan_acceleration_structure_util.h decodes
// it; the builder itself is external to the reference).  Method: binned SAH
// binary BVH (16 bins, one triangle per leaf), collapsed to at most six
// children per node by repeatedly opening the child with the largest surface
// area, laid out breadth first so that every node's children are contiguous.
// Child boxes are quantized to 8 bits per bound with a per-axis power-of-two
// scale and rounded outward, checked with the decode the traversal uses
// (set_child_bounds, util.h:499-508).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "layout.h"
#include "rtx_scene.h"

using namespace rtxl;

namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = INFINITY;
      hi[a] = -INFINITY;
    }
  }
  void grow(const float* p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  double area() const {
    double dx = std::max(0.0, (double)hi[0] - lo[0]);
    double dy = std::max(0.0, (double)hi[1] - lo[1]);
    double dz = std::max(0.0, (double)hi[2] - lo[2]);
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct BNode {
  Box box;
  int32_t left = -1, right = -1;
  int32_t item = -1;  // >= 0: leaf holding item
};

// Binned SAH over item boxes; one item per leaf.  The split is searched on the
// axis of the largest centroid extent, or with all_axes on each of the three.
void build_binary(const std::vector<Box>& items, std::vector<BNode>& nodes, bool all_axes) {
  const int n = (int)items.size();
  nodes.clear();
  if (n == 0) return;
  nodes.reserve(2 * n);
  std::vector<int32_t> idx(n);
  std::vector<float> cen(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    idx[i] = i;
    for (int a = 0; a < 3; ++a) cen[3 * (size_t)i + a] = 0.5f * (items[i].lo[a] + items[i].hi[a]);
  }
  struct Job {
    int32_t node, begin, end;
  };
  std::vector<Job> stack;
  nodes.emplace_back();
  stack.push_back({0, 0, n});
  constexpr int kBins = 16;
  while (!stack.empty()) {
    Job j = stack.back();
    stack.pop_back();
    Box b;
    b.reset();
    Box cb;
    cb.reset();
    for (int i = j.begin; i < j.end; ++i) {
      b.grow(items[idx[i]]);
      cb.grow(&cen[3 * (size_t)idx[i]]);
    }
    nodes[j.node].box = b;
    const int count = j.end - j.begin;
    if (count == 1) {
      nodes[j.node].item = idx[j.begin];
      continue;
    }
    int axis = 0;
    float ext = cb.hi[0] - cb.lo[0];
    for (int a = 1; a < 3; ++a)
      if (cb.hi[a] - cb.lo[a] > ext) {
        ext = cb.hi[a] - cb.lo[a];
        axis = a;
      }
    int mid = -1;
    double best = INFINITY;
    int best_k = -1, best_a = -1;
    for (int a = all_axes ? 0 : axis; a < (all_axes ? 3 : axis + 1); ++a) {
      const float ea = cb.hi[a] - cb.lo[a];
      if (!(ea > 0.0f)) continue;
      Box bin_box[kBins];
      int bin_cnt[kBins] = {0};
      for (int k = 0; k < kBins; ++k) bin_box[k].reset();
      const double scale = kBins / (double)ea;
      for (int i = j.begin; i < j.end; ++i) {
        int k = (int)((cen[3 * (size_t)idx[i] + a] - cb.lo[a]) * scale);
        k = std::min(kBins - 1, std::max(0, k));
        bin_cnt[k]++;
        bin_box[k].grow(items[idx[i]]);
      }
      double left_area[kBins];
      int left_cnt[kBins];
      Box acc;
      acc.reset();
      int c = 0;
      for (int k = 0; k < kBins; ++k) {
        acc.grow(bin_box[k]);
        c += bin_cnt[k];
        left_area[k] = acc.area();
        left_cnt[k] = c;
      }
      acc.reset();
      c = 0;
      for (int k = kBins - 1; k >= 1; --k) {
        acc.grow(bin_box[k]);
        c += bin_cnt[k];
        if (left_cnt[k - 1] == 0 || c == 0) continue;
        double cost = left_area[k - 1] * left_cnt[k - 1] + acc.area() * c;
        if (cost < best) {
          best = cost;
          best_k = k;
          best_a = a;
        }
      }
    }
    if (best_k > 0) {
      const double scale = kBins / (double)(cb.hi[best_a] - cb.lo[best_a]);
      const float lo_a = cb.lo[best_a];
      auto it = std::partition(idx.begin() + j.begin, idx.begin() + j.end, [&](int32_t item) {
        int k = (int)((cen[3 * (size_t)item + best_a] - lo_a) * scale);
        return std::min(kBins - 1, std::max(0, k)) < best_k;
      });
      mid = (int)(it - idx.begin());
      if (mid == j.begin || mid == j.end) mid = -1;
    }
    if (mid < 0) {
      // Coincident centroids: split by a stable median on the axis.
      mid = j.begin + count / 2;
      std::nth_element(idx.begin() + j.begin, idx.begin() + mid, idx.begin() + j.end,
                       [&](int32_t x, int32_t y) {
                         float cx = cen[3 * (size_t)x + axis], cy = cen[3 * (size_t)y + axis];
                         return cx < cy || (cx == cy && x < y);
                       });
    }
    int32_t l = (int32_t)nodes.size();
    nodes.emplace_back();
    int32_t r = (int32_t)nodes.size();
    nodes.emplace_back();
    nodes[j.node].left = l;
    nodes[j.node].right = r;
    stack.push_back({r, mid, j.end});
    stack.push_back({l, j.begin, mid});
  }
}

// SAH-optimal collapse of a binary tree to 6-wide nodes.  D(n, j) is the least
// cost of covering binary node n's subtree with at most j roots, where a root
// costs its box area (a ray's chance of testing what lies under it) plus, for
// an internal root, the cost of covering its own subtree with at most six
// roots.  Leaves are one item each and every leaf is a root exactly once, so
// only internal nodes change the cost.
struct CollapsePlan {
  std::vector<int8_t> inner;  // per node: roots given to the left child when it is a 6-wide node
  std::vector<int8_t> split;  // per node and j = 2..6: 0 keeps the node as one root, else roots to the left
};

void plan_collapse(const std::vector<BNode>& nodes, CollapsePlan& p) {
  const size_t n = nodes.size();
  p.inner.assign(n, 0);
  p.split.assign(7 * n, 0);
  std::vector<double> D(7 * n, 0.0);
  // children are created after their parent, so a backward sweep is bottom-up
  for (size_t q = n; q-- > 0;) {
    const BNode& b = nodes[q];
    const double a = b.box.area();
    if (b.item >= 0) {
      for (int j = 1; j <= 6; ++j) D[7 * q + j] = a;
      continue;
    }
    const size_t l = (size_t)b.left, r = (size_t)b.right;
    double best = INFINITY;
    int bk = 1;
    for (int k = 1; k <= 5; ++k) {
      const double c = D[7 * l + k] + D[7 * r + 6 - k];
      if (c < best) {
        best = c;
        bk = k;
      }
    }
    p.inner[q] = (int8_t)bk;
    D[7 * q + 1] = a + best;
    for (int j = 2; j <= 6; ++j) {
      double bj = D[7 * q + 1];
      int sk = 0;
      for (int k = 1; k < j; ++k) {
        const double c = D[7 * l + k] + D[7 * r + j - k];
        if (c < bj) {
          bj = c;
          sk = k;
        }
      }
      D[7 * q + j] = bj;
      p.split[7 * q + j] = (int8_t)sk;
    }
  }
}

// The roots that cover node x's subtree with at most j of them, in tree order.
void plan_roots(const std::vector<BNode>& nodes, const CollapsePlan& p, int32_t x, int j, int32_t* out, int& n) {
  const int k = (j >= 2 && nodes[x].item < 0) ? p.split[7 * (size_t)x + j] : 0;
  if (k == 0) {
    out[n++] = x;
    return;
  }
  plan_roots(nodes, p, nodes[x].left, k, out, n);
  plan_roots(nodes, p, nodes[x].right, j - k, out, n);
}

// Up to six binary-tree nodes that become the children of 6-wide node `id`:
// from the plan when there is one, else by opening the child with the largest
// surface area until six are open.
int collapse_children(const std::vector<BNode>& nodes, const CollapsePlan* plan, int32_t id, int32_t out[6]) {
  if (nodes[id].item >= 0) {
    out[0] = id;
    return 1;
  }
  if (plan) {
    int n = 0;
    const int k = plan->inner[id];
    plan_roots(nodes, *plan, nodes[id].left, k, out, n);
    plan_roots(nodes, *plan, nodes[id].right, 6 - k, out, n);
    return n;
  }
  int n = 2;
  out[0] = nodes[id].left;
  out[1] = nodes[id].right;
  while (n < 6) {
    int best = -1;
    double best_area = -1.0;
    for (int k = 0; k < n; ++k) {
      if (nodes[out[k]].item >= 0) continue;
      double a = nodes[out[k]].box.area();
      if (a > best_area) {
        best_area = a;
        best = k;
      }
    }
    if (best < 0) break;
    int32_t open = out[best];
    for (int k = n; k > best + 1; --k) out[k] = out[k - 1];
    out[best] = nodes[open].left;
    out[best + 1] = nodes[open].right;
    ++n;
  }
  return n;
}

inline float decode_bound(float origin, int q, int e) { return origin + ldexpf((float)q, e - 8); }

bool quantize_axis(float origin, const float* lo, const float* hi, int n, int e, uint8_t* qlo, uint8_t* qhi) {
  const double s = std::ldexp(1.0, e - 8);
  for (int k = 0; k < n; ++k) {
    double dl = ((double)lo[k] - (double)origin) / s;
    long q = (long)std::floor(dl);
    q = std::max(0L, std::min(255L, q));
    while (q > 0 && decode_bound(origin, (int)q, e) > lo[k]) --q;
    if (decode_bound(origin, (int)q, e) > lo[k]) return false;
    double dh = ((double)hi[k] - (double)origin) / s;
    long qh = (long)std::ceil(dh);
    qh = std::max(0L, qh);
    if (qh > 255) return false;
    while (decode_bound(origin, (int)qh, e) < hi[k]) {
      if (++qh > 255) return false;
    }
    qlo[k] = (uint8_t)q;
    qhi[k] = (uint8_t)qh;
  }
  return true;
}

// Writes one 64-byte internal node.  Returns false if a box cannot be quantized.
bool write_internal(uint8_t* dst, const Box* child_boxes, const uint32_t* child_types, const uint32_t* child_sizes,
                    int n_children, int32_t child_offset_blocks) {
  memset(dst, 0, kInternalBytes);
  float origin[3];
  for (int a = 0; a < 3; ++a) {
    origin[a] = INFINITY;
    for (int k = 0; k < n_children; ++k) origin[a] = std::min(origin[a], child_boxes[k].lo[a]);
  }
  uint8_t q[6][6];  // [LowerX, UpperX, LowerY, UpperY, LowerZ, UpperZ][child]
  int8_t exps[3];
  for (int a = 0; a < 3; ++a) {
    float lo[6], hi[6];
    double extent = 0.0;
    for (int k = 0; k < n_children; ++k) {
      lo[k] = child_boxes[k].lo[a];
      hi[k] = child_boxes[k].hi[a];
      extent = std::max(extent, (double)hi[k] - (double)origin[a]);
    }
    int e = -128;
    if (extent > 0.0) e = std::max(-128, std::min(127, (int)std::ceil(std::log2(extent)) - 1));
    while (!quantize_axis(origin[a], lo, hi, n_children, e, q[2 * a], q[2 * a + 1])) {
      if (++e > 127) return false;
    }
    exps[a] = (int8_t)e;
  }
  store_as<float>(dst + kInOrigin + 0, origin[0]);
  store_as<float>(dst + kInOrigin + 4, origin[1]);
  store_as<float>(dst + kInOrigin + 8, origin[2]);
  store_as<int32_t>(dst + kInChildOffset, child_offset_blocks);
  dst[kInNodeType] = (uint8_t)kTypeInternal;
  for (int a = 0; a < 3; ++a) dst[kInExp + a] = (uint8_t)exps[a];
  dst[kInMask] = 0xff;
  for (int k = 0; k < 6; ++k) {
    uint8_t info = 0;
    if (k < n_children) info = (uint8_t)((child_sizes[k] & 3u) | (child_types[k] << 2));
    dst[kInChildInfo + k] = info;
  }
  const uint32_t field[6] = {kInLowerX, kInUpperX, kInLowerY, kInUpperY, kInLowerZ, kInUpperZ};
  for (int f = 0; f < 6; ++f)
    for (int k = 0; k < 6; ++k) dst[field[f] + k] = k < n_children ? q[f][k] : 0;
  return true;
}

void write_header(uint8_t* dst, uint64_t root_offset, const Box& b) {
  memset(dst, 0, kHeaderBytes);
  store_as<uint64_t>(dst + kHdrRootOffset, root_offset);
  for (int a = 0; a < 3; ++a) {
    store_as<float>(dst + kHdrBoundsMin + 4 * a, b.lo[a]);
    store_as<float>(dst + kHdrBoundsMax + 4 * a, b.hi[a]);
  }
}

// Lays out a BVH (header + nodes + leaves) into a position-independent blob.
// leaf_bytes: 64 (quad) or 128 (instance).  write_leaf(dst, item) fills a leaf.
// sah_collapse: 6-wide nodes from the SAH-optimal plan, else the largest-area opening.
template <typename LeafWriter>
bool emit_bvh(const std::vector<BNode>& nodes, uint32_t leaf_type, uint32_t leaf_bytes, LeafWriter write_leaf,
              std::vector<uint8_t>& blob, uint64_t* n_internal, bool sah_collapse) {
  blob.clear();
  *n_internal = 0;
  if (nodes.empty()) return false;
  CollapsePlan plan_store;
  if (sah_collapse) plan_collapse(nodes, plan_store);
  const CollapsePlan* plan = sah_collapse ? &plan_store : nullptr;
  // Pass 1: assign offsets breadth first.
  struct Item {
    int32_t bnode;
    uint64_t off;
  };
  std::vector<Item> queue;
  queue.push_back({0, kHeaderBytes});
  uint64_t cursor = kHeaderBytes + kInternalBytes;
  std::vector<uint64_t> block_of;  // children block offset per queue entry
  for (size_t qi = 0; qi < queue.size(); ++qi) {
    int32_t ch[6];
    int nc = collapse_children(nodes, plan, queue[qi].bnode, ch);
    block_of.push_back(cursor);
    for (int k = 0; k < nc; ++k) {
      bool leaf = nodes[ch[k]].item >= 0;
      if (!leaf) queue.push_back({ch[k], cursor});
      cursor += leaf ? leaf_bytes : kInternalBytes;
    }
  }
  blob.assign(cursor, 0);
  write_header(blob.data(), kHeaderBytes, nodes[0].box);
  for (size_t qi = 0; qi < queue.size(); ++qi) {
    int32_t ch[6];
    int nc = collapse_children(nodes, plan, queue[qi].bnode, ch);
    Box boxes[6];
    uint32_t types[6], sizes[6];
    uint64_t coff = block_of[qi];
    for (int k = 0; k < nc; ++k) {
      bool leaf = nodes[ch[k]].item >= 0;
      boxes[k] = nodes[ch[k]].box;
      types[k] = leaf ? leaf_type : kTypeInternal;
      sizes[k] = (leaf ? leaf_bytes : kInternalBytes) / 64;
      if (leaf) write_leaf(blob.data() + coff, nodes[ch[k]].item);
      coff += sizes[k] * 64;
    }
    int64_t rel = ((int64_t)block_of[qi] - (int64_t)queue[qi].off) / 64;
    if (rel > INT32_MAX) return false;
    if (!write_internal(blob.data() + queue[qi].off, boxes, types, sizes, nc, (int32_t)rel)) return false;
  }
  *n_internal = queue.size();
  return true;
}

struct Blas {
  std::vector<float> tris;
  uint32_t geometry_index = 0;
  std::vector<uint8_t> blob;
  Box box;
  uint64_t n_internal = 0;
  uint64_t offset = 0;
  bool ok = false;
};

struct Instance {
  uint32_t blas;
  double o2w[12];
  float w2o_rows[12];  // m[0..3][0..2] as stored at leaf +16
  float o2w_rows[12];  // m[0..3][0..2] as stored at leaf +80
  uint32_t instance_id;
  uint32_t hit_group;
  Box world_box;
};

}  // namespace

struct rtx_builder {
  std::vector<Blas> blas;
  std::vector<Instance> inst;
  std::vector<uint8_t> arena;
  uint64_t tlas_offset = 0;
  uint64_t tlas_internal = 0;
  uint32_t quality = RTX_BUILD_DEFAULT;
  bool finished = false;
};

namespace {

void build_blas(Blas& b, uint32_t quality) {
  const uint32_t n = (uint32_t)(b.tris.size() / 9);
  std::vector<Box> boxes(n);
  for (uint32_t i = 0; i < n; ++i) {
    boxes[i].reset();
    for (int v = 0; v < 3; ++v) boxes[i].grow(&b.tris[9 * (size_t)i + 3 * v]);
  }
  std::vector<BNode> nodes;
  const bool sah = quality == RTX_BUILD_SAH_OPTIMAL;
  build_binary(boxes, nodes, sah);
  const float* tris = b.tris.data();
  const uint32_t geom = b.geometry_index;
  auto leaf = [tris, geom](uint8_t* dst, int32_t item) {
    memset(dst, 0, kQuadLeafBytes);
    store_as<uint32_t>(dst + 0, 0xff000000u);
    store_as<uint32_t>(dst + kQlDescriptor1, (geom & 0x0fffffffu) | (1u << 30));
    store_as<uint32_t>(dst + kQlPrim0, (uint32_t)item);
    store_as<uint32_t>(dst + kQlPrim1Delta, (1u << 19) | (2u << 21));
    const float* t = tris + 9 * (size_t)item;
    memcpy(dst + kQlV0, t, 36);
    memcpy(dst + kQlV0 + 36, t + 6, 12);  // vertex 3 repeats vertex 2 (one triangle per leaf)
  };
  b.ok = emit_bvh(nodes, kTypeQuad, kQuadLeafBytes, leaf, b.blob, &b.n_internal, sah);
  if (!nodes.empty()) b.box = nodes[0].box;
}

bool invert3x4(const double m[12], double inv[12]) {
  const double a = m[0], bb = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  const double det = a * (e * i - f * h) - bb * (d * i - f * g) + c * (d * h - e * g);
  if (det == 0.0 || !std::isfinite(det)) return false;
  const double r = 1.0 / det;
  double A[9] = {(e * i - f * h) * r, (c * h - bb * i) * r, (bb * f - c * e) * r,
                 (f * g - d * i) * r, (a * i - c * g) * r, (c * d - a * f) * r,
                 (d * h - e * g) * r, (bb * g - a * h) * r, (a * e - bb * d) * r};
  for (int row = 0; row < 3; ++row) {
    for (int col = 0; col < 3; ++col) inv[4 * row + col] = A[3 * row + col];
    double t = -(A[3 * row + 0] * m[3] + A[3 * row + 1] * m[7] + A[3 * row + 2] * m[11]);
    inv[4 * row + 3] = t + 0.0;
  }
  return true;
}

// Stored rows m[j][i]: res[i] = sum_j m[j][i] * v[j] (float4x4::operator*, util.h:47-55),
// so for p' = A p + t the rows are m[j][i] = A[i][j] (j < 3) and m[3][i] = t[i].
void matrix_rows(const double m[12], float rows[12]) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) rows[3 * j + i] = (float)m[4 * i + j];
  for (int i = 0; i < 3; ++i) rows[9 + i] = (float)m[4 * i + 3];
}

}  // namespace

extern "C" {

rtx_status rtx_builder_create(rtx_builder** out) {
  if (!out) return RTX_ERR_INVALID_ARGUMENT;
  *out = new (std::nothrow) rtx_builder();
  return *out ? RTX_OK : RTX_ERR_OUT_OF_MEMORY;
}

rtx_status rtx_builder_destroy(rtx_builder* b) {
  delete b;
  return RTX_OK;
}

rtx_status rtx_builder_set_quality(rtx_builder* b, uint32_t quality) {
  if (!b || b->finished || (quality != RTX_BUILD_DEFAULT && quality != RTX_BUILD_SAH_OPTIMAL))
    return RTX_ERR_INVALID_ARGUMENT;
  b->quality = quality;
  return RTX_OK;
}

rtx_status rtx_builder_add_blas(rtx_builder* b, const float* tris, uint32_t n_tris, uint32_t geometry_index,
                                uint32_t* blas_id) {
  if (!b || !tris || n_tris == 0 || b->finished || geometry_index > 0x0fffffffu) return RTX_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < 9 * (size_t)n_tris; ++i)
    if (!std::isfinite(tris[i])) return RTX_ERR_INVALID_ARGUMENT;
  Blas bl;
  bl.tris.assign(tris, tris + 9 * (size_t)n_tris);
  bl.geometry_index = geometry_index;
  b->blas.push_back(std::move(bl));
  if (blas_id) *blas_id = (uint32_t)(b->blas.size() - 1);
  return RTX_OK;
}

rtx_status rtx_builder_add_instance(rtx_builder* b, uint32_t blas_id, const float object_to_world[12],
                                    uint32_t instance_id, uint32_t hit_group_offset) {
  if (!b || !object_to_world || blas_id >= b->blas.size() || b->finished || hit_group_offset > 0xffffffu)
    return RTX_ERR_INVALID_ARGUMENT;
  Instance in;
  in.blas = blas_id;
  for (int k = 0; k < 12; ++k) in.o2w[k] = object_to_world[k];
  double w2o[12];
  if (!invert3x4(in.o2w, w2o)) return RTX_ERR_INVALID_ARGUMENT;
  matrix_rows(w2o, in.w2o_rows);
  matrix_rows(in.o2w, in.o2w_rows);
  in.instance_id = instance_id;
  in.hit_group = hit_group_offset;
  b->inst.push_back(in);
  return RTX_OK;
}

}  // extern "C"

namespace rtx_internal {
rtx_status ctx_stream(rtx_ctx* c, hipStream_t* s, int* device);
rtx_status gpu_build_blas_pool(hipStream_t s, const float* tris_host, const uint32_t* pofs_host,
                               const uint32_t* geom_host, uint32_t n_blas, uint8_t** pool_dev, uint64_t* pool_bytes,
                               std::vector<float>& boxes_out);
}  // namespace rtx_internal

namespace {

// TLAS over the instances and the arena: TLAS at 0, then the bottom levels.
// With `pool_dev` set, the bottom levels are one pool of pool_bytes in device
// memory whose BLAS b header sits at byte 64 b, copied straight into the
// arena on stream s; otherwise each BLAS's blob follows the previous one.
rtx_status assemble(rtx_builder* b, const uint8_t* pool_dev, uint64_t pool_bytes, hipStream_t s, uint64_t* arena_bytes,
                    uint64_t* tlas_offset) {
  // Instance world boxes: the 8 transformed corners of the BLAS box, rounded outward.
  std::vector<Box> boxes(b->inst.size());
  for (size_t k = 0; k < b->inst.size(); ++k) {
    const Instance& in = b->inst[k];
    const Box& ob = b->blas[in.blas].box;
    Box wb;
    wb.reset();
    for (int c = 0; c < 8; ++c) {
      double p[3] = {(c & 1) ? ob.hi[0] : ob.lo[0], (c & 2) ? ob.hi[1] : ob.lo[1], (c & 4) ? ob.hi[2] : ob.lo[2]};
      for (int a = 0; a < 3; ++a) {
        double w = in.o2w[4 * a] * p[0] + in.o2w[4 * a + 1] * p[1] + in.o2w[4 * a + 2] * p[2] + in.o2w[4 * a + 3];
        float lo = std::nextafter((float)w, -INFINITY), hi = std::nextafter((float)w, INFINITY);
        wb.lo[a] = std::min(wb.lo[a], lo);
        wb.hi[a] = std::max(wb.hi[a], hi);
      }
    }
    boxes[k] = wb;
    b->inst[k].world_box = wb;
  }
  std::vector<BNode> nodes;
  const bool sah = b->quality == RTX_BUILD_SAH_OPTIMAL;
  build_binary(boxes, nodes, sah);
  std::vector<uint8_t> tlas;
  std::vector<uint64_t> leaf_pos(b->inst.size(), 0);
  const std::vector<Instance>& insts = b->inst;
  auto leaf = [&](uint8_t* dst, int32_t item) {
    const Instance& in = insts[item];
    memset(dst, 0, kInstanceLeafBytes);
    store_as<uint32_t>(dst + 0, 0xff000000u);
    store_as<uint32_t>(dst + kIlHitGroup, (in.hit_group & 0xffffffu) | (1u << 30));
    memcpy(dst + kIlWorldToObject, in.w2o_rows, 48);
    store_as<uint32_t>(dst + kIlInstanceId, in.instance_id);
    store_as<uint32_t>(dst + kIlInstanceIndex, (uint32_t)item);
    memcpy(dst + kIlObjectToWorld, in.o2w_rows, 48);
  };
  if (!emit_bvh(nodes, kTypeInstance, kInstanceLeafBytes, leaf, tlas, &b->tlas_internal, sah))
    return RTX_ERR_INVALID_ARGUMENT;
  // Locate each instance leaf by walking the TLAS (InstanceIndex names the instance).
  {
    std::vector<uint64_t> q = {kHeaderBytes};
    for (size_t qi = 0; qi < q.size(); ++qi) {
      const uint8_t* nd = tlas.data() + q[qi];
      uint64_t c = q[qi] + 64 * (int64_t)load_as<int32_t>(nd + kInChildOffset);
      for (int k = 0; k < 6; ++k) {
        uint8_t info = nd[kInChildInfo + k];
        uint32_t sz = child_size(info);
        if (sz == 0) continue;
        if (child_type(info) == kTypeInternal)
          q.push_back(c);
        else
          leaf_pos[load_as<uint32_t>(tlas.data() + c + kIlInstanceIndex)] = c;
        c += 64 * sz;
      }
    }
  }
  uint64_t cursor = tlas.size();
  if (pool_dev) {
    for (size_t k = 0; k < b->blas.size(); ++k) b->blas[k].offset = cursor + 64 * k;
    cursor += pool_bytes;
  } else {
    for (auto& bl : b->blas) {
      bl.offset = cursor;
      cursor += bl.blob.size();
    }
  }
  try {
    b->arena.assign(cursor, 0);
  } catch (...) {
    return RTX_ERR_OUT_OF_MEMORY;
  }
  memcpy(b->arena.data(), tlas.data(), tlas.size());
  if (pool_dev) {
    if (hipMemcpyAsync(b->arena.data() + tlas.size(), pool_dev, pool_bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      (void)hipGetLastError();
      return RTX_ERR_HIP;
    }
  } else {
    for (auto& bl : b->blas) {
      memcpy(b->arena.data() + bl.offset, bl.blob.data(), bl.blob.size());
      std::vector<uint8_t>().swap(bl.blob);
    }
  }
  for (size_t k = 0; k < b->inst.size(); ++k) {
    uint64_t lp = leaf_pos[k];
    uint64_t rel = b->blas[b->inst[k].blas].offset - lp;
    store_as<uint64_t>(b->arena.data() + lp + kIlBvhAddress, rel);
  }
  b->tlas_offset = 0;
  b->finished = true;
  if (arena_bytes) *arena_bytes = b->arena.size();
  if (tlas_offset) *tlas_offset = 0;
  return RTX_OK;
}

}  // namespace

extern "C" {

rtx_status rtx_builder_finish(rtx_builder* b, uint32_t threads, uint64_t* arena_bytes, uint64_t* tlas_offset) {
  if (!b || b->finished || b->blas.empty() || b->inst.empty()) return RTX_ERR_INVALID_ARGUMENT;
  unsigned nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
  nt = std::min<unsigned>(nt, (unsigned)b->blas.size());
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (size_t i = next++; i < b->blas.size(); i = next++) build_blas(b->blas[i], b->quality);
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  for (auto& bl : b->blas)
    if (!bl.ok) return RTX_ERR_INVALID_ARGUMENT;
  return assemble(b, nullptr, 0, nullptr, arena_bytes, tlas_offset);
}

rtx_status rtx_builder_finish_device(rtx_builder* b, rtx_ctx* ctx, uint64_t* arena_bytes, uint64_t* tlas_offset) {
  if (!b || !ctx || b->finished || b->blas.empty() || b->inst.empty()) return RTX_ERR_INVALID_ARGUMENT;
  hipStream_t s;
  int dev = 0;
  if (rtx_internal::ctx_stream(ctx, &s, &dev) != RTX_OK) return RTX_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(dev) != hipSuccess) return RTX_ERR_HIP;
  const uint32_t B = (uint32_t)b->blas.size();
  std::vector<uint32_t> pofs(B + 1, 0), geom(B);
  uint64_t n = 0;
  for (uint32_t k = 0; k < B; ++k) {
    pofs[k] = (uint32_t)n;
    const uint64_t nt = b->blas[k].tris.size() / 9;
    if (nt == 0) return RTX_ERR_INVALID_ARGUMENT;
    n += nt;
    geom[k] = b->blas[k].geometry_index;
  }
  if (n >= (1ull << 31)) return RTX_ERR_INVALID_ARGUMENT;
  pofs[B] = (uint32_t)n;
  // gathered without zero-filling first: every float is written below
  std::unique_ptr<float[]> tris(new (std::nothrow) float[n * 9]);
  if (!tris) return RTX_ERR_OUT_OF_MEMORY;
  for (uint32_t k = 0; k < B; ++k) memcpy(tris.get() + 9ull * pofs[k], b->blas[k].tris.data(), b->blas[k].tris.size() * 4);
  struct DevPool {
    uint8_t* p = nullptr;
    ~DevPool() {
      if (p) (void)hipFree(p);
    }
  } pool;
  uint64_t pool_bytes = 0;
  std::vector<float> boxes;
  rtx_status st = rtx_internal::gpu_build_blas_pool(s, tris.get(), pofs.data(), geom.data(), B, &pool.p, &pool_bytes, boxes);
  tris.reset();
  if (st != RTX_OK) return st;
  const uint64_t slots = pool_bytes / 64;
  for (uint32_t k = 0; k < B; ++k) {
    Blas& bl = b->blas[k];
    for (int a = 0; a < 3; ++a) {
      bl.box.lo[a] = boxes[6 * k + a];
      bl.box.hi[a] = boxes[6 * k + 3 + a];
    }
    bl.n_internal = 0;
    bl.ok = true;
  }
  // every slot of the pool is a header, an internal node or a quad leaf
  b->blas[0].n_internal = slots - B - n;
  return assemble(b, pool.p, pool_bytes, s, arena_bytes, tlas_offset);
}

rtx_status rtx_builder_copy_arena(const rtx_builder* b, void* dst, size_t capacity) {
  if (!b || !dst || !b->finished || capacity < b->arena.size()) return RTX_ERR_INVALID_ARGUMENT;
  memcpy(dst, b->arena.data(), b->arena.size());
  return RTX_OK;
}

rtx_status rtx_builder_arena_info(const rtx_builder* b, uint64_t* arena_bytes, uint64_t* tlas_offset) {
  if (!b || !b->finished) return RTX_ERR_INVALID_ARGUMENT;
  if (arena_bytes) *arena_bytes = b->arena.size();
  if (tlas_offset) *tlas_offset = b->tlas_offset;
  return RTX_OK;
}

rtx_status rtx_builder_counts(const rtx_builder* b, uint64_t* n_triangles, uint64_t* n_blas, uint64_t* n_instances,
                              uint64_t* n_internal_nodes) {
  if (!b) return RTX_ERR_INVALID_ARGUMENT;
  uint64_t tris = 0, internal = b->tlas_internal;
  for (auto& bl : b->blas) {
    tris += bl.tris.size() / 9;
    internal += bl.n_internal;
  }
  if (n_triangles) *n_triangles = tris;
  if (n_blas) *n_blas = b->blas.size();
  if (n_instances) *n_instances = b->inst.size();
  if (n_internal_nodes) *n_internal_nodes = internal;
  return RTX_OK;
}

rtx_status rtx_builder_blas_offset(const rtx_builder* b, uint32_t blas_id, uint64_t* offset) {
  if (!b || !offset || !b->finished || blas_id >= b->blas.size()) return RTX_ERR_INVALID_ARGUMENT;
  *offset = b->blas[blas_id].offset;
  return RTX_OK;
}

}  // extern "C"

// Accessors used by the workload generator (scene.cpp).
namespace rtx_internal {
bool builder_triangle(const rtx_builder* b, uint32_t instance_id, uint32_t prim, float v[9], double o2w[12]) {
  // Instance ids of generated workloads equal their index; fall back to a search otherwise.
  size_t k = instance_id;
  if (k >= b->inst.size() || b->inst[k].instance_id != instance_id) {
    k = b->inst.size();
    for (size_t i = 0; i < b->inst.size(); ++i)
      if (b->inst[i].instance_id == instance_id) {
        k = i;
        break;
      }
    if (k == b->inst.size()) return false;
  }
  const Blas& bl = b->blas[b->inst[k].blas];
  if ((size_t)prim * 9 + 9 > bl.tris.size()) return false;
  memcpy(v, &bl.tris[9 * (size_t)prim], 36);
  memcpy(o2w, b->inst[k].o2w, sizeof(double) * 12);
  return true;
}
}  // namespace rtx_internal
