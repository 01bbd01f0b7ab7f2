/*
 * oracle.c — CPU restatement of the reference traversal path (test infrastructure).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library.  Each function names the reference lines it restates;
 * the loop structure of traceRay (nested top/bottom loops) and of
 * traceRayWithTreelets (two lists) is kept as the reference writes it, so the
 * flattened GPU kernels are checked against the original control flow.
 *
 * Build: gcc -O3 -ffp-contract=off -fno-fast-math -shared -fPIC (see Makefile).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* float3 helpers: src/gpgpu-sim/vector-math.cc:28-108                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  float x, y, z;
} f3;
typedef struct {
  float x, y, z, w;
} f4;
typedef struct {
  float m[4][4];
} f4x4;

static f3 mk3(float a, float b, float c) {
  f3 r;
  r.x = a;
  r.y = b;
  r.z = c;
  return r;
}
static f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 cross3(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* float4x4::operator*, vulkan_acceleration_structure_util.h:47-55:
 * res[i] starts at 0 and accumulates m[j][i]*vec[j] for j = 0..3. */
static f4 mat_mul(const f4x4* M, f4 v) {
  float vec[4] = {v.x, v.y, v.z, v.w};
  float res[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) res[i] += M->m[j][i] * vec[j];
  f4 r = {res[0], res[1], res[2], res[3]};
  return r;
}

/* instance_leaf_matrix_to_float4x4, util.h:210-224 */
static f4x4 leaf_matrix(const float* a) {
  f4x4 M;
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 3; j++) M.m[i][j] = a[3 * i + j];
    M.m[i][3] = 0;
  }
  M.m[3][3] = 1;
  return M;
}

/* get_norm / normalized, vulkan_ray_tracing.cc:148-167 */
static float get_norm3(f3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
static f3 normalized3(f3 v) {
  float n = get_norm3(v);
  return mk3(v.x / n, v.y / n, v.z / n);
}

typedef struct {
  f3 o;
  float tmin;
  f3 d;
  float tmax;
} oray;

/* make_transformed_ray, vulkan_ray_tracing.cc:169-181 */
static oray make_transformed_ray(const oray* r, const f4x4* M, float* tmult) {
  f4 o4 = mat_mul(M, (f4){r->o.x, r->o.y, r->o.z, 1});
  f4 d4 = mat_mul(M, (f4){r->d.x, r->d.y, r->d.z, 0});
  oray t;
  t.o = mk3(o4.x / o4.w, o4.y / o4.w, o4.z / o4.w);
  f3 d = mk3(d4.x, d4.y, d4.z);
  *tmult = get_norm3(d);
  t.d = normalized3(d);
  t.tmin = r->tmin * (*tmult);
  t.tmax = r->tmax * (*tmult);
  return t;
}

/* MIN/MAX/MIN_MAX/MAX_MIN macros, vulkan_ray_tracing.h:53-56 (compare-select). */
static float MINf(float a, float b) { return a < b ? a : b; }
static float MAXf(float a, float b) { return a > b ? a : b; }
static float MIN_MAXf(float a, float b, float c) { return MAXf(MINf(a, b), c); }
static float MAX_MINf(float a, float b, float c) { return MINf(MAXf(a, b), c); }

/* magic_max7 / magic_min7, vulkan_ray_tracing.cc:183-197 */
static float magic_max7(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
  float t1 = MIN_MAXf(a0, a1, d);
  float t2 = MIN_MAXf(b0, b1, t1);
  return MIN_MAXf(c0, c1, t2);
}
static float magic_min7(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
  float t1 = MAX_MINf(a0, a1, d);
  float t2 = MAX_MINf(b0, b1, t1);
  return MAX_MINf(c0, c1, t2);
}

/* calculate_idir, vulkan_ray_tracing.cc:221-237; exp2f(-80.0f) is exactly 2^-80. */
static f3 calculate_idir(f3 d) {
  const float ooeps = 0x1p-80f;
  f3 r;
  r.x = 1.0f / (fabsf(d.x) > ooeps ? d.x : copysignf(ooeps, d.x));
  r.y = 1.0f / (fabsf(d.y) > ooeps ? d.y : copysignf(ooeps, d.y));
  r.z = 1.0f / (fabsf(d.z) > ooeps ? d.z : copysignf(ooeps, d.z));
  return r;
}

/* get_t_bound + ray_box_test, vulkan_ray_tracing.cc:199-257 */
static int ray_box_test(f3 low, f3 high, f3 idir, f3 o, float tmin, float tmax, float* thit) {
  f3 lo = mk3((low.x - o.x) * idir.x, (low.y - o.y) * idir.y, (low.z - o.z) * idir.z);
  f3 hi = mk3((high.x - o.x) * idir.x, (high.y - o.y) * idir.y, (high.z - o.z) * idir.z);
  float mn = magic_max7(lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, tmin);
  float mx = magic_min7(lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, tmax);
  *thit = mn;
  return mn <= mx;
}

/* mt_ray_triangle_test, vulkan_ray_tracing.cc:3089-3111 */
static int mt_test(f3 p0, f3 p1, f3 p2, const oray* r, float* thit) {
  f3 v0v1 = sub3(p1, p0);
  f3 v0v2 = sub3(p2, p0);
  f3 pvec = cross3(r->d, v0v2);
  float det = dot3(v0v1, pvec);
  float idet = 1 / det;
  f3 tvec = sub3(r->o, p0);
  float u = dot3(tvec, pvec) * idet;
  if (u < 0 || u > 1) return 0;
  f3 qvec = cross3(tvec, v0v1);
  float v = dot3(r->d, qvec) * idet;
  if (v < 0 || (u + v) > 1) return 0;
  *thit = dot3(v0v2, qvec) * idet;
  return 1;
}

/* Barycentric, vulkan_ray_tracing.cc:3113-3130; returns {v, w, u}. */
static f3 barycentric(f3 p, f3 a, f3 b, f3 c) {
  f3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
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

/* ------------------------------------------------------------------------ */
/* Node decoders: vulkan_acceleration_structure_util.h                         */
/* ------------------------------------------------------------------------ */
static uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}
static uint64_t rd64(const uint8_t* p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}
static float rdf(const uint8_t* p) {
  float v;
  memcpy(&v, p, 4);
  return v;
}

typedef struct {
  uint64_t root_offset;
  f3 bmin, bmax;
} hdr_t; /* GEN_RT_BVH_unpack, util.h:96-107 */

typedef struct {
  f3 origin;
  int32_t child_offset;
  int32_t ex, ey, ez;
  uint32_t size[6], type[6];
  uint32_t lx[6], ux[6], ly[6], uy[6], lz[6], uz[6];
} inode_t; /* GEN_RT_BVH_INTERNAL_NODE_unpack, util.h:134-207 */

typedef struct {
  uint32_t hit_group;
  float w2o[12]; /* WorldToObjectm00..m22 followed by the +52 row (util.h:240-252 field order) */
  uint64_t bvh_address;
  uint32_t instance_id;
  float o2w[12];
} ileaf_t; /* GEN_RT_BVH_INSTANCE_LEAF_unpack, util.h:271-386 */

typedef struct {
  uint32_t geometry_index, leaf_type;
  uint32_t prim0, prim1_delta;
  f3 v[4];
} qleaf_t; /* descriptor + GEN_RT_BVH_QUAD_LEAF_unpack, util.h:402-466 */

static hdr_t unpack_hdr(const uint8_t* p) {
  hdr_t h;
  h.root_offset = rd64(p);
  h.bmin = mk3(rdf(p + 8), rdf(p + 12), rdf(p + 16));
  h.bmax = mk3(rdf(p + 20), rdf(p + 24), rdf(p + 28));
  return h;
}

static inode_t unpack_inode(const uint8_t* p) {
  inode_t n;
  n.origin = mk3(rdf(p), rdf(p + 4), rdf(p + 8));
  n.child_offset = (int32_t)rd32(p + 12);
  n.ex = (int32_t)(int8_t)p[18];
  n.ey = (int32_t)(int8_t)p[19];
  n.ez = (int32_t)(int8_t)p[20];
  for (int i = 0; i < 6; i++) {
    uint8_t t = p[22 + i] & 0x3f;
    n.size[i] = t & 0x03;
    n.type[i] = t >> 2;
  }
  for (int i = 0; i < 6; i++) {
    n.lx[i] = p[28 + i];
    n.ux[i] = p[34 + i];
    n.ly[i] = p[40 + i];
    n.uy[i] = p[46 + i];
    n.lz[i] = p[52 + i];
    n.uz[i] = p[58 + i];
  }
  return n;
}

/* set_child_bounds, util.h:499-508 */
static void child_bounds(const inode_t* n, int i, f3* lo, f3* hi) {
  lo->x = n->origin.x + ldexpf((float)n->lx[i], n->ex - 8);
  lo->y = n->origin.y + ldexpf((float)n->ly[i], n->ey - 8);
  lo->z = n->origin.z + ldexpf((float)n->lz[i], n->ez - 8);
  hi->x = n->origin.x + ldexpf((float)n->ux[i], n->ex - 8);
  hi->y = n->origin.y + ldexpf((float)n->uy[i], n->ey - 8);
  hi->z = n->origin.z + ldexpf((float)n->uz[i], n->ez - 8);
}

static ileaf_t unpack_ileaf(const uint8_t* p) {
  ileaf_t l;
  l.hit_group = rd32(p + 4) & 0x00ffffff;
  for (int k = 0; k < 12; k++) l.w2o[k] = rdf(p + 16 + 4 * k);
  l.bvh_address = rd64(p + 64);
  l.instance_id = rd32(p + 72);
  for (int k = 0; k < 12; k++) l.o2w[k] = rdf(p + 80 + 4 * k);
  return l;
}

static qleaf_t unpack_qleaf(const uint8_t* p) {
  qleaf_t q;
  uint32_t d1 = rd32(p + 4);
  q.geometry_index = d1 & 0x0fffffff;
  q.leaf_type = (d1 >> 29) & 1;
  q.prim0 = rd32(p + 8);
  q.prim1_delta = rd32(p + 12) & 0x0001ffff;
  for (int k = 0; k < 4; k++) q.v[k] = mk3(rdf(p + 16 + 12 * k), rdf(p + 20 + 12 * k), rdf(p + 24 + 12 * k));
  return q;
}

/* ------------------------------------------------------------------------ */
/* Partition (createTreelets + buildNodeToRootMap)                            */
/* ------------------------------------------------------------------------ */
struct orc_partition {
  uint32_t budget;
  uint32_t n;
  uint64_t arena_bytes;
  uint64_t* roots;
  uint64_t* list_offsets;
  uint64_t* list_nodes;
  uint64_t n_list;
  uint32_t* owner; /* arena_bytes / 64 */
  uint64_t total_bytes;
};

typedef struct {
  uint64_t addr;
  uint8_t top, leaf;
  int32_t size;
} pent; /* StackEntry fields used by createTreelets */

typedef struct {
  pent* e;
  int* nodesize;
  size_t head, tail, cap;
} pdeque;

static int pdq_push(pdeque* q, pent e, int nodesize) {
  if (q->tail == q->cap) {
    if (q->head > 0) {
      memmove(q->e, q->e + q->head, (q->tail - q->head) * sizeof(pent));
      memmove(q->nodesize, q->nodesize + q->head, (q->tail - q->head) * sizeof(int));
      q->tail -= q->head;
      q->head = 0;
    }
    if (q->tail == q->cap) {
      size_t nc = q->cap ? 2 * q->cap : 64;
      pent* ne = (pent*)realloc(q->e, nc * sizeof(pent));
      if (!ne) return 0;
      q->e = ne;
      int* ns = (int*)realloc(q->nodesize, nc * sizeof(int));
      if (!ns) return 0;
      q->nodesize = ns;
      q->cap = nc;
    }
  }
  q->e[q->tail] = e;
  q->nodesize[q->tail] = nodesize;
  q->tail++;
  return 1;
}
static int pdq_empty(const pdeque* q) { return q->head == q->tail; }

typedef struct {
  uint64_t* v;
  size_t n, cap;
} u64vec;
static int u64_push(u64vec* a, uint64_t x) {
  if (a->n == a->cap) {
    size_t nc = a->cap ? 2 * a->cap : 64;
    uint64_t* nv = (uint64_t*)realloc(a->v, nc * sizeof(uint64_t));
    if (!nv) return 0;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  return 1;
}

static int in_arena(uint64_t off, uint64_t len, uint64_t bytes) { return off <= bytes && len <= bytes - off; }

void orc_free_partition(orc_partition* p) {
  if (!p) return;
  free(p->roots);
  free(p->list_offsets);
  free(p->list_nodes);
  free(p->owner);
  free(p);
}

int orc_create_treelets(const uint8_t* A, uint64_t bytes, uint64_t tlas, uint32_t budget, int expand_duplicates,
                        orc_partition** out) {
  if (!A || !out || (bytes % 64) != 0 || !in_arena(tlas, 64, bytes)) return ORC_ERR_INVALID_ARGUMENT;
  const uint64_t slots = bytes / 64;
  int st = ORC_OK;
  /* completed_treelet_roots_addr_only: root slot -> node list (latest formation wins) */
  u64vec* lists = (u64vec*)calloc(slots, sizeof(u64vec));
  uint8_t* has_list = (uint8_t*)calloc(slots, 1);
  pdeque P = {0}, S = {0};
  u64vec cur = {0};
  uint64_t total = 0;
  if (!lists || !has_list) {
    st = ORC_ERR_OUT_OF_MEMORY;
    goto done;
  }

  /* vulkan_ray_tracing.cc:846-866 */
  int remaining = (int)budget;
  if (!pdq_push(&P, (pent){tlas, 1, 0, 64}, 128)) goto oom;
  hdr_t top = unpack_hdr(A + tlas);
  remaining -= 64;
  if (remaining < 0) {
    st = ORC_ERR_NODE_LARGER_THAN_BUDGET;
    goto done;
  }
  total += 64;
  if (!u64_push(&cur, tlas)) goto oom;
  uint64_t top_root = tlas + top.root_offset;
  if (!pdq_push(&S, (pent){top_root, 1, 0, 64}, 64)) goto oom;

  while (!pdq_empty(&P)) {
    /* vulkan_ray_tracing.cc:870-887 */
    pent next = S.e[S.head];
    S.head++;
    int have_next = 1;
    while (have_next) {
      uint64_t a = next.addr;
      if (next.top && !next.leaf) { /* :889-940 */
        if (!in_arena(a, 64, bytes)) goto range;
        inode_t n = unpack_inode(A + a);
        remaining -= 64;
        if (remaining < 0) goto budget_err;
        total += 64;
        if (!u64_push(&cur, a)) goto oom;
        uint64_t c = a + (int64_t)n.child_offset * 64;
        for (int i = 0; i < 6; i++) {
          if (n.size[i] > 0) {
            if (n.type[i] != 0) {
              if (n.type[i] != 1) {
                st = ORC_ERR_BAD_NODE_TYPE;
                goto done;
              }
              if (!pdq_push(&S, (pent){c, 1, 1, 128}, 128 + 64)) goto oom;
            } else {
              if (!pdq_push(&S, (pent){c, 1, 0, 64}, 64)) goto oom;
            }
          }
          c += n.size[i] * 64;
        }
      } else if (next.top && next.leaf) { /* :942-996 */
        if (!in_arena(a, 128, bytes)) goto range;
        ileaf_t l = unpack_ileaf(A + a);
        remaining -= 128;
        if (remaining < 0) goto budget_err;
        total += 128;
        if (!u64_push(&cur, a)) goto oom;
        uint64_t h = a + l.bvh_address;
        if (!in_arena(h, 64, bytes)) goto range;
        hdr_t bh = unpack_hdr(A + h);
        remaining -= 64;
        if (!u64_push(&cur, h)) goto oom;
        if (remaining < 0) goto budget_err;
        total += 64;
        if (!pdq_push(&S, (pent){h + bh.root_offset, 0, 0, 64}, 64)) goto oom;
      } else if (!next.top && !next.leaf) { /* :997-1072 */
        if (!in_arena(a, 64, bytes)) goto range;
        inode_t n = unpack_inode(A + a);
        remaining -= 64;
        if (remaining < 0) goto budget_err;
        total += 64;
        if (!u64_push(&cur, a)) goto oom;
        uint64_t c = a + (int64_t)n.child_offset * 64;
        for (int i = 0; i < 6; i++) {
          if (n.size[i] > 0) {
            if (n.type[i] != 0) {
              if (!pdq_push(&S, (pent){c, 0, 1, 64}, 64)) goto oom;
            } else {
              if (!pdq_push(&S, (pent){c, 0, 0, 64}, 64)) goto oom;
            }
          }
          c += n.size[i] * 64;
        }
      } else { /* bottom leaf, :1074-1115 (quad and procedural both subtract 64) */
        if (!in_arena(a, 64, bytes)) goto range;
        if (!u64_push(&cur, a)) goto oom;
        remaining -= 64;
        if (remaining < 0) goto budget_err;
        total += 64;
      }

      /* Fit check, :1117-1271.  The reference reads front() before empty();
       * the empty case reaches the third branch either way. */
      if (!pdq_empty(&S) && remaining - S.nodesize[S.head] >= 0) {
        next = S.e[S.head];
        S.head++;
      } else {
        /* Close the treelet rooted at P.front(). */
        uint64_t root = P.e[P.head].addr;
        uint64_t rs = root / 64;
        if (root % 64 != 0 || rs >= slots) goto range;
        free(lists[rs].v);
        lists[rs] = cur;
        has_list[rs] = 1;
        memset(&cur, 0, sizeof(cur));
        P.head++;
        if (!pdq_empty(&S)) {
          /* :1159-1181 every unvisited BFS entry becomes a pending root */
          for (size_t k = S.head; k < S.tail; k++)
            if (!pdq_push(&P, S.e[k], S.nodesize[k])) goto oom;
        }
        S.head = S.tail = 0;
        if (!expand_duplicates) {
          /* A treelet is a function of its root alone, so forming it again
           * (a BLAS reached through another instance) reproduces the same
           * list and the same pending roots. */
          while (!pdq_empty(&P)) {
            uint64_t r = P.e[P.head].addr;
            if (r % 64 == 0 && r / 64 < slots && has_list[r / 64])
              P.head++;
            else
              break;
          }
        }
        if (!pdq_empty(&P))
          if (!pdq_push(&S, P.e[P.head], P.nodesize[P.head])) goto oom;
        remaining = (int)budget;
        have_next = 0;
      }
    }
  }

  {
    /* De-duplicate each list (keep first occurrence) and number the roots in
     * ascending address order (:1310-1336), then build node->root (:475-489). */
    orc_partition* p = (orc_partition*)calloc(1, sizeof(orc_partition));
    if (!p) goto oom;
    p->budget = budget;
    p->arena_bytes = bytes;
    p->total_bytes = total;
    uint32_t n = 0;
    uint64_t entries = 0;
    uint32_t* stamp = (uint32_t*)calloc(slots, sizeof(uint32_t));
    p->owner = (uint32_t*)malloc(slots * sizeof(uint32_t));
    if (!stamp || !p->owner) {
      free(stamp);
      orc_free_partition(p);
      goto oom;
    }
    for (uint64_t s = 0; s < slots; s++) {
      p->owner[s] = 0xFFFFFFFFu;
      if (has_list[s]) {
        n++;
        entries += lists[s].n;
      }
    }
    p->n = n;
    p->roots = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    p->list_offsets = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    p->list_nodes = (uint64_t*)malloc((entries ? entries : 1) * sizeof(uint64_t));
    if (!p->roots || !p->list_offsets || !p->list_nodes) {
      free(stamp);
      orc_free_partition(p);
      goto oom;
    }
    uint32_t k = 0;
    uint64_t cursor = 0;
    for (uint64_t s = 0; s < slots; s++) {
      if (!has_list[s]) continue;
      p->roots[k] = s * 64;
      p->list_offsets[k] = cursor;
      for (size_t i = 0; i < lists[s].n; i++) {
        uint64_t a = lists[s].v[i];
        uint64_t as = a / 64;
        if (a % 64 != 0 || as >= slots) {
          free(stamp);
          orc_free_partition(p);
          goto range;
        }
        if (stamp[as] == k + 1) continue;
        stamp[as] = k + 1;
        p->list_nodes[cursor++] = a;
      }
      k++;
    }
    p->list_offsets[n] = cursor;
    p->n_list = cursor;
    for (uint32_t t = 0; t < n; t++) {
      p->owner[p->roots[t] / 64] = t;
      for (uint64_t i = p->list_offsets[t]; i < p->list_offsets[t + 1]; i++) p->owner[p->list_nodes[i] / 64] = t;
    }
    free(stamp);
    *out = p;
  }
  goto done;
budget_err:
  st = ORC_ERR_NODE_LARGER_THAN_BUDGET;
  goto done;
range:
  st = ORC_ERR_OFFSET_OUT_OF_ARENA;
  goto done;
oom:
  st = ORC_ERR_OUT_OF_MEMORY;
done:
  if (lists) {
    for (uint64_t s = 0; s < slots; s++) free(lists[s].v);
    free(lists);
  }
  free(has_list);
  free(P.e);
  free(P.nodesize);
  free(S.e);
  free(S.nodesize);
  free(cur.v);
  return st;
}

uint32_t orc_partition_count(const orc_partition* p) { return p ? p->n : 0; }
uint64_t orc_partition_list_entries(const orc_partition* p) { return p ? p->n_list : 0; }
uint64_t orc_partition_total_bytes(const orc_partition* p) { return p ? p->total_bytes : 0; }

int orc_partition_copy(const orc_partition* p, uint64_t* roots, uint64_t* list_offsets, uint64_t* list_nodes,
                       uint32_t* owner) {
  if (!p) return ORC_ERR_INVALID_ARGUMENT;
  if (roots) memcpy(roots, p->roots, p->n * sizeof(uint64_t));
  if (list_offsets) memcpy(list_offsets, p->list_offsets, (p->n + 1) * sizeof(uint64_t));
  if (list_nodes) memcpy(list_nodes, p->list_nodes, p->n_list * sizeof(uint64_t));
  if (owner) memcpy(owner, p->owner, (p->arena_bytes / 64) * sizeof(uint32_t));
  return ORC_OK;
}

/* Slots holding instance leaves: the typed-1 children of TLAS internal nodes. */
static int instance_slots(const uint8_t* A, uint64_t bytes, uint64_t tlas, uint8_t* is_inst) {
  const uint64_t slots = bytes / 64;
  uint8_t* seen = (uint8_t*)calloc(slots, 1);
  uint64_t* q = (uint64_t*)malloc(slots * sizeof(uint64_t));
  if (!seen || !q) {
    free(seen);
    free(q);
    return ORC_ERR_OUT_OF_MEMORY;
  }
  size_t head = 0, tail = 0;
  hdr_t h = unpack_hdr(A + tlas);
  uint64_t root = tlas + h.root_offset;
  int st = ORC_OK;
  if (!in_arena(root, 64, bytes) || root % 64) {
    st = ORC_ERR_OFFSET_OUT_OF_ARENA;
    goto out;
  }
  q[tail++] = root;
  seen[root / 64] = 1;
  while (head < tail) {
    const uint64_t a = q[head++];
    inode_t n = unpack_inode(A + a);
    uint64_t c = a + (int64_t)n.child_offset * 64;
    for (int i = 0; i < 6; i++) {
      if (n.size[i] > 0) {
        if (c % 64 || !in_arena(c, 64, bytes)) {
          st = ORC_ERR_OFFSET_OUT_OF_ARENA;
          goto out;
        }
        if (n.type[i] == 1) {
          is_inst[c / 64] = 1;
        } else if (n.type[i] == 0 && !seen[c / 64]) {
          seen[c / 64] = 1;
          q[tail++] = c;
        }
      }
      c += n.size[i] * 64;
    }
  }
out:
  free(seen);
  free(q);
  return st;
}

int orc_treelet_layout(const uint8_t* A, uint64_t bytes, uint64_t tlas, const orc_partition* p, uint32_t stride,
                       uint64_t* remap, int32_t* metadata) {
  if (!A || !p || !remap || bytes != p->arena_bytes) return ORC_ERR_INVALID_ARGUMENT;
  const uint64_t slots = bytes / 64;
  uint8_t* is_inst = (uint8_t*)calloc(slots, 1);
  if (!is_inst) return ORC_ERR_OUT_OF_MEMORY;
  int st = instance_slots(A, bytes, tlas, is_inst);
  if (st != ORC_OK) {
    free(is_inst);
    return st;
  }
  for (uint64_t s = 0; s < slots; s++) remap[s] = ~0ull;
  const uint64_t block = (uint64_t)p->budget + stride;
  const uint32_t per = p->budget / 64;
  /* for (auto root : treelet_roots): ascending root address (:1484) */
  for (uint32_t t = 0; t < p->n && st == ORC_OK; t++) {
    const uint64_t root = p->roots[t];
    const uint64_t base = (uint64_t)t * block; /* treelet_layout_bvh + i * (max_treelet_size + stride) */
    if (remap[root / 64] != ~0ull) { /* assert(mapping.count(root) == 0), :1483 */
      st = ORC_ERR_INVALID_ARGUMENT;
      break;
    }
    remap[root / 64] = base;
    uint64_t next = base + (is_inst[root / 64] ? 128 : 64); /* + root.first.size */
    for (uint64_t i = p->list_offsets[t]; i < p->list_offsets[t + 1]; i++) {
      const uint64_t a = p->list_nodes[i];
      if (a == root) continue;
      if (remap[a / 64] == ~0ull) remap[a / 64] = next; /* the first mapping stays (:1496-1502) */
      next += is_inst[a / 64] ? 128 : 64;
    }
  }
  if (st == ORC_OK && metadata) {
    for (uint32_t t = 0; t < p->n; t++) {
      int32_t* m = metadata + (uint64_t)t * per;
      const uint64_t r = remap[p->roots[t] / 64];
      uint64_t j = 0;
      for (uint64_t i = p->list_offsets[t]; i < p->list_offsets[t + 1] && j < per; i++, j++)
        m[j] = (int32_t)((int64_t)remap[p->list_nodes[i] / 64] - (int64_t)r);
      for (; j < per; j++) m[j] = -1;
    }
  }
  free(is_inst);
  return st;
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                  */
/* ------------------------------------------------------------------------ */
enum { TX_STRUCT = 0, TX_INTERNAL = 1, TX_INSTANCE = 2, TX_DESC = 3, TX_QUAD = 4, TX_QUAD_HIT = 5, TX_PROC = 6 };

/* ---- tree_level_map (vulkan_ray_tracing.cc:2424, 1662) --------------------
 * A per-ray map from node address (arena offset here) to tree level.  The
 * reference assigns a child `map[node] + 1` with the node's value when the
 * node is processed, a BLAS root its instance leaf's value, and takes the ray's
 * level as the maximum over the final map (:2274-2279, :3062-3067): a node that
 * a later assignment overwrites keeps only its last value.  Open addressing
 * with the keys of the current ray listed, so clearing and the final maximum
 * cost one step per key. */
typedef struct {
  uint64_t* key; /* offset + 1, 0 = empty */
  uint32_t* val;
  uint32_t* used; /* table positions written by the current ray */
  size_t cap, n;
} lmap;

static size_t lm_pos(const lmap* m, uint64_t k) {
  size_t h = (size_t)((k * 0x9E3779B97F4A7C15ull) >> 20) & (m->cap - 1);
  while (m->key[h] != 0 && m->key[h] != k + 1) h = (h + 1) & (m->cap - 1);
  return h;
}
static int lm_grow(lmap* m) {
  size_t nc = m->cap ? 2 * m->cap : 1024;
  uint64_t* nk = (uint64_t*)calloc(nc, sizeof(uint64_t));
  uint32_t* nv = (uint32_t*)malloc(nc * sizeof(uint32_t));
  uint32_t* nu = (uint32_t*)malloc(nc / 2 * sizeof(uint32_t));
  if (!nk || !nv || !nu) {
    free(nk);
    free(nv);
    free(nu);
    return 0;
  }
  lmap o = *m;
  m->key = nk;
  m->val = nv;
  m->used = nu;
  m->cap = nc;
  m->n = 0;
  for (size_t i = 0; i < o.n; i++) {
    size_t h = lm_pos(m, o.key[o.used[i]] - 1);
    m->key[h] = o.key[o.used[i]];
    m->val[h] = o.val[o.used[i]];
    m->used[m->n++] = (uint32_t)h;
  }
  free(o.key);
  free(o.val);
  free(o.used);
  return 1;
}
static int lm_set(lmap* m, uint64_t k, uint32_t v) {
  if (2 * (m->n + 1) > m->cap && !lm_grow(m)) return 0;
  size_t h = lm_pos(m, k);
  if (m->key[h] == 0) {
    m->key[h] = k + 1;
    m->used[m->n++] = (uint32_t)h;
  }
  m->val[h] = v;
  return 1;
}
static uint32_t lm_get(const lmap* m, uint64_t k) { /* the reference asserts the key is present */
  if (m->cap == 0) return 0;
  size_t h = lm_pos(m, k);
  return m->key[h] ? m->val[h] : 0;
}
static uint32_t lm_max_and_clear(lmap* m) {
  uint32_t mx = 0;
  for (size_t i = 0; i < m->n; i++) {
    if (m->val[m->used[i]] > mx) mx = m->val[m->used[i]];
    m->key[m->used[i]] = 0;
  }
  m->n = 0;
  return mx;
}
static void lm_free(lmap* m) {
  free(m->key);
  free(m->val);
  free(m->used);
  memset(m, 0, sizeof(*m));
}

typedef struct {
  const uint8_t* A;
  uint64_t bytes;
  const orc_partition* part;
  orc_record* out; /* NULL: count only */
  uint64_t cap;
  uint64_t base; /* first record index of this ray */
  uint32_t count;
  uint32_t fetch_bytes;
  uint32_t accessed_bytes; /* accessedDataSize += transaction.size (vulkan_ray_tracing.cc:2260) */
  int overflow;
  lmap* lm; /* the ray's tree_level_map */
} sink_t;

static void emit(sink_t* s, uint64_t off, uint32_t size, uint32_t type, uint32_t level) {
  if (type != TX_DESC) s->fetch_bytes += size;
  s->accessed_bytes += size;
  if (s->out) {
    uint64_t idx = s->base + s->count;
    if (idx < s->cap) {
      orc_record r;
      r.offset = off;
      r.treelet = 0xFFFFFFFFu;
      if (s->part && off / 64 < s->part->arena_bytes / 64) r.treelet = s->part->owner[off / 64];
      r.size = (uint16_t)size;
      r.type = (uint8_t)type;
      r.level = (uint8_t)(level > 255 ? 255 : level);
      s->out[idx] = r;
    } else {
      s->overflow = 1;
    }
  }
  s->count++;
}

/* ---- plain traversal, traceRay (vulkan_ray_tracing.cc:2309-3076) -------- */
typedef struct {
  uint64_t addr;
  uint8_t top, leaf;
  uint32_t level;
} sent;

typedef struct {
  sent* e;
  size_t n, cap;
} sstack;
static int ss_push(sstack* s, sent e) {
  if (s->n == s->cap) {
    size_t nc = s->cap ? 2 * s->cap : 64;
    sent* ne = (sent*)realloc(s->e, nc * sizeof(sent));
    if (!ne) return 0;
    s->e = ne;
    s->cap = nc;
  }
  s->e[s->n++] = e;
  return 1;
}

typedef struct {
  int err;
  uint32_t max_level; /* deepest path level assigned (checked against the map's maximum in tests) */
} tstate;


#define CHECK_RANGE(off, len)                  \
  do {                                         \
    if (!in_arena((off), (len), s->bytes)) {   \
      ts->err = ORC_ERR_OFFSET_OUT_OF_ARENA;   \
      goto fail;                               \
    }                                          \
  } while (0)

static void level_set(tstate* ts, uint32_t lv) {
  if (lv > ts->max_level) ts->max_level = lv;
}

static void finish_hit(const uint8_t* A, orc_hit* h, const oray* ray, float min_thit, uint64_t closest_leaf,
                       uint64_t closest_inst, const oray* closest_obj, float min_thit_object) {
  if (min_thit < ray->tmax) {
    qleaf_t q = unpack_qleaf(A + closest_leaf);
    ileaf_t il = unpack_ileaf(A + closest_inst);
    h->hit = 1;
    h->geometry_index = q.geometry_index;
    h->primitive_index = q.prim0;
    h->instance_index = il.instance_id;
    f3 ip = add3(ray->o, mk3(ray->d.x * min_thit, ray->d.y * min_thit, ray->d.z * min_thit));
    h->t = min_thit;
    h->p[0] = ip.x;
    h->p[1] = ip.y;
    h->p[2] = ip.z;
    f3 op = add3(closest_obj->o, mk3(closest_obj->d.x * min_thit_object, closest_obj->d.y * min_thit_object,
                                     closest_obj->d.z * min_thit_object));
    f3 b = barycentric(op, q.v[0], q.v[1], q.v[2]);
    h->bary[0] = b.x;
    h->bary[1] = b.y;
    h->bary[2] = b.z;
    /* Hit_data.worldToObjectMatrix / objectToWorldMatrix come from this leaf (:3002-3003) */
    h->instance_leaf = closest_inst;
  }
}

static int trace_plain(sink_t* s, const uint64_t tlas, const orc_ray* in, orc_hit* h) {
  tstate tsv = {ORC_OK, 0}, *ts = &tsv;
  const uint8_t* A = s->A;
  sstack st = {0};
  memset(h, 0, sizeof(*h));
  h->instance_leaf = ~0ull; /* no closest instance until a hit */
  oray ray;
  ray.o = mk3(in->org[0], in->org[1], in->org[2]);
  ray.d = mk3(in->dir[0], in->dir[1], in->dir[2]);
  ray.tmin = in->tmin;
  ray.tmax = in->tmax;
  const float Tmin = in->tmin, Tmax = in->tmax;
  const int terminateOnFirstHit = (in->flags & 0x4) != 0;
  const int skipAnyHitShader = (in->flags & 0x1) != 0;
  float min_thit = ray.tmax;
  uint64_t closest_leaf = 0, closest_inst = 0;
  oray closest_obj;
  memset(&closest_obj, 0, sizeof(closest_obj));
  float min_thit_object = 0;
  uint32_t n_nodes = 0, n_any = 0;

  CHECK_RANGE(tlas, 64);
  hdr_t top = unpack_hdr(A + tlas);
  emit(s, tlas, 64, TX_STRUCT, 0); /* :2444-2449 */
  uint64_t topRoot = tlas + top.root_offset;
  level_set(ts, 1);
  if (!lm_set(s->lm, topRoot, 1)) goto oom; /* tree_level_map[topRootAddr] = 1 (:2470) */
  {
    float thit;
    if (ray_box_test(top.bmin, top.bmax, calculate_idir(ray.d), ray.o, ray.tmin, ray.tmax, &thit))
      if (!ss_push(&st, (sent){topRoot, 1, 0, 1})) goto oom;
  }

  /* The instance whose BLAS is being traversed (the reference keeps it in
   * locals of the top-level leaf loop, :2602-2668). */
  ileaf_t instanceLeaf;
  uint64_t inst_addr = 0;
  oray objectRay;
  float tMult = 1.0f;

  while (st.n > 0) { /* :2486 */
    uint64_t next = 0;
    uint32_t next_level = 0;
    int have_next = 0;
    if (!st.e[st.n - 1].top) {
      ts->err = ORC_ERR_INVALID_ARGUMENT;
      goto fail;
    }
    if (!st.e[st.n - 1].leaf) {
      next = st.e[st.n - 1].addr;
      next_level = st.e[st.n - 1].level;
      have_next = 1;
      st.n--;
    }
    while (have_next) { /* top-level internal chain, :2499-2599 */
      uint64_t node_addr = next;
      uint32_t node_level = next_level;
      have_next = 0;
      CHECK_RANGE(node_addr, 64);
      inode_t node = unpack_inode(A + node_addr);
      emit(s, node_addr, 64, TX_INTERNAL, node_level);
      n_nodes++;
      int child_hit[6];
      float thit[6];
      for (int i = 0; i < 6; i++) {
        if (node.size[i] > 0) {
          f3 idir = calculate_idir(ray.d);
          f3 lo, hi;
          child_bounds(&node, i, &lo, &hi);
          child_hit[i] = ray_box_test(lo, hi, idir, ray.o, ray.tmin, ray.tmax, &thit[i]);
          if (child_hit[i] && thit[i] >= min_thit) child_hit[i] = 0;
        } else
          child_hit[i] = 0;
      }
      uint64_t child = node_addr + (int64_t)node.child_offset * 64;
      const uint32_t node_map = lm_get(s->lm, node_addr);
      for (int i = 0; i < 6; i++) {
        if (child_hit[i]) {
          if (node.type[i] != 0) {
            if (node.type[i] != 1) {
              ts->err = ORC_ERR_BAD_NODE_TYPE;
              goto fail;
            }
            if (!ss_push(&st, (sent){child, 1, 1, node_level + 1})) goto oom;
          } else if (!have_next) {
            next = child;
            next_level = node_level + 1;
            have_next = 1;
          } else {
            if (!ss_push(&st, (sent){child, 1, 0, node_level + 1})) goto oom;
          }
          level_set(ts, node_level + 1);
          if (!lm_set(s->lm, child, node_map + 1)) goto oom; /* :2569, 2576, 2581 */
        }
        child += node.size[i] * 64;
      }
    }

    while (st.n > 0 && st.e[st.n - 1].leaf) { /* top-level leaves, :2602-2668 */
      sent le = st.e[--st.n];
      uint64_t leaf_addr = le.addr;
      CHECK_RANGE(leaf_addr, 128);
      instanceLeaf = unpack_ileaf(A + leaf_addr);
      inst_addr = leaf_addr;
      emit(s, leaf_addr, 128, TX_INSTANCE, le.level);
      n_nodes++;
      f4x4 w2o = leaf_matrix(instanceLeaf.w2o);
      uint64_t blas = leaf_addr + instanceLeaf.bvh_address;
      CHECK_RANGE(blas, 64);
      hdr_t bh = unpack_hdr(A + blas);
      emit(s, blas, 64, TX_STRUCT, le.level);
      objectRay = make_transformed_ray(&ray, &w2o, &tMult);
      uint64_t botRoot = blas + bh.root_offset;
      if (!ss_push(&st, (sent){botRoot, 0, 0, le.level})) goto oom;
      level_set(ts, le.level);
      if (!lm_set(s->lm, botRoot, lm_get(s->lm, leaf_addr))) goto oom; /* :2668 */

      while (st.n > 0 && !st.e[st.n - 1].top) { /* bottom level, :2679-2986 */
        sent be = st.e[--st.n];
        uint64_t bnext = be.addr;
        uint32_t bnext_level = be.level;
        int bhave = 1;
        while (bhave) { /* bottom internal chain, :2686-2786 */
          uint64_t node_addr = bnext;
          uint32_t node_level = bnext_level;
          bhave = 0;
          CHECK_RANGE(node_addr, 64);
          inode_t node = unpack_inode(A + node_addr);
          emit(s, node_addr, 64, TX_INTERNAL, node_level);
          n_nodes++;
          int child_hit[6];
          float thit[6];
          for (int i = 0; i < 6; i++) {
            if (node.size[i] > 0) {
              f3 idir = calculate_idir(objectRay.d);
              f3 lo, hi;
              child_bounds(&node, i, &lo, &hi);
              child_hit[i] = ray_box_test(lo, hi, idir, objectRay.o, objectRay.tmin, objectRay.tmax, &thit[i]);
              if (child_hit[i] && thit[i] >= min_thit * tMult) child_hit[i] = 0;
            } else
              child_hit[i] = 0;
          }
          uint64_t child = node_addr + (int64_t)node.child_offset * 64;
          const uint32_t node_map = lm_get(s->lm, node_addr);
          for (int i = 0; i < 6; i++) {
            if (child_hit[i]) {
              if (node.type[i] != 0) {
                if (!ss_push(&st, (sent){child, 0, 1, node_level + 1})) goto oom;
              } else if (!bhave) {
                bnext = child;
                bnext_level = node_level + 1;
                bhave = 1;
              } else {
                if (!ss_push(&st, (sent){child, 0, 0, node_level + 1})) goto oom;
              }
              level_set(ts, node_level + 1);
              if (!lm_set(s->lm, child, node_map + 1)) goto oom; /* :2756, 2763, 2767 */
            }
            child += node.size[i] * 64;
          }
        }

        while (st.n > 0 && !st.e[st.n - 1].top && st.e[st.n - 1].leaf) { /* bottom leaves, :2789-2985 */
          const sent leaf_e = st.e[--st.n];
          const uint64_t leaf = leaf_e.addr;
          const uint32_t leaf_level = leaf_e.level;
          CHECK_RANGE(leaf, 64);
          emit(s, leaf, 8, TX_DESC, leaf_level);
          qleaf_t q = unpack_qleaf(A + leaf);
          if (q.leaf_type == 0) {
            if (q.prim1_delta != 0) {
              ts->err = ORC_ERR_BAD_NODE_TYPE;
              goto fail;
            }
            float thit = 0;
            int hit = mt_test(q.v[0], q.v[1], q.v[2], &objectRay, &thit);
            float world_thit = thit / tMult;
            if (hit && Tmin <= world_thit && world_thit <= Tmax) { /* :2842-2975 */
              if (skipAnyHitShader && world_thit < min_thit) min_thit = thit / tMult;
              min_thit_object = thit;
              closest_leaf = leaf;
              closest_inst = inst_addr;
              closest_obj = objectRay;
              emit(s, leaf, 64, TX_QUAD_HIT, leaf_level);
              n_nodes++;
              if (!skipAnyHitShader) n_any++;
              if (terminateOnFirstHit) st.n = 0;
            } else {
              emit(s, leaf, 64, TX_QUAD, leaf_level);
              n_nodes++;
            }
          } else {
            emit(s, leaf, 64, TX_PROC, leaf_level); /* procedural leaf: intersection shader not modelled */
            n_nodes++;
          }
        }
      }
    }
  }
  finish_hit(A, h, &ray, min_thit, closest_leaf, closest_inst, &closest_obj, min_thit_object);
  (void)instanceLeaf;
  h->n_records = s->count;
  h->n_nodes = n_nodes;
  h->fetch_bytes = s->fetch_bytes;
  h->accessed_bytes = s->accessed_bytes;
  h->n_any_hits = n_any;
  h->max_level = lm_max_and_clear(s->lm);
  h->path_level = ts->max_level;
  free(st.e);
  return ORC_OK;
oom:
  ts->err = ORC_ERR_OUT_OF_MEMORY;
fail:
  free(st.e);
  return ts->err;
}

/* ---- treelet traversal, traceRayWithTreelets (:1522-2307) ---------------- */
typedef struct {
  uint64_t addr;
  uint8_t top, leaf;
  uint32_t level;
  float tmult;
  uint64_t inst; /* instance leaf offset (StackEntry::instanceLeaf) */
  oray obj;      /* StackEntry::objectRay */
} tent;

typedef struct {
  tent* e;
  size_t head, tail, cap; /* deque with room at both ends */
} tdeque;

static int td_grow(tdeque* d) {
  size_t n = d->tail - d->head;
  size_t nc = d->cap ? 2 * d->cap : 128;
  tent* ne = (tent*)malloc(nc * sizeof(tent));
  if (!ne) return 0;
  size_t nh = (nc - n) / 2;
  if (n) memcpy(ne + nh, d->e + d->head, n * sizeof(tent));
  free(d->e);
  d->e = ne;
  d->head = nh;
  d->tail = nh + n;
  d->cap = nc;
  return 1;
}
static int td_push_front(tdeque* d, tent e) {
  if (d->head == 0 && !td_grow(d)) return 0;
  d->e[--d->head] = e;
  return 1;
}
static int td_push_back(tdeque* d, tent e) {
  if (d->tail == d->cap && !td_grow(d)) return 0;
  d->e[d->tail++] = e;
  return 1;
}
static int td_empty(const tdeque* d) { return d->head == d->tail; }

static uint64_t root_of(const orc_partition* p, uint64_t off) {
  /* addrToTreeletID (:468-472): the node's root address */
  return p->roots[p->owner[off / 64]];
}

static int trace_treelet(sink_t* s, const uint64_t tlas, const orc_ray* in, orc_hit* h, tdeque* cur,
                         tdeque* other) {
  tstate tsv = {ORC_OK, 0}, *ts = &tsv;
  const uint8_t* A = s->A;
  const orc_partition* P = s->part;
  memset(h, 0, sizeof(*h));
  h->instance_leaf = ~0ull; /* no closest instance until a hit */
  cur->head = cur->tail = cur->cap / 2;
  other->head = other->tail = other->cap / 2;
  oray ray;
  ray.o = mk3(in->org[0], in->org[1], in->org[2]);
  ray.d = mk3(in->dir[0], in->dir[1], in->dir[2]);
  ray.tmin = in->tmin;
  ray.tmax = in->tmax;
  const float Tmin = in->tmin, Tmax = in->tmax;
  const int terminateOnFirstHit = (in->flags & 0x4) != 0;
  float min_thit = ray.tmax;
  uint64_t closest_leaf = 0, closest_inst = 0;
  oray closest_obj;
  memset(&closest_obj, 0, sizeof(closest_obj));
  float min_thit_object = 0;
  uint32_t n_nodes = 0;

  CHECK_RANGE(tlas, 64);
  hdr_t top = unpack_hdr(A + tlas);
  emit(s, tlas, 64, TX_STRUCT, 0);
  uint64_t topRoot = tlas + top.root_offset;
  /* current_treelet_root starts as the TLAS header's device address (:1707).
   * Once a node is moved from the other list, the reference stores that
   * node's host address (:1752), which never equals a device treelet root;
   * `host_root` models that state. */
  uint64_t current_root = tlas;
  int host_root = 0;
  level_set(ts, 1);
  if (!lm_set(s->lm, topRoot, 1)) goto oom; /* tree_level_map[topRootAddr] = 1 (:1710) */
  tent z;
  memset(&z, 0, sizeof(z));
  {
    float thit;
    if (ray_box_test(top.bmin, top.bmax, calculate_idir(ray.d), ray.o, ray.tmin, ray.tmax, &thit)) {
      CHECK_RANGE(topRoot, 64);
      tent e = z;
      e.addr = topRoot;
      e.top = 1;
      e.level = 1;
      if (!host_root && current_root == root_of(P, topRoot)) {
        if (!td_push_back(cur, e)) goto oom;
      } else if (!td_push_back(other, e)) {
        goto oom;
      }
    }
  }

  while (!td_empty(cur) || !td_empty(other)) { /* :1742 */
    if (td_empty(cur)) {
      if (!td_push_front(cur, other->e[other->head])) goto oom;
      host_root = 1; /* current_treelet_root = other_treelet_stack.front().addr (host address) */
      other->head++;
    }
    tent cn = cur->e[cur->head++];
    if (cn.top && !cn.leaf) { /* :1758-1870 */
      CHECK_RANGE(cn.addr, 64);
      inode_t node = unpack_inode(A + cn.addr);
      emit(s, cn.addr, 64, TX_INTERNAL, cn.level);
      n_nodes++;
      int child_hit[6];
      float thit[6];
      for (int i = 0; i < 6; i++) {
        if (node.size[i] > 0) {
          f3 idir = calculate_idir(ray.d);
          f3 lo, hi;
          child_bounds(&node, i, &lo, &hi);
          child_hit[i] = ray_box_test(lo, hi, idir, ray.o, ray.tmin, ray.tmax, &thit[i]);
          if (child_hit[i] && thit[i] >= min_thit) child_hit[i] = 0;
        } else
          child_hit[i] = 0;
      }
      uint64_t child = cn.addr + (int64_t)node.child_offset * 64;
      const uint32_t node_map = lm_get(s->lm, cn.addr); /* the node's value now, not when it was pushed */
      for (int i = 0; i < 6; i++) {
        if (child_hit[i]) {
          if (node.type[i] != 0 && node.type[i] != 1) {
            ts->err = ORC_ERR_BAD_NODE_TYPE;
            goto fail;
          }
          CHECK_RANGE(child, 64);
          tent e = z;
          e.addr = child;
          e.top = 1;
          e.leaf = node.type[i] != 0;
          e.level = cn.level + 1;
          if (!host_root && current_root == root_of(P, child)) {
            if (!td_push_front(cur, e)) goto oom;
          } else if (!td_push_front(other, e)) {
            goto oom;
          }
          level_set(ts, cn.level + 1);
          if (!lm_set(s->lm, child, node_map + 1)) goto oom; /* :1837, 1858 */
        }
        child += node.size[i] * 64;
      }
    } else if (cn.top && cn.leaf) { /* :1872-1955 */
      CHECK_RANGE(cn.addr, 128);
      ileaf_t il = unpack_ileaf(A + cn.addr);
      emit(s, cn.addr, 128, TX_INSTANCE, cn.level);
      n_nodes++;
      f4x4 w2o = leaf_matrix(il.w2o);
      uint64_t blas = cn.addr + il.bvh_address;
      CHECK_RANGE(blas, 64);
      hdr_t bh = unpack_hdr(A + blas);
      emit(s, blas, 64, TX_STRUCT, cn.level);
      tent e = z;
      e.obj = make_transformed_ray(&ray, &w2o, &e.tmult);
      e.addr = blas + bh.root_offset;
      CHECK_RANGE(e.addr, 64);
      e.top = 0;
      e.leaf = 0;
      e.level = cn.level;
      e.inst = cn.addr;
      if (!host_root && current_root == root_of(P, e.addr)) {
        if (!td_push_front(cur, e)) goto oom;
      } else if (!td_push_front(other, e)) {
        goto oom;
      }
      level_set(ts, cn.level);
      if (!lm_set(s->lm, e.addr, lm_get(s->lm, cn.addr))) goto oom; /* :1944 */
    } else if (!cn.top && !cn.leaf) { /* :1957-2072 */
      CHECK_RANGE(cn.addr, 64);
      inode_t node = unpack_inode(A + cn.addr);
      emit(s, cn.addr, 64, TX_INTERNAL, cn.level);
      n_nodes++;
      int child_hit[6];
      float thit[6];
      for (int i = 0; i < 6; i++) {
        if (node.size[i] > 0) {
          f3 idir = calculate_idir(cn.obj.d);
          f3 lo, hi;
          child_bounds(&node, i, &lo, &hi);
          child_hit[i] = ray_box_test(lo, hi, idir, cn.obj.o, cn.obj.tmin, cn.obj.tmax, &thit[i]);
          if (child_hit[i] && thit[i] >= min_thit * cn.tmult) child_hit[i] = 0;
        } else
          child_hit[i] = 0;
      }
      uint64_t child = cn.addr + (int64_t)node.child_offset * 64;
      const uint32_t node_map = lm_get(s->lm, cn.addr);
      for (int i = 0; i < 6; i++) {
        if (child_hit[i]) {
          CHECK_RANGE(child, 64);
          tent e = cn;
          e.addr = child;
          e.leaf = node.type[i] != 0;
          e.level = cn.level + 1;
          if (!host_root && current_root == root_of(P, child)) {
            if (!td_push_front(cur, e)) goto oom;
          } else if (!td_push_front(other, e)) {
            goto oom;
          }
          level_set(ts, cn.level + 1);
          if (!lm_set(s->lm, child, node_map + 1)) goto oom; /* :2034, 2055 */
        }
        child += node.size[i] * 64;
      }
    } else { /* bottom leaf, :2073-2203 */
      CHECK_RANGE(cn.addr, 64);
      emit(s, cn.addr, 8, TX_DESC, cn.level);
      qleaf_t q = unpack_qleaf(A + cn.addr);
      if (q.leaf_type == 0) {
        if (q.prim1_delta != 0) {
          ts->err = ORC_ERR_BAD_NODE_TYPE;
          goto fail;
        }
        float thit = 0;
        int hit = mt_test(q.v[0], q.v[1], q.v[2], &cn.obj, &thit);
        float world_thit = thit / cn.tmult;
        if (hit && Tmin <= world_thit && world_thit <= Tmax && world_thit < min_thit) {
          min_thit = thit / cn.tmult;
          min_thit_object = thit;
          closest_leaf = cn.addr;
          closest_inst = cn.inst;
          closest_obj = cn.obj;
          emit(s, cn.addr, 64, TX_QUAD_HIT, cn.level);
          n_nodes++;
          if (terminateOnFirstHit) {
            cur->head = cur->tail;
            other->head = other->tail;
          }
        } else {
          emit(s, cn.addr, 64, TX_QUAD, cn.level);
          n_nodes++;
        }
      } else {
        emit(s, cn.addr, 64, TX_PROC, cn.level);
        n_nodes++;
      }
    }
  }
  finish_hit(A, h, &ray, min_thit, closest_leaf, closest_inst, &closest_obj, min_thit_object);
  h->n_records = s->count;
  h->n_nodes = n_nodes;
  h->fetch_bytes = s->fetch_bytes;
  h->accessed_bytes = s->accessed_bytes;
  h->n_any_hits = 0;
  h->max_level = lm_max_and_clear(s->lm);
  h->path_level = ts->max_level;
  return ORC_OK;
oom:
  ts->err = ORC_ERR_OUT_OF_MEMORY;
fail:
  return ts->err;
}

static int trace_range(const uint8_t* A, uint64_t bytes, uint64_t tlas, const orc_partition* part, int mode,
                       const orc_ray* rays, uint32_t begin, uint32_t end, orc_hit* hits, orc_record* records,
                       uint64_t cap, uint64_t* offsets, uint64_t* cursor_io, int* overflow) {
  tdeque cur = {0}, other = {0};
  lmap lm = {0};
  int st = ORC_OK;
  uint64_t cursor = cursor_io ? *cursor_io : 0;
  if (mode == 1 && (!td_grow(&cur) || !td_grow(&other))) {
    st = ORC_ERR_OUT_OF_MEMORY;
    goto out;
  }
  for (uint32_t i = begin; i < end; i++) {
    sink_t s;
    memset(&s, 0, sizeof(s));
    s.A = A;
    s.bytes = bytes;
    s.part = part;
    s.out = records;
    s.cap = cap;
    s.base = cursor;
    s.lm = &lm;
    (void)lm_max_and_clear(&lm); /* a fresh map per ray */
    st = mode == 0 ? trace_plain(&s, tlas, &rays[i], &hits[i]) : trace_treelet(&s, tlas, &rays[i], &hits[i], &cur, &other);
    if (st != ORC_OK) goto out;
    if (s.overflow) *overflow = 1;
    if (offsets) offsets[i] = cursor;
    cursor += s.count;
  }
out:
  if (cursor_io) *cursor_io = cursor;
  free(cur.e);
  free(other.e);
  lm_free(&lm);
  return st;
}

int orc_trace(const uint8_t* A, uint64_t bytes, uint64_t tlas, const orc_partition* part, int mode,
              const orc_ray* rays, uint32_t n, orc_hit* hits, orc_record* records, uint64_t cap,
              uint64_t* offsets) {
  if (!A || !rays || !hits || (mode != 0 && mode != 1) || (mode == 1 && !part)) return ORC_ERR_INVALID_ARGUMENT;
  if (part && part->arena_bytes != bytes) return ORC_ERR_INVALID_ARGUMENT;
  uint64_t cursor = 0;
  int overflow = 0;
  int st = trace_range(A, bytes, tlas, part, mode, rays, 0, n, hits, records, cap, offsets, &cursor, &overflow);
  if (st != ORC_OK) return st;
  if (offsets) offsets[n] = cursor;
  return overflow ? ORC_ERR_TRACE_OVERFLOW : ORC_OK;
}

typedef struct {
  const uint8_t* A;
  uint64_t bytes, tlas;
  const orc_partition* part;
  int mode;
  const orc_ray* rays;
  orc_hit* hits;
  uint32_t begin, end;
  int status;
} job_t;

static void* job_run(void* arg) {
  job_t* j = (job_t*)arg;
  int overflow = 0;
  j->status = trace_range(j->A, j->bytes, j->tlas, j->part, j->mode, j->rays, j->begin, j->end, j->hits, NULL, 0,
                          NULL, NULL, &overflow);
  return NULL;
}

int orc_trace_parallel(const uint8_t* A, uint64_t bytes, uint64_t tlas, const orc_partition* part, int mode,
                       const orc_ray* rays, uint32_t n, orc_hit* hits, int threads) {
  if (!A || !rays || !hits || threads < 1 || (mode != 0 && mode != 1) || (mode == 1 && !part))
    return ORC_ERR_INVALID_ARGUMENT;
  if (threads > 256) threads = 256;
  job_t jobs[256];
  pthread_t tid[256];
  int created[256];
  /* Contiguous blocks of rays, one per thread. */
  uint32_t per = (n + (uint32_t)threads - 1) / (uint32_t)threads;
  for (int t = 0; t < threads; t++) {
    uint32_t b = (uint32_t)t * per, e = b + per;
    if (b > n) b = n;
    if (e > n) e = n;
    jobs[t] = (job_t){A, bytes, tlas, part, mode, rays, hits, b, e, ORC_OK};
    created[t] = 0;
    if (b >= e) continue;
    if (pthread_create(&tid[t], NULL, job_run, &jobs[t]) == 0)
      created[t] = 1;
    else
      job_run(&jobs[t]);
  }
  int st = ORC_OK;
  for (int t = 0; t < threads; t++) {
    if (created[t]) pthread_join(tid[t], NULL);
    if (jobs[t].status != ORC_OK) st = jobs[t].status;
  }
  return st;
}

/* ------------------------------------------------------------------------ */
/* Exposed math                                                               */
/* ------------------------------------------------------------------------ */
int orc_ray_box_test(const float lo[3], const float hi[3], const float idir[3], const float org[3], float tmin,
                     float tmax, float* thit) {
  return ray_box_test(mk3(lo[0], lo[1], lo[2]), mk3(hi[0], hi[1], hi[2]), mk3(idir[0], idir[1], idir[2]),
                      mk3(org[0], org[1], org[2]), tmin, tmax, thit);
}

void orc_calculate_idir(const float dir[3], float idir[3]) {
  f3 r = calculate_idir(mk3(dir[0], dir[1], dir[2]));
  idir[0] = r.x;
  idir[1] = r.y;
  idir[2] = r.z;
}

int orc_mt_ray_triangle_test(const float p0[3], const float p1[3], const float p2[3], const float org[3],
                             const float dir[3], float* thit) {
  oray r;
  r.o = mk3(org[0], org[1], org[2]);
  r.d = mk3(dir[0], dir[1], dir[2]);
  r.tmin = 0;
  r.tmax = 0;
  return mt_test(mk3(p0[0], p0[1], p0[2]), mk3(p1[0], p1[1], p1[2]), mk3(p2[0], p2[1], p2[2]), &r, thit);
}

void orc_barycentric(const float p[3], const float a[3], const float b[3], const float c[3], float out[3]) {
  f3 r = barycentric(mk3(p[0], p[1], p[2]), mk3(a[0], a[1], a[2]), mk3(b[0], b[1], b[2]), mk3(c[0], c[1], c[2]));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

void orc_set_child_bounds(const uint8_t* node, int child, float lo[3], float hi[3]) {
  inode_t n = unpack_inode(node);
  f3 l, h;
  child_bounds(&n, child, &l, &h);
  lo[0] = l.x;
  lo[1] = l.y;
  lo[2] = l.z;
  hi[0] = h.x;
  hi[1] = h.y;
  hi[2] = h.z;
}

void orc_make_transformed_ray(const float rows[12], const float org[3], const float dir[3], float tmin, float tmax,
                              float out_org[3], float out_dir[3], float* out_tmin, float* out_tmax,
                              float* t_multiplier) {
  f4x4 M = leaf_matrix(rows);
  oray r;
  r.o = mk3(org[0], org[1], org[2]);
  r.d = mk3(dir[0], dir[1], dir[2]);
  r.tmin = tmin;
  r.tmax = tmax;
  oray t = make_transformed_ray(&r, &M, t_multiplier);
  out_org[0] = t.o.x;
  out_org[1] = t.o.y;
  out_org[2] = t.o.z;
  out_dir[0] = t.d.x;
  out_dir[1] = t.d.y;
  out_dir[2] = t.d.z;
  *out_tmin = t.tmin;
  *out_tmax = t.tmax;
}

int orc_validate(const uint8_t* A, uint64_t bytes, uint64_t tlas, uint64_t* n_nodes) {
  if (!A || bytes % 64 || !in_arena(tlas, 64, bytes)) return ORC_ERR_INVALID_ARGUMENT;
  uint64_t slots = bytes / 64;
  uint8_t* seen = (uint8_t*)calloc(slots, 1);
  u64vec q = {0};
  int st = ORC_OK;
  uint64_t count = 0;
  if (!seen) return ORC_ERR_OUT_OF_MEMORY;
  /* queue entries: offset | kind<<62 (0 internal top, 1 internal bottom) */
  hdr_t th = unpack_hdr(A + tlas);
  seen[tlas / 64] = 1;
  count++;
  uint64_t r = tlas + th.root_offset;
  if (r % 64 || !in_arena(r, 64, bytes)) {
    st = ORC_ERR_OFFSET_OUT_OF_ARENA;
    goto out;
  }
  if (!u64_push(&q, r)) goto oom;
  for (size_t qi = 0; qi < q.n; qi++) {
    uint64_t a = q.v[qi] & ((1ull << 62) - 1);
    int bottom = (q.v[qi] >> 62) & 1;
    if (seen[a / 64]) continue;
    seen[a / 64] = 1;
    count++;
    inode_t n = unpack_inode(A + a);
    uint64_t c = a + (int64_t)n.child_offset * 64;
    for (int i = 0; i < 6; i++) {
      if (n.size[i] == 0) continue;
      uint64_t len = (uint64_t)n.size[i] * 64;
      if (c % 64 || !in_arena(c, len, bytes)) {
        st = ORC_ERR_OFFSET_OUT_OF_ARENA;
        goto out;
      }
      if (n.type[i] == 0) {
        if (!u64_push(&q, c | ((uint64_t)bottom << 62))) goto oom;
      } else if (!bottom) {
        if (n.type[i] != 1 || n.size[i] != 2) {
          st = ORC_ERR_BAD_NODE_TYPE;
          goto out;
        }
        if (!seen[c / 64]) {
          seen[c / 64] = 1;
          count++;
        }
        ileaf_t il = unpack_ileaf(A + c);
        uint64_t h = c + il.bvh_address;
        if (h % 64 || !in_arena(h, 64, bytes)) {
          st = ORC_ERR_OFFSET_OUT_OF_ARENA;
          goto out;
        }
        hdr_t bh = unpack_hdr(A + h);
        uint64_t br = h + bh.root_offset;
        if (br % 64 || !in_arena(br, 64, bytes)) {
          st = ORC_ERR_OFFSET_OUT_OF_ARENA;
          goto out;
        }
        if (!seen[h / 64]) {
          seen[h / 64] = 1;
          count++;
        }
        if (!u64_push(&q, br | (1ull << 62))) goto oom;
      } else {
        if (n.type[i] == 1) {
          st = ORC_ERR_BAD_NODE_TYPE;
          goto out;
        }
        if (!seen[c / 64]) {
          seen[c / 64] = 1;
          count++;
        }
      }
      c += len;
    }
  }
  goto out;
oom:
  st = ORC_ERR_OUT_OF_MEMORY;
out:
  if (n_nodes) *n_nodes = count;
  free(seen);
  free(q.v);
  return st;
}
