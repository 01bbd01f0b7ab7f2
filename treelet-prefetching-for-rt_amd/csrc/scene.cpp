// scene.cpp — seeded procedural scenes and ray batches.
//
// Stand-in for the LumiBench / RayTracingInVulkan scenes and raygen shaders
// that drive the reference (README.md:87-108); those are external to it.
// Every generator is a pure function of its seed (SplitMix64 streams), so a
// workload is reproduced bit for bit on any host.
#include <cmath>
#include <cstring>
#include <vector>

#include "rtx_scene.h"

namespace rtx_internal {
bool builder_triangle(const rtx_builder* b, uint32_t instance_id, uint32_t prim, float v[9], double o2w[12]);
}

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t mix_seed(uint64_t a, uint64_t b) {
  SplitMix64 g(a ^ (b * 0xd1b54a32d192ed03ull));
  return g.next();
}

const double kPi = 3.14159265358979323846;

void sphere_triangles(uint32_t slices, uint32_t stacks, double radius, double disp, uint64_t seed, float* out) {
  SplitMix64 rng(seed);
  const uint32_t rows = stacks + 1;
  std::vector<double> pos(3 * (size_t)slices * rows);
  for (uint32_t j = 0; j < rows; ++j) {
    double theta = kPi * j / stacks;
    for (uint32_t i = 0; i < slices; ++i) {
      double phi = 2.0 * kPi * i / slices;
      double r = radius * (1.0 + disp * (2.0 * rng.uniform() - 1.0));
      double* p = &pos[3 * ((size_t)j * slices + i)];
      p[0] = r * std::sin(theta) * std::cos(phi);
      p[1] = r * std::cos(theta);
      p[2] = r * std::sin(theta) * std::sin(phi);
    }
  }
  size_t t = 0;
  auto put = [&](uint32_t i, uint32_t j) {
    const double* p = &pos[3 * ((size_t)j * slices + (i % slices))];
    out[t++] = (float)p[0];
    out[t++] = (float)p[1];
    out[t++] = (float)p[2];
  };
  for (uint32_t j = 0; j < stacks; ++j)
    for (uint32_t i = 0; i < slices; ++i) {
      put(i, j);
      put(i, j + 1);
      put(i + 1, j);
      put(i + 1, j);
      put(i, j + 1);
      put(i + 1, j + 1);
    }
}

struct Camera {
  double eye[3], look[3], up[3], fov_deg;
};

Camera workload_camera(const rtx_workload_desc* d) {
  Camera c{};
  if (d->kind == RTX_WORKLOAD_SPHERE_AND_QUAD) {
    c = {{0.0, 0.6, 3.6}, {0.0, -0.15, 0.0}, {0.0, 1.0, 0.0}, 45.0};
  } else {
    // In front of the box, close enough that the box fills the view.
    double h = 0.5 * d->box_size;
    c = {{h, h, -0.4 * d->box_size}, {h, h, h}, {0.0, 1.0, 0.0}, 80.0};
  }
  return c;
}

void normalize3(double v[3]) {
  double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n > 0) {
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
  }
}

void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

extern "C" {

rtx_status rtx_mesh_displaced_sphere(uint32_t slices, uint32_t stacks, float radius, float disp, uint64_t seed,
                                     float* tris_out) {
  if (!tris_out || slices < 3 || stacks < 2 || !(radius > 0.0f) || !(disp >= 0.0f) || disp >= 1.0f)
    return RTX_ERR_INVALID_ARGUMENT;
  sphere_triangles(slices, stacks, radius, disp, seed, tris_out);
  return RTX_OK;
}

}  // extern "C"

namespace {

// Adds the workload's meshes and instances to a fresh builder.
rtx_status populate_workload(const rtx_workload_desc* d, rtx_builder* b) {
  rtx_status st = RTX_OK;
  const float identity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  if (d->kind == RTX_WORKLOAD_SPHERE_AND_QUAD) {
    uint32_t n = 2 * d->sphere_slices * d->sphere_stacks;
    std::vector<float> tris(9 * (size_t)n);
    st = rtx_mesh_displaced_sphere(d->sphere_slices, d->sphere_stacks, 1.0f, d->displacement, d->seed, tris.data());
    uint32_t sphere = 0, quad = 0;
    if (st == RTX_OK) st = rtx_builder_add_blas(b, tris.data(), n, 0, &sphere);
    const float g = -1.1f, e = 4.0f;
    const float qt[18] = {-e, g, -e, -e, g, e, e, g, -e, e, g, -e, -e, g, e, e, g, e};
    if (st == RTX_OK) st = rtx_builder_add_blas(b, qt, 2, 0, &quad);
    if (st == RTX_OK) st = rtx_builder_add_instance(b, sphere, identity, 0, 0);
    if (st == RTX_OK) st = rtx_builder_add_instance(b, quad, identity, 1, 0);
  } else if (d->kind == RTX_WORKLOAD_INSTANCED_SPHERES) {
    if (d->n_instances == 0 || d->box_size <= 0.0f || d->scale_min <= 0.0f || d->scale_max < d->scale_min)
      return RTX_ERR_INVALID_ARGUMENT;
    uint32_t n = 2 * d->sphere_slices * d->sphere_stacks;
    std::vector<float> tris(9 * (size_t)n);
    uint32_t shared = 0;
    if (d->shared_blas) {
      st = rtx_mesh_displaced_sphere(d->sphere_slices, d->sphere_stacks, 1.0f, d->displacement, d->seed, tris.data());
      if (st == RTX_OK) st = rtx_builder_add_blas(b, tris.data(), n, 0, &shared);
    }
    SplitMix64 place(mix_seed(d->seed, 0x5ca77e7ull));
    for (uint32_t k = 0; k < d->n_instances && st == RTX_OK; ++k) {
      uint32_t id = shared;
      if (!d->shared_blas) {
        st = rtx_mesh_displaced_sphere(d->sphere_slices, d->sphere_stacks, 1.0f, d->displacement,
                                       mix_seed(d->seed, k + 1), tris.data());
        if (st == RTX_OK) st = rtx_builder_add_blas(b, tris.data(), n, 0, &id);
      }
      double px = place.uniform() * d->box_size, py = place.uniform() * d->box_size,
             pz = place.uniform() * d->box_size;
      double s = d->scale_min + place.uniform() * (d->scale_max - d->scale_min);
      float m[12] = {(float)s, 0, 0, (float)px, 0, (float)s, 0, (float)py, 0, 0, (float)s, (float)pz};
      if (st == RTX_OK) st = rtx_builder_add_instance(b, id, m, k, 0);
    }
  } else {
    st = RTX_ERR_INVALID_ARGUMENT;
  }
  return st;
}

}  // namespace

extern "C" {

rtx_status rtx_workload_scene(const rtx_workload_desc* d, uint32_t threads, rtx_builder** out) {
  if (!d || !out) return RTX_ERR_INVALID_ARGUMENT;
  rtx_builder* b = nullptr;
  rtx_status st = rtx_builder_create(&b);
  if (st != RTX_OK) return st;
  st = populate_workload(d, b);
  if (st == RTX_OK) st = rtx_builder_finish(b, threads, nullptr, nullptr);
  if (st != RTX_OK) {
    rtx_builder_destroy(b);
    return st;
  }
  *out = b;
  return RTX_OK;
}

rtx_status rtx_workload_scene_device(const rtx_workload_desc* d, rtx_ctx* ctx, rtx_builder** out) {
  if (!d || !ctx || !out) return RTX_ERR_INVALID_ARGUMENT;
  rtx_builder* b = nullptr;
  rtx_status st = rtx_builder_create(&b);
  if (st != RTX_OK) return st;
  st = populate_workload(d, b);
  if (st == RTX_OK) st = rtx_builder_finish_device(b, ctx, nullptr, nullptr);
  if (st != RTX_OK) {
    rtx_builder_destroy(b);
    return st;
  }
  *out = b;
  return RTX_OK;
}

rtx_status rtx_workload_builder(const rtx_workload_desc* d, rtx_builder** out) {
  if (!d || !out) return RTX_ERR_INVALID_ARGUMENT;
  rtx_builder* b = nullptr;
  rtx_status st = rtx_builder_create(&b);
  if (st != RTX_OK) return st;
  st = populate_workload(d, b);
  if (st != RTX_OK) {
    rtx_builder_destroy(b);
    return st;
  }
  *out = b;
  return RTX_OK;
}

rtx_status rtx_workload_primary_rays(const rtx_workload_desc* d, rtx_ray* rays) {
  if (!d || !rays || d->width == 0 || d->height == 0) return RTX_ERR_INVALID_ARGUMENT;
  Camera c = workload_camera(d);
  double fwd[3] = {c.look[0] - c.eye[0], c.look[1] - c.eye[1], c.look[2] - c.eye[2]};
  normalize3(fwd);
  double right[3], up[3];
  cross3(fwd, c.up, right);
  normalize3(right);
  cross3(right, fwd, up);
  const double tan_half = std::tan(0.5 * c.fov_deg * kPi / 180.0);
  const double aspect = (double)d->width / d->height;
  // Rays are ordered by 8x8 pixel tiles (row-major tiles, row-major pixels
  // inside a tile), so the 64 lanes of a wave trace one compact tile, as the
  // reference groups 8x4 pixels per 32-wide warp (WARP_8X4,
  // vulkan_ray_tracing.cc:3503-3592).  Partial tiles at the right and bottom
  // edges keep the same order over the pixels that exist.
  const uint32_t tiles_x = (d->width + 7) / 8;
  size_t k = 0;
  for (uint32_t ty = 0; ty < (d->height + 7) / 8; ++ty)
    for (uint32_t tx = 0; tx < tiles_x; ++tx)
      for (uint32_t iy = 0; iy < 8; ++iy)
        for (uint32_t ix = 0; ix < 8; ++ix) {
      const uint32_t x = tx * 8 + ix, y = ty * 8 + iy;
      if (x >= d->width || y >= d->height) continue;
      double u = (2.0 * (x + 0.5) / d->width - 1.0) * tan_half * aspect;
      double v = (1.0 - 2.0 * (y + 0.5) / d->height) * tan_half;
      double dir[3];
      for (int a = 0; a < 3; ++a) dir[a] = fwd[a] + u * right[a] + v * up[a];
      normalize3(dir);
      rtx_ray& r = rays[k++];
      memset(&r, 0, sizeof(r));
      for (int a = 0; a < 3; ++a) {
        r.org[a] = (float)c.eye[a];
        r.dir[a] = (float)dir[a];
      }
      r.tmin = 1e-3f;
      r.tmax = 1e30f;
      r.flags = d->ray_flags;
      r.cull_mask = 0xff;
    }
  return RTX_OK;
}

rtx_status rtx_workload_secondary_rays(const rtx_builder* scene, const rtx_ray* rays, const rtx_hit* hits, uint32_t n,
                                       uint32_t kind, uint64_t seed, const float light_pos[3], rtx_ray* out,
                                       uint32_t* src_index, uint32_t* n_out) {
  if (!scene || !rays || !hits || !out || !n_out || kind > 1 || (kind == 1 && !light_pos))
    return RTX_ERR_INVALID_ARGUMENT;
  uint32_t m = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!hits[i].hit) continue;
    float v[9];
    double o2w[12];
    if (!rtx_internal::builder_triangle(scene, hits[i].instance_index, hits[i].primitive_index, v, o2w))
      return RTX_ERR_INVALID_ARGUMENT;
    double e1[3] = {(double)v[3] - v[0], (double)v[4] - v[1], (double)v[5] - v[2]};
    double e2[3] = {(double)v[6] - v[0], (double)v[7] - v[1], (double)v[8] - v[2]};
    double no[3];
    cross3(e1, e2, no);
    // World normal = inverse-transpose of the linear part; cofactor matrix of M is proportional to it.
    const double* M = o2w;
    double C[9] = {M[5] * M[10] - M[6] * M[9], M[6] * M[8] - M[4] * M[10], M[4] * M[9] - M[5] * M[8],
                   M[2] * M[9] - M[1] * M[10], M[0] * M[10] - M[2] * M[8], M[1] * M[8] - M[0] * M[9],
                   M[1] * M[6] - M[2] * M[5], M[2] * M[4] - M[0] * M[6], M[0] * M[5] - M[1] * M[4]};
    double nw[3] = {C[0] * no[0] + C[1] * no[1] + C[2] * no[2], C[3] * no[0] + C[4] * no[1] + C[5] * no[2],
                    C[6] * no[0] + C[7] * no[1] + C[8] * no[2]};
    normalize3(nw);
    const double d[3] = {rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]};
    if (nw[0] * d[0] + nw[1] * d[1] + nw[2] * d[2] > 0) {
      nw[0] = -nw[0];
      nw[1] = -nw[1];
      nw[2] = -nw[2];
    }
    double p[3] = {hits[i].p[0], hits[i].p[1], hits[i].p[2]};
    double scale = 1.0 + std::fabs(p[0]) + std::fabs(p[1]) + std::fabs(p[2]);
    double eps = 1e-4 * scale;
    double o[3] = {p[0] + eps * nw[0], p[1] + eps * nw[1], p[2] + eps * nw[2]};
    rtx_ray& r = out[m];
    memset(&r, 0, sizeof(r));
    for (int a = 0; a < 3; ++a) r.org[a] = (float)o[a];
    r.cull_mask = 0xff;
    if (kind == 0) {
      SplitMix64 g(mix_seed(seed, i));
      double u1 = g.uniform(), u2 = g.uniform();
      double rr = std::sqrt(u1), phi = 2.0 * kPi * u2;
      double lx = rr * std::cos(phi), ly = rr * std::sin(phi), lz = std::sqrt(std::max(0.0, 1.0 - u1));
      double t[3] = {1, 0, 0};
      if (std::fabs(nw[0]) > 0.9) t[0] = 0, t[1] = 1;
      double bx[3], by[3];
      cross3(nw, t, bx);
      normalize3(bx);
      cross3(nw, bx, by);
      double dir[3];
      for (int a = 0; a < 3; ++a) dir[a] = lx * bx[a] + ly * by[a] + lz * nw[a];
      normalize3(dir);
      for (int a = 0; a < 3; ++a) r.dir[a] = (float)dir[a];
      r.tmin = 0.0f;
      r.tmax = 1e30f;
      r.flags = RTX_RAY_FLAG_OPAQUE;
    } else {
      double dir[3] = {light_pos[0] - o[0], light_pos[1] - o[1], light_pos[2] - o[2]};
      double dist = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
      normalize3(dir);
      for (int a = 0; a < 3; ++a) r.dir[a] = (float)dir[a];
      r.tmin = 0.0f;
      r.tmax = (float)dist;
      r.flags = RTX_RAY_FLAG_OPAQUE | RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER;
    }
    if (src_index) src_index[m] = i;
    ++m;
  }
  *n_out = m;
  return RTX_OK;
}

}  // extern "C"
