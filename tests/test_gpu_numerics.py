"""Numerical edge cases of the traversal against the oracle, bit for bit, in
both orders and with one and eight lanes per ray: degenerate triangles (the
Möller-Trumbore test has no determinant epsilon, so 1/det can be infinite),
sheared / mirrored / non-uniformly scaled instances, extreme coordinate
scales, and empty or negative ray intervals."""
import types

import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal

pytestmark = pytest.mark.gpu


def _rays_towards(rtx, centers, n, seed, spread=0.3, dist=6.0):
    """n rays from random points around the scene towards jittered targets
    near the given centers."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, rtx.RAY_DTYPE)
    tgt = centers[rng.integers(0, len(centers), n)] + rng.normal(0, spread, (n, 3)) * np.abs(centers).max() * 0.05
    d = rng.normal(0, 1, (n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org = tgt - d * dist
    r["org"] = org.astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = np.float32(1e30)
    r["flags"] = 1
    return r


def _check_all(gpu, oracle, scene, rays, what):
    """Both orders; records from the one-lane kernel, hits from both kernels."""
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
    for mode in (0, 1):
        # with a partition built, plain-order records carry the node's treelet too
        h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part)
        h, o, r = gpu.trace(rays, mode=mode, records=True)
        assert_hits_equal(h, h_ref, f"{what}, mode {mode}")
        assert_records_equal(o, r, o_ref, r_ref, f"{what}, mode {mode}")
        for lanes in (1, 8):
            assert_hits_equal(gpu.trace(rays, mode=mode, lanes_per_ray=lanes), h_ref,
                              f"{what}, mode {mode}, {lanes} lanes per ray")
    return h_ref


def _degenerate_scene(rtx):
    """A mesh with zero-area triangles (repeated and collinear vertices, points)
    and rays aimed at them."""
    tris = rtx.displaced_sphere(16, 8, seed=11).copy()
    rng = np.random.default_rng(3)
    pick = rng.choice(len(tris), 96, replace=False)
    t = tris.reshape(-1, 3, 3)
    t[pick[:32], 1] = t[pick[:32], 0]                                   # repeated vertex
    t[pick[32:64], 2] = 0.5 * (t[pick[32:64], 0] + t[pick[32:64], 1])   # collinear
    t[pick[64:], :] = t[pick[64:], :1]                                  # a point
    b = rtx.SceneBuilder()
    bid = b.add_blas(tris, geometry_index=2)
    b.add_instance(bid, instance_id=5)
    scene = b.finish()
    cen = t[pick].mean(axis=1)
    return scene, _rays_towards(rtx, cen.astype(np.float64), 4096, seed=5, spread=0.02, dist=3.0)


def _affine_scene(rtx):
    """Instances under shear, non-uniform scale and mirroring."""
    rng = np.random.default_rng(17)
    b = rtx.SceneBuilder()
    blas = [b.add_blas(rtx.displaced_sphere(12, 6, seed=s), geometry_index=s) for s in range(3)]
    centers = []
    for k in range(24):
        m = np.eye(3) + rng.normal(0, 0.35, (3, 3))       # shear and scale
        m[:, 0] *= rng.choice([-1.0, 1.0]) * rng.uniform(0.3, 3.0)
        m[:, 2] *= rng.uniform(0.3, 3.0)
        pos = rng.uniform(-12, 12, 3)
        o2w = np.hstack([m, pos[:, None]]).astype(np.float32)
        b.add_instance(blas[k % 3], o2w, instance_id=100 + k)
        centers.append(pos)
    scene = b.finish()
    return scene, _rays_towards(rtx, np.array(centers), 8192, seed=9, spread=1.0, dist=30.0)


def _scaled_scene(rtx, scale):
    """Eight instances of one mesh with every coordinate multiplied by scale."""
    b = rtx.SceneBuilder()
    tris = (rtx.displaced_sphere(12, 6, seed=4).astype(np.float64) * scale).astype(np.float32)
    bid = b.add_blas(tris, geometry_index=1)
    rng = np.random.default_rng(2)
    centers = []
    for k in range(8):
        pos = rng.uniform(-5, 5, 3) * scale
        o2w = np.hstack([np.eye(3), pos[:, None]]).astype(np.float32)
        b.add_instance(bid, o2w, instance_id=k)
        centers.append(pos)
    scene = b.finish()
    rays = _rays_towards(rtx, np.array(centers), 4096, seed=4, spread=1.0, dist=20.0 * scale)
    rays["tmax"] = np.float32(np.finfo(np.float32).max)
    return scene, rays


def test_degenerate_triangles(gpu, rtx, oracle):
    """Zero-area triangles mixed into a mesh: det = 0 makes 1/det infinite and
    u, v, t NaN or infinite."""
    scene, rays = _degenerate_scene(rtx)
    h = _check_all(gpu, oracle, scene, rays, "degenerate triangles")
    assert h["hit"].sum() > 100


def test_affine_instances(gpu, rtx, oracle):
    """Instances under shear, non-uniform scale and mirroring (negative
    determinant): the object-space ray has a non-unit direction before
    normalisation and tMult differs per instance."""
    scene, rays = _affine_scene(rtx)
    h = _check_all(gpu, oracle, scene, rays, "affine instances")
    assert h["hit"].sum() > 1000


@pytest.mark.parametrize("scale", [1e-9, 1e12])
def test_extreme_coordinate_scales(gpu, rtx, oracle, scale):
    """The same instanced geometry at 1e-9 and 1e12: the Möller-Trumbore
    products range down to about 1e-28 and up to about 1e35 (near the top of
    the float range), and the child bound exponents sit far from 0."""
    scene, rays = _scaled_scene(rtx, scale)
    h = _check_all(gpu, oracle, scene, rays, f"coordinates x {scale:g}")
    assert h["hit"].sum() > 1000


def test_ray_intervals(gpu, rtx, oracle, c1):
    """Empty (tmin > tmax), point (tmin == tmax), negative-start and zero-length
    intervals, and a tmax that ends inside the geometry."""
    scene, rays = c1
    r = rays[:4096].copy()
    r["tmin"][0::5] = 50.0
    r["tmax"][0::5] = 10.0
    r["tmin"][1::5] = r["tmax"][1::5] = 4.0
    r["tmin"][2::5] = -100.0
    r["tmax"][3::5] = 0.0
    r["tmax"][4::5] = 2.5
    _check_all(gpu, oracle, scene, r, "ray intervals")


def test_cull_distances_at_float_boundaries(gpu, rtx, oracle, c1):
    """The closest-hit cull (a child box is skipped when its entry distance is
    at or beyond the closest hit so far, vulkan_ray_tracing.cc:2537, 2725) at
    the edges of float order: tmax equal to the ray's own hit distance (the
    ground quad's leaf box is entered exactly there), one float either side,
    signed zeros, denormals, -FLT_MAX, -inf, FLT_MAX, +inf, with tmin 0, -1 or
    -inf and opaque and non-opaque rays."""
    scene, rays = c1
    hall = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, records=False)
    pick = np.concatenate([np.flatnonzero(hall["hit"] != 0)[::4][:1536], np.flatnonzero(hall["hit"] == 0)[:512]])
    base = rays[pick].copy()
    h0 = hall[pick]
    hit = h0["hit"] != 0
    t = np.where(hit, h0["t"], np.float32(10.0)).astype(np.float32)
    f32 = np.float32
    tiny = np.finfo(np.float32).smallest_subnormal
    big = np.finfo(np.float32).max
    fixed = np.array([-np.inf, -big, -1.0, -0.0, 0.0, tiny, -tiny, big, np.inf], np.float32)
    groups = []
    for k, tm in enumerate((t, np.nextafter(t, f32(np.inf)), np.nextafter(t, f32(-np.inf)))):
        r = base.copy()
        r["tmax"] = tm
        r["tmin"] = np.array([0.0, -1.0, -np.inf], np.float32)[(np.arange(len(r)) + k) % 3]
        groups.append(r)
    r = base.copy()
    r["tmax"] = fixed[np.arange(len(r)) % len(fixed)]
    r["tmin"] = np.where(np.arange(len(r)) % 2 == 0, f32(-np.inf), f32(0.0))
    groups.append(r)
    rr = np.concatenate(groups)
    rr["flags"] = np.where(np.arange(len(rr)) % 3 == 0, 0, 1)  # some rays non-opaque
    h = _check_all(gpu, oracle, scene, rr, "cull boundaries")
    assert h["hit"].sum() > 500


def _overflowing_tlas_arena():
    """The hand-assembled tiny arena with the TLAS root's x origin at 3e38 and
    x exponent 127: its child box decodes to x in [3e38, +inf] (q 128 * 2^119
    + 3e38 overflows), and the TLAS header's x bounds become [3e38, +inf] too."""
    import struct
    from golden import tiny_scene

    A = bytearray(tiny_scene.tiny_arena().tobytes())
    struct.pack_into("<f", A, 8, 3e38)          # header BoundsMin.x
    struct.pack_into("<f", A, 20, np.inf)       # header BoundsMax.x
    struct.pack_into("<f", A, 64, 3e38)         # TLAS root Origin.x
    struct.pack_into("<b", A, 64 + 18, 127)     # TLAS root exponent x
    return np.frombuffer(bytes(A), np.uint8).copy()


def test_infinite_direction_against_overflowing_bounds(gpu, rtx, oracle):
    """Rays with infinite direction components (idir 0 on that axis) against a
    TLAS node whose x bound decodes to +inf: (inf - o) * 0 is NaN, which the
    reference's compare-selects drop and a sign-picked near bound would not.
    The child boxes are ordered (lo <= hi), so only the ray decides which slab
    test runs."""
    A = _overflowing_tlas_arena()
    rng = np.random.default_rng(21)
    n = 2048
    r = np.zeros(n, rtx.RAY_DTYPE)
    comp = np.array([np.inf, -np.inf, 0.0, 1.0, -1.0, 0.5], np.float32)
    d = comp[rng.integers(0, len(comp), (n, 3))]
    d[:, 2] = np.where(rng.random(n) < 0.5, 1.0, d[:, 2])
    r["dir"] = d
    r["org"] = np.stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 2.5, n), rng.uniform(-1, 4, n)], 1)
    r["tmin"] = rng.choice(np.array([-10.0, 0.0], np.float32), n)
    r["tmax"] = rng.choice(np.array([-0.5, 100.0, 1e30], np.float32), n)
    r["flags"] = 1
    r[0] = tuple([(0.25, 0.25, 2.0), -10.0, (np.inf, 0.0, 1.0), -0.5, 1, 0xff, 0, 0, 0, (0, 0, 0)])
    _check_all(gpu, oracle, types.SimpleNamespace(arena=A, tlas_offset=0), r, "infinite directions")
    # the hand-derived ray enters the instance leaf (its TLAS child box is hit at thit = -1)
    h0, o0, r0 = oracle.trace(A, 0, r[:1], mode=0)
    assert 2 in set(r0["type"].tolist()), "hand-derived ray should visit the instance leaf"


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("prefetch,slot", [(1, 0), (2, 0), (3, 0), (5, 0), (1, 1024), (5, 1024)])
def test_staged_kernels_on_affine_instances(gpu, rtx, oracle, mode, prefetch, slot):
    """The LDS staging kernels on mirrored and sheared instances: staging
    changes where node bytes are read, never the hits or records."""
    scene, rays = _affine_scene(rtx)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part)
    h, o, r = gpu.trace(rays, mode=mode, prefetch=prefetch, records=True, stage_bytes=8192, slot_bytes=slot)
    assert_hits_equal(h, h_ref, f"staged {prefetch}/{slot}, mode {mode}")
    assert_records_equal(o, r, o_ref, r_ref, f"staged {prefetch}/{slot}, mode {mode}")


@pytest.mark.parametrize("mode", [0, 1])
def test_numerics_scenes_in_one_batch(gpu, rtx, oracle, mode):
    """The scenes above as jobs of one multi-scene batch (and a graph replay)."""
    items = [_degenerate_scene(rtx), _affine_scene(rtx), _scaled_scene(rtx, 1e12), _scaled_scene(rtx, 1e-9)]
    jobs, refs = [], []
    for scene, rays in items:
        sid = gpu.add_scene(scene.arena, scene.tlas_offset)
        part = None
        if mode == 1:
            gpu.build_scene_treelets(sid, 1024)
            part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
        jobs.append((sid, gpu.to_device(rays), len(rays), gpu.buffer(len(rays) * rtx.HIT_DTYPE.itemsize)))
        refs.append(oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part, records=False))
    assert gpu.trace_batch(jobs, mode=mode) == 0
    for k, ((_, _, n, hb), ref) in enumerate(zip(jobs, refs)):
        assert_hits_equal(hb.download(rtx.HIT_DTYPE, n), ref, f"batch job {k}, mode {mode}")
    g = gpu.batch_graph(jobs, mode=mode)
    st, _ = g.launch(3)
    g.close()
    assert st == 0
    for k, ((_, _, n, hb), ref) in enumerate(zip(jobs, refs)):
        assert_hits_equal(hb.download(rtx.HIT_DTYPE, n), ref, f"batch graph job {k}, mode {mode}")
