"""The eight-lanes-per-ray traversal kernels (rtx_trace_options.lanes_per_ray = 8,
chosen by default for batches of at most 128 rays per CU), plain and treelet
order: hits bit-exact against the oracle and against the one-lane-per-ray
kernels."""
import os
import sys

import numpy as np
import pytest

from conftest import assert_hits_equal

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import tiny_scene  # noqa: E402

pytestmark = pytest.mark.gpu


def _both(gpu, rays, mode=0, **kw):
    """Hits of the same batch from both kernels, after checking they agree bit for bit."""
    h8 = gpu.trace(rays, mode=mode, lanes_per_ray=8, **kw)
    h1 = gpu.trace(rays, mode=mode, lanes_per_ray=1, **kw)
    assert h8.tobytes() == h1.tobytes()
    return h8


@pytest.mark.parametrize("lds_stack", [4, 8, 16])
def test_wide_c1(gpu, oracle, c1, lds_stack):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, records=False)
    assert_hits_equal(_both(gpu, rays, lds_stack=lds_stack), ref, f"eight lanes per ray, C1, stack {lds_stack}")
    # the default picks the cooperative kernel for this batch size
    assert_hits_equal(gpu.trace(rays, mode=0), ref, "default kernel choice, C1")


def test_wide_instanced(gpu, oracle, instanced, instanced_shared):
    for scene, rays in (instanced, instanced_shared):
        gpu.set_scene(scene.arena, scene.tlas_offset)
        ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, records=False)
        assert_hits_equal(_both(gpu, rays), ref, "eight lanes per ray, instanced")


@pytest.mark.parametrize("shape", [dict(waves_per_cu=4), dict(waves_per_cu=64), dict(refill=1), dict(refill=64),
                                   dict(waves_per_cu=4, refill=8, lds_stack=4)])
def test_wide_launch_shapes(gpu, oracle, c1, shape):
    """Few groups refilled many times, more workgroups than fit at once, a
    refill on every idle group and only when all eight groups are idle."""
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, records=False)
    assert_hits_equal(gpu.trace(rays, mode=0, lanes_per_ray=8, **shape), ref, f"eight lanes per ray, {shape}")


def test_wide_nonfinite_rays_and_flags(gpu, oracle, c1):
    """NaN / infinite ray components (compare-select slab tests) and the
    non-opaque and terminate-on-first-hit flags in one batch."""
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    r = rays[:6144].copy()
    rng = np.random.default_rng(23)
    pick = rng.choice(len(r), 96, replace=False)
    r["dir"][pick[:16], 0] = np.nan
    r["org"][pick[16:32], 1] = np.inf
    r["tmax"][pick[32:48]] = np.inf
    r["tmin"][pick[48:64]] = np.nan
    r["dir"][pick[64:80]] = 0.0
    r["org"][pick[80:96], 2] = -np.inf
    r["flags"][1::3] = 0  # non-opaque: every candidate counted, closest-hit quirk of the reference
    r["flags"][2::3] = 1 | 4  # opaque, terminate on first hit
    ref = oracle.trace(scene.arena, scene.tlas_offset, r, mode=0, records=False)
    assert_hits_equal(_both(gpu, r), ref, "eight lanes per ray, non-finite rays and flags")


def test_wide_large_batch_forced(gpu, oracle, rtx):
    """A batch of many refills per group (64 rays per lane slot) forced onto the
    cooperative kernel, on the instanced scene with bounce rays."""
    d = rtx.workload_desc(rtx.WORKLOAD_INSTANCED_SPHERES, n_instances=64, slices=16, stacks=8, width=256,
                          height=256, seed=5, box_size=30.0)
    scene = rtx.workload_scene(d)
    prim = rtx.primary_rays(d)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    hits = gpu.trace(prim, mode=0)
    bounce, _ = scene.secondary_rays(prim, hits, kind=0, seed=3)
    rays = np.concatenate([prim, bounce])
    ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, records=False)
    assert_hits_equal(_both(gpu, rays, waves_per_cu=4), ref, "eight lanes per ray, forced, bounce rays")


@pytest.mark.parametrize("case", ["plain_opaque", "plain_non_opaque", "plain_terminate"])
def test_wide_tiny_hand_derived(gpu, rtx, case):
    exp = tiny_scene.expected()[case]
    gpu.set_scene(tiny_scene.tiny_arena(), 0)
    h = gpu.trace(tiny_scene.tiny_ray(rtx.RAY_DTYPE, flags=exp["flags"]), mode=0, lanes_per_ray=8)
    for k, v in exp["hit"].items():
        if isinstance(v, list):
            assert list(h[0][k]) == v, k
        else:
            assert h[0][k] == v, k
    assert int(h[0]["n_records"]) == len(exp["records"])


def test_wide_malformed_arena_and_bad_options(gpu, rtx):
    a = tiny_scene.tiny_arena()
    ray = tiny_scene.tiny_ray(rtx.RAY_DTYPE)
    bad = a.copy()
    bad[64 + 12:64 + 16] = np.frombuffer(np.int32(1000).tobytes(), np.uint8)
    gpu.set_scene(bad, 0)
    with pytest.raises(rtx.RtxError) as e:
        gpu.trace(ray, mode=0, lanes_per_ray=8)
    assert e.value.status == 3
    gpu.set_scene(a, 0)
    for kw in (dict(lanes_per_ray=3), dict(lanes_per_ray=8, records=True)):
        with pytest.raises(rtx.RtxError) as e:
            gpu.trace(ray, mode=0, **kw)
        assert e.value.status == 1
    gpu.build_treelets(8192)
    with pytest.raises(rtx.RtxError) as e:
        gpu.trace(ray, mode=1, lanes_per_ray=8, prefetch=1)
    assert e.value.status == 1
    # the hand-derived closest hits of the two orders (prim 10 plain, 11 treelet)
    for mode, prim in ((0, 10), (1, 11)):
        h = gpu.trace(ray, mode=mode, lanes_per_ray=8)
        assert h[0]["hit"] == 1 and h[0]["primitive_index"] == prim


@pytest.mark.parametrize("budget", [512, 8192, 65536])
def test_wide_treelet_c1(gpu, oracle, c1, budget):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(budget)
    part = oracle.Partition(scene.arena, scene.tlas_offset, budget)
    ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=1, partition=part, records=False)
    assert_hits_equal(_both(gpu, rays, mode=1), ref, f"eight lanes per ray, treelet order, budget {budget}")
    assert_hits_equal(gpu.trace(rays, mode=1), ref, "default kernel choice, treelet order")


def test_wide_treelet_instanced(gpu, oracle, instanced, instanced_shared):
    """Instances opened and reopened across the two lists (shared BLAS too)."""
    for scene, rays in (instanced, instanced_shared):
        gpu.set_scene(scene.arena, scene.tlas_offset)
        gpu.build_treelets(4096)
        part = oracle.Partition(scene.arena, scene.tlas_offset, 4096)
        ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=1, partition=part, records=False)
        assert_hits_equal(_both(gpu, rays, mode=1), ref, "eight lanes per ray, treelet order, instanced")


def test_wide_treelet_nonfinite_rays_flags_and_shapes(gpu, oracle, c1):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(8192)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    r = rays[:6144].copy()
    rng = np.random.default_rng(29)
    pick = rng.choice(len(r), 64, replace=False)
    r["dir"][pick[:16], 0] = np.nan
    r["org"][pick[16:32], 1] = np.inf
    r["tmax"][pick[32:48]] = np.inf
    r["dir"][pick[48:64]] = 0.0
    r["flags"][1::3] = 0
    r["flags"][2::3] = 1 | 4
    ref = oracle.trace(scene.arena, scene.tlas_offset, r, mode=1, partition=part, records=False)
    assert_hits_equal(_both(gpu, r, mode=1), ref, "eight lanes per ray, treelet order, non-finite rays and flags")
    for shape in (dict(waves_per_cu=4), dict(refill=64), dict(waves_per_cu=4, refill=8, lds_stack=16)):
        assert_hits_equal(gpu.trace(r, mode=1, lanes_per_ray=8, **shape), ref, f"treelet order, {shape}")


@pytest.mark.parametrize("case", ["treelet_opaque", "treelet_terminate"])
@pytest.mark.parametrize("budget", [8192, 256, 192])
def test_wide_tiny_treelet_hand_derived(gpu, rtx, case, budget):
    exp = tiny_scene.expected()[case]
    gpu.set_scene(tiny_scene.tiny_arena(), 0)
    gpu.build_treelets(budget)
    h = gpu.trace(tiny_scene.tiny_ray(rtx.RAY_DTYPE, flags=exp["flags"]), mode=1, lanes_per_ray=8)
    for k, v in exp["hit"].items():
        if isinstance(v, list):
            assert list(h[0][k]) == v, k
        else:
            assert h[0][k] == v, k
    assert int(h[0]["n_records"]) == len(exp["records"])
