"""GPU traversal and partition vs the CPU oracle, bit-exact (integer, index and FP32 bit patterns)."""
import types

import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal

pytestmark = pytest.mark.gpu


def _scene_on(gpu, scene):
    gpu.set_scene(scene.arena, scene.tlas_offset)


@pytest.mark.parametrize("lds_stack", [4, 8, 16])
def test_plain_c1_hits_and_records(gpu, oracle, c1, lds_stack):
    scene, rays = c1
    _scene_on(gpu, scene)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0)
    h, o, r = gpu.trace(rays, mode=0, records=True, lds_stack=lds_stack)
    assert_hits_equal(h, h_ref, "plain C1")
    assert_records_equal(o, r, o_ref, r_ref, "plain C1")
    assert h["hit"].sum() > 0.5 * len(rays)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("shape", [dict(waves_per_cu=4), dict(waves_per_cu=64), dict(refill=1), dict(refill=64),
                                   dict(waves_per_cu=4, refill=1, lds_stack=4)])
def test_launch_shapes_bit_exact(gpu, oracle, c1, mode, shape):
    """Persistent-wave corner cases: a few waves, more workgroups than fit on
    the GPU at once (later ones start after earlier ones exit), a refill on
    every idle lane, and a refill only when the whole wave is idle."""
    scene, rays = c1
    _scene_on(gpu, scene)
    part = None
    if mode == 1:
        gpu.build_treelets(8192)
        part = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part)
    h, o, r = gpu.trace(rays, mode=mode, records=True, **shape)
    assert_hits_equal(h, h_ref, f"mode {mode} {shape}")
    assert_records_equal(o, r, o_ref, r_ref, f"mode {mode} {shape}")


@pytest.mark.parametrize("budget", [512, 8192, 65536])
def test_partition_c1(gpu, oracle, c1, budget):
    scene, _ = c1
    _scene_on(gpu, scene)
    s = gpu.build_treelets(budget)
    p = oracle.Partition(scene.arena, scene.tlas_offset, budget)
    t = gpu.get_treelets()
    assert s["n_treelets"] == p.n_treelets
    assert np.array_equal(t["roots"], p.roots)
    assert np.array_equal(t["list_offsets"], p.list_offsets)
    assert np.array_equal(t["list_nodes"], p.list_nodes)
    assert np.array_equal(t["owner"], p.owner)
    assert s["total_bytes"] == p.total_bytes


@pytest.mark.parametrize("budget", [512, 8192, 65536])
def test_treelet_c1_hits_and_records(gpu, oracle, c1, budget):
    scene, rays = c1
    _scene_on(gpu, scene)
    gpu.build_treelets(budget)
    p = oracle.Partition(scene.arena, scene.tlas_offset, budget)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=1, partition=p)
    h, o, r = gpu.trace(rays, mode=1, records=True)
    assert_hits_equal(h, h_ref, f"treelet C1 budget {budget}")
    assert_records_equal(o, r, o_ref, r_ref, f"treelet C1 budget {budget}")


def test_plain_records_carry_treelets(gpu, oracle, c1):
    scene, rays = c1
    _scene_on(gpu, scene)
    gpu.build_treelets(8192)
    p = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, partition=p)
    h, o, r = gpu.trace(rays, mode=0, records=True)
    assert_hits_equal(h, h_ref, "plain with partition")
    assert_records_equal(o, r, o_ref, r_ref, "plain with partition")


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("prefetch,stage", [(1, 4096), (1, 8192), (2, 8192), (3, 16384), (1, 65536), (5, 8192), (5, 16384)])
def test_staging_never_changes_results(gpu, oracle, c1, mode, prefetch, stage):
    """Treelet staging into LDS only moves where node bytes are read from."""
    scene, rays = c1
    _scene_on(gpu, scene)
    gpu.build_treelets(8192)
    p = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=p)
    h, o, r = gpu.trace(rays, mode=mode, prefetch=prefetch, threshold=0.5, records=True, stage_bytes=stage)
    assert_hits_equal(h, h_ref, f"staged mode {mode} prefetch {prefetch} stage {stage}")
    assert_records_equal(o, r, o_ref, r_ref, f"staged mode {mode} prefetch {prefetch} stage {stage}")
    s = gpu.prefetch_stats()
    # one node read per internal node, instance leaf and primitive leaf visited
    assert s["node_reads"] == int(h["n_nodes"].sum())
    assert s["lds_hits"] <= s["node_reads"]
    assert s["wasted_bytes"] <= s["staged_bytes"]
    assert s["staged_treelets"] > 0 and s["lds_hits"] > 0
    # effectiveness classes: every staged unit ends TIMELY or NEVER_USED (all
    # slots are evicted when the waves end); TOO_EARLY is part of NEVER_USED;
    # a TIMELY unit served at least one LDS read; LATE cannot happen
    assert s["timely_bytes"] + s["wasted_bytes"] == s["staged_bytes"]
    assert s["too_early_bytes"] <= s["wasted_bytes"] and s["timely_bytes"] // 128 <= s["lds_hits"]
    assert s["late_bytes"] == 0 and s["timely_bytes"] > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_instanced_scenes(gpu, oracle, instanced, instanced_shared, mode):
    for scene, rays in (instanced, instanced_shared):
        _scene_on(gpu, scene)
        p = None
        if mode == 1:
            gpu.build_treelets(4096)
            p = oracle.Partition(scene.arena, scene.tlas_offset, 4096)
            t = gpu.get_treelets()
            assert np.array_equal(t["owner"], p.owner)
            assert np.array_equal(t["list_nodes"], p.list_nodes)
        h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=p)
        h, o, r = gpu.trace(rays, mode=mode, records=True)
        assert_hits_equal(h, h_ref, f"instanced mode {mode}")
        assert_records_equal(o, r, o_ref, r_ref, f"instanced mode {mode}")


def test_nonfinite_rays_bit_exact(gpu, oracle, c1):
    """Rays with NaN / infinite components mixed into ordinary waves: the slab
    tests of those waves take the compare-select form, and every lane still
    matches the oracle."""
    scene, rays = c1
    _scene_on(gpu, scene)
    r = rays[:4096].copy()
    rng = np.random.default_rng(17)
    pick = rng.choice(len(r), 96, replace=False)
    r["dir"][pick[:16], 0] = np.nan
    r["org"][pick[16:32], 1] = np.inf
    r["tmax"][pick[32:48]] = np.inf
    r["tmin"][pick[48:64]] = np.nan
    r["dir"][pick[64:80]] = 0.0
    r["org"][pick[80:96], 2] = -np.inf
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, r, mode=0)
    h, o, rec = gpu.trace(r, mode=0, records=True)
    assert_hits_equal(h, h_ref, "plain, non-finite rays")
    assert_records_equal(o, rec, o_ref, r_ref, "plain, non-finite rays")
    gpu.build_treelets(8192)
    p = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, r, mode=1, partition=p)
    h, o, rec = gpu.trace(r, mode=1, records=True)
    assert_hits_equal(h, h_ref, "treelet, non-finite rays")
    assert_records_equal(o, rec, o_ref, r_ref, "treelet, non-finite rays")


def _layout_scenes(rtx, c1, instanced):
    from golden import tiny_scene
    tiny = types.SimpleNamespace(arena=tiny_scene.tiny_arena(), tlas_offset=0)
    shared = types.SimpleNamespace(arena=tiny_scene.tiny_shared_arena(), tlas_offset=0)
    return [("tiny", tiny, (256, 192, 8192)), ("tiny shared", shared, (256, 8192)),
            ("c1", c1[0], (1024, 8192)), ("instanced", instanced[0], (1024, 4096))]


@pytest.mark.parametrize("stride", [0, 256])
def test_treelet_layout_matches_the_oracle(gpu, rtx, oracle, c1, instanced, stride):
    """Device remapBVHToTreeletLayout: the same place for every node and the same
    metadata table as the oracle (itself pinned by hand-derived layouts)."""
    for name, scene, budgets in _layout_scenes(rtx, c1, instanced):
        gpu.set_scene(scene.arena, scene.tlas_offset)
        for budget in budgets:
            gpu.build_treelets(budget)
            s = gpu.build_treelet_layout(stride)
            remap, meta = gpu.get_treelet_layout()
            part = oracle.Partition(scene.arena, scene.tlas_offset, budget)
            r_ref, m_ref = part.layout(stride)
            assert np.array_equal(remap, r_ref), f"{name} @ {budget}: layout places differ"
            assert np.array_equal(meta, m_ref), f"{name} @ {budget}: metadata differs"
            assert s["block_bytes"] == budget + stride and s["layout_bytes"] == part.n_treelets * (budget + stride)
            assert s["mapped_nodes"] == int((r_ref != np.uint64(0xFFFFFFFFFFFFFFFF)).sum())


def test_treelet_layout_refuses_where_the_reference_asserts(gpu, rtx, oracle, instanced_shared):
    """A shared BLAS whose treelet roots are also listed by earlier treelets: the
    reference asserts (vulkan_ray_tracing.cc:1483); both oracle and device refuse."""
    scene, _ = instanced_shared
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    with pytest.raises(oracle.OracleError):
        oracle.Partition(scene.arena, scene.tlas_offset, 1024).layout(0)
    with pytest.raises(rtx.RtxError) as e:
        gpu.build_treelet_layout(0)
    assert e.value.status == 11


@pytest.mark.parametrize("mode", [0, 1])
def test_records_in_treelet_layout_addresses(gpu, rtx, oracle, c1, instanced, mode):
    """-remap_to_treelet_layout: every record names its node's layout place
    (vulkan_ray_tracing.cc:2446-2450, 2509-2512); everything else is unchanged."""
    for name, scene, budgets in _layout_scenes(rtx, c1, instanced):
        rays = c1[1] if name == "c1" else instanced[1] if name == "instanced" else \
            np.concatenate([tiny_scene_ray(rtx, f) for f in (1, 0, 5)])
        gpu.set_scene(scene.arena, scene.tlas_offset)
        budget = budgets[0]
        gpu.build_treelets(budget)
        gpu.build_treelet_layout(128)
        part = oracle.Partition(scene.arena, scene.tlas_offset, budget)
        remap, _ = part.layout(128)
        h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part)
        r_ref = r_ref.copy()
        r_ref["offset"] = remap[(r_ref["offset"] // 64).astype(np.int64)]
        h, o, r = gpu.trace(rays, mode=mode, records=True, record_layout=1)
        assert_hits_equal(h, h_ref, f"{name} layout records, mode {mode}")
        assert_records_equal(o, r, o_ref, r_ref, f"{name} layout records, mode {mode}")


def tiny_scene_ray(rtx, flags):
    from golden import tiny_scene
    return tiny_scene.tiny_ray(rtx.RAY_DTYPE, flags=flags)
