"""The reference's tree_level_map on the GPU.

A ray's level is the maximum over its final tree_level_map
(vulkan_ray_tracing.cc:2274-2279): a node assigned again keeps only its last
value.  Where a node is reachable at two levels (a BLAS under instance leaves
of different depths) that differs from the deepest assignment, and the library
traces such scenes a second time with the map kept per lane.  Checked against
hand-derived values and the oracle, through every entry point that writes hits.
"""
import os
import sys

import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import tiny_scene  # noqa: E402

pytestmark = pytest.mark.gpu


def _two_depth_rays(rtx, n=4096, seed=5):
    """Rays down the z axis over the two-depth arena's footprint: some hit a
    triangle under the first (deeper) instance, which culls the second
    instance's boxes; the rest miss and visit every BLAS node twice."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=rtx.RAY_DTYPE)
    r["org"][:, 0] = rng.uniform(-1.2, 3.2, n)
    r["org"][:, 1] = rng.uniform(-1.2, 3.2, n)
    r["dir"] = (0.0, 0.0, 1.0)
    r["tmax"] = rng.choice([100.0, 4.5, 2.5], n)
    r["flags"] = rng.choice([0, 1, 1 | 4], n)
    r["cull_mask"] = 0xFF
    return r


@pytest.mark.parametrize("lanes_per_ray", [1, 8])
@pytest.mark.parametrize("mode,case", [(0, "plain_miss"), (1, "treelet_miss_8192")])
def test_two_depth_hand_derived(gpu, rtx, mode, case, lanes_per_ray):
    exp = tiny_scene.expected()["two_depth"]
    arena = tiny_scene.tiny_two_depth_arena()
    gpu.set_scene(arena, 0)
    if mode:
        gpu.build_treelets(8192)
    ray = tiny_scene.tiny_ray(rtx.RAY_DTYPE)
    ray["org"] = exp["ray_miss"]["org"]
    h = gpu.trace(ray, mode=mode, lanes_per_ray=lanes_per_ray)
    assert h[0]["max_level"] == exp[case]["max_level"] == 4
    assert h[0]["n_nodes"] == exp[case]["n_nodes"] and h[0]["fetch_bytes"] == exp[case]["fetch_bytes"]
    h, o, r = gpu.trace(ray, mode=mode, records=True)
    assert h[0]["max_level"] == 4
    got = [[int(x["offset"]), int(x["size"]), int(x["type"]), int(x["level"])] for x in r]
    assert got == exp[case]["records"]
    c = gpu.trace(tiny_scene.tiny_ray(rtx.RAY_DTYPE), mode=0, lanes_per_ray=lanes_per_ray)
    assert c[0]["max_level"] == exp["plain_opaque_centre"]["max_level"]


@pytest.mark.parametrize("mode,budget", [(0, 0), (1, 8192), (1, 256)])
def test_two_depth_rays_match_the_oracle(gpu, rtx, oracle, mode, budget):
    """At budget 256 the instances' BLAS nodes fall in several treelets, so
    treelet order interleaves the two instances and a node's value is read
    after the other instance may have reassigned it (:1837, :2034)."""
    arena = tiny_scene.tiny_two_depth_arena()
    rays = _two_depth_rays(rtx)
    gpu.set_scene(arena, 0)
    part = None
    if mode:
        gpu.build_treelets(budget)
        part = oracle.Partition(arena, 0, budget)
    h_ref, o_ref, r_ref = oracle.trace(arena, 0, rays, mode=mode, partition=part)
    # the case the level pass exists for occurs, and so does the other
    assert (h_ref["max_level"] < h_ref["path_level"]).sum() > 100
    assert (h_ref["max_level"] == h_ref["path_level"]).sum() > 100
    for lanes in (1, 8):
        assert_hits_equal(gpu.trace(rays, mode=mode, lanes_per_ray=lanes), h_ref, f"mode {mode} lanes {lanes}")
    h, o, r = gpu.trace(rays, mode=mode, records=True)
    assert_hits_equal(h, h_ref, f"mode {mode} records")
    assert_records_equal(o, r, o_ref, r_ref, f"mode {mode} records")
    # the step statistics take the map's maximum
    hb = gpu.to_device(h)
    assert gpu.trace_stats(hb, len(rays))["max_tree_depth"] == int(h_ref["max_level"].max())


@pytest.mark.parametrize("mode", [0, 1])
def test_level_pass_without_allocation(gpu, rtx, oracle, mode):
    arena = tiny_scene.tiny_two_depth_arena()
    rays = _two_depth_rays(rtx, n=3000, seed=9)
    gpu.set_scene(arena, 0)
    part = None
    if mode:
        gpu.build_treelets(8192)
        part = oracle.Partition(arena, 0, 8192)
    h_ref = oracle.trace(arena, 0, rays, mode=mode, partition=part, records=False)
    rb, hb = gpu.to_device(rays), gpu.buffer(len(rays) * rtx.HIT_DTYPE.itemsize)
    for lanes in (1, 8):
        assert gpu.trace_static_device(rb, len(rays), hb, mode=mode, lanes_per_ray=lanes) == 0
        assert_hits_equal(hb.download(rtx.HIT_DTYPE, len(rays)), h_ref, f"static mode {mode} lanes {lanes}")


def test_level_pass_in_batches_and_graphs(gpu, rtx, oracle, c1):
    """A batch with a two-depth scene between two tree-shaped ones: the level
    pass runs for its job only, eagerly and in every graph replay."""
    two = tiny_scene.tiny_two_depth_arena()
    scene, c1_rays = c1
    s0 = gpu.add_scene(scene.arena, scene.tlas_offset)
    s1 = gpu.add_scene(two, 0)
    s2 = gpu.add_scene(scene.arena, scene.tlas_offset)
    rays1 = _two_depth_rays(rtx, n=5000, seed=13)
    items = [(s0, c1_rays[:3000], scene), (s1, rays1, None), (s2, c1_rays[3000:7000], scene)]
    jobs, refs = [], []
    for sid, rays, sc in items:
        jobs.append((sid, gpu.to_device(rays), len(rays), gpu.buffer(len(rays) * rtx.HIT_DTYPE.itemsize)))
        arena, tlas = (two, 0) if sc is None else (sc.arena, sc.tlas_offset)
        refs.append(oracle.trace(arena, tlas, rays, mode=0, records=False))
    assert (refs[1]["max_level"] < refs[1]["path_level"]).any()

    def check(what):
        for k, ((_, _, n, hb), ref) in enumerate(zip(jobs, refs)):
            assert_hits_equal(hb.download(rtx.HIT_DTYPE, n), ref, f"{what} job {k}")

    assert gpu.trace_batch(jobs, mode=0) == 0
    check("eager")
    g = gpu.batch_graph(jobs, mode=0)
    try:
        for rep in range(3):
            for _, _, n, hb in jobs:
                hb.upload(np.zeros(n, dtype=rtx.HIT_DTYPE))
            st, _ = g.launch(1)
            assert st == 0
            check(f"replay {rep}")
            assert gpu.trace_batch(jobs, mode=0) == 0  # eager work between replays
    finally:
        g.close()


def _cyclic_arena():
    """The tiny arena with internal node A's children moved to the BLAS root
    (internal) and L0: root -> A -> root."""
    a = tiny_scene.tiny_arena().copy()
    node = bytearray(a[448:512].tobytes())
    node[12:16] = np.int32(-2).tobytes()  # children at 448 - 128 = 320
    node[22] = 1 | (0 << 2)  # child 0: the BLAS root, internal
    node[23] = 1 | (4 << 2)  # child 1: L0 at 384, quad leaf
    a[448:512] = np.frombuffer(bytes(node), dtype=np.uint8)
    return a


def test_cycles_are_reported_by_validation(gpu, rtx):
    gpu.set_scene(_cyclic_arena(), 0)
    with pytest.raises(rtx.RtxError) as e:
        gpu.validate_scene()
    assert e.value.status == 15  # RTX_ERR_CYCLE
    # a DAG (one BLAS under two depths) and a tree are not cycles
    for arena in (tiny_scene.tiny_two_depth_arena(), tiny_scene.tiny_shared_arena(), tiny_scene.tiny_arena()):
        gpu.set_scene(arena, 0)
        assert gpu.validate_scene() > 0
