"""The device BLAS builder (rtx_builder_finish_device): layout invariants, and
traversal of device-built arenas against the oracle and a float64 brute force."""
import struct
import time

import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal, c1_desc, small_instanced_desc

pytestmark = pytest.mark.gpu


def _walk(A, n, oracle, check_boxes=True):
    """(triangle leaves below internal node n, number of internal nodes), asserting
    that every decoded child box holds the triangles below it."""
    c = n + 64 * struct.unpack_from("<i", A, n + 12)[0]
    leaves, internal = [], 1
    lo, hi = np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)
    for i in range(6):
        info = int(A[n + 22 + i]) & 0x3F
        size, typ = info & 3, info >> 2
        if not size:
            continue
        assert size == 1 and typ in (0, 4)
        if typ == 0:
            sub, k, clo, chi = _walk(A, c, oracle, check_boxes)
            leaves += sub
            internal += k
        else:
            v = np.frombuffer(A[c + 16:c + 52].tobytes(), np.float32).reshape(3, 3)
            clo, chi = v.min(0), v.max(0)
            leaves.append(c)
        if check_boxes:
            blo, bhi = oracle.set_child_bounds(A[n:n + 64], i)
            assert np.all(blo <= clo) and np.all(bhi >= chi), (n, i)
        lo, hi = np.minimum(lo, clo), np.maximum(hi, chi)
        c += 64 * size
    return leaves, internal, lo, hi


def test_device_blas_layout(gpu, rtx, oracle):
    tris = rtx.displaced_sphere(32, 16, seed=5)
    b = rtx.SceneBuilder()
    bid = b.add_blas(tris, geometry_index=5)
    b.add_instance(bid, instance_id=42)
    s = b.finish(ctx=gpu)
    A = s.arena
    assert oracle.validate(A, s.tlas_offset) > len(tris)
    hdr = s.blas_offset(0)
    root = hdr + struct.unpack_from("<Q", A, hdr)[0]
    leaves, internal, lo, hi = _walk(A, root, oracle)
    prims = sorted(struct.unpack_from("<I", A, l + 8)[0] for l in leaves)
    assert prims == list(range(len(tris)))
    for l in leaves:
        p = struct.unpack_from("<I", A, l + 8)[0]
        assert np.array_equal(np.frombuffer(A[l + 16:l + 52].tobytes(), np.float32), tris[p])
        d1 = struct.unpack_from("<I", A, l + 4)[0]
        assert d1 & 0x0FFFFFFF == 5 and (d1 >> 29) & 1 == 0
    # header bounds are the mesh bounds
    hb = np.frombuffer(A[hdr + 8:hdr + 32].tobytes(), np.float32)
    v = tris.reshape(-1, 3)
    assert np.array_equal(hb[:3], v.min(0)) and np.array_equal(hb[3:], v.max(0))
    assert internal <= len(tris)


def test_device_blas_layout_mixed_sizes(gpu, rtx, oracle):
    """BLAS above and below 2,048 triangles in one build: the boxes of the
    small ones are formed per workgroup in LDS, those of the large one through
    device-wide counters; both must give the mesh bounds and hold their leaves."""
    meshes = [rtx.displaced_sphere(64, 40, seed=9), rtx.displaced_sphere(32, 16, seed=5),
              rtx.displaced_sphere(32, 32, seed=2), rtx.displaced_sphere(3, 2, seed=1)]
    assert [len(m) for m in meshes] == [5120, 1024, 2048, 12]
    b = rtx.SceneBuilder()
    for k, m in enumerate(meshes):
        b.add_instance(b.add_blas(m, geometry_index=k), np.array([[1, 0, 0, 4 * k], [0, 1, 0, 0], [0, 0, 1, 0]],
                                                                   np.float32), instance_id=k)
    s = b.finish(ctx=gpu)
    A = s.arena
    assert oracle.validate(A, s.tlas_offset) > sum(len(m) for m in meshes)
    for k, m in enumerate(meshes):
        hdr = s.blas_offset(k)
        root = hdr + struct.unpack_from("<Q", A, hdr)[0]
        leaves, internal, lo, hi = _walk(A, root, oracle)
        assert sorted(struct.unpack_from("<I", A, l + 8)[0] for l in leaves) == list(range(len(m)))
        hb = np.frombuffer(A[hdr + 8:hdr + 32].tobytes(), np.float32)
        v = m.reshape(-1, 3)
        assert np.array_equal(hb[:3], v.min(0)) and np.array_equal(hb[3:], v.max(0)), k
        assert np.array_equal(lo, v.min(0)) and np.array_equal(hi, v.max(0)), k


def test_device_blas_single_triangle_and_many_blas(gpu, rtx, oracle):
    b = rtx.SceneBuilder()
    one = b.add_blas(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), geometry_index=1)
    two = b.add_blas(rtx.displaced_sphere(4, 2, seed=3))
    b.add_instance(one, instance_id=0)
    b.add_instance(two, np.array([[1, 0, 0, 3], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32), instance_id=1)
    s = b.finish(ctx=gpu)
    assert oracle.validate(s.arena, s.tlas_offset) > 0
    rays = np.zeros(2, rtx.RAY_DTYPE)
    rays["org"] = [[0.2, 0.2, -1], [3, 0, -5]]
    rays["dir"] = [[0, 0, 1], [0, 0, 1]]
    rays["tmax"] = 100
    rays["flags"] = 1
    gpu.set_scene(s.arena, s.tlas_offset)
    h = gpu.trace(rays)
    ref = oracle.trace(s.arena, s.tlas_offset, rays, records=False)
    assert_hits_equal(h, ref, "tiny device-built scene")
    assert h["hit"][0] == 1 and h["instance_index"][0] == 0 and h["t"][0] == np.float32(1.0)


@pytest.mark.parametrize("which", ["c1", "instanced"])
def test_device_built_scene_traversal(gpu, rtx, oracle, which):
    d = c1_desc(rtx) if which == "c1" else small_instanced_desc(rtx)
    s = rtx.workload_scene(d, ctx=gpu)
    rays = rtx.primary_rays(d)
    gpu.set_scene(s.arena, s.tlas_offset)
    for mode in (0, 1):
        part = None
        if mode == 1:
            gpu.build_treelets(4096)
            part = oracle.Partition(s.arena, s.tlas_offset, 4096)
        h, o, r = gpu.trace(rays, mode=mode, records=True)
        h_ref, o_ref, r_ref = oracle.trace(s.arena, s.tlas_offset, rays, mode=mode, partition=part)
        assert_hits_equal(h, h_ref, f"{which} device-built, mode {mode}")
        assert_records_equal(o, r, o_ref, r_ref, f"{which} device-built, mode {mode}")
    # the same closest t as the host-built scene (same triangles, another tree)
    host = rtx.workload_scene(d)
    h_host = oracle.trace(host.arena, host.tlas_offset, rays, records=False)
    hit = (h_ref["hit"] == 1) & (h_host["hit"] == 1)
    assert np.array_equal(h_ref["hit"], h_host["hit"])
    assert np.array_equal(h_ref["t"][hit], h_host["t"][hit])


@pytest.mark.parametrize("which", ["c1", "instanced"])
def test_sah_optimal_scene_traversal(gpu, rtx, oracle, which):
    """Arenas of the SAH-optimal host build (three-axis binned SAH, SAH-optimal
    6-wide collapse) traverse bit-exact against the oracle, hits and records, in
    both orders, and hit the same rays at the same distances as the default
    build through fewer nodes."""
    d = c1_desc(rtx) if which == "c1" else small_instanced_desc(rtx)
    s = rtx.workload_builder(d).finish(quality=rtx.BUILD_SAH_OPTIMAL)
    rays = rtx.primary_rays(d)
    print(f"{which}: SAH-optimal arena {s.arena.nbytes} B, {len(rays)} rays", flush=True)
    gpu.set_scene(s.arena, s.tlas_offset)
    for mode in (0, 1):
        part = None
        if mode == 1:
            gpu.build_treelets(8192)
            part = oracle.Partition(s.arena, s.tlas_offset, 8192)
        h, o, r = gpu.trace(rays, mode=mode, records=True)
        h_ref, o_ref, r_ref = oracle.trace(s.arena, s.tlas_offset, rays, mode=mode, partition=part)
        assert_hits_equal(h, h_ref, f"{which} SAH-optimal, mode {mode}")
        assert_records_equal(o, r, o_ref, r_ref, f"{which} SAH-optimal, mode {mode}")
        hw = gpu.trace(rays, mode=mode, lanes_per_ray=8)
        assert_hits_equal(hw, h_ref, f"{which} SAH-optimal, eight lanes per ray, mode {mode}")
        if mode == 0:
            base = rtx.workload_scene(d)
            h_base = oracle.trace(base.arena, base.tlas_offset, rays, records=False)
            assert np.array_equal(h["hit"], h_base["hit"])
            assert np.array_equal(h["t"].view(np.uint32), h_base["t"].view(np.uint32))
            assert h["n_nodes"].mean() < h_base["n_nodes"].mean()


def test_sah_optimal_tlas_over_device_built_blas(gpu, rtx, oracle):
    """The device build with the SAH-optimal setting (it applies to the host-built
    TLAS there): bit-exact traversal against the oracle, and the same hits and
    distances as the default device build."""
    d = small_instanced_desc(rtx, n_instances=64)
    s = rtx.workload_builder(d).finish(ctx=gpu, quality=rtx.BUILD_SAH_OPTIMAL)
    base = rtx.workload_builder(d).finish(ctx=gpu)
    rays = rtx.primary_rays(d)
    gpu.set_scene(s.arena, s.tlas_offset)
    for mode in (0, 1):
        part = None
        if mode == 1:
            gpu.build_treelets(4096)
            part = oracle.Partition(s.arena, s.tlas_offset, 4096)
        h, o, r = gpu.trace(rays, mode=mode, records=True)
        h_ref, o_ref, r_ref = oracle.trace(s.arena, s.tlas_offset, rays, mode=mode, partition=part)
        assert_hits_equal(h, h_ref, f"SAH-optimal TLAS over device BLAS, mode {mode}")
        assert_records_equal(o, r, o_ref, r_ref, f"SAH-optimal TLAS over device BLAS, mode {mode}")
    h_base = oracle.trace(base.arena, base.tlas_offset, rays, records=False)
    assert np.array_equal(h_ref["hit"], h_base["hit"])
    assert np.array_equal(h_ref["t"][h_ref["hit"] == 1], h_base["t"][h_base["hit"] == 1])


def test_device_builder_speed_and_quality(gpu, rtx, oracle):
    import bench

    d = bench.c2_desc(rtx, 256, 256)
    t0 = time.perf_counter()
    host = rtx.workload_scene(d)
    t_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev = rtx.workload_scene(d, ctx=gpu)
    t_dev = time.perf_counter() - t0
    rays = rtx.primary_rays(d)
    nodes = []
    for s in (host, dev):
        gpu.set_scene(s.arena, s.tlas_offset)
        nodes.append(float(gpu.trace(rays)["n_nodes"].mean()))
    print(f"1M triangles: host SAH build {t_host:.2f} s, {nodes[0]:.2f} nodes/ray; "
          f"device build {t_dev:.2f} s, {nodes[1]:.2f} nodes/ray")
    assert nodes[1] < 2.0 * nodes[0]
