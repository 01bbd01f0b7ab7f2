"""The C-ABI boundary of rtx_trace beyond plain parity: claim order, records
written in one traversal (slabs), graph staleness, partition installation
checks and arenas whose instance leaves lie far from offset 0 or from the
TLAS header.  Every GPU result is compared with the oracle bit for bit."""
import struct

import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal
from golden import tiny_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("lanes", [1, 8])
def test_claim_start_keeps_outputs_at_input_indices(gpu, oracle, c1, mode, lanes):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(8192)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 8192) if mode == 1 else None
    h_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part, records=False)
    n = len(rays)
    for start in (1, n // 3, n - 1):
        h = gpu.trace(rays, mode=mode, lanes_per_ray=lanes, claim_start=start)
        assert_hits_equal(h, h_ref, f"claim_start {start}, mode {mode}, {lanes} lanes")


def test_claim_start_out_of_range(gpu, rtx, c1):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    rb = gpu.to_device(rays[:100])
    hb = gpu.buffer(100 * rtx.HIT_DTYPE.itemsize)
    assert gpu.trace_device(rb, 100, hb, claim_start=100) == 1
    assert gpu.trace_device(rb, 0, hb, claim_start=1) == 1
    assert gpu.trace_device(rb, 100, hb, claim_start=99) == 0


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("slab", [1, 5, 16, 64, 1024])
def test_slab_records_bit_exact(gpu, oracle, instanced, mode, slab):
    """Records from per-ray slabs plus a second pass over the rays that
    overflowed them: identical to the oracle for slabs that fit every ray, none
    and some of them."""
    scene, rays = instanced
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part)
    h, o, r = gpu.trace(rays, mode=mode, records=True, record_slab=slab, claim_start=len(rays) // 2)
    assert_hits_equal(h, h_ref, f"slab {slab}, mode {mode}")
    assert_records_equal(o, r, o_ref, r_ref, f"slab {slab}, mode {mode}")
    if slab == 5:
        assert (h_ref["n_records"] > 5).any() and (h_ref["n_records"] <= 5).any()


def test_slab_records_capacity_and_staging(gpu, rtx, oracle, c1):
    """A buffer one record short reports RTX_ERR_TRACE_OVERFLOW with the
    offsets filled; staged kernels write the same records through slabs."""
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, partition=part)
    n, total = len(rays), int(o_ref[-1])
    rb, hb, ob = gpu.to_device(rays), gpu.buffer(n * rtx.HIT_DTYPE.itemsize), gpu.buffer((n + 1) * 8)
    recb = gpu.buffer(total * rtx.RECORD_DTYPE.itemsize)
    assert gpu.trace_device(rb, n, hb, records_buf=recb, record_capacity=total - 1, offsets_buf=ob) == 5
    assert np.array_equal(ob.download(np.uint64, n + 1), o_ref)
    assert gpu.trace_device(rb, n, hb, records_buf=recb, record_capacity=total, offsets_buf=ob, record_slab=8) == 0
    assert_records_equal(ob.download(np.uint64, n + 1), recb.download(rtx.RECORD_DTYPE, total), o_ref, r_ref, "slab 8")
    h, o, r = gpu.trace(rays, mode=0, prefetch=1, records=True, stage_bytes=8192, slot_bytes=1024, record_slab=12)
    assert_hits_equal(h, h_ref, "staged, slab 12")
    assert_records_equal(o, r, o_ref, r_ref, "staged, slab 12")


def test_records_past_the_slab_limit(gpu, rtx, oracle, c1):
    """2,359,296 rays with 1,024-record slabs would need 18 GiB of slabs (8
    bytes per record), past the 16 GiB cap: rtx_trace traces them in two ray ranges that fit (one
    traversal), or, with a claim order, in a counting pass and a writing pass.
    Both give the records of the default 256-record slabs (one range)."""
    rtx_slab_limit = 16 << 30
    assert 2048 * 1152 * 1024 * 8 > rtx_slab_limit
    scene, _ = c1
    from conftest import c1_desc

    d = c1_desc(rtx, 2048, 1152)
    rays = rtx.primary_rays(d)
    n = len(rays)
    assert n == 2048 * 1152
    gpu.set_scene(scene.arena, scene.tlas_offset)
    h1, o1, r1 = gpu.trace(rays, mode=0, records=True)
    # two ranges, the claim starting inside the second
    h2, o2, r2 = gpu.trace(rays, mode=0, records=True, record_slab=1024, claim_start=n - 1000)
    assert_hits_equal(h2, h1, "ranges vs one slab")
    assert_records_equal(o2, r2, o1, r1, "ranges vs one slab")
    # a claim order: counting pass and writing pass
    order = gpu.to_device(np.arange(n, dtype=np.uint32)[::-1].copy())
    rb, hb, ob = gpu.to_device(rays), gpu.buffer(n * rtx.HIT_DTYPE.itemsize), gpu.buffer((n + 1) * 8)
    total = int(o1[-1])
    recb = gpu.buffer(total * rtx.RECORD_DTYPE.itemsize)
    assert gpu.trace_device(rb, n, hb, records_buf=recb, record_capacity=total, offsets_buf=ob, record_slab=1024,
                            claim_order=order) == 0
    assert_hits_equal(hb.download(rtx.HIT_DTYPE, n), h1, "two-pass vs one slab")
    assert_records_equal(ob.download(np.uint64, n + 1), recb.download(rtx.RECORD_DTYPE, total), o1, r1,
                         "two-pass vs one slab")
    # a capacity one short past the first range: reported, offsets filled
    assert gpu.trace_device(rb, n, hb, records_buf=recb, record_capacity=total - 1, offsets_buf=ob,
                            record_slab=1024) == 5
    assert np.array_equal(ob.download(np.uint64, n + 1), o1)
    sample = np.arange(0, n, 293)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays[sample], mode=0)
    assert_hits_equal(h1[sample], h_ref, "slab vs oracle sample")
    got = np.concatenate([r1[o1[i]:o1[i + 1]] for i in sample])
    assert got.tobytes() == r_ref.tobytes()


def test_stale_graphs_report_instead_of_running(rtx, instanced):
    """A treelet-order graph reads its scenes' partitions: rebuilding one makes
    the graph stale, and destroying the context makes every graph stale; both
    are reported before any launch."""
    scene, rays = instanced
    ctx = rtx.Context(0)
    sid = ctx.add_scene(scene.arena, scene.tlas_offset)
    ctx.build_scene_treelets(sid, 1024)
    n = len(rays)
    jobs = [(sid, ctx.to_device(rays), n, ctx.buffer(n * rtx.HIT_DTYPE.itemsize))]
    g_tree = ctx.batch_graph(jobs, mode=1)
    g_plain = ctx.batch_graph(jobs, mode=0)
    assert g_tree.launch(1)[0] == 0 and g_plain.launch(1)[0] == 0
    ctx.build_scene_treelets(sid, 2048)
    assert g_tree.launch(1)[0] == 14
    assert g_plain.launch(1)[0] == 0  # plain order reads no partition
    g_tree2 = ctx.batch_graph(jobs, mode=1)
    assert g_tree2.launch(2)[0] == 0
    for _, rb, _, hb in jobs:
        rb.free()
        hb.free()
    ctx.close()
    assert g_plain.launch(1)[0] == 14 and g_tree2.launch(1)[0] == 14
    for g in (g_tree, g_plain, g_tree2):
        g.close()


def test_set_treelets_rejects_duplicated_list_entries(gpu, rtx, oracle, instanced):
    scene, _ = instanced
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    t = gpu.get_treelets()
    lo, ln = t["list_offsets"], t["list_nodes"].copy()
    k = int(np.argmax(np.diff(lo) >= 2))
    assert lo[k + 1] - lo[k] >= 2
    ln[lo[k] + 1] = ln[lo[k]]  # the list's first node twice
    with pytest.raises(rtx.RtxError) as e:
        gpu.set_treelets(1024, t["roots"], lo, ln, t["owner"])
    assert e.value.status == 1
    gpu.set_treelets(1024, t["roots"], lo, t["list_nodes"], t["owner"])  # the original installs


def _far_tlas_arena(tlas_at):
    """The tiny arena with its BLAS (bytes 256..832) at offset 0 and its TLAS
    (bytes 0..256) at `tlas_at`: the instance leaf lies at tlas_at + 128."""
    src = tiny_scene.tiny_arena()
    A = np.zeros(tlas_at + 256, np.uint8)
    A[0:576] = src[256:832]
    A[tlas_at:tlas_at + 256] = src[0:256]
    leaf = tlas_at + 128
    struct.pack_into("<q", A, leaf + 64, 0 - leaf)  # BVHAddress: leaf -> BLAS header at 0
    return A


def _far_instance_arena(leaf_at):
    """The tiny arena with its instance leaf moved to `leaf_at` and the BLAS
    after it: the TLAS header stays at 0, the leaf lies leaf_at bytes away."""
    src = tiny_scene.tiny_arena()
    A = np.zeros(leaf_at + 128 + 576, np.uint8)
    A[0:128] = src[0:128]
    struct.pack_into("<i", A, 64 + 12, (leaf_at - 64) // 64)  # root ChildOffset -> leaf
    A[leaf_at:leaf_at + 128] = src[128:256]
    A[leaf_at + 128:leaf_at + 128 + 576] = src[256:832]  # BVHAddress 128 still reaches the BLAS
    return A


@pytest.mark.parametrize("mode", [0, 1])
def test_instance_leaves_beyond_256_mib(gpu, rtx, oracle, mode):
    """BLAS below a TLAS that sits 300 MiB into the arena: instance leaves are
    named relative to the TLAS header, so both orders trace it."""
    A = _far_tlas_arena(300 << 20)
    tlas = 300 << 20
    rays = np.concatenate([tiny_scene.tiny_ray(rtx.RAY_DTYPE, flags=f) for f in (1, 0, 5)])
    gpu.set_scene(A, tlas)
    part = None
    if mode == 1:
        gpu.build_treelets(256)
        part = oracle.Partition(A, tlas, 256)
    h_ref, o_ref, r_ref = oracle.trace(A, tlas, rays, mode=mode, partition=part)
    h, o, r = gpu.trace(rays, mode=mode, records=True)
    assert_hits_equal(h, h_ref, f"far TLAS, mode {mode}")
    assert_records_equal(o, r, o_ref, r_ref, f"far TLAS, mode {mode}")
    assert h["hit"][0] == 1 and h["instance_leaf"][0] == tlas + 128
    assert_hits_equal(gpu.trace(rays, mode=mode, lanes_per_ray=8), h_ref, f"far TLAS, mode {mode}, 8 lanes")


def test_instance_leaf_far_from_the_tlas_header(gpu, rtx, oracle):
    """An instance leaf 130 MiB from the TLAS header: plain order traces it;
    treelet order, whose stack entries name the leaf within 128 MiB of the
    header, reports RTX_ERR_UNSUPPORTED."""
    leaf_at = 130 << 20
    A = _far_instance_arena(leaf_at)
    rays = tiny_scene.tiny_ray(rtx.RAY_DTYPE)
    gpu.set_scene(A, 0)
    h_ref, o_ref, r_ref = oracle.trace(A, 0, rays, mode=0)
    assert h_ref["hit"][0] == 1 and h_ref["instance_leaf"][0] == leaf_at
    h, o, r = gpu.trace(rays, mode=0, records=True)
    assert_hits_equal(h, h_ref, "far instance, plain")
    assert_records_equal(o, r, o_ref, r_ref, "far instance, plain")
    gpu.build_treelets(256)
    with pytest.raises(rtx.RtxError) as e:
        gpu.trace(rays, mode=1)
    assert e.value.status == 11


def test_single_lane_adapter_matches_the_oracle(rtx, oracle, instanced):
    """rtx_adapter_trace_ray, the per-lane call with traceRay's operands
    (vulkan_ray_tracing.h:329-356), and rtx_adapter_trace_lanes (a warp's lanes
    in one call) give the oracle's hits and records, in both orders."""
    import ctypes

    lib = rtx._lib
    scene, rays = instanced
    a = ctypes.c_void_p()
    arena = np.ascontiguousarray(scene.arena)
    assert lib.rtx_adapter_create(0, arena.ctypes.data, arena.nbytes, scene.tlas_offset, 1024, ctypes.byref(a)) == 0
    try:
        part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
        pick = np.arange(0, len(rays), max(1, len(rays) // 48))[:48]
        for mode in (0, 1):
            h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays[pick], mode=mode, partition=part)
            recs = np.zeros(4096, rtx.RECORD_DTYPE)
            for j, i in enumerate(pick):
                r = rays[i]
                hit = np.zeros(1, rtx.HIT_DTYPE)
                nrec = ctypes.c_uint64(0)
                f3 = ctypes.c_float * 3
                st = lib.rtx_adapter_trace_ray(a, mode, int(r["flags"]), int(r["cull_mask"]),
                                               int(r["sbt_record_offset"]), int(r["sbt_record_stride"]),
                                               int(r["miss_index"]), f3(*r["org"]), float(r["tmin"]),
                                               f3(*r["dir"]), float(r["tmax"]), hit.ctypes.data, recs.ctypes.data,
                                               len(recs), ctypes.byref(nrec))
                assert st == 0
                assert_hits_equal(hit, h_ref[j:j + 1], f"adapter lane {i}, mode {mode}")
                exp = r_ref[o_ref[j]:o_ref[j + 1]]
                assert nrec.value == len(exp) and recs[:len(exp)].tobytes() == exp.tobytes()
            # one warp of lanes in one call, and a record buffer one short of the total
            warp = np.ascontiguousarray(rays[pick[:32]])
            hits = np.zeros(32, rtx.HIT_DTYPE)
            offs = np.zeros(33, np.uint64)
            out = np.zeros(8192, rtx.RECORD_DTYPE)
            st = lib.rtx_adapter_trace_lanes(a, mode, warp.ctypes.data, 32, hits.ctypes.data, out.ctypes.data,
                                             len(out), offs.ctypes.data)
            assert st == 0
            assert_hits_equal(hits, h_ref[:32], f"adapter warp, mode {mode}")
            assert np.array_equal(offs, o_ref[:33] - o_ref[0])
            assert out[:int(offs[-1])].tobytes() == r_ref[:int(offs[-1])].tobytes()
            st = lib.rtx_adapter_trace_lanes(a, mode, warp.ctypes.data, 32, hits.ctypes.data, out.ctypes.data,
                                             int(offs[-1]) - 1, offs.ctypes.data)
            assert st == 5
    finally:
        assert lib.rtx_adapter_destroy(a) == 0


@pytest.mark.parametrize("hash,world", [("d", 0), ("f", 0), ("g", 1), ("t", 1), ("t", 0)])
def test_coherence_order_matches_the_oracle(gpu, rtx, oracle, instanced, hash, world):
    """rtx_coherence_order: each ray's hash and the packet order equal the CPU
    restatement of the reference's engine; tracing in that claim order leaves
    every hit unchanged.  The spherical hashes go through acosf / atan2f, whose
    GPU and C-library results may differ in the last bit: a ray whose CPU angle
    lies within 1e-4 degrees of a bin edge may land in the neighbouring bin."""
    import coherence

    scene, rays = instanced
    gpu.set_scene(scene.arena, scene.tlas_offset)
    bounce, _ = scene.secondary_rays(rays, gpu.trace(rays), kind=0, seed=3)
    batch = np.concatenate([rays, bounce])
    cfg = dict(hash=hash, world=world, francois_bits=2, grid_bits=4, sphere_bits=4, two_point_ratio=0.5, packet=64)
    order, hashes = gpu.coherence_order(batch, **cfg)
    # the world box the engine used
    A = scene.arena
    if world == 1:
        wmin = np.frombuffer(A[scene.tlas_offset + 8: scene.tlas_offset + 20].tobytes(), np.float32)
        wmax = np.frombuffer(A[scene.tlas_offset + 20: scene.tlas_offset + 32].tobytes(), np.float32)
    else:
        root = scene.tlas_offset + int(np.frombuffer(A[scene.tlas_offset:scene.tlas_offset + 8].tobytes(), np.uint64)[0])
        wmin, wmax = np.zeros(3, np.float32), np.zeros(3, np.float32)
        for i in range(6):
            if int(A[root + 22 + i]) & 3 == 0:
                continue
            lo, hi = oracle.set_child_bounds(A[root:root + 64], i)
            wmin = np.where(lo < wmin, lo, wmin).astype(np.float32)
            wmax = np.where(hi < wmax, hi, wmax).astype(np.float32)  # the reference's running min
    ref = coherence.ray_hash(batch, hash, 2, 4, 4, 0.5, wmin, wmax).astype(np.uint64)
    diff = np.nonzero(hashes.astype(np.uint64) != ref)[0]
    assert len(diff) <= (len(batch) // 1000 if hash in "gd" else 0), f"{len(diff)} hashes differ"
    # the schedule of the GPU's own hashes, restated on the CPU
    assert np.array_equal(order, coherence.packet_order_sorted(hashes, 64))
    assert np.array_equal(np.sort(order), np.arange(len(batch), dtype=np.uint32))
    # tracing in that claim order: the same hits at the same indices
    ob = gpu.to_device(order)
    rb, hb = gpu.to_device(batch), gpu.buffer(len(batch) * rtx.HIT_DTYPE.itemsize)
    assert gpu.trace_device(rb, len(batch), hb, claim_order=ob) == 0
    assert_hits_equal(hb.download(rtx.HIT_DTYPE, len(batch)), gpu.trace(batch), f"coherence order {hash}")


def test_claim_order_rejects_rays_outside_the_batch(gpu, rtx, c1):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    n = 500
    rb, hb = gpu.to_device(rays[:n]), gpu.buffer(n * rtx.HIT_DTYPE.itemsize)
    bad = np.arange(n, dtype=np.uint32)
    bad[17] = n + 3
    assert gpu.trace_device(rb, n, hb, claim_order=gpu.to_device(bad), lanes_per_ray=1) == 3
    assert gpu.trace_device(rb, n, hb, claim_order=gpu.to_device(bad), lanes_per_ray=8) == 3
    good = np.arange(n, dtype=np.uint32)[::-1].copy()
    assert gpu.trace_device(rb, n, hb, claim_order=gpu.to_device(good)) == 0
    assert gpu.trace_device(rb, n, hb, claim_order=gpu.to_device(good), claim_start=2) == 1


@pytest.mark.parametrize("mode", [0, 1])
def test_run_statistics_match_the_oracle(gpu, rtx, oracle, instanced, mode):
    """rtx_trace_stats_of: the reference's end-of-kernel RT counters
    (gpu-sim.cc:2168-2182) from the hits and records, against the same sums
    over the oracle's output; without records the per-type counts follow from
    the per-ray counters."""
    scene, rays = instanced
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=part)
    n, total = len(rays), int(o_ref[-1])
    rb, hb, ob = gpu.to_device(rays), gpu.buffer(n * rtx.HIT_DTYPE.itemsize), gpu.buffer((n + 1) * 8)
    recb = gpu.buffer(total * rtx.RECORD_DTYPE.itemsize)
    assert gpu.trace_device(rb, n, hb, mode=mode, records_buf=recb, record_capacity=total, offsets_buf=ob) == 0
    s = gpu.trace_stats(hb, n, recb, total)
    assert s["access_type"] == [int((r_ref["type"] == t).sum()) for t in range(7)]
    assert s["num_hits"] == int(h_ref["hit"].sum()) and s["num_any_hits"] == int(h_ref["n_any_hits"].sum())
    assert s["tot_nodes"] == int(h_ref["n_nodes"].sum()) and s["max_nodes_per_ray"] == int(h_ref["n_nodes"].max())
    assert s["max_tree_depth"] == int(h_ref["max_level"].max()) and s["n_rays"] == n
    assert s["accessed_bytes"] == int(r_ref["size"].astype(np.int64).sum())
    assert s["fetch_bytes"] == int(h_ref["fetch_bytes"].astype(np.int64).sum()) and s["from_records"] == 1
    d = gpu.trace_stats(hb, n)
    leaves = [int((r_ref["type"] == t).sum()) for t in (4, 5, 6)]
    assert d["access_type"][:4] == s["access_type"][:4] and d["access_type"][4] == sum(leaves)
    assert d["from_records"] == 0 and d["accessed_bytes"] == s["accessed_bytes"]


@pytest.mark.parametrize("mode", [0, 1])
def test_async_traces_queue_and_report_at_the_wait(gpu, rtx, oracle, c1, mode):
    """rtx_trace_async queues launches and returns; after rtx_synchronize every
    queued batch holds the hits rtx_trace gives, and a ray error of one queued
    launch is reported by that wait and not by the next."""
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    if mode:
        gpu.build_treelets(8192)
    parts = [rays[:5000], rays[5000:], rays[::3], rays[:40]]  # the last one takes the eight-lane kernel
    bufs = [(gpu.to_device(r), gpu.buffer(len(r) * rtx.HIT_DTYPE.itemsize)) for r in parts]
    for r, (rb, hb) in zip(parts, bufs):
        assert gpu.trace_async_device(rb, len(r), hb, mode=mode) == 0
    assert gpu.wait() == 0
    for k, (r, (rb, hb)) in enumerate(zip(parts, bufs)):
        assert_hits_equal(hb.download(rtx.HIT_DTYPE, len(r)), gpu.trace(r, mode=mode), f"async batch {k}")
    # a claim order naming a ray outside the batch: reported by the wait
    bad = gpu.to_device(np.array([0, 1, len(parts[3]) + 7], dtype=np.uint32))
    rb, hb = bufs[3]
    assert gpu.trace_async_device(rb, 3, hb, mode=mode, claim_order=bad) == 0
    assert gpu.trace_async_device(bufs[0][0], len(parts[0]), bufs[0][1], mode=mode) == 0
    assert gpu.wait() == 3  # RTX_ERR_OFFSET_OUT_OF_ARENA
    assert gpu.wait() == 0
    # records and staging are not queued
    o = rtx._TraceOptions(mode, 0, 0.0, 1, 0, 0, 0, 0, 0, 0.0, 0.0, 0, 0, 0, 0, None)
    assert rtx._lib.rtx_trace_async(gpu._h, rb.ptr, 3, rtx.ctypes.byref(o), hb.ptr) == 11
