"""Parity at the benchmark sizes: the full batches are traced on the GPU and a
seeded sample of rays is checked bit for bit against the oracle (hits and node
records), plus size-independent properties over every ray."""
import os
import sys

import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _sample_check(gpu, oracle, scene, rays, hits, mode, part=None, n=6000, seed=0, records=True):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(rays), size=min(n, len(rays)), replace=False))
    sub = rays[idx]
    ref = oracle.trace(scene.arena, scene.tlas_offset, sub, mode=mode, partition=part, records=records)
    if records:
        h_ref, o_ref, r_ref = ref
        h, o, r = gpu.trace(sub, mode=mode, records=True)
        assert_records_equal(o, r, o_ref, r_ref, f"records, mode {mode}")
    else:
        h_ref = ref
    assert_hits_equal(hits[idx], h_ref, f"full batch sample, mode {mode}")


def _props(rays, hits):
    # every ray: a hit lies inside its interval and on its line; the node count is
    # bounded by the records; fetch bytes are whole 64-byte units
    ok = hits["hit"] == 1
    t = hits["t"][ok]
    assert np.all(t >= rays["tmin"][ok]) and np.all(t <= rays["tmax"][ok])
    p = rays["org"][ok] + rays["dir"][ok] * t[:, None]
    assert np.allclose(hits["p"][ok], p, rtol=1e-6, atol=1e-4)
    assert np.all(hits["n_nodes"] <= hits["n_records"])
    assert np.all(hits["fetch_bytes"] % 64 == 0)


def test_diffuse_workload_full_batch(gpu, rtx, oracle):
    import bench

    d = bench.c2_desc(rtx)
    scene = rtx.workload_scene(d)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    prim = rtx.primary_rays(d)
    h0 = gpu.trace(prim)
    bounce, _ = scene.secondary_rays(prim, h0, kind=0, seed=99)
    batch = np.concatenate([prim, bounce])
    print(f"diffuse workload: {scene.n_triangles} triangles, {len(batch)} rays", flush=True)
    hits = gpu.trace(batch)
    _props(batch, hits)
    _sample_check(gpu, oracle, scene, batch, hits, 0)
    gpu.build_treelets(8192)
    part = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    t = gpu.get_treelets()
    assert np.array_equal(t["owner"], part.owner) and np.array_equal(t["roots"], part.roots)
    hits1 = gpu.trace(batch, mode=1)
    _props(batch, hits1)
    _sample_check(gpu, oracle, scene, batch, hits1, 1, part, seed=1)
    # learned-policy staging with a random table over the full batch: identical hits
    s = gpu.build_treelets(1024)
    rng = np.random.default_rng(5)
    tab = (rng.integers(0, s["n_treelets"], (s["n_treelets"] * 8, 4)).astype(np.uint32) + 1) | (200 << 24)
    gpu.set_prefetch_table(tab.astype(np.uint32))
    hits2 = gpu.trace(batch, prefetch=rtx.PREFETCH_LEARNED, stage_bytes=8192, slot_bytes=1024, min_score=1.0)
    assert_hits_equal(hits2, hits, "learned staging over the full batch")


@pytest.mark.parametrize("budget", [16384, 32768, 65536])
def test_diffuse_workload_treelet_budgets(gpu, rtx, oracle, budget):
    """The 1M-triangle scene's partition at the larger budgets equals the
    oracle's (roots, lists, owner), and treelet order over a seeded sample of
    the full batch is bit-exact in hits and records."""
    import bench

    d = bench.c2_desc(rtx)
    scene = rtx.workload_scene(d)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    prim = rtx.primary_rays(d)
    bounce, _ = scene.secondary_rays(prim, gpu.trace(prim), kind=0, seed=99)
    batch = np.concatenate([prim, bounce])
    s = gpu.build_treelets(budget)
    part = oracle.Partition(scene.arena, scene.tlas_offset, budget)
    t = gpu.get_treelets()
    assert s["n_treelets"] == part.n_treelets
    for k in ("roots", "list_offsets", "list_nodes", "owner"):
        assert np.array_equal(t[k], getattr(part, k)), k
    print(f"budget {budget}: {part.n_treelets} treelets", flush=True)
    hits = gpu.trace(batch, mode=1)
    _props(batch, hits)
    _sample_check(gpu, oracle, scene, batch, hits, 1, part, n=3000, seed=budget)


def test_large_workload_sample(gpu, rtx, oracle):
    import bench

    d = bench.c5_desc(rtx, 1920, 1080)  # a quarter of the samples; the same 21M-triangle scene
    scene = rtx.workload_scene(d)
    print(f"large workload: {scene.n_triangles} triangles, arena {scene.arena.nbytes / 2**30:.2f} GiB", flush=True)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    prim = rtx.primary_rays(d)
    hits = gpu.trace(prim)
    _props(prim, hits)
    _sample_check(gpu, oracle, scene, prim, hits, 0, n=3000, records=False)
    b1, _ = scene.secondary_rays(prim, hits, kind=0, seed=1000)
    hb = gpu.trace(b1)
    _props(b1, hb)
    _sample_check(gpu, oracle, scene, b1, hb, 0, n=3000, seed=2, records=False)


def _exact_props(rays, hits):
    """Every ray, bit for bit where the reference's arithmetic fixes it: a hit's
    point is origin + direction * t in FP32 (separate multiply and add,
    vulkan_ray_tracing.cc:2997), a miss leaves the hit fields zero and names no
    instance leaf."""
    _props(rays, hits)
    ok = hits["hit"] == 1
    t = hits["t"][ok].astype(np.float32)
    p = (rays["org"][ok].astype(np.float32) + rays["dir"][ok].astype(np.float32) * t[:, None]).astype(np.float32)
    assert np.array_equal(hits["p"][ok].view(np.uint32), p.view(np.uint32))
    miss = ~ok
    assert np.all(hits["t"][miss] == 0) and np.all(hits["instance_leaf"][miss] == np.uint64(0xFFFFFFFFFFFFFFFF))
    assert np.all(hits["accessed_bytes"] >= hits["fetch_bytes"])


def test_c3_multi_scene_batch_at_full_size(rtx, oracle):
    """64 scenes with log-uniform 1k-2M triangles and log-uniform 4k-1M mixed
    rays (primary / shadow / diffuse) in one hipGraph replay of the persistent
    batch launch, both traversal orders: 2,000 seeded rays of every scene bit
    for bit against the oracle, exact properties over every ray."""
    from rtx import workloads

    ctx = rtx.Context(0)
    try:
        items = workloads.multi_scene_batch(ctx)
        jobs = []
        for sid, scene, rays in items:
            jobs.append((sid, ctx.to_device(rays), len(rays), ctx.buffer(len(rays) * rtx.HIT_DTYPE.itemsize)))
        n_rays = sum(j[2] for j in jobs)
        n_tris = sum(s.n_triangles for _, s, _ in items)
        print(f"c3: {len(items)} scenes, {n_tris} triangles, {n_rays} rays", flush=True)
        assert len(items) == 64 and n_tris > 5_000_000 and n_rays > 2_000_000
        results = {}
        for mode in (0, 1):
            if mode == 1:
                for k, (sid, scene, _) in enumerate(items):
                    ctx.build_scene_treelets(sid, 8192)
            order = sorted(range(len(jobs)), key=lambda k: -jobs[k][2])
            g = ctx.batch_graph([jobs[k] for k in order], mode=mode)
            st, ms = g.launch(1)
            g.close()
            assert st == 0
            print(f"c3 mode {mode}: one replay in {ms:.3f} ms", flush=True)
            rng = np.random.default_rng(100 + mode)
            out = []
            for k, ((sid, scene, rays), (_, _, n, hb)) in enumerate(zip(items, jobs)):
                h = hb.download(rtx.HIT_DTYPE, n)
                _exact_props(rays, h)
                idx = np.sort(rng.choice(n, size=min(2000, n), replace=False))
                part = oracle.Partition(scene.arena, scene.tlas_offset, 8192) if mode == 1 else None
                ref = oracle.trace(scene.arena, scene.tlas_offset, rays[idx], mode=mode, partition=part,
                                   records=False)
                assert_hits_equal(h[idx], ref, f"c3 scene {k}, mode {mode}")
                out.append(h)
            results[mode] = out
        # both orders find the same closest hit for opaque rays that do not stop at their first hit
        for (sid, scene, rays), h0, h1 in zip(items, results[0], results[1]):
            full = rays["flags"] == 1
            assert np.array_equal(h0["t"][full].view(np.uint32), h1["t"][full].view(np.uint32))
    finally:
        for _, rb, _, hb in locals().get("jobs", []):
            rb.free()
            hb.free()
        ctx.close()


def test_c5_full_frame_in_one_batch_launch(rtx, oracle):
    """The large workload at its full size: 21M triangles, 3840x2160 pixels x 4
    samples, camera rays plus three diffuse bounces (53M rays), traced as one
    persistent batch launch over the four segments and as four rtx_trace
    launches.  Both give the same bits for every ray; 2,000 seeded rays of
    every segment match the oracle bit for bit."""
    import bench

    ctx = rtx.Context(0)
    try:
        d = bench.c5_desc(rtx)
        scene = rtx.workload_scene(d)
        print(f"c5: {scene.n_triangles} triangles, arena {scene.arena.nbytes / 2**30:.2f} GiB", flush=True)
        ctx.set_scene(scene.arena, scene.tlas_offset)
        segs = [rtx.primary_rays(d)]
        hits = [ctx.trace(segs[0])]
        for b in range(3):
            nxt, _ = scene.secondary_rays(segs[-1], hits[-1], kind=0, seed=1000 + b)
            segs.append(nxt)
            hits.append(ctx.trace(nxt))
        sizes = [len(s) for s in segs]
        print(f"c5 segments: {sizes} rays", flush=True)
        assert sizes[0] == 3840 * 2160 * 4 and sum(sizes) > 45_000_000
        sid = ctx.add_scene(scene.arena, scene.tlas_offset)
        bufs = [(ctx.to_device(s), ctx.buffer(len(s) * rtx.HIT_DTYPE.itemsize)) for s in segs]
        jobs = [(sid, rb, len(s), hb) for s, (rb, hb) in zip(segs, bufs)]
        order = sorted(range(4), key=lambda k: -sizes[k])
        assert ctx.trace_batch([jobs[k] for k in order], mode=0) == 0
        ms = ctx.last_kernel_ms()
        print(f"c5 batch launch: {ms:.2f} ms, {sum(sizes) / ms / 1e3:.0f} Mrays/s", flush=True)
        for k, (s, (rb, hb)) in enumerate(zip(segs, bufs)):
            hb_hits = hb.download(rtx.HIT_DTYPE, len(s))
            assert hb_hits.tobytes() == hits[k].tobytes(), f"segment {k}: batch launch differs from rtx_trace"
            _exact_props(s, hits[k])
            rng = np.random.default_rng(k)
            idx = np.sort(rng.choice(len(s), size=2000, replace=False))
            ref = oracle.trace(scene.arena, scene.tlas_offset, s[idx], mode=0, records=False)
            assert_hits_equal(hits[k][idx], ref, f"c5 segment {k}")
            rb.free()
            hb.free()
    finally:
        ctx.close()
