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
