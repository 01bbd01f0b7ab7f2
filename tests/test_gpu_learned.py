"""Slot-cache staging and the learned prefetch policy: hits and records stay
bit-exact vs the oracle whatever the prediction table holds."""
import numpy as np
import pytest

from conftest import assert_hits_equal, assert_records_equal, small_instanced_desc

pytestmark = pytest.mark.gpu


def _random_table(rng, n_treelets, fill=0.7):
    t = rng.integers(0, n_treelets, (n_treelets * 8, 4)).astype(np.uint32)
    p = rng.integers(1, 256, (n_treelets * 8, 4)).astype(np.uint32)
    tab = (t + 1) | (p << 24)
    tab[rng.uniform(size=tab.shape) > fill] = 0
    return tab


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("budget,stage,slot", [(1024, 8192, 1024), (512, 8192, 512), (2048, 16384, 2048),
                                               (1024, 4096, 0)])
def test_slot_cache_bit_exact(gpu, oracle, instanced, rtx, mode, budget, stage, slot):
    scene, rays = instanced
    print(f"slots mode={mode} budget={budget} stage={stage} slot={slot}", flush=True)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(budget)
    p = oracle.Partition(scene.arena, scene.tlas_offset, budget)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=p)
    h, o, r = gpu.trace(rays, mode=mode, prefetch=rtx.PREFETCH_ALWAYS, records=True, stage_bytes=stage,
                        slot_bytes=slot)
    st = gpu.prefetch_stats()
    assert_hits_equal(h, h_ref, "slot cache")
    assert_records_equal(o, r, o_ref, r_ref, "slot cache")
    assert st["lds_hits"] > 0 and st["staged_treelets"] > 0
    assert st["wasted_bytes"] <= st["staged_bytes"]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("min_score", [0.01, 2.0])
def test_learned_random_table_bit_exact(gpu, oracle, instanced, rtx, mode, min_score):
    scene, rays = instanced
    print(f"learned random table mode={mode} min_score={min_score}", flush=True)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    s = gpu.build_treelets(1024)
    rng = np.random.default_rng(11 + mode)
    gpu.set_prefetch_table(_random_table(rng, s["n_treelets"]))
    p = oracle.Partition(scene.arena, scene.tlas_offset, 1024)
    h_ref, o_ref, r_ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=p)
    h, o, r = gpu.trace(rays, mode=mode, prefetch=rtx.PREFETCH_LEARNED, records=True, stage_bytes=8192,
                        slot_bytes=1024, min_score=min_score)
    st = gpu.prefetch_stats()
    assert_hits_equal(h, h_ref, "learned, random table")
    assert_records_equal(o, r, o_ref, r_ref, "learned, random table")
    assert st["predictions"] > 0 and st["prediction_hits"] <= st["predictions"]


def test_learned_without_table_falls_back(gpu, instanced, rtx):
    scene, rays = instanced
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    h0 = gpu.trace(rays, prefetch=rtx.PREFETCH_ALWAYS, stage_bytes=8192, slot_bytes=1024)
    a = gpu.prefetch_stats()
    h1 = gpu.trace(rays, prefetch=rtx.PREFETCH_LEARNED, stage_bytes=8192, slot_bytes=1024)
    b = gpu.prefetch_stats()
    assert_hits_equal(h1, h0, "learned without table")
    assert b["predictions"] == 0 and b["fallbacks"] > 0
    assert a["predictions"] == 0 and a["fallbacks"] == 0 and b["lds_hits"] > 0


def test_table_is_dropped_with_its_partition(gpu, instanced, rtx):
    scene, _ = instanced
    gpu.set_scene(scene.arena, scene.tlas_offset)
    s = gpu.build_treelets(1024)
    gpu.set_prefetch_table(np.zeros((s["n_treelets"] * 8, 4), np.uint32))
    with pytest.raises(rtx.RtxError):
        gpu.set_prefetch_table(np.zeros((s["n_treelets"] * 8 - 8, 4), np.uint32))
    bad = np.zeros((s["n_treelets"] * 8, 4), np.uint32)
    bad[0, 0] = s["n_treelets"] + 1  # names treelet n_treelets: out of range
    with pytest.raises(rtx.RtxError):
        gpu.set_prefetch_table(bad)


def test_trained_network_gains_over_heuristic(gpu, rtx):
    """The network trained on this library's own traces (80/20 split: held-out
    pixels and a held-out bounce seed) against the heuristic it falls back to,
    same slots, on the held-out rays: more node reads served from LDS and no
    more wasted staging bytes, with identical hits."""
    from rtx import learned

    n_inst = 256
    d = rtx.workload_desc(rtx.WORKLOAD_INSTANCED_SPHERES, n_instances=n_inst, slices=32, stacks=16, shared_blas=0,
                          width=512, height=512, seed=2023, box_size=100.0 * (n_inst / 1024) ** (1 / 3),
                          scale_min=0.5, scale_max=2.0, displacement=0.05)
    scene = rtx.workload_scene(d)
    gpu.set_scene(scene.arena, scene.tlas_offset)
    gpu.build_treelets(1024)
    prim = rtx.primary_rays(d)
    hits = gpu.trace(prim)
    idx = np.arange(len(prim))
    tr, te = idx % 5 != 0, idx % 5 == 0
    train = [prim[tr][::2]]
    for seed in (1, 2, 3, 4):
        train.append(scene.secondary_rays(prim[tr][::4], hits[tr][::4], kind=0, seed=seed)[0])
    test = np.concatenate([prim[te], scene.secondary_rays(prim[te], hits[te], kind=0, seed=5)[0]])
    lp = learned.LearnedPrefetcher(gpu, scene.arena, scene.tlas_offset, n_classes=16384)
    try:
        losses = lp.fit(np.concatenate(train), epochs=2, max_samples=1_200_000)
        top1, top4, n_tr = lp.accuracy(test)
        lp.install()
    finally:
        lp.close()
    kw = dict(stage_bytes=8192, slot_bytes=1024)
    h_heur = gpu.trace(test, prefetch=rtx.PREFETCH_ALWAYS, **kw)
    heur = gpu.prefetch_stats()
    h_learn = gpu.trace(test, prefetch=rtx.PREFETCH_LEARNED, **kw)
    lrn = gpu.prefetch_stats()
    hr = lambda s: s["lds_hits"] / s["node_reads"]  # noqa: E731
    wb = lambda s: s["wasted_bytes"] / s["node_reads"]  # noqa: E731
    print(f"losses {losses}; held-out next-treelet accuracy top1 {top1:.3f} top4 {top4:.3f} over {n_tr}; "
          f"LDS hit rate heuristic {hr(heur):.4f} learned {hr(lrn):.4f}; wasted B/read heuristic {wb(heur):.3f} "
          f"learned {wb(lrn):.3f}; predicted stagings {lrn['predictions']} used {lrn['prediction_hits']}")
    assert_hits_equal(h_learn, h_heur, "learned vs heuristic")
    assert losses[-1] < losses[0] and top4 > 0.4
    assert lrn["predictions"] > 0 and lrn["prediction_hits"] > 0.3 * lrn["predictions"]
    assert hr(lrn) > hr(heur)
    assert wb(lrn) <= 1.03 * wb(heur)
