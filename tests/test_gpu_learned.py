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


def _split_rays(rtx, ctx, scene, desc, bounce_seeds_train=(1, 2, 3, 4), bounce_seed_test=5):
    """80/20 split of a camera's pixels (every fifth pixel held out), each side with
    diffuse bounces of its own hits; the held-out side's bounces use another seed."""
    prim = rtx.primary_rays(desc)
    ctx.set_scene(scene.arena, scene.tlas_offset)
    hits = ctx.trace(prim)
    idx = np.arange(len(prim))
    tr, te = idx % 5 != 0, idx % 5 == 0
    train = [prim[tr][::2]]
    for seed in bounce_seeds_train:
        train.append(scene.secondary_rays(prim[tr][::4], hits[tr][::4], kind=0, seed=seed)[0])
    test = np.concatenate([prim[te], scene.secondary_rays(prim[te], hits[te], kind=0, seed=bounce_seed_test)[0]])
    return np.concatenate(train), test


def test_network_trained_on_c1_and_c2_gains_on_held_out_c2(gpu, rtx):
    """The next-treelet network trained on this library's own traces of both
    benchmark scenes (the 16K-triangle sphere over a quad and the 1M-triangle
    instanced spheres; 80/20 pixel split, held-out bounce seed), judged against
    the heuristic it falls back to on held-out rays of the 1M-triangle scene,
    same LDS slots: more node reads served from LDS and no more wasted staging
    bytes per node read, with identical hits."""
    import bench
    from conftest import c1_desc
    from rtx import learned

    budget, kw = 1024, dict(stage_bytes=8192, slot_bytes=1024)
    d1, d2 = c1_desc(rtx), bench.c2_desc(rtx, 1024, 1024)
    s1, s2 = rtx.workload_scene(d1), rtx.workload_scene(d2)
    tr1, _ = _split_rays(rtx, gpu, s1, d1)
    tr2, te2 = _split_rays(rtx, gpu, s2, d2)
    src1 = learned.scene_transitions(gpu, s1.arena, s1.tlas_offset, budget, tr1)
    src2 = learned.scene_transitions(gpu, s2.arena, s2.tlas_offset, budget, tr2)  # leaves the C2 scene set
    jp = learned.JointPrefetcher(gpu, [src1, src2], n_classes=16384)
    try:
        losses = jp.fit(epochs=2, max_samples=3_000_000)
        test = learned.scene_transitions(gpu, s2.arena, s2.tlas_offset, budget, te2)
        top1, top4 = jp.accuracy(1, test["cur"], test["nxt"], test["oct"])
        gpu.set_prefetch_table(jp.table(1))
    finally:
        jp.close()
    h_heur = gpu.trace(te2, prefetch=rtx.PREFETCH_ALWAYS, **kw)
    heur = gpu.prefetch_stats()
    h_learn = gpu.trace(te2, prefetch=rtx.PREFETCH_LEARNED, **kw)
    lrn = gpu.prefetch_stats()
    hr = lambda s: s["lds_hits"] / s["node_reads"]  # noqa: E731
    wb = lambda s: s["wasted_bytes"] / s["node_reads"]  # noqa: E731
    print(f"trained on {jp.n_train} transitions of both scenes; losses {losses}; held-out next-treelet "
          f"accuracy top1 {top1:.3f} top4 {top4:.3f}; LDS hit rate heuristic {hr(heur):.4f} learned "
          f"{hr(lrn):.4f}; wasted B/read heuristic {wb(heur):.3f} learned {wb(lrn):.3f}; predicted stagings "
          f"{lrn['predictions']} used {lrn['prediction_hits']}", flush=True)
    assert_hits_equal(h_learn, h_heur, "learned vs heuristic")
    assert losses[-1] < losses[0] and top4 > 0.4
    assert lrn["predictions"] > 0 and lrn["prediction_hits"] > 0.3 * lrn["predictions"]
    assert hr(lrn) > hr(heur)
    assert wb(lrn) <= wb(heur)
