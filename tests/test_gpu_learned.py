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
