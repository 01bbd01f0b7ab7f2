"""Multi-scene persistent batches, their HIP-graph capture and the autotuner, vs the oracle."""
import numpy as np
import pytest

from conftest import assert_hits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(rtx, gpu):
    from rtx import workloads

    items = workloads.multi_scene_batch(gpu, n_scenes=6, tri_range=(1_000, 40_000), ray_range=(1_000, 12_000),
                                        seed=11)
    jobs, refs = [], []
    for sid, scene, rays in items:
        rb = gpu.to_device(rays)
        hb = gpu.buffer(len(rays) * rtx.HIT_DTYPE.itemsize)
        jobs.append((sid, rb, len(rays), hb))
        refs.append((scene, rays))
    return jobs, refs


def _check(rtx, oracle, jobs, refs, mode, parts=None):
    for k, ((sid, rb, n, hb), (scene, rays)) in enumerate(zip(jobs, refs)):
        h = hb.download(rtx.HIT_DTYPE, n)
        p = parts[k] if parts else None
        ref = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=p, records=False)
        assert_hits_equal(h, ref, f"batch job {k} mode {mode}")


def test_batch_mixed_rays_present(rtx, batch):
    jobs, refs = batch
    flags = np.concatenate([r["flags"] for _, r in refs])
    assert np.any(flags == (1 | 4 | 8)) and np.any(flags == 1)


def test_batch_plain(rtx, oracle, gpu, batch):
    jobs, refs = batch
    assert gpu.trace_batch(jobs, mode=0) == 0
    _check(rtx, oracle, jobs, refs, 0)


def test_batch_treelet(rtx, oracle, gpu, batch):
    jobs, refs = batch
    parts = []
    for (sid, *_), (scene, _) in zip(jobs, refs):
        gpu.build_scene_treelets(sid, 4096)
        parts.append(oracle.Partition(scene.arena, scene.tlas_offset, 4096))
    assert gpu.trace_batch(jobs, mode=1, lds_stack=4, refill=48) == 0
    _check(rtx, oracle, jobs, refs, 1, parts)


def test_batch_graph_replays_match_eager(rtx, oracle, gpu, batch):
    jobs, refs = batch
    assert gpu.trace_batch(jobs, mode=0) == 0
    eager = [hb.download(rtx.HIT_DTYPE, n) for (_, _, n, hb) in jobs]
    g = gpu.batch_graph(jobs, mode=0, waves_per_cu=8)
    for replay in range(3):
        for (_, _, n, hb) in jobs:  # eager work between replays: clear the outputs
            hb.upload(np.zeros(n, rtx.HIT_DTYPE))
        gpu.synchronize()
        st, ms = g.launch(1)
        assert st == 0 and ms > 0
        for k, (_, _, n, hb) in enumerate(jobs):
            assert hb.download(rtx.HIT_DTYPE, n).tobytes() == eager[k].tobytes(), f"replay {replay} job {k}"
    st, ms = g.launch(4)
    assert st == 0
    g.close()
    _check(rtx, oracle, jobs, refs, 0)


def test_autotune(rtx, oracle, gpu, batch):
    jobs, refs = batch
    best, ms = gpu.autotune_batch(jobs, mode=0)
    assert best["waves_per_cu"] in (0, 8, 12, 16, 20) and best["lds_stack"] in (4, 8, 16)
    assert best["refill"] in (24, 32, 48) and ms > 0
    assert gpu.trace_batch(jobs, **best) == 0
    _check(rtx, oracle, jobs, refs, 0)


def test_autotune_treelet_budget(rtx, oracle, gpu, batch):
    jobs, refs = batch
    best, budget, ms = gpu.autotune_batch_treelets(jobs, budgets=(2048, 8192))
    assert budget in (2048, 8192) and ms > 0 and best["mode"] == 1
    assert gpu.trace_batch(jobs, **best) == 0
    parts = [oracle.Partition(scene.arena, scene.tlas_offset, budget) for scene, _ in refs]
    _check(rtx, oracle, jobs, refs, 1, parts)
