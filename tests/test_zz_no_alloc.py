"""rtx_trace_static, the entry point that allocates nothing, against rtx_trace:
two calls in a row on the same GPU (the second on a smaller batch, so the
static control words and spill area are reused), in both traversal orders,
with the per-lane kernel (large batch) and the eight-lanes-per-ray kernel
(small batch).  Hits must be bit-identical to rtx_trace's."""
import numpy as np
import pytest

from conftest import assert_hits_equal, c1_desc

pytestmark = pytest.mark.gpu


def _hits(ctx, rtx, rays, fn, **kw):
    n = len(rays)
    rb = ctx.to_device(rays)
    hb = ctx.buffer(n * rtx.HIT_DTYPE.itemsize)
    hb.upload(np.full(n * rtx.HIT_DTYPE.itemsize, 0xAB, np.uint8))  # stale bytes a write must replace
    st = fn(rb, n, hb, **kw)
    h = hb.download(rtx.HIT_DTYPE, n)
    rb.free()
    hb.free()
    return st, h


@pytest.mark.parametrize("mode", [0, 1])
def test_static_entry_matches_rtx_trace(rtx, oracle, mode):
    ctx = rtx.Context(0)
    try:
        d = c1_desc(rtx, 320, 256)  # 81,920 rays: the per-lane kernel
        scene = rtx.workload_scene(d)
        rays = rtx.primary_rays(d)
        ctx.set_scene(scene.arena, scene.tlas_offset)
        if mode == 1:
            ctx.build_treelets(8192)
        small = rays[7:7 + 3001]  # 3,001 rays: the eight-lanes-per-ray kernel
        for batch, start in ((rays, len(rays) // 4), (small, 0)):
            st0, ref = _hits(ctx, rtx, batch, ctx.trace_device, mode=mode, claim_start=start)
            assert st0 == 0
            st1, got = _hits(ctx, rtx, batch, ctx.trace_static_device, mode=mode, claim_start=start)
            assert st1 == 0, f"rtx_trace_static status {st1}"
            assert_hits_equal(got, ref, f"static vs rtx_trace, mode {mode}, {len(batch)} rays")
        # and against the oracle on a sample
        part = oracle.Partition(scene.arena, scene.tlas_offset, 8192) if mode == 1 else None
        sample = rays[::41]
        h_ref = oracle.trace(scene.arena, scene.tlas_offset, sample, mode=mode, partition=part, records=False)
        st, got = _hits(ctx, rtx, sample, ctx.trace_static_device, mode=mode)
        assert st == 0
        assert_hits_equal(got, h_ref, f"static vs oracle, mode {mode}")
    finally:
        ctx.close()


def test_static_entry_rejects_what_it_cannot_do(rtx, c1):
    scene, rays = c1
    ctx = rtx.Context(0)
    try:
        rb = ctx.to_device(rays[:64])
        hb = ctx.buffer(64 * rtx.HIT_DTYPE.itemsize)
        assert ctx.trace_static_device(rb, 64, hb) == 9  # no scene
        ctx.set_scene(scene.arena, scene.tlas_offset)
        assert ctx.trace_static_device(rb, 64, hb, mode=1) == 10  # no partition
        assert ctx.trace_static_device(rb, 64, hb, lds_stack=5) == 1
        assert ctx.trace_static_device(rb, 64, hb, claim_start=64) == 1
        assert ctx.trace_static_device(rb, 0, hb) == 0
        assert ctx.trace_static_device(rb, 64, hb) == 0
    finally:
        ctx.close()
