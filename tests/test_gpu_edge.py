"""GPU traversal against the hand-derived vectors of the tiny arena, and the
edge cases: malformed arenas, empty batches, record-buffer overflow, budgets
smaller than a node."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import tiny_scene  # noqa: E402
from conftest import assert_hits_equal, assert_records_equal  # noqa: E402

pytestmark = pytest.mark.gpu


def _check(h, off, rec, exp, levels):
    got = [[int(r["offset"]), int(r["size"]), int(r["type"])] for r in rec]
    assert got == exp["records"]
    assert [int(r["level"]) for r in rec] == [levels[str(o)] for o, _, _ in exp["records"]]
    for k, v in exp["hit"].items():
        if isinstance(v, list):
            assert list(h[0][k]) == v, k
        else:
            assert h[0][k] == v, k
    assert int(off[1]) == len(exp["records"])


@pytest.mark.parametrize("case", ["plain_opaque", "plain_non_opaque", "plain_terminate"])
def test_tiny_plain_hand_derived(gpu, rtx, case):
    exp, levels = tiny_scene.expected()[case], tiny_scene.expected()["levels"]
    gpu.set_scene(tiny_scene.tiny_arena(), 0)
    h, off, rec = gpu.trace(tiny_scene.tiny_ray(rtx.RAY_DTYPE, flags=exp["flags"]), mode=0, records=True)
    _check(h, off, rec, exp, levels)


@pytest.mark.parametrize("case", ["treelet_opaque", "treelet_terminate"])
@pytest.mark.parametrize("budget", [8192, 256, 192])
def test_tiny_treelet_hand_derived(gpu, rtx, case, budget):
    exp, levels = tiny_scene.expected()[case], tiny_scene.expected()["levels"]
    part = tiny_scene.expected()["partitions"][str(budget)]
    gpu.set_scene(tiny_scene.tiny_arena(), 0)
    gpu.build_treelets(budget)
    t = gpu.get_treelets()
    assert list(t["roots"]) == part["roots"]
    h, off, rec = gpu.trace(tiny_scene.tiny_ray(rtx.RAY_DTYPE, flags=exp["flags"]), mode=1, records=True)
    _check(h, off, rec, exp, levels)
    for r in rec:
        owner_root = max(root for root, lst in zip(part["roots"], part["lists"]) if int(r["offset"]) in lst)
        assert part["roots"][int(r["treelet"])] == owner_root


@pytest.mark.parametrize("budget", [128, 64])
def test_tiny_budget_smaller_than_a_node(gpu, rtx, budget):
    gpu.set_scene(tiny_scene.tiny_arena(), 0)
    with pytest.raises(rtx.RtxError) as e:
        gpu.build_treelets(budget)
    assert e.value.status in (1, 4)  # 64 is below the accepted minimum, 128 cannot hold instance leaf + header
    if budget == 128:
        assert e.value.status == 4


def test_malformed_arenas_report_like_the_oracle(gpu, rtx, oracle):
    a = tiny_scene.tiny_arena()
    ray = tiny_scene.tiny_ray(rtx.RAY_DTYPE)
    bad = a.copy()
    bad[64 + 12:64 + 16] = np.frombuffer(np.int32(1000).tobytes(), np.uint8)  # child offset out of the arena
    gpu.set_scene(bad, 0)
    with pytest.raises(rtx.RtxError) as e:
        gpu.trace(ray, mode=0)
    assert e.value.status == 3
    bad = a.copy()
    bad[64 + 22] = 2 | (4 << 2)  # TLAS child typed as a quad leaf
    gpu.set_scene(bad, 0)
    with pytest.raises(rtx.RtxError) as e:
        gpu.trace(ray, mode=0)
    assert e.value.status == 2
    # the context still works afterwards
    gpu.set_scene(a, 0)
    h = gpu.trace(ray, mode=0)
    assert h[0]["hit"] == 1 and h[0]["primitive_index"] == 10


def test_empty_batch_and_record_overflow(gpu, rtx, c1):
    scene, rays = c1
    gpu.set_scene(scene.arena, scene.tlas_offset)
    h, off, rec = gpu.trace(rays[:0], records=True)
    assert len(h) == 0 and len(rec) == 0 and int(off[0]) == 0
    n = 256
    rb, hb, ob = gpu.to_device(rays[:n]), gpu.buffer(n * rtx.HIT_DTYPE.itemsize), gpu.buffer((n + 1) * 8)
    recb = gpu.buffer(16 * 8)
    st = gpu.trace_device(rb, n, hb, records_buf=recb, record_capacity=8, offsets_buf=ob)
    assert st == 5  # RTX_ERR_TRACE_OVERFLOW: the count pass found more than 8 records
    total = int(ob.download(np.uint64, n + 1)[-1])
    assert total > 8
    recb = gpu.buffer(total * 16)
    assert gpu.trace_device(rb, n, hb, records_buf=recb, record_capacity=total, offsets_buf=ob) == 0


def test_many_calls_and_errors_reuse_control_state(gpu, rtx, oracle, c1):
    """Successive calls share per-context control state (ray counter, error
    word, statistics): results stay exact over more calls than it holds slots,
    and an error in one call does not leak into the next."""
    scene, rays = c1
    sub = rays[:2048]
    ref = oracle.trace(scene.arena, scene.tlas_offset, sub, records=False)
    bad = tiny_scene.tiny_arena().copy()
    bad[64 + 12:64 + 16] = np.frombuffer(np.int32(1000).tobytes(), np.uint8)
    ray = tiny_scene.tiny_ray(rtx.RAY_DTYPE)
    rb, hb = gpu.to_device(sub), gpu.buffer(len(sub) * rtx.HIT_DTYPE.itemsize)
    for k in range(150):
        if k in (40, 41, 100):
            gpu.set_scene(bad, 0)
            with pytest.raises(rtx.RtxError) as e:
                gpu.trace(ray, mode=0)
            assert e.value.status == 3
        gpu.set_scene(scene.arena, scene.tlas_offset)
        assert gpu.trace_device(rb, len(sub), hb) == 0
        if k % 25 == 0 or k in (40, 41, 100):
            from conftest import assert_hits_equal
            assert_hits_equal(hb.download(rtx.HIT_DTYPE, len(sub)), ref, f"call {k}")


def _invert_child_boxes(arena, tlas, limit):
    """Copy of the arena where up to `limit` internal-node children reached from
    the TLAS (breadth first, so the top levels every ray visits) have their
    lower and upper x bound bytes swapped (lo > hi): the reference's slab test
    still reads such a box as the span between the two."""
    import struct
    A = arena.copy()
    root = tlas + struct.unpack_from("<Q", A, tlas)[0]
    queue, done, seen = [root], 0, set()
    while queue and done < limit:
        n = queue.pop(0)
        if n in seen:
            continue
        seen.add(n)
        c = n + 64 * struct.unpack_from("<i", A, n + 12)[0]
        for k in range(6):
            info = int(A[n + 22 + k]) & 0x3F
            size, typ = info & 3, info >> 2
            if not size:
                continue
            lo, hi = int(A[n + 28 + k]), int(A[n + 34 + k])
            if lo < hi and done < limit:
                A[n + 28 + k], A[n + 34 + k] = hi, lo
                done += 1
            if typ == 0:
                queue.append(c)
            c += 64 * size
    assert done > 0
    return A


@pytest.mark.parametrize("mode", [0, 1])
def test_inverted_child_boxes_match_the_oracle(gpu, rtx, oracle, c1, mode):
    """An arena whose child boxes are not all ordered keeps the compare-select
    slab tests (the sign-ordered form would read an inverted box as empty)."""
    scene, rays = c1
    A = _invert_child_boxes(scene.arena, scene.tlas_offset, 40)
    gpu.set_scene(A, scene.tlas_offset)
    part = None
    if mode == 1:
        gpu.build_treelets(8192)
        part = oracle.Partition(A, scene.tlas_offset, 8192)
    h_ref, o_ref, r_ref = oracle.trace(A, scene.tlas_offset, rays, mode=mode, partition=part)
    h, o, r = gpu.trace(rays, mode=mode, records=True)
    from conftest import assert_hits_equal, assert_records_equal
    assert_hits_equal(h, h_ref, f"inverted boxes, mode {mode}")
    assert_records_equal(o, r, o_ref, r_ref, f"inverted boxes, mode {mode}")
    assert_hits_equal(gpu.trace(rays, mode=mode, lanes_per_ray=8), h_ref, f"inverted boxes, mode {mode}, 8 lanes")
    # the reference reads a swapped pair as the same span, so the rays visit
    # what they visit in the original arena; a test that took swapped bounds
    # as an empty box would miss those children
    h_orig = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=mode, partition=None if mode == 0 else
                          oracle.Partition(scene.arena, scene.tlas_offset, 8192), records=False)
    assert np.array_equal(h_orig["n_nodes"], h_ref["n_nodes"]) and h_ref["n_nodes"].sum() > 0


@pytest.mark.parametrize("lds_stack", [4, 8, 16])
def test_stack_capacity_edge(gpu, rtx, oracle, lds_stack):
    """A chain BLAS whose plain-order stack peaks at 5 * depth + 6 entries: at
    the deepest depth that fits the lane's K LDS entries plus 128 spilled ones
    the hits and records equal the oracle's, one level deeper the call reports
    RTX_ERR_STACK_OVERFLOW, with one lane per ray (the per-node pushes, their
    spills and the per-child path's overflow check) and with eight."""
    cap = lds_stack + 128
    fits = (cap - 6) // 5
    assert 5 * fits + 6 <= cap < 5 * (fits + 1) + 6
    ray = tiny_scene.deep_chain_ray(rtx.RAY_DTYPE, n=64)
    A = tiny_scene.deep_chain_arena(fits)
    gpu.set_scene(A, 0)
    h_ref, o_ref, r_ref = oracle.trace(A, 0, ray, mode=0)
    for lanes in (1, 8):
        assert_hits_equal(gpu.trace(ray, mode=0, lds_stack=lds_stack, lanes_per_ray=lanes), h_ref,
                          f"depth {fits}, stack {lds_stack}, {lanes} lanes")
    h, o, r = gpu.trace(ray, mode=0, records=True, lds_stack=lds_stack)
    assert_hits_equal(h, h_ref, "records")
    assert_records_equal(o, r, o_ref, r_ref, f"depth {fits}, stack {lds_stack}")
    gpu.set_scene(tiny_scene.deep_chain_arena(fits + 1), 0)
    for lanes in (1, 8):
        with pytest.raises(rtx.RtxError) as e:
            gpu.trace(ray, mode=0, lds_stack=lds_stack, lanes_per_ray=lanes)
        assert e.value.status == 6, (lanes, e.value.status)


def test_stack_capacity_treelet_order(gpu, rtx, oracle):
    """Treelet order on the chain BLAS: at depth 20 the two lists hold 106
    entries at most, however they split, and fit; at depth 60 the first
    treelet (1 KiB: 16 nodes) is left early and the other list alone has to
    take some 290 entries, past its 136."""
    ray = tiny_scene.deep_chain_ray(rtx.RAY_DTYPE, n=64)
    A = tiny_scene.deep_chain_arena(20)
    gpu.set_scene(A, 0)
    gpu.build_treelets(1024)
    part = oracle.Partition(A, 0, 1024)
    h_ref, o_ref, r_ref = oracle.trace(A, 0, ray, mode=1, partition=part)
    for lanes in (1, 8):
        assert_hits_equal(gpu.trace(ray, mode=1, lanes_per_ray=lanes), h_ref, f"treelet, {lanes} lanes")
    h, o, r = gpu.trace(ray, mode=1, records=True)
    assert_records_equal(o, r, o_ref, r_ref, "treelet records")
    gpu.set_scene(tiny_scene.deep_chain_arena(60), 0)
    gpu.build_treelets(1024)
    for lanes in (1, 8):
        with pytest.raises(rtx.RtxError) as e:
            gpu.trace(ray, mode=1, lanes_per_ray=lanes)
        assert e.value.status == 6, (lanes, e.value.status)
