"""CPU checks of the learned-prefetch host logic (rtx.learned): BVH parents and
levels, treelet ancestry, transitions from record streams, table packing."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import tiny_scene  # noqa: E402


@pytest.fixture(scope="module")
def learned(rtx):
    from rtx import learned as _l

    return _l


def test_node_tree_tiny_matches_hand_levels(learned):
    arena = tiny_scene.tiny_arena()
    parent, level = learned.node_tree(arena, 0)
    levels = {int(k): v for k, v in tiny_scene.expected()["levels"].items() if k != "_about"}
    for off, lv in levels.items():
        assert level[off // 64] == lv, off
    # hand-read parents of the layout: header<-none, root<-header, leaf<-root, BLAS header/root<-leaf,
    # L0/A/B<-BLAS root, L1/L2<-A, L3/L4<-B
    want = {64: 0, 128: 64, 256: 128, 320: 128, 384: 320, 448: 320, 512: 320, 576: 448, 640: 448,
            704: 512, 768: 512}
    for off, p in want.items():
        assert parent[off // 64] == p // 64, off
    assert parent[0] == 0xFFFFFFFF


def test_node_tree_levels_match_oracle_records(learned, oracle, instanced):
    scene, rays = instanced
    parent, level = learned.node_tree(scene.arena, scene.tlas_offset)
    _, off, rec = oracle.trace(scene.arena, scene.tlas_offset, rays[::7], mode=0)
    slots = (rec["offset"] // 64).astype(np.int64)
    assert np.array_equal(level[slots], rec["level"].astype(np.uint32))
    # every visited node except the TLAS header has a parent, itself visited before it by some ray
    nz = slots[slots != scene.tlas_offset // 64]
    assert np.all(parent[nz] != 0xFFFFFFFF)


def test_treelet_context_tiny(learned, oracle):
    arena = tiny_scene.tiny_arena()
    p = oracle.Partition(arena, 0, 256)  # roots 0,128,384,448,512 (hand-derived)
    parent, level = learned.node_tree(arena, 0)
    root_level, anc = learned.treelet_context(p.roots, p.owner, parent, level)
    assert list(root_level) == [0, 2, 3, 3, 3]
    N = 0xFFFFFFFF
    # 128's parent (64) is in treelet 0; 384/448/512 hang off the BLAS root 320, owned by treelet 1
    assert anc.tolist() == [[N, N, N, N], [0, N, N, N], [1, 0, N, N], [1, 0, N, N], [1, 0, N, N]]


def test_transitions_from_records(learned, rtx):
    rays = np.zeros(3, rtx.RAY_DTYPE)
    rays["dir"] = [[1, 1, 1], [-1, 1, -1], [0, -1, 0]]
    rec = np.zeros(9, rtx.RECORD_DTYPE)
    rec["treelet"] = [0, 0, 5, 5, 2, 0, 7, 7, 3]
    rec["type"] = [0, 1, 3, 4, 1, 0, 1, 1, 1]
    # ray0: records 0..4 -> 0,0,(5 desc skipped),5,2 ; ray1: 5..7 -> 0,7,7 ; ray2: 8 -> 3
    offsets = np.array([0, 5, 8, 9], np.uint64)
    cur, nxt, rid = learned.transitions(rays, offsets, rec)
    assert cur.tolist() == [0, 5, 0] and nxt.tolist() == [5, 2, 7] and rid.tolist() == [0, 0, 1]
    assert learned.octants(rays).tolist() == [0, 5, 2]


def test_mlp_sample_layout(rtx):
    assert rtx.MLP_SAMPLE_DTYPE.itemsize == 32
