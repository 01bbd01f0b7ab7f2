"""The reference's interchange formats: per-ray treelet trace text and the
split-file acceleration-structure dump (CPU only)."""
import os
import struct
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import tiny_scene  # noqa: E402


def test_treelet_trace_tiny_hand_derived(rtx, oracle):
    # traceRayWithTreelets on the tiny arena at 256 B: treelets rooted at 0, 128,
    # 384, 448, 512 own [0,64], [128,256,320], [384], [448,576,640], [512,704,768]
    arena = tiny_scene.tiny_arena()
    ray = tiny_scene.tiny_ray(oracle.RAY_DTYPE, flags=1)
    p = oracle.Partition(arena, 0, 256)
    _, off, rec = oracle.trace(arena, 0, ray, mode=1, partition=p)
    text = rtx.format_treelet_trace(off, rec, p.roots)
    assert text == ("RayID,1,0x0,0x0,0x80,0x80,0x80,0x200,0x200,0x200,0x200,0x200,"
                    "0x1c0,0x1c0,0x1c0,0x1c0,0x1c0,0x180,0x180\n")
    # the root address is printed with %x: base + offset, low 32 bits
    text = rtx.format_treelet_trace(off, rec, p.roots, base_address=0x7F0012345000, first_ray_id=41)
    assert text.startswith("RayID,41,0x12345000,0x12345000,0x12345080,")


def test_treelet_trace_many_rays(rtx, oracle, instanced):
    scene, rays = instanced
    p = oracle.Partition(scene.arena, scene.tlas_offset, 4096)
    _, off, rec = oracle.trace(scene.arena, scene.tlas_offset, rays[:200], mode=1, partition=p)
    lines = rtx.format_treelet_trace(off, rec, p.roots).splitlines()
    assert len(lines) == 200
    for i, line in enumerate(lines):
        parts = line.split(",")
        assert parts[0] == "RayID" and int(parts[1]) == i + 1
        roots = [int(x, 16) for x in parts[2:]]
        r = rec[off[i]:off[i + 1]]
        assert roots == [int(p.roots[t]) for t in r["treelet"]]


def test_treelet_trace_needs_treelets(rtx):
    rec = np.zeros(1, rtx.RECORD_DTYPE)
    rec["treelet"] = 0xFFFFFFFF
    with pytest.raises(rtx.RtxError):
        rtx.format_treelet_trace(np.array([0, 1], np.uint64), rec, np.zeros(1, np.uint64))


def _meta(prefix):
    with open(prefix + ".asmetadata") as f:
        return [int(x) for x in f.read().split(",")]


def test_as_dump_round_trip_forward(rtx, oracle, instanced, tmp_path):
    scene, rays = instanced
    prefix = str(tmp_path / "0_1")
    rtx.as_dump_write(scene.arena, scene.tlas_offset, prefix)
    m = _meta(prefix)
    assert m[1] == 1000150000 and m[8] == 0 and m[9] == 1  # BLAS all above the TLAS
    assert os.path.getsize(prefix + ".asmain") == m[0]
    assert not os.path.exists(prefix + ".asback")
    arena, tlas = rtx.as_dump_read(prefix)
    assert tlas == scene.tlas_offset
    assert np.array_equal(arena[: scene.arena.nbytes], scene.arena)
    h0 = oracle.trace(scene.arena, scene.tlas_offset, rays, records=False)
    h1 = oracle.trace(arena, tlas, rays, records=False)
    assert h0.tobytes() == h1.tobytes()


def test_as_dump_round_trip_backward(rtx, oracle, tmp_path):
    # move the tiny arena's BLAS (bytes 256..832) below its TLAS (bytes 0..256)
    src = tiny_scene.tiny_arena()
    blas, top = src[256:832].copy(), src[0:256].copy()
    arena = np.concatenate([blas, top])
    tlas = len(blas)
    leaf = tlas + 128
    struct.pack_into("<q", arena, leaf + 64, 0 - leaf)  # BVHAddress: leaf -> BLAS header at 0
    ray = tiny_scene.tiny_ray(oracle.RAY_DTYPE, flags=1)
    h_ref = oracle.trace(src, 0, ray, records=False)
    h_mov = oracle.trace(arena, tlas, ray, records=False)
    # the same hit, except that the closest instance leaf now sits at its moved offset
    assert h_ref["instance_leaf"][0] == 128 and h_mov["instance_leaf"][0] == leaf
    h_exp = h_ref.copy()
    h_exp["instance_leaf"] = leaf
    assert h_exp.tobytes() == h_mov.tobytes()
    prefix = str(tmp_path / "2_3")
    rtx.as_dump_write(arena, tlas, prefix)
    m = _meta(prefix)
    assert m[8] == 1 and m[9] == 0 and m[2] == m[3] == -tlas
    back, t2 = rtx.as_dump_read(prefix)
    assert t2 == tlas and np.array_equal(back[: arena.nbytes], arena)
    assert oracle.trace(back, t2, ray, records=False).tobytes() == h_exp.tobytes()


def test_as_dump_rejects_bad_metadata(rtx, tmp_path):
    prefix = str(tmp_path / "9_9")
    with open(prefix + ".asmetadata", "w") as f:
        f.write("64,1000150000,0,0")
    with pytest.raises(rtx.RtxError):
        rtx.as_dump_read(prefix)
