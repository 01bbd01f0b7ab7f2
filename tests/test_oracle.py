"""Pins the CPU oracle without running the reference (which cannot be built here).

Each check is chosen so that a plausible restatement mistake fails it:
  * closed forms of the intersection math and of the child-bound decode;
  * float64 brute force for the slab and Moller-Trumbore tests away from ties;
  * a hand-assembled BVH with hand-derived traces, hits and partitions;
  * brute-force nearest hit over every triangle for whole scenes;
  * structural invariants of traces and partitions.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import tiny_scene  # noqa: E402

F32 = np.float32


# ---- intersection math: closed forms and brute force ------------------------
def test_calculate_idir_closed_forms(oracle):
    assert np.array_equal(oracle.calculate_idir([1.0, 2.0, -4.0]), np.array([1.0, 0.5, -0.25], F32))
    big = F32(2.0**80)
    r = oracle.calculate_idir([0.0, -0.0, 2.0**-81])
    assert r[0] == big and r[1] == -big and r[2] == big
    # |d| exactly 2^-80 is not greater than ooeps: replaced by +-2^-80
    r = oracle.calculate_idir([2.0**-80, -(2.0**-80), 2.0**-79])
    assert r[0] == big and r[1] == -big and r[2] == F32(2.0**79)
    # NaN fails the > test, so copysign(2^-80, NaN) is used
    r = oracle.calculate_idir([np.nan, 1.0, 1.0])
    assert abs(r[0]) == big


def _slab64(lo, hi, o, d, tmin, tmax):
    t0, t1 = tmin, tmax
    for a in range(3):
        inv = 1.0 / d[a]
        ta, tb = (lo[a] - o[a]) * inv, (hi[a] - o[a]) * inv
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return t0 <= t1, t0


def test_ray_box_matches_float64_slab(oracle):
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(3000):
        lo = rng.uniform(-2, 1, 3)
        hi = lo + rng.uniform(0.05, 2, 3)
        o = rng.uniform(-4, 4, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        lo32, hi32, o32, d32 = (x.astype(F32) for x in (lo, hi, o, d))
        ref_hit, ref_t = _slab64(lo32.astype(np.float64), hi32.astype(np.float64), o32.astype(np.float64),
                                 d32.astype(np.float64), 0.0, 1e30)
        # skip near-ties where float32 rounding may legitimately flip the answer
        margins = []
        for a in range(3):
            inv = 1.0 / float(d32[a])
            margins += [(float(lo32[a]) - float(o32[a])) * inv, (float(hi32[a]) - float(o32[a])) * inv]
        hit, th = oracle.ray_box_test(lo32, hi32, oracle.calculate_idir(d32), o32, 0.0, 1e30)
        entry = max([0.0] + [min(margins[2 * a], margins[2 * a + 1]) for a in range(3)])
        exit_ = min([1e30] + [max(margins[2 * a], margins[2 * a + 1]) for a in range(3)])
        if abs(entry - exit_) < 1e-4 * max(1.0, abs(entry)):
            continue
        assert hit == ref_hit
        if hit:
            assert th == pytest.approx(ref_t, rel=1e-5, abs=1e-5)
        checked += 1
    assert checked > 2500


def test_ray_box_axis_parallel_and_nan(oracle):
    lo, hi = [0, 0, 0], [1, 1, 1]
    # axis-parallel ray inside the x/y slab, entering along +z
    hit, th = oracle.ray_box_test(lo, hi, oracle.calculate_idir([0, 0, 1]), [0.5, 0.5, -2], 0.0, 10.0)
    assert hit and th == 2.0
    # axis-parallel ray outside the x slab: (0 - 1.5) * 2^80 ... both bounds negative -> miss
    hit, _ = oracle.ray_box_test(lo, hi, oracle.calculate_idir([0, 0, 1]), [1.5, 0.5, -2], 0.0, 10.0)
    assert not hit
    # tmax cuts the interval: box entered at t=2, tmax=1 -> miss
    hit, _ = oracle.ray_box_test(lo, hi, oracle.calculate_idir([0, 0, 1]), [0.5, 0.5, -2], 0.0, 1.0)
    assert not hit
    # NaN bounds: MIN(NaN, hi) selects hi, MAX(NaN, t) selects t (compare-select semantics)
    hit, th = oracle.ray_box_test([np.nan, 0, 0], [1, 1, 1], oracle.calculate_idir([1, 1, 1]), [-1, -1, -1],
                                  0.0, 10.0)
    # x: lo=NaN, hi=2 -> MIN(NaN,2)=2, MAX(2,0)=2; MAX(NaN,2)=2, MIN(2,10)=2; y,z: [1,2]
    assert hit and th == 2.0


def test_ray_box_infinite_direction_against_overflowing_bound(oracle):
    """An infinite direction component gives idir 0 on that axis; a bound that
    decoded to +inf then makes an (inf - o) * 0 = NaN slab product.  Derived by
    hand from magic_max7 / magic_min7 (VRT:183-197) and MIN/MAX (VRTH:53-56):
      x: lo (3e38 - 0) * 0 = 0, hi (inf - 0) * 0 = NaN
         min side MAX(MIN(0, NaN), -10): 0 < NaN is false -> NaN; NaN > -10 is false -> -10
         max side MIN(MAX(0, NaN), -0.5): 0 > NaN is false -> NaN; NaN < -0.5 is false -> -0.5
      y: idir 2^80: lo -2^79, hi 2^79 -> min side -10, max side -0.5
      z: lo (1 - 2) = -1, hi (5 - 2) = 3 -> min side MAX(-1, -10) = -1, max side MIN(3, -0.5) = -0.5
    so the box is hit at thit = -1.  IEEE max/min with the near bound picked by
    the sign of idir (+0: near = lo) would keep the x product 0 and miss
    (0 > -0.5)."""
    idir = oracle.calculate_idir([np.inf, 0.0, 1.0])
    assert idir[0] == 0.0 and idir[1] == 2.0 ** 80 and idir[2] == 1.0
    hit, th = oracle.ray_box_test([3e38, 0.0, 1.0], [np.inf, 1.0, 5.0], idir, [0.0, 0.5, 2.0], -10.0, -0.5)
    assert hit and th == -1.0


def _mt64(p0, p1, p2, o, d):
    e1, e2 = p1 - p0, p2 - p0
    m = np.array([-d, e1, e2]).T
    try:
        t, u, v = np.linalg.solve(m, o - p0)
    except np.linalg.LinAlgError:
        return None
    return t, u, v


def test_mt_triangle_matches_float64(oracle):
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(3000):
        p = rng.uniform(-1, 1, (3, 3)).astype(F32)
        o = rng.uniform(-3, 3, 3).astype(F32)
        target = (p[0] * 0.2 + p[1] * 0.3 + p[2] * 0.5) + rng.normal(scale=0.5, size=3)
        d = (target - o)
        d = (d / np.linalg.norm(d)).astype(F32)
        ref = _mt64(*(x.astype(np.float64) for x in (p[0], p[1], p[2], o, d)))
        if ref is None:
            continue
        t, u, v = ref
        if min(abs(u), abs(u - 1), abs(v), abs(u + v - 1)) < 1e-4:
            continue
        inside = u >= 0 and u <= 1 and v >= 0 and u + v <= 1
        hit, th = oracle.mt_ray_triangle_test(p[0], p[1], p[2], o, d)
        assert hit == inside
        if hit:
            # no t > 0 test inside the routine: negative t is returned as-is
            assert th == pytest.approx(t, rel=1e-4, abs=1e-4)
        checked += 1
    assert checked > 2500


def test_mt_triangle_inclusive_edges_and_degenerate(oracle):
    p0, p1, p2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    d = [0, 0, 1]
    assert oracle.mt_ray_triangle_test(p0, p1, p2, [0, 0, -1], d) == (True, 1.0)  # vertex: u=v=0 inclusive
    assert oracle.mt_ray_triangle_test(p0, p1, p2, [0.5, 0.5, -1], d)[0]  # hypotenuse: u+v=1 inclusive
    assert not oracle.mt_ray_triangle_test(p0, p1, p2, [0.75, 0.5, -1], d)[0]
    hit, th = oracle.mt_ray_triangle_test(p0, p1, p2, [0.25, 0.25, 3], d)  # behind the origin: t = -3 returned
    assert hit and th == -3.0
    # degenerate triangle: det = 0 -> idet = inf, u = 0*inf = NaN passes the u test, then v = NaN passes too
    hit, th = oracle.mt_ray_triangle_test([0, 0, 0], [1, 0, 0], [2, 0, 0], [0.5, 0, -1], d)
    assert hit and np.isnan(th)


def test_barycentric_closed_form(oracle):
    rng = np.random.default_rng(3)
    for _ in range(500):
        a, b, c = rng.uniform(-2, 2, (3, 3))
        w = rng.dirichlet([1, 1, 1])
        p = w[0] * a + w[1] * b + w[2] * c
        r = oracle.barycentric(p, a, b, c)
        # returns (v, w, u): weights of b, c and a
        assert r[0] == pytest.approx(w[1], abs=2e-3)
        assert r[1] == pytest.approx(w[2], abs=2e-3)
        assert r[2] == pytest.approx(w[0], abs=2e-3)
    r = oracle.barycentric([0.25, 0.25, 1], [-1, -1, 1], [3, -1, 1], [-1, 3, 1])
    assert list(r) == [0.3125, 0.3125, 0.375]


def test_set_child_bounds_closed_form():
    import oracle

    node = bytearray(64)
    import struct

    struct.pack_into("<3f", node, 0, 1.5, -2.0, 0.125)
    struct.pack_into("<3b", node, 18, 3, -2, 9)
    for i in range(6):
        for f in range(6):
            node[28 + 6 * f + i] = (17 * i + 31 * f + 5) % 256
    for i in range(6):
        lo, hi = oracle.set_child_bounds(np.frombuffer(bytes(node), np.uint8), i)
        q = [(17 * i + 31 * f + 5) % 256 for f in range(6)]
        exp_lo = [F32(1.5 + q[0] * 2.0**-5), F32(-2.0 + q[2] * 2.0**-10), F32(0.125 + q[4] * 2.0**1)]
        exp_hi = [F32(1.5 + q[1] * 2.0**-5), F32(-2.0 + q[3] * 2.0**-10), F32(0.125 + q[5] * 2.0**1)]
        assert list(lo) == exp_lo and list(hi) == exp_hi


def test_make_transformed_ray(oracle):
    ident = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    o, d, tmin, tmax, m = oracle.make_transformed_ray(ident, [1, 2, 3], [0, 3, 4], 0.5, 10.0)
    assert list(o) == [1, 2, 3] and list(d) == [0, 0.6, 0.8] and m == 5.0 and tmin == 2.5 and tmax == 50.0
    # uniform scale 1/2 and translation (-1, 0, 2): rows hold A^T with translation last
    rows = [0.5, 0, 0, 0, 0.5, 0, 0, 0, 0.5, -1, 0, 2]
    o, d, tmin, tmax, m = oracle.make_transformed_ray(rows, [4, 2, -2], [2, 0, 0], 0.0, 8.0)
    assert list(o) == [1, 1, 1] and list(d) == [1, 0, 0] and m == 1.0 and tmax == 8.0
    # -0 origin becomes +0: res[i] starts at +0 and accumulates
    o, *_ = oracle.make_transformed_ray(ident, [-0.0, 1, 1], [1, 0, 0], 0.0, 1.0)
    assert np.signbit(o[0]) == np.False_


# ---- hand-assembled BVH ---------------------------------------------------------
def _check_tiny(oracle, mode, case, partition=None):
    exp = tiny_scene.expected()[case]
    arena = tiny_scene.tiny_arena()
    ray = tiny_scene.tiny_ray(oracle.RAY_DTYPE, flags=exp["flags"])
    h, off, rec = oracle.trace(arena, 0, ray, mode=mode, partition=partition)
    got = [[int(r["offset"]), int(r["size"]), int(r["type"])] for r in rec]
    assert got == exp["records"]
    levels = tiny_scene.expected()["levels"]
    assert [int(r["level"]) for r in rec] == [levels[str(o)] for o, _, _ in exp["records"]]
    for k, v in exp["hit"].items():
        if isinstance(v, list):
            assert list(h[0][k]) == v, k
        else:
            assert h[0][k] == v, k
    assert h[0]["n_records"] == len(exp["records"])
    # accessedDataSize (VRT:2260) sums every hand-derived record, descriptors included
    assert h[0]["accessed_bytes"] == sum(size for _, size, _ in exp["records"])
    # the closest hit's instance: the arena's only instance leaf (offset 128); none on a miss
    assert h[0]["instance_leaf"] == (128 if exp["hit"].get("hit", 0) else 0xFFFFFFFFFFFFFFFF)
    return rec


@pytest.mark.parametrize("case", ["plain_opaque", "plain_non_opaque", "plain_terminate"])
def test_tiny_plain(oracle, case):
    _check_tiny(oracle, 0, case)


@pytest.mark.parametrize("case", ["treelet_opaque", "treelet_terminate"])
@pytest.mark.parametrize("budget", [8192, 256, 192])
def test_tiny_treelet(oracle, case, budget):
    p = oracle.Partition(tiny_scene.tiny_arena(), 0, budget)
    rec = _check_tiny(oracle, 1, case, p)
    roots = tiny_scene.expected()["partitions"][str(budget)]["roots"]
    for r in rec:
        owner_root = max(root for root, lst in zip(roots, tiny_scene.expected()["partitions"][str(budget)]["lists"])
                         if int(r["offset"]) in lst)
        assert roots[int(r["treelet"])] == owner_root


@pytest.mark.parametrize("mode,case", [(0, "plain_miss"), (1, "treelet_miss_8192")])
def test_tree_level_map_keeps_the_last_assignment(oracle, mode, case):
    """One BLAS under instance leaves at levels 3 and 2: the second instance
    reassigns every BLAS node one level lower, and the ray's level is the
    maximum over the final map (vulkan_ray_tracing.cc:2274-2279), 4, not the
    deepest assignment, 5.  Records and counters are hand-derived."""
    exp = tiny_scene.expected()["two_depth"]
    arena = tiny_scene.tiny_two_depth_arena()
    ray = tiny_scene.tiny_ray(oracle.RAY_DTYPE)
    ray["org"] = exp["ray_miss"]["org"]
    part = oracle.Partition(arena, 0, 8192) if mode else None
    if part is not None:
        assert part.n_treelets == 1  # every push goes to the current list
    h, off, rec = oracle.trace(arena, 0, ray, mode=mode, partition=part)
    got = [[int(r["offset"]), int(r["size"]), int(r["type"]), int(r["level"])] for r in rec]
    e = exp[case]
    assert got == e["records"]
    for k in ("hit", "n_nodes", "fetch_bytes", "max_level", "path_level"):
        assert h[0][k] == e[k], k
    # the opaque centre ray culls every box under the second instance
    h, _, _ = oracle.trace(arena, 0, tiny_scene.tiny_ray(oracle.RAY_DTYPE), mode=0)
    assert h[0]["max_level"] == exp["plain_opaque_centre"]["max_level"]
    assert h[0]["path_level"] == exp["plain_opaque_centre"]["path_level"]


def test_map_level_is_the_path_level_when_every_node_has_one_level(oracle, c1, instanced, instanced_shared):
    """Where each node is reachable at one level only (tree-shaped scenes, and a
    shared BLAS whose instance leaves sit at one depth), every assignment of a
    node carries the same value, so the final map's maximum is the deepest path
    level; the two-depth arena is where they part."""
    for scene, rays in (c1, instanced, instanced_shared):
        for mode in (0, 1):
            part = oracle.Partition(scene.arena, scene.tlas_offset, 1024) if mode else None
            h, _, _ = oracle.trace(scene.arena, scene.tlas_offset, rays[:2000], mode=mode, partition=part)
            assert np.array_equal(h["max_level"], h["path_level"])
    arena = tiny_scene.tiny_two_depth_arena()
    rng = np.random.default_rng(5)
    rays = np.zeros(400, dtype=oracle.RAY_DTYPE)
    rays["org"][:, 0] = rng.uniform(-1, 3, 400)
    rays["org"][:, 1] = rng.uniform(-1, 3, 400)
    rays["dir"] = (0, 0, 1)
    rays["tmax"] = 100
    rays["flags"] = rng.integers(0, 2, 400)
    rays["cull_mask"] = 0xFF
    h, _, _ = oracle.trace(arena, 0, rays, mode=0)
    assert (h["max_level"] <= h["path_level"]).all()
    assert (h["max_level"] < h["path_level"]).any() and (h["max_level"] == h["path_level"]).any()


@pytest.mark.parametrize("budget", ["8192", "256", "192"])
def test_tiny_partition(oracle, budget):
    exp = tiny_scene.expected()["partitions"][budget]
    for expand in (False, True):
        p = oracle.Partition(tiny_scene.tiny_arena(), 0, int(budget), expand_duplicates=expand)
        assert list(p.roots) == exp["roots"]
        assert [list(map(int, p.nodes_of(t))) for t in range(p.n_treelets)] == exp["lists"]
        assert p.total_bytes == exp["total_bytes"]


@pytest.mark.parametrize("budget,status", [(128, 4), (64, 4)])
def test_tiny_partition_budget_errors(oracle, budget, status):
    with pytest.raises(oracle.OracleError) as e:
        oracle.Partition(tiny_scene.tiny_arena(), 0, budget)
    assert e.value.status == status


def test_malformed_arenas(oracle):
    a = tiny_scene.tiny_arena()
    ray = tiny_scene.tiny_ray(oracle.RAY_DTYPE)
    bad = a.copy()
    bad[64 + 12:64 + 16] = np.frombuffer(np.int32(1000).tobytes(), np.uint8)  # child offset out of the arena
    with pytest.raises(oracle.OracleError) as e:
        oracle.trace(bad, 0, ray, mode=0)
    assert e.value.status == 3
    bad = a.copy()
    bad[64 + 22] = 2 | (4 << 2)  # TLAS child typed as a quad leaf
    with pytest.raises(oracle.OracleError) as e:
        oracle.trace(bad, 0, ray, mode=0)
    assert e.value.status == 2
    # header, TLAS root, instance leaf, BLAS header, BLAS root, A, B and five quad leaves
    assert oracle.validate(a, 0) == 12


# ---- whole scenes: brute force and invariants ----------------------------------
@pytest.mark.parametrize("budget", [256, 192])
def test_tiny_treelet_layout_hand_derived(oracle, budget):
    p = oracle.Partition(tiny_scene.tiny_arena(), 0, budget)
    remap, meta = p.layout(256)
    exp = tiny_scene.expected()["layouts"][str(budget)]
    got = {str(i * 64): int(v) for i, v in enumerate(remap) if v != np.uint64(0xFFFFFFFFFFFFFFFF)}
    assert got == exp
    # metadata: each list entry's place relative to its root's place, -1 past the list
    for t in range(p.n_treelets):
        nodes = p.nodes_of(t)
        assert list(meta[t, :len(nodes)]) == [exp[str(int(a))] - exp[str(int(p.roots[t]))] for a in nodes]
        assert np.all(meta[t, len(nodes):] == -1)


def test_tiny_shared_blas_partition_and_layout_hand_derived(oracle):
    """Two instance leaves of one BLAS: the BLAS header and root are listed by
    both instance treelets; the later root owns them (VRT:479-488) while the
    layout keeps the first treelet's place (VRT:1496-1502)."""
    exp = tiny_scene.expected()["shared"]
    A = tiny_scene.tiny_shared_arena()
    p = oracle.Partition(A, 0, 256)
    assert p.roots.tolist() == exp["partition_256"]["roots"]
    assert [p.nodes_of(t).tolist() for t in range(p.n_treelets)] == exp["partition_256"]["lists"]
    assert {str(i * 64): int(o) for i, o in enumerate(p.owner) if o != 0xFFFFFFFF} == exp["partition_256"]["owner"]
    remap, meta = p.layout(256)
    assert {str(i * 64): int(v) for i, v in enumerate(remap) if v != np.uint64(0xFFFFFFFFFFFFFFFF)} == exp["layout_256"]
    assert meta[2].tolist() == exp["metadata_256_treelet_2"]


def _layout_invariants(p, remap, stride, sizes):
    block = p.budget + stride
    none = np.uint64(0xFFFFFFFFFFFFFFFF)
    placed = remap != none
    # every listed node is placed; roots at their block starts
    for t in range(p.n_treelets):
        assert remap[int(p.roots[t]) // 64] == t * block
        for a in p.nodes_of(t):
            r = int(remap[int(a) // 64])
            # a node lies in the block of the first treelet listing it: never a later one
            assert r < (t + 1) * block
    # placed nodes never overlap and stay inside their block
    idx = np.nonzero(placed)[0]
    order = np.argsort(remap[idx])
    pos = remap[idx][order].astype(np.int64)
    sz = sizes[idx][order]
    assert np.all(pos[1:] >= pos[:-1] + sz[:-1])
    assert np.all(pos % block + sz <= block)


def test_treelet_layout_invariants(oracle, c1, instanced, instanced_shared):
    """Shared BLAS nodes keep the place of the first treelet that lists them
    (vulkan_ray_tracing.cc:1496-1502); places never overlap."""
    for scene, _ in (c1, instanced, instanced_shared):
        A = scene.arena
        kinds = _node_kinds(A, scene.tlas_offset)
        sizes = np.full(A.nbytes // 64, 64, np.int64)
        for off, (kind, _) in kinds.items():
            if kind == "inst":
                sizes[off // 64] = 128
        for budget, stride in ((1024, 0), (2048, 256), (8192, 64)):
            p = oracle.Partition(A, scene.tlas_offset, budget)
            try:
                remap, meta = p.layout(stride)
            except oracle.OracleError:
                continue  # a root listed by an earlier treelet: the reference asserts (:1483)
            _layout_invariants(p, remap, stride, sizes)
            assert meta.shape == (p.n_treelets, budget // 64)


def _brute_nearest(scene, rays):
    """float64 nearest hit over every triangle of every instance (world t)."""
    import rtx

    # walk the arena to get (instance id, world->object rows, triangles) per instance
    A = scene.arena
    out_t = np.full(len(rays), np.inf)
    out_second = np.full(len(rays), np.inf)
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    for inst_off, rows, tris in _instances(A):
        M = rows.reshape(4, 3).astype(np.float64)
        oo = o @ M[:3] + M[3]
        dd = d @ M[:3]
        tm = np.linalg.norm(dd, axis=1)
        dd = dd / tm[:, None]
        p0, e1, e2 = tris[:, 0], tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
        for k in range(0, len(tris), 512):
            P0, E1, E2 = p0[k:k + 512], e1[k:k + 512], e2[k:k + 512]
            pv = np.cross(dd[:, None, :], E2[None])
            det = np.einsum("tj,rtj->rt", E1, pv)
            tv = oo[:, None, :] - P0[None]
            with np.errstate(all="ignore"):
                u = np.einsum("rtj,rtj->rt", tv, pv) / det
                qv = np.cross(tv, E1[None])
                v = np.einsum("rj,rtj->rt", dd, qv) / det
                t = np.einsum("tj,rtj->rt", E2, qv) / det / tm[:, None]
            ok = (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t >= rays["tmin"][:, None]) & (t <= rays["tmax"][:, None])
            t = np.where(ok, t, np.inf)
            srt = np.sort(t, axis=1)[:, :2]
            both = np.concatenate([out_t[:, None], out_second[:, None], srt], axis=1)
            both.sort(axis=1)
            out_t, out_second = both[:, 0], both[:, 1]
    return out_t, out_second


def _instances(A):
    import struct

    root = struct.unpack_from("<Q", A, 0)[0]
    stack = [root]
    while stack:
        n = stack.pop()
        co = struct.unpack_from("<i", A, n + 12)[0]
        c = n + 64 * co
        for i in range(6):
            info = int(A[n + 22 + i]) & 0x3F
            size, typ = info & 3, info >> 2
            if size == 0:
                continue
            if typ == 0:
                stack.append(c)
            else:
                rows = np.frombuffer(A[c + 16:c + 64].tobytes(), np.float32)
                blas = c + struct.unpack_from("<Q", A, c + 64)[0]
                yield c, rows, _blas_tris(A, blas)
            c += 64 * size


def _blas_tris(A, hdr):
    import struct

    stack = [hdr + struct.unpack_from("<Q", A, hdr)[0]]
    tris = []
    while stack:
        n = stack.pop()
        c = n + 64 * struct.unpack_from("<i", A, n + 12)[0]
        for i in range(6):
            info = int(A[n + 22 + i]) & 0x3F
            size, typ = info & 3, info >> 2
            if size == 0:
                continue
            if typ == 0:
                stack.append(c)
            else:
                tris.append(np.frombuffer(A[c + 16:c + 52].tobytes(), np.float32).reshape(3, 3))
            c += 64 * size
    return np.array(tris, dtype=np.float64)


def test_brute_force_nearest_hit_c1(oracle, c1):
    scene, rays = c1
    sub = rays[::8]
    t_ref, t2 = _brute_nearest(scene, sub)
    for mode in (0, 1):
        part = oracle.Partition(scene.arena, scene.tlas_offset, 8192) if mode else None
        h = oracle.trace(scene.arena, scene.tlas_offset, sub, mode=mode, partition=part, records=False)
        clear = np.isfinite(t_ref) == (t2 > t_ref + 1e-4) | ~np.isfinite(t_ref)
        hit_ref = np.isfinite(t_ref)
        assert np.array_equal(h["hit"][clear].astype(bool), hit_ref[clear])
        m = clear & hit_ref
        assert np.allclose(h["t"][m], t_ref[m], rtol=1e-5, atol=1e-6)


def test_brute_force_nearest_hit_instanced(oracle, instanced, instanced_shared):
    for scene, rays in (instanced, instanced_shared):
        sub = rays[::4]
        t_ref, t2 = _brute_nearest(scene, sub)
        part = oracle.Partition(scene.arena, scene.tlas_offset, 4096)
        h = oracle.trace(scene.arena, scene.tlas_offset, sub, mode=1, partition=part, records=False)
        hit_ref = np.isfinite(t_ref)
        assert np.array_equal(h["hit"].astype(bool), hit_ref)
        assert np.allclose(h["t"][hit_ref], t_ref[hit_ref], rtol=1e-5, atol=1e-6)


def test_hit_instance_leaf_maps_the_object_hit_to_the_world_hit(oracle, instanced):
    """instance_leaf names the instance whose matrices Hit_data carries
    (VRT:2856, 3002-3003): its object->world rows (+80) take the object-space
    hit point, the barycentric combination of the closest leaf's vertices (the
    last QUAD_LEAF_HIT record), to the world hit point o + d t.  In plain order
    every accepted hit overwrites the closest leaf while t only shrinks
    (VRT:2850-2859), so there the check takes rays with one accepted hit."""
    scene, rays = instanced
    A = scene.arena
    for mode, part in ((0, None), (1, oracle.Partition(A, scene.tlas_offset, 1024))):
        h, off, rec = oracle.trace(A, scene.tlas_offset, rays, mode=mode, partition=part)
        n_acc = np.add.reduceat((rec["type"] == 5).astype(np.int64), off[:-1].astype(np.int64))
        idx = np.nonzero(h["hit"] & ((n_acc == 1) if mode == 0 else True))[0]
        assert len(idx) > 200
        for i in idx[:: max(1, len(idx) // 400)]:
            r = rec[off[i]:off[i + 1]]
            leaf = int(r["offset"][np.nonzero(r["type"] == 5)[0][-1]])
            v = np.frombuffer(A[leaf + 16: leaf + 52].tobytes(), np.float32).reshape(3, 3).astype(np.float64)
            bv, bw, bu = h["bary"][i].astype(np.float64)
            obj = bu * v[0] + bv * v[1] + bw * v[2]
            il = int(h["instance_leaf"][i])
            rows = np.frombuffer(A[il + 80: il + 128].tobytes(), np.float32).reshape(4, 3).astype(np.float64)
            world = obj @ rows[:3] + rows[3]
            assert np.allclose(world, h["p"][i], rtol=1e-4, atol=1e-4), (mode, i, world, h["p"][i])
            assert np.frombuffer(A[il + 72: il + 76].tobytes(), np.uint32)[0] == h["instance_index"][i]
        assert (h["instance_leaf"][h["hit"] == 0] == 0xFFFFFFFFFFFFFFFF).all() or mode == 1


def test_plain_and_treelet_agree_on_t(oracle, c1):
    scene, rays = c1
    p = oracle.Partition(scene.arena, scene.tlas_offset, 8192)
    hp = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=0, records=False)
    ht = oracle.trace(scene.arena, scene.tlas_offset, rays, mode=1, partition=p, records=False)
    assert np.array_equal(hp["hit"], ht["hit"])
    assert np.array_equal(hp["t"].view(np.uint32), ht["t"].view(np.uint32))


def test_trace_invariants(oracle, c1):
    scene, rays = c1
    sub = rays[::16]
    for mode in (0, 1):
        part = oracle.Partition(scene.arena, scene.tlas_offset, 8192) if mode else None
        h, off, rec = oracle.trace(scene.arena, scene.tlas_offset, sub, mode=mode, partition=part)
        for i in range(len(sub)):
            r = rec[off[i]:off[i + 1]]
            assert r[0]["offset"] == scene.tlas_offset and r[0]["type"] == 0 and r[0]["size"] == 64
            types = r["type"]
            sizes = r["size"].astype(np.int64)
            assert h[i]["fetch_bytes"] == sizes[types != 3].sum()
            assert h[i]["n_nodes"] == np.isin(types, [1, 2, 4, 5, 6]).sum()
            desc = np.nonzero(types == 3)[0]
            assert np.all(r["offset"][desc + 1] == r["offset"][desc])
            assert np.all(np.isin(types[desc + 1], [4, 5, 6]))
            inst = np.nonzero(types == 2)[0]
            assert np.all(types[inst + 1] == 0)
            if mode == 0 and h[i]["hit"]:
                # plain order reports the last accepted leaf (vulkan_ray_tracing.cc:2853)
                last_hit = r["offset"][types == 5][-1]
                import struct

                prim = struct.unpack_from("<I", scene.arena, int(last_hit) + 8)[0]
                assert prim == h[i]["primitive_index"]


def _node_kinds(A, tlas):
    """offset -> (kind, breadth-first children as createTreelets expands them)."""
    import struct

    kinds = {}
    top_root = tlas + struct.unpack_from("<Q", A, tlas)[0]
    kinds[tlas] = ("hdr_top", [top_root])
    stack = [(top_root, True)]
    while stack:
        n, top = stack.pop()
        if n in kinds:
            continue
        c = n + 64 * struct.unpack_from("<i", A, n + 12)[0]
        ch = []
        for i in range(6):
            info = int(A[n + 22 + i]) & 0x3F
            size, typ = info & 3, info >> 2
            if size == 0:
                continue
            ch.append(c)
            if typ == 0:
                stack.append((c, top))
            elif top:
                hdr = c + struct.unpack_from("<Q", A, c + 64)[0]
                broot = hdr + struct.unpack_from("<Q", A, hdr)[0]
                kinds[c] = ("inst", [broot])
                kinds[hdr] = ("hdr_blas", [])
                stack.append((broot, False))
            else:
                kinds[c] = ("leaf", [])
            c += 64 * size
        kinds[n] = ("internal", ch)
    return kinds


def _partition_invariants(arena, tlas, p, budget):
    kinds = _node_kinds(arena, tlas)
    n = p.n_treelets
    roots = [int(x) for x in p.roots]
    assert roots == sorted(roots) and len(set(roots)) == n
    owner_from_lists = {}
    for t in range(n):
        lst = [int(x) for x in p.nodes_of(t)]
        assert len(set(lst)) == len(lst)
        assert lst[0] == roots[t]
        for x in lst:
            owner_from_lists[x] = max(owner_from_lists.get(x, -1), t)
    # node -> treelet is the largest root whose list holds the node (:479-488)
    for x, t in owner_from_lists.items():
        assert p.owner[x // 64] == t
    assert np.count_nonzero(p.owner != 0xFFFFFFFF) == len(owner_from_lists)
    # every reachable node is covered
    assert set(kinds) == set(owner_from_lists)
    root_set = set(roots)
    for t in range(n):
        lst = [int(x) for x in p.nodes_of(t)]
        members = set(lst)
        # bytes: instance leaf 128, every other node (headers included) 64
        total = sum(128 if kinds[x][0] == "inst" else 64 for x in lst)
        assert total <= budget
        # an instance leaf brings its BLAS header into the same treelet
        for x in lst:
            if kinds[x][0] == "inst":
                hdrs = [h for h, k in kinds.items() if k[0] == "hdr_blas" and h in members]
                assert hdrs
            # every breadth-first child of a member is in the treelet or roots its own
            for c in kinds[x][1]:
                if c not in members:
                    assert c in root_set


def test_partition_invariants(oracle, c1, instanced, instanced_shared):
    for scene in (c1[0], instanced[0], instanced_shared[0]):
        for budget in (512, 4096, 65536):
            p = oracle.Partition(scene.arena, scene.tlas_offset, budget)
            _partition_invariants(scene.arena, scene.tlas_offset, p, budget)
            assert p.total_bytes >= scene.arena.nbytes or scene is instanced_shared[0]


def test_expand_duplicates_same_partition(oracle, instanced_shared):
    scene, _ = instanced_shared
    a = oracle.Partition(scene.arena, scene.tlas_offset, 1024, expand_duplicates=False)
    b = oracle.Partition(scene.arena, scene.tlas_offset, 1024, expand_duplicates=True)
    assert np.array_equal(a.roots, b.roots)
    assert np.array_equal(a.list_offsets, b.list_offsets)
    assert np.array_equal(a.list_nodes, b.list_nodes)
    assert np.array_equal(a.owner, b.owner)
    assert b.total_bytes > a.total_bytes  # the reference re-forms shared treelets
