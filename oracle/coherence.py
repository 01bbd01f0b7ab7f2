"""CPU restatement of the reference's ray coherence engine (test infrastructure).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of rtx_coherence_order,
never by the product path.  Restates src/gpgpu-sim/ray_coherency_engine.cc:
  hash_comp                  :393-402
  hash_direction_spherical   :408-421
  hash_origin_grid           :423-430
  hash_francois              :432-445
  hash_grid_spherical        :447-456
  hash_two_point             :462-469
  hash_direction_only        :471-474
  pools and packets          :43-54 (insert), :155-175 (schedule), :182-193 (largest pool)
with FP32 arithmetic (numpy float32, acosf / atan2f from the C library).
"""
import numpy as np

PI = np.float32(3.14159265358979323846)


def hash_comp(x, b):
    o = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    mask = np.uint64((0xFFFFFFFF >> (32 - b)))
    sign = o >> np.uint64(31)
    ex = (o >> np.uint64(31 - b)) & mask
    mant = (o >> np.uint64(23 - b)) & mask
    return (sign << np.uint64(2 * b)) | (ex << np.uint64(b)) | mant


def _to_u64(v):
    v = np.asarray(v, np.float32)
    out = np.zeros(v.shape, np.uint64)
    ok = ~np.isnan(v)
    out[ok] = v[ok].astype(np.float64).astype(np.uint64)
    return out


def _clamp(x, lo, hi):
    # std::max(std::min(x, hi), lo)
    m = np.where(np.float32(hi) < x, np.float32(hi), x).astype(np.float32)
    return np.where(m < np.float32(lo), np.float32(lo), m).astype(np.float32)


def hash_direction_spherical(d, b):
    d = np.asarray(d, np.float32)
    with np.errstate(invalid="ignore"):
        theta = _to_u64(np.arccos(_clamp(d[:, 2], -1.0, 1.0)) / PI * np.float32(180))
        phi = _to_u64((np.arctan2(d[:, 1], d[:, 0]) + PI) / PI * np.float32(180))
    return ((phi >> np.uint64(9 - (b + 1))) << np.uint64(b)) | (theta >> np.uint64(8 - b))


def hash_origin_grid(o, b, wmin, wmax):
    o = np.asarray(o, np.float32)
    g = np.float32(1 << b)
    top = np.float32((1 << b) - 1)
    parts = []
    with np.errstate(divide="ignore", invalid="ignore"):
        for a in range(3):
            v = (o[:, a] - np.float32(wmin[a])) / (np.float32(wmax[a]) - np.float32(wmin[a])) * g
            parts.append(_to_u64(_clamp(v.astype(np.float32), 0.0, top)))
    return (parts[0] << np.uint64(2 * b)) | (parts[1] << np.uint64(b)) | parts[2]


def ray_hash(rays, hash="d", francois_bits=2, grid_bits=4, sphere_bits=4, two_point_ratio=0.5,
             world_min=(0, 0, 0), world_max=(1, 1, 1)):
    o = np.asarray(rays["org"], np.float32)
    d = np.asarray(rays["dir"], np.float32)
    if hash == "f":
        b, cb = francois_bits, 2 * francois_bits + 1
        hd = (hash_comp(d[:, 2], b) << np.uint64(2 * cb)) | (hash_comp(d[:, 1], b) << np.uint64(cb)) | hash_comp(d[:, 0], b)
        ho = (hash_comp(o[:, 0], b) << np.uint64(2 * cb)) | (hash_comp(o[:, 1], b) << np.uint64(cb)) | hash_comp(o[:, 2], b)
        return ho ^ hd
    if hash == "g":
        return hash_origin_grid(o, grid_bits, world_min, world_max) ^ hash_direction_spherical(d, sphere_bits)
    if hash == "t":
        e = np.asarray(world_max, np.float32) - np.asarray(world_min, np.float32)
        mx = e[0] if e[0] > e[1] else e[1]
        ext = mx if mx > e[2] else e[2]
        s = np.float32(two_point_ratio) * np.float32(ext)
        tgt = (o + (s * d).astype(np.float32)).astype(np.float32)
        return hash_origin_grid(o, grid_bits, world_min, world_max) ^ hash_origin_grid(tgt, grid_bits, world_min,
                                                                                        world_max)
    return hash_direction_spherical(d, sphere_bits)


def packet_order(hashes, packet=64):
    """Ray indices in the order the engine schedules them: packets of `packet` rays
    from the largest remaining pool first, ties to the lowest hash (the first maximum
    of the ordered pool map), each pool's rays in input order."""
    h = np.asarray(hashes, np.uint64)
    n = len(h)
    left = np.empty(n, np.int64)
    pools = {}
    for i in range(n):
        pools.setdefault(int(h[i]), []).append(i)
    order = []
    remaining = {k: list(v) for k, v in pools.items()}
    while remaining:
        best, size = None, -1
        for k in sorted(remaining):
            if len(remaining[k]) > size:
                best, size = k, len(remaining[k])
        take = remaining[best][:packet]
        order.extend(take)
        remaining[best] = remaining[best][packet:]
        if not remaining[best]:
            del remaining[best]
    return np.asarray(order, np.uint32)


def packet_order_sorted(hashes, packet=64):
    """The same schedule as one sort (for large batches): chunks by rays left in their
    pool before the chunk (descending), then hash (ascending)."""
    h = np.asarray(hashes, np.uint64)
    idx = np.argsort(h, kind="stable")
    hs = h[idx]
    n = len(h)
    head = np.ones(n, bool)
    head[1:] = hs[1:] != hs[:-1]
    start = np.maximum.accumulate(np.where(head, np.arange(n), 0))
    ends = np.append(np.nonzero(head)[0][1:], n)
    pool_end = ends[np.cumsum(head) - 1]
    size = pool_end - start
    left = size - ((np.arange(n) - start) // packet) * packet
    key = np.lexsort((hs, -left))  # primary: -left ascending, secondary: hash ascending (stable)
    return idx[key].astype(np.uint32)
