"""Hand-assembled 832-byte arena with hand-derived expected outputs.

The bytes are written field by field from the node layout of
src/cuda-sim/vulkan_acceleration_structure_util.h (not by the product
builder), and the expected traces/partitions in tiny_expected.json were
derived by hand from traceRay (vulkan_ray_tracing.cc:2309-3076),
traceRayWithTreelets (:1522-2307) and createTreelets (:823-1470).

Layout (byte offsets):
    0   TLAS header     RootNodeOffset 64, bounds [-1,-1,0]..[3,3,6]
   64   TLAS root       one child: instance leaf at 128 (ChildSize 2, type 1)
  128   instance leaf   identity world->object, BVHAddress 128 (-> 256), InstanceID 7
  256   BLAS header     RootNodeOffset 64 (-> 320)
  320   BLAS root       children at 384: L0 (quad, z=3), A (internal, z 1..2), B (internal, z 4..5)
  448   A               children at 576: L1 (z=1), L2 (z=2)
  512   B               children at 704: L3 (z=4), L4 (z=5)
  384/576/640/704/768   quad leaves L0..L4, PrimitiveIndex0 10..14, GeometryIndex 3
Every triangle is (-1,-1,z) (3,-1,z) (-1,3,z); all boxes quantize exactly
(origin -1 + q/32).  The ray is o=(0.25,0.25,0), d=(0,0,1), t in [0,100].
"""
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _internal(origin, exps, child_offset, children):
    """children: list of (size_blocks, type, (lx,ux,ly,uy,lz,uz))"""
    b = bytearray(64)
    struct.pack_into("<3f", b, 0, *origin)
    struct.pack_into("<i", b, 12, child_offset)
    b[16] = 0
    for a in range(3):
        struct.pack_into("<b", b, 18 + a, exps[a])
    b[21] = 0xFF
    for i, (size, typ, q) in enumerate(children):
        b[22 + i] = size | (typ << 2)
        for f in range(6):
            b[28 + 6 * f + i] = q[f]
    return b


def _quad(prim, geom, z):
    b = bytearray(64)
    struct.pack_into("<I", b, 0, 0xFF000000)
    struct.pack_into("<I", b, 4, (geom & 0x0FFFFFFF) | (1 << 30))
    struct.pack_into("<I", b, 8, prim)
    struct.pack_into("<I", b, 12, 0)
    struct.pack_into("<12f", b, 16, -1, -1, z, 3, -1, z, -1, 3, z, -1, 3, z)
    return b


def _header(root_offset, lo, hi):
    b = bytearray(64)
    struct.pack_into("<Q3f3f", b, 0, root_offset, *lo, *hi)
    return b


def tiny_arena():
    A = bytearray(832)
    A[0:64] = _header(64, (-1, -1, 0), (3, 3, 6))
    # TLAS root: instance child box x,y [-1,3], z [1,5] with origin (-1,-1,1), scale 1/32
    A[64:128] = _internal((-1, -1, 1), (3, 3, 3), 1, [(2, 1, (0, 128, 0, 128, 0, 128))])
    leaf = bytearray(128)
    struct.pack_into("<I", leaf, 0, 0xFF000000)
    struct.pack_into("<I", leaf, 4, 1 << 30)
    struct.pack_into("<12f", leaf, 16, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)  # world->object rows
    struct.pack_into("<QII", leaf, 64, 128, 7, 0)
    struct.pack_into("<12f", leaf, 80, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)  # object->world rows
    A[128:256] = leaf
    A[256:320] = _header(64, (-1, -1, 1), (3, 3, 5))
    # BLAS root children (origin z=1, scale 1/32): L0 z 3 -> 64; A z 1..2 -> 0..32; B z 4..5 -> 96..128
    A[320:384] = _internal((-1, -1, 1), (3, 3, 3), 1, [
        (1, 4, (0, 128, 0, 128, 64, 64)),
        (1, 0, (0, 128, 0, 128, 0, 32)),
        (1, 0, (0, 128, 0, 128, 96, 128)),
    ])
    A[384:448] = _quad(10, 3, 3.0)
    A[448:512] = _internal((-1, -1, 1), (3, 3, 3), 2, [(1, 4, (0, 128, 0, 128, 0, 0)), (1, 4, (0, 128, 0, 128, 32, 32))])
    A[512:576] = _internal((-1, -1, 4), (3, 3, 3), 3, [(1, 4, (0, 128, 0, 128, 0, 0)), (1, 4, (0, 128, 0, 128, 32, 32))])
    A[576:640] = _quad(11, 3, 1.0)
    A[640:704] = _quad(12, 3, 2.0)
    A[704:768] = _quad(13, 3, 4.0)
    A[768:832] = _quad(14, 3, 5.0)
    return np.frombuffer(bytes(A), dtype=np.uint8).copy()


def tiny_shared_arena():
    """Two instance leaves of one BLAS (960 bytes):
        0   TLAS header     RootNodeOffset 64
       64   TLAS root       children at 128: I1 (instance, ChildSize 2), I2 (instance, ChildSize 2)
      128   I1              identity, BVHAddress 256 (-> 384), InstanceID 7
      256   I2              identity, BVHAddress 128 (-> 384), InstanceID 8
      384.. the tiny arena's BLAS (bytes 256..832 of tiny_arena), moved up by 128
    """
    src = tiny_arena()
    A = bytearray(960)
    A[0:64] = _header(64, (-1, -1, 0), (3, 3, 6))
    A[64:128] = _internal((-1, -1, 1), (3, 3, 3), 1, [(2, 1, (0, 128, 0, 128, 0, 128)), (2, 1, (0, 128, 0, 128, 0, 128))])
    for at, bvh, iid in ((128, 256, 7), (256, 128, 8)):
        leaf = bytearray(src[128:256].tobytes())
        struct.pack_into("<QII", leaf, 64, bvh, iid, 0)
        A[at:at + 128] = leaf
    A[384:960] = src[256:832].tobytes()
    return np.frombuffer(bytes(A), dtype=np.uint8).copy()


def tiny_two_depth_arena():
    """One BLAS under instance leaves at two tree levels (1,024 bytes):
        0   TLAS header     RootNodeOffset 64
       64   TLAS root       children at 128: I1 (instance, ChildSize 2), T1 (internal, ChildSize 1)
      128   I1              identity, BVHAddress 320 (-> 448), InstanceID 7   (level 2)
      256   T1              one child at 320: I2 (instance, ChildSize 2)
      320   I2              identity, BVHAddress 128 (-> 448), InstanceID 8   (level 3)
      448.. the tiny arena's BLAS (bytes 256..832 of tiny_arena), moved up by 192
    Every child box is x, y in [-1, 3], z in [1, 5].
    """
    src = tiny_arena()
    A = bytearray(1024)
    A[0:64] = _header(64, (-1, -1, 0), (3, 3, 6))
    full = (0, 128, 0, 128, 0, 128)
    A[64:128] = _internal((-1, -1, 1), (3, 3, 3), 1, [(2, 1, full), (1, 0, full)])
    A[256:320] = _internal((-1, -1, 1), (3, 3, 3), 1, [(2, 1, full)])
    for at, bvh, iid in ((128, 320, 7), (320, 128, 8)):
        leaf = bytearray(src[128:256].tobytes())
        struct.pack_into("<QII", leaf, 64, bvh, iid, 0)
        A[at:at + 128] = leaf
    A[448:1024] = src[256:832].tobytes()
    return np.frombuffer(bytes(A), dtype=np.uint8).copy()


def tiny_ray(ray_dtype, flags=1):
    r = np.zeros(1, dtype=ray_dtype)
    r["org"] = (0.25, 0.25, 0.0)
    r["dir"] = (0.0, 0.0, 1.0)
    r["tmin"] = 0.0
    r["tmax"] = 100.0
    r["flags"] = flags
    r["cull_mask"] = 0xFF
    return r


def expected():
    with open(os.path.join(HERE, "tiny_expected.json")) as f:
        return json.load(f)


def deep_chain_arena(depth):
    """A BLAS that is a chain of `depth` + 1 internal nodes, each with five quad
    leaves and the next node as children (the last has six leaves), under the
    tiny arena's TLAS and instance leaf; every child box spans x, y in [-1, 3]
    and z in [0, 4], every triangle lies at x, y in [2, 3] (off the ray of
    deep_chain_ray).  Plain order keeps the five leaves of every level on the
    stack while it descends: 5 * depth + 6 entries at the deepest node.
        0   TLAS header, 64 TLAS root, 128 instance leaf (BVHAddress -> 256)
      256   BLAS header (root at 320)
      320 + 384 j   chain node j, its children from 384 + 384 j: five leaves,
                    then node j + 1 at 704 + 384 j
    """
    src = tiny_arena()
    size = 320 + 384 * depth + 64 + 6 * 64
    A = bytearray(size)
    A[0:64] = _header(64, (-1, -1, -1), (3, 3, 6))
    full = (0, 128, 0, 128, 0, 128)
    A[64:128] = _internal((-1, -1, 0), (3, 3, 3), 1, [(2, 1, full)])
    A[128:256] = src[128:256].tobytes()
    A[256:320] = _header(64, (-1, -1, 0), (3, 3, 4))

    def leaf(prim):
        b = bytearray(_quad(prim, 3, 2.0))
        struct.pack_into("<12f", b, 16, 2, 2, 2, 3, 2, 2, 2, 3, 2, 2, 3, 2)
        return b

    prim = 0
    for j in range(depth + 1):
        at = 320 + 384 * j
        last = j == depth
        kids = [(1, 4, full)] * 5 + [(1, 4 if last else 0, full)]
        A[at:at + 64] = _internal((-1, -1, 0), (3, 3, 3), 1, kids)
        for k in range(6 if last else 5):
            A[at + 64 + 64 * k:at + 128 + 64 * k] = leaf(prim)
            prim += 1
    return np.frombuffer(bytes(A), dtype=np.uint8).copy()


def deep_chain_ray(ray_dtype, n=1):
    r = np.zeros(n, dtype=ray_dtype)
    r["org"] = (0.25, 0.25, -1.0)
    r["dir"] = (0.0, 0.0, 1.0)
    r["tmin"] = 0.0
    r["tmax"] = 100.0
    r["flags"] = 1
    r["cull_mask"] = 0xFF
    return r
