import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "treelet-prefetching-for-rt_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")


@pytest.fixture(scope="session")
def rtx():
    import rtx as _rtx

    return _rtx


@pytest.fixture(scope="session")
def oracle():
    import oracle as _oracle

    return _oracle


def c1_desc(rtx, width=128, height=128, seed=7):
    """Sphere (128x64 quads = 16,384 triangles) over a 2-triangle ground quad."""
    return rtx.workload_desc(rtx.WORKLOAD_SPHERE_AND_QUAD, slices=128, stacks=64, width=width, height=height,
                             seed=seed, displacement=0.05)


@pytest.fixture(scope="session")
def c1(rtx):
    d = c1_desc(rtx)
    scene = rtx.workload_scene(d)
    rays = rtx.primary_rays(d)
    return scene, rays


def small_instanced_desc(rtx, n_instances=16, shared=0, width=64, height=64, seed=3):
    return rtx.workload_desc(rtx.WORKLOAD_INSTANCED_SPHERES, n_instances=n_instances, slices=16, stacks=8,
                             shared_blas=shared, width=width, height=height, seed=seed, box_size=20.0)


@pytest.fixture(scope="session")
def instanced(rtx):
    d = small_instanced_desc(rtx)
    scene = rtx.workload_scene(d)
    return scene, rtx.primary_rays(d)


@pytest.fixture(scope="session")
def instanced_shared(rtx):
    d = small_instanced_desc(rtx, shared=1, seed=5)
    scene = rtx.workload_scene(d)
    return scene, rtx.primary_rays(d)


@pytest.fixture(scope="session")
def gpu(rtx):
    ctx = rtx.Context(0)
    yield ctx
    ctx.close()


def assert_hits_equal(a, b, what=""):
    """Bit-exact equality of two hit arrays: floats compare as bit patterns,
    except that a NaN equals any NaN.  IEEE 754 leaves the sign and payload of
    a NaN made by an invalid operation (inf - inf, 0 * inf) open: x86 makes it
    negative, the GPU positive.  Fields of one array the other lacks (the
    oracle's path_level, the library's reserved word) are not compared."""
    for f in a.dtype.names:
        if f not in b.dtype.names:
            continue
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            both_nan = np.isnan(x) & np.isnan(y)
            x, y = x.view(np.uint32), y.view(np.uint32)
            diff = (x != y) & ~both_nan
        else:
            diff = x != y
        bad = np.nonzero(diff.reshape(len(a), -1).any(axis=1))[0]
        assert len(bad) == 0, f"{what}: field {f} differs at {len(bad)} rays, first {bad[:5]}: {a[bad[:3]]} vs {b[bad[:3]]}"


def assert_records_equal(oa, ra, ob, rb, what=""):
    assert np.array_equal(oa, ob), f"{what}: record offsets differ"
    for f in ("offset", "treelet", "size", "type", "level"):
        bad = np.nonzero(ra[f] != rb[f])[0]
        assert len(bad) == 0, f"{what}: record field {f} differs at {len(bad)} records, first {bad[:5]}"
