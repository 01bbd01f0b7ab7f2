"""Synthetic workloads built on the library's scene and ray generators.

multi_scene_batch(): many scenes of instanced displaced spheres with
log-uniform triangle counts, each with a log-uniform ray batch mixing 50%
primary, 25% shadow (Opaque | TerminateOnFirstHit | SkipClosestHitShader)
and 25% cosine-weighted diffuse rays.  The shadow and diffuse rays start at
the hits of the scene's primary rays, traced with the plain kernel.
"""
import math

import numpy as np

from . import (MODE_PLAIN, WORKLOAD_INSTANCED_SPHERES, primary_rays, workload_desc, workload_scene)

SPHERE_TRIS = 1024  # 32 x 16 quads


def _log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


def scene_desc(n_tris, width, height, seed):
    n_inst = max(1, int(round(n_tris / SPHERE_TRIS)))
    box = 100.0 * (n_inst / 1024.0) ** (1.0 / 3.0)
    return workload_desc(WORKLOAD_INSTANCED_SPHERES, n_instances=n_inst, slices=32, stacks=16, shared_blas=0,
                         width=width, height=height, seed=seed, box_size=max(box, 4.0))


def mixed_rays(ctx, scene, desc, n_rays, seed):
    """n_rays rays: half primary, a quarter shadow, a quarter diffuse."""
    n_p = max(1, n_rays // 2)
    prim = primary_rays(desc)[:n_p]
    hits = ctx.trace(prim, mode=MODE_PLAIN)
    h = 0.5 * desc.box_size
    light = (h, 2.5 * desc.box_size, h)
    shadow, _ = scene.secondary_rays(prim, hits, kind=1, seed=seed, light=light)
    diffuse, _ = scene.secondary_rays(prim, hits, kind=0, seed=seed + 1)
    n_s = n_rays // 4
    n_d = n_rays - n_p - n_s
    parts = [prim]
    for src, k in ((shadow, n_s), (diffuse, n_d)):
        if k == 0:
            continue
        if len(src) == 0:  # no hits: fall back to more primary rays
            src = prim
        reps = int(math.ceil(k / len(src)))
        parts.append(np.tile(src, reps)[:k])
    return np.concatenate(parts)


def multi_scene_batch(ctx, n_scenes=64, tri_range=(1_000, 2_000_000), ray_range=(4_096, 1_048_576), seed=2024,
                      threads=0):
    """Returns (scenes, ray arrays) with scenes registered on ctx (ctx.add_scene)."""
    rng = np.random.default_rng(seed)
    tris = _log_uniform(rng, *tri_range, n_scenes)
    rays = _log_uniform(rng, *ray_range, n_scenes).astype(np.int64)
    out = []
    for k in range(n_scenes):
        n = int(rays[k])
        n_p = max(1, n // 2)
        w = int(math.ceil(math.sqrt(n_p)))
        d = scene_desc(tris[k], w, int(math.ceil(n_p / w)), seed + 17 * k)
        scene = workload_scene(d, threads)
        ctx.set_scene(scene.arena, scene.tlas_offset)
        batch = mixed_rays(ctx, scene, d, n, seed + k)
        sid = ctx.add_scene(scene.arena, scene.tlas_offset)
        out.append((sid, scene, batch))
    return out
