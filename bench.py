#!/usr/bin/env python3
"""Ray-traversal benchmark on one MI355X.

Workload (1,048,576 triangles, diffuse): 1,024 instances of a 1,024-triangle
displaced sphere (32x16 quads, radial displacement U[-0.05, 0.05], a distinct
seed and BLAS per instance) at uniform positions in a 100^3 box with uniform
scale in [0.5, 2]; 1024x1024 pinhole primary rays plus one cosine-weighted
diffuse bounce ray per primary hit, all FP32, all resident in HBM.

A step traces the whole batch (primary + bounce rays) once.  `value` is the
batch's rays per second of the headline mode over the timed steps (wall clock
around synchronous rtx_trace calls); every mode is measured and reported.

  python bench.py [--steps K] [--warmup W] [--budget BYTES] [--mode NAME]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "treelet-prefetching-for-rt_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s plain vs treelet vs learned, 1M-tri diffuse; node-fetch GB/s % of HBM peak"

# name -> (traversal order, prefetch policy); policies: 1 always, 2 popularity, 3 partial
MODES = {
    "plain": (0, 0),
    "treelet": (1, 0),
    "plain+stage": (0, 1),
    "treelet+stage": (1, 1),
    "treelet+stage_pop": (1, 2),
    "treelet+stage_partial": (1, 3),
}
HEADLINE = "plain"


def c2_desc(rtx, width=1024, height=1024):
    return rtx.workload_desc(rtx.WORKLOAD_INSTANCED_SPHERES, n_instances=1024, slices=32, stacks=16, shared_blas=0,
                             width=width, height=height, seed=2023, box_size=100.0, scale_min=0.5, scale_max=2.0,
                             displacement=0.05)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, batch, target_s=10.0):
    """The oracle (C restatement of traceRay) on host cores, plain order, over a strided sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    probe = batch[:: max(1, len(batch) // 8192)]
    t0 = time.perf_counter()
    oracle.trace_parallel(scene.arena, scene.tlas_offset, probe, mode=0, threads=threads)
    rate = len(probe) / max(time.perf_counter() - t0, 1e-6)
    n = int(min(len(batch), max(len(probe), rate * target_s)))
    sample = batch[:: max(1, len(batch) // n)]
    t0 = time.perf_counter()
    oracle.trace_parallel(scene.arena, scene.tlas_offset, sample, mode=0, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(len(sample) / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} rays (every {max(1, len(batch) // n)}th ray of the primary+bounce batch), "
                      f"plain order, {dt:.1f} s on {threads} host threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--budget", type=int, default=8192, help="treelet byte budget")
    ap.add_argument("--stage", type=int, default=8192, help="LDS bytes per wave for a staged treelet")
    ap.add_argument("--threshold", type=float, default=0.5, help="popularity share for the popularity policy")
    ap.add_argument("--mode", default=HEADLINE, choices=sorted(MODES))
    ap.add_argument("--modes", default=",".join(MODES), help="modes to measure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--width", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (rays shard trivially across "
              "GPUs, see DESIGN.md)", file=sys.stderr)
        return 2

    import rtx

    desc = c2_desc(rtx, args.width, args.width)
    t0 = time.perf_counter()
    scene = rtx.workload_scene(desc)
    log(f"scene: {scene.n_triangles} tris, {scene.n_blas} BLAS, {scene.arena.nbytes / 2**20:.1f} MiB arena, "
        f"built in {time.perf_counter() - t0:.1f} s")
    ctx = rtx.Context(0)
    info = ctx.device_info()
    ctx.set_scene(scene.arena, scene.tlas_offset)
    t0 = time.perf_counter()
    summ = ctx.build_treelets(args.budget)
    log(f"treelets: {summ['n_treelets']} at {args.budget} B budget, built in {time.perf_counter() - t0:.2f} s")

    primary = rtx.primary_rays(desc)
    hits = ctx.trace(primary, mode=0)
    bounce, _ = scene.secondary_rays(primary, hits, kind=0, seed=99)
    batch = np.concatenate([primary, bounce])
    n = len(batch)
    log(f"batch: {len(primary)} primary ({int(hits['hit'].sum())} hits) + {len(bounce)} diffuse = {n} rays")
    rays_d = ctx.to_device(batch)
    hits_d = ctx.buffer(n * rtx.HIT_DTYPE.itemsize)

    results = {}
    for name in [m for m in args.modes.split(",") if m in MODES]:
        order, prefetch = MODES[name]
        kw = dict(mode=order, prefetch=prefetch, threshold=args.threshold, stage_bytes=args.stage if prefetch else 0)
        for _ in range(args.warmup):
            st = ctx.trace_device(rays_d, n, hits_d, **kw)
            if st != 0:
                raise rtx.RtxError(st, f"rtx_trace {name}")
        kms = []
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = ctx.trace_device(rays_d, n, hits_d, **kw)
            if st != 0:
                raise rtx.RtxError(st, f"rtx_trace {name}")
            kms.append(ctx.last_kernel_ms())
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        h = hits_d.download(rtx.HIT_DTYPE, n)
        fetch = float(h["fetch_bytes"].astype(np.float64).sum())
        kernel_ms = float(np.mean(kms))
        results[name] = {
            "mrays_s": n / wall / 1e6,
            "ms_per_step": wall * 1e3,
            "kernel_ms": kernel_ms,
            "node_fetch_gbs": fetch / (kernel_ms * 1e-3) / 1e9,
            "fetch_bytes": fetch,
            "nodes_per_ray": float(h["n_nodes"].mean()),
            "hits": int(h["hit"].sum()),
            "prefetch": ctx.prefetch_stats(),
        }
        log(f"{name:>18}: {results[name]['mrays_s']:8.1f} Mrays/s  kernel {kernel_ms:.3f} ms  "
            f"node fetch {results[name]['node_fetch_gbs']:.0f} GB/s")

    head = results[args.mode]
    out = {
        "metric": METRIC,
        "value": round(head["mrays_s"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"1,024 instances x 1,024-tri displaced spheres ({scene.n_triangles - 0} tris), "
                        f"{args.width}x{args.width} primary + 1 cosine diffuse bounce per hit ({n} rays/step)",
            "mode": args.mode,
            "treelet_budget_bytes": args.budget,
            "n_treelets": summ["n_treelets"],
            "arena_mib": round(scene.arena.nbytes / 2**20, 2),
            "device": info["name"],
            "n_cus": info["n_cus"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["node_fetch_gbs"], 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["node_fetch_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "trace_kernel",
            "kernel_ms": round(head["kernel_ms"], 4),
            "bytes_per_launch": head["fetch_bytes"],
        },
        "modes": {k: {"mrays_s": round(v["mrays_s"], 3), "kernel_ms": round(v["kernel_ms"], 4),
                      "node_fetch_gbs": round(v["node_fetch_gbs"], 2), "nodes_per_ray": round(v["nodes_per_ray"], 3),
                      "prefetch": v["prefetch"]} for k, v in results.items()},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(scene, batch, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
