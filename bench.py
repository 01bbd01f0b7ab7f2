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

# name -> traversal order, prefetch policy (0 none, 1 always, 2 popularity, 3 partial, 4 learned),
# treelet budget, LDS staging bytes per wave, LDS slot bytes (0: one treelet resident)
MODES = {
    "plain": dict(order=0, prefetch=0, budget=8192, stage=0, slot=0),
    "treelet": dict(order=1, prefetch=0, budget=8192, stage=0, slot=0),
    "plain+stage": dict(order=0, prefetch=1, budget=8192, stage=8192, slot=0),
    "treelet+stage": dict(order=1, prefetch=1, budget=8192, stage=8192, slot=0),
    "treelet+stage_pop": dict(order=1, prefetch=2, budget=8192, stage=8192, slot=0),
    "treelet+stage_partial": dict(order=1, prefetch=3, budget=8192, stage=8192, slot=0),
    "plain+slots": dict(order=0, prefetch=1, budget=1024, stage=8192, slot=1024),
    "plain+learned": dict(order=0, prefetch=4, budget=1024, stage=8192, slot=1024),
}
DEFAULT_MODES = "plain,treelet,plain+stage,treelet+stage,plain+slots,plain+learned"
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


def train_learned(rtx, ctx, scene, primary, hits, args):
    """Train the next-treelet network on traces of rays other than the timed batch
    (a fifth of the pixels held out, diffuse bounces of other seeds) and install its table."""
    from rtx import learned

    idx = np.arange(len(primary))
    keep = idx % 5 != 0
    tp, th = primary[keep], hits[keep]
    train = [tp[::2]]
    for seed in (1, 2, 3, 4):
        b, _ = scene.secondary_rays(tp[::4], th[::4], kind=0, seed=seed)
        train.append(b)
    train = np.concatenate(train)
    t0 = time.perf_counter()
    lp = learned.LearnedPrefetcher(ctx, scene.arena, scene.tlas_offset, n_classes=16384)
    losses = lp.fit(train, epochs=args.learn_epochs, max_samples=args.learn_samples)
    fit_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tab = lp.install()
    info = {"train_rays": len(train), "samples": lp.n_train, "classes": lp.n_classes, "epochs": args.learn_epochs,
            "losses": [round(x, 4) for x in losses], "fit_s": round(fit_s, 2),
            "table_s": round(time.perf_counter() - t0, 2), "table_rows_filled": int((tab[:, 0] != 0).sum())}
    lp.close()
    log(f"learned prefetcher: {info}")
    return info


def load_traffic(mode):
    """HBM bytes per launch of the traversal kernel from the committed counter
    profile (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel), if present."""
    path = os.path.join(ROOT, "profiles", "traffic_c2_diffuse.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["modes"][mode]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--budget", type=int, default=0, help="override every mode's treelet byte budget")
    ap.add_argument("--stage", type=int, default=0, help="override every staging mode's LDS bytes per wave")
    ap.add_argument("--learn-epochs", type=int, default=2)
    ap.add_argument("--learn-samples", type=int, default=3_000_000)
    ap.add_argument("--threshold", type=float, default=0.5, help="popularity share for the popularity policy")
    ap.add_argument("--mode", default=HEADLINE, choices=sorted(MODES))
    ap.add_argument("--modes", default=DEFAULT_MODES, help="modes to measure")
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

    primary = rtx.primary_rays(desc)
    hits = ctx.trace(primary, mode=0)
    bounce, _ = scene.secondary_rays(primary, hits, kind=0, seed=99)
    batch = np.concatenate([primary, bounce])
    n = len(batch)
    log(f"batch: {len(primary)} primary ({int(hits['hit'].sum())} hits) + {len(bounce)} diffuse = {n} rays")
    rays_d = ctx.to_device(batch)
    hits_d = ctx.buffer(n * rtx.HIT_DTYPE.itemsize)

    modes = [m for m in args.modes.split(",") if m in MODES]
    if args.mode not in modes:
        modes.insert(0, args.mode)
    partitions = {}
    cur_budget = None
    learned_info = None
    results = {}
    for name in modes:
        cfg = dict(MODES[name])
        if args.budget:
            cfg["budget"] = args.budget
        if args.stage and cfg["stage"]:
            cfg["stage"] = args.stage
        if (cfg["order"] == 1 or cfg["prefetch"]) and cur_budget != cfg["budget"]:
            t0 = time.perf_counter()
            partitions[cfg["budget"]] = ctx.build_treelets(cfg["budget"])["n_treelets"]
            cur_budget = cfg["budget"]
            log(f"treelets: {partitions[cur_budget]} at {cur_budget} B budget, built in "
                f"{time.perf_counter() - t0:.2f} s")
        if cfg["prefetch"] == 4:
            learned_info = train_learned(rtx, ctx, scene, primary, hits, args)
        kw = dict(mode=cfg["order"], prefetch=cfg["prefetch"], threshold=args.threshold, stage_bytes=cfg["stage"],
                  slot_bytes=cfg["slot"])
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
        pst = ctx.prefetch_stats()
        results[name] = {
            "mrays_s": n / wall / 1e6,
            "ms_per_step": wall * 1e3,
            "kernel_ms": kernel_ms,
            "node_fetch_gbs": fetch / (kernel_ms * 1e-3) / 1e9,
            "fetch_bytes": fetch,
            "nodes_per_ray": float(h["n_nodes"].mean()),
            "hits": int(h["hit"].sum()),
            "config": cfg,
            "prefetch": pst,
        }
        if cfg["prefetch"]:
            reads = max(pst["node_reads"], 1)
            results[name]["lds_hit_rate"] = pst["lds_hits"] / reads
            results[name]["wasted_bytes_per_read"] = pst["wasted_bytes"] / reads
        log(f"{name:>22}: {results[name]['mrays_s']:8.1f} Mrays/s  kernel {kernel_ms:.3f} ms  "
            f"node fetch {results[name]['node_fetch_gbs']:.0f} GB/s"
            + (f"  LDS hit rate {results[name]['lds_hit_rate']:.3f}  wasted {results[name]['wasted_bytes_per_read']:.2f}"
               f" B/read" if cfg["prefetch"] else ""))

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
            "partitions": {str(k): v for k, v in partitions.items()},
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
            "traffic": load_traffic(args.mode),
            "kernel": "trace_kernel",
            "kernel_ms": round(head["kernel_ms"], 4),
            "bytes_per_launch": head["fetch_bytes"],
        },
        "modes": {k: {"mrays_s": round(v["mrays_s"], 3), "kernel_ms": round(v["kernel_ms"], 4),
                      "node_fetch_gbs": round(v["node_fetch_gbs"], 2), "nodes_per_ray": round(v["nodes_per_ray"], 3),
                      "config": v["config"], "prefetch": v["prefetch"],
                      **({"lds_hit_rate": round(v["lds_hit_rate"], 4),
                          "wasted_bytes_per_read": round(v["wasted_bytes_per_read"], 3)} if "lds_hit_rate" in v else {})}
                  for k, v in results.items()},
        "learned": learned_info,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(scene, batch, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
