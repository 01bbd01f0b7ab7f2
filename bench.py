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

--workload large: 20,480 instances of the same sphere (20,971,520 triangles,
1.7 GiB arena) in a 271^3 box, 3840x2160 pixels x 4 samples (a stratified 2x2
grid per pixel) and 3 cosine-weighted diffuse bounces per path without Russian
roulette; a step traces the four segments (camera rays, then each bounce) once.

--workload c1: the small scene the reference's CPU path runs, a displaced UV sphere
of 16,384 triangles over a 2-triangle ground quad, 128x128 primary rays.

  python bench.py [--steps K] [--warmup W] [--budget BYTES] [--mode NAME] [--workload diffuse|large|c1]
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


def c5_desc(rtx, width=7680, height=4320):
    n_inst = 20480
    return rtx.workload_desc(rtx.WORKLOAD_INSTANCED_SPHERES, n_instances=n_inst, slices=32, stacks=16, shared_blas=0,
                             width=width, height=height, seed=2025, box_size=100.0 * (n_inst / 1024) ** (1 / 3),
                             scale_min=0.5, scale_max=2.0, displacement=0.05)


def c1_desc(rtx, width=128, height=128):
    return rtx.workload_desc(rtx.WORKLOAD_SPHERE_AND_QUAD, slices=128, stacks=64, width=width, height=height, seed=7,
                             displacement=0.05)


def c2_desc(rtx, width=1024, height=1024):
    return rtx.workload_desc(rtx.WORKLOAD_INSTANCED_SPHERES, n_instances=1024, slices=32, stacks=16, shared_blas=0,
                             width=width, height=height, seed=2023, box_size=100.0, scale_min=0.5, scale_max=2.0,
                             displacement=0.05)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, sample, what, target_s=10.0):
    """The oracle (C restatement of traceRay) on host cores, plain order: passes over
    `sample` until target_s seconds of work have run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    rays = done = 0
    t0 = time.perf_counter()
    while True:
        oracle.trace_parallel(scene.arena, scene.tlas_offset, sample, mode=0, threads=threads)
        rays += len(sample)
        done += 1
        dt = time.perf_counter() - t0
        if dt >= target_s:
            break
    return {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {len(sample)} rays x {done} passes, plain order, {dt:.1f} s on {threads} host threads"}


def training_rays_diffuse(scene, primary, hits):
    """Rays other than the timed batch: a fifth of the pixels held out, diffuse
    bounces of other seeds."""
    idx = np.arange(len(primary))
    keep = idx % 5 != 0
    tp, th = primary[keep], hits[keep]
    train = [tp[::2]]
    for seed in (1, 2, 3, 4):
        b, _ = scene.secondary_rays(tp[::4], th[::4], kind=0, seed=seed)
        train.append(b)
    return np.concatenate(train)


def train_learned(ctx, scene, train, args):
    """Train the next-treelet network on traces of `train` and install its table."""
    from rtx import learned

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


def load_traffic(workload, mode):
    """HBM bytes per launch of the traversal kernel from the committed counter
    profile (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel), if present."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
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
    ap.add_argument("--width", type=int, default=0, help="override the image width (height follows the aspect)")
    ap.add_argument("--workload", default="diffuse", choices=["diffuse", "large", "c1"])
    ap.add_argument("--build", default="host", choices=["host", "device"],
                    help="BLAS builder: host SAH (default) or the device radix-tree builder")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="copy MIB MiB device to device once before tracing: a dispatch of known size "
                         "that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this library traces on one "
              "device (rays are independent, so a multi-GPU run would split the batch with the arena "
              "replicated)", file=sys.stderr)
        return 2

    import rtx

    if args.workload == "diffuse":
        w = args.width or 1024
        desc = c2_desc(rtx, w, w)
    elif args.workload == "c1":
        w = args.width or 128
        desc = c1_desc(rtx, w, w)
    else:
        w = args.width or 7680
        desc = c5_desc(rtx, w, w * 9 // 16)
    ctx = rtx.Context(0)
    info = ctx.device_info()
    t0 = time.perf_counter()
    builder = rtx.workload_builder(desc)
    mesh_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    scene = builder.finish(ctx=ctx if args.build == "device" else None)
    build_s = time.perf_counter() - t0
    log(f"scene: {scene.n_triangles} tris, {scene.n_blas} BLAS, {scene.arena.nbytes / 2**20:.1f} MiB arena, "
        f"meshes in {mesh_s:.2f} s, BVH built ({args.build}) in {build_s:.2f} s")
    ctx.set_scene(scene.arena, scene.tlas_offset)
    if args.copy_probe:
        a, b = ctx.buffer(args.copy_probe << 20), ctx.buffer(args.copy_probe << 20)
        b.copy_from(a, args.copy_probe << 20)
        ctx.synchronize()
        log(f"copy probe: {args.copy_probe} MiB device to device")
        a.free()
        b.free()

    t0 = time.perf_counter()
    primary = rtx.primary_rays(desc)
    hits = ctx.trace(primary, mode=0)
    if args.workload == "diffuse":
        bounce, _ = scene.secondary_rays(primary, hits, kind=0, seed=99)
        segments = [np.concatenate([primary, bounce])]
        log(f"batch: {len(primary)} primary ({int(hits['hit'].sum())} hits) + {len(bounce)} diffuse = "
            f"{len(segments[0])} rays")
        train_rays = training_rays_diffuse(scene, primary, hits)
        workload = (f"1,024 instances x 1,024-tri displaced spheres ({scene.n_triangles} tris), {w}x{w} primary + "
                    f"1 cosine diffuse bounce per hit ({len(segments[0])} rays/step)")
    elif args.workload == "c1":
        segments = [primary]
        train_rays = primary
        log(f"batch: {len(primary)} primary rays ({int(hits['hit'].sum())} hits)")
        workload = (f"displaced UV sphere of 16,384 tris over a 2-tri ground quad ({scene.n_triangles} tris), "
                    f"{w}x{w} primary rays ({len(primary)} rays/step)")
    else:
        # camera rays, then three diffuse bounces from each segment's hits; the
        # training traces come from an earlier frame (other bounce seeds)
        segments = [primary]
        train = [primary[::16]]
        seg, sh = primary, hits
        tseg, tsh = primary[::16], hits[::16]
        for b in range(3):
            nxt, _ = scene.secondary_rays(seg, sh, kind=0, seed=1000 + b)
            segments.append(nxt)
            tn, _ = scene.secondary_rays(tseg, tsh, kind=0, seed=2000 + b)
            train.append(tn)
            if b < 2:
                sh = ctx.trace(nxt, mode=0)
                tsh = ctx.trace(tn, mode=0)
            seg, tseg = nxt, tn
        train_rays = np.concatenate(train)
        log("segments: " + ", ".join(str(len(x)) for x in segments) + f" rays; set up in "
            f"{time.perf_counter() - t0:.1f} s")
        workload = (f"20,480 instances x 1,024-tri displaced spheres ({scene.n_triangles} tris), {w // 2}x"
                    f"{w * 9 // 32} pixels x 4 spp (2x2 stratified) x (camera + 3 diffuse bounces, no Russian "
                    f"roulette) ({sum(len(x) for x in segments)} rays/step)")
    del hits
    n = sum(len(x) for x in segments)
    dev = [(ctx.to_device(x), len(x), ctx.buffer(len(x) * rtx.HIT_DTYPE.itemsize)) for x in segments]
    batch = np.concatenate(segments) if args.workload != "large" else None

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
            learned_info = train_learned(ctx, scene, train_rays, args)
        kw = dict(mode=cfg["order"], prefetch=cfg["prefetch"], threshold=args.threshold, stage_bytes=cfg["stage"],
                  slot_bytes=cfg["slot"])

        def step():
            ms = 0.0
            for rays_d, k, hits_d in dev:
                st = ctx.trace_device(rays_d, k, hits_d, **kw)
                if st != 0:
                    raise rtx.RtxError(st, f"rtx_trace {name}")
                ms += ctx.last_kernel_ms()
            return ms

        for _ in range(args.warmup):
            step()
        kms = []
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            kms.append(step())
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        fetch = 0.0
        nodes = 0.0
        pst = {}
        for rays_d, k, hits_d in dev:
            ctx.trace_device(rays_d, k, hits_d, **kw)
            for key, v in ctx.prefetch_stats().items():
                pst[key] = pst.get(key, 0) + v
            h = hits_d.download(rtx.HIT_DTYPE, k)
            fetch += float(h["fetch_bytes"].astype(np.float64).sum())
            nodes += float(h["n_nodes"].astype(np.float64).sum())
        kernel_ms = float(np.mean(kms))
        results[name] = {
            "mrays_s": n / wall / 1e6,
            "ms_per_step": wall * 1e3,
            "kernel_ms": kernel_ms,
            "node_fetch_gbs": fetch / (kernel_ms * 1e-3) / 1e9,
            "fetch_bytes": fetch,
            "nodes_per_ray": nodes / n,
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
    # host-buffer rate of the headline mode: upload, trace, download (PCIe-inclusive; not `value`)
    hcfg = MODES[args.mode]
    t0 = time.perf_counter()
    for seg in segments:
        ctx.trace(seg, mode=hcfg["order"], prefetch=hcfg["prefetch"], stage_bytes=hcfg["stage"],
                  slot_bytes=hcfg["slot"])
    pcie_mrays = n / (time.perf_counter() - t0) / 1e6
    log(f"host buffers in and out ({args.mode}): {pcie_mrays:.1f} Mrays/s")

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
            "workload": workload,
            "mode": args.mode,
            "partitions": {str(k): v for k, v in partitions.items()},
            "arena_mib": round(scene.arena.nbytes / 2**20, 2),
            "bvh_build": args.build,
            "bvh_build_s": round(build_s, 3),
            "mesh_gen_s": round(mesh_s, 3),
            "device": info["name"],
            "n_cus": info["n_cus"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["node_fetch_gbs"], 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["node_fetch_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(args.workload, args.mode),
            # rtx_trace's default kernel choice: eight lanes per ray without staging for
            # batches of at most 128 (plain) or 256 (treelet) rays per CU; each segment
            # is one launch
            "kernel": ("trace_wide" if not MODES[args.mode]["stage"] and max(len(x) for x in segments) <=
                       (128 if MODES[args.mode]["order"] == 0 else 256) * info["n_cus"] else "trace_kernel"),
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
        "pcie_inclusive_mrays_s": round(pcie_mrays, 2),
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        if batch is not None:
            smp, what = batch, ("the whole primary+bounce batch" if args.workload == "diffuse"
                                else "the whole primary batch")
        else:
            smp = np.concatenate([x[:: max(1, len(x) // 250000)] for x in segments])
            what = "about 250,000 evenly strided rays of each segment"
        out["cpu_baseline"] = cpu_baseline(scene, smp, what, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
