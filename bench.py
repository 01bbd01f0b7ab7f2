#!/usr/bin/env python3
"""Ray-traversal benchmark on one MI355X.

Workload (default, 1,048,576 triangles, diffuse): 1,024 instances of a
1,024-triangle displaced sphere (32x16 quads, radial displacement U[-0.05,
0.05], a distinct seed and BLAS per instance) at uniform positions in a 100^3
box with uniform scale in [0.5, 2]; 1024x1024 pinhole primary rays plus one
cosine-weighted diffuse bounce ray per primary hit, all FP32, all resident in
HBM.  A step traces the whole batch (primary + bounce rays) once.

--workload large: 20,480 instances of the same sphere (20,971,520 triangles,
1.7 GiB arena) in a 271^3 box, 3840x2160 pixels x 4 samples (a stratified 2x2
grid per pixel) and 3 cosine-weighted diffuse bounces per path without Russian
roulette; a step traces the four segments (camera rays, then each bounce) once,
as four launches or as one persistent batch launch.

--workload c3: 64 scenes of instanced displaced spheres with log-uniform
triangle counts in [1k, 2M], each with a log-uniform batch of [4k, 1M] rays
(50 % primary, 25 % shadow with TerminateOnFirstHit, 25 % diffuse); a step is
one replay of the hipGraph that holds the persistent batch launch over all 64.

--workload c1: the small scene the reference's CPU path runs, a displaced UV
sphere of 16,384 triangles over a 2-triangle ground quad, 128x128 primary rays.

`value` is the headline mode's rays per second over the timed steps (wall
clock around synchronous calls); every mode is measured and reported.

  python bench.py [--steps K] [--warmup W] [--workload diffuse|large|c3|c1] [--mode NAME] [--modes a,b,...]
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

# name -> traversal order, prefetch policy (0 none, 1 always, 2 popularity, 3 partial, 4 learned, 5 latter part),
# treelet budget, LDS staging bytes per wave, LDS slot bytes (0: one treelet resident),
# launch ("trace": rtx_trace per segment, "batch": one graph-captured persistent launch over
# every segment, longest segments first), records (hits + node-access records, as traceRay
# returns them) and claim (segments of camera then secondary rays: claim the secondary
# rays first)
MODES = {
    "plain": dict(order=0, prefetch=0, budget=8192, stage=0, slot=0, launch="trace", records=False, claim=True),
    "plain+input_order": dict(order=0, prefetch=0, budget=8192, stage=0, slot=0, launch="trace", records=False,
                              claim=False),
    "treelet": dict(order=1, prefetch=0, budget=8192, stage=0, slot=0, launch="trace", records=False, claim=False),
    "plain+records": dict(order=0, prefetch=0, budget=8192, stage=0, slot=0, launch="trace", records=True,
                          claim=True),
    "treelet+records": dict(order=1, prefetch=0, budget=8192, stage=0, slot=0, launch="trace", records=True,
                            claim=False),
    "plain+stage": dict(order=0, prefetch=1, budget=8192, stage=8192, slot=0, launch="trace", records=False,
                        claim=False),
    "treelet+stage": dict(order=1, prefetch=1, budget=8192, stage=8192, slot=0, launch="trace", records=False,
                          claim=False),
    "treelet+stage_pop": dict(order=1, prefetch=2, budget=8192, stage=8192, slot=0, launch="trace", records=False,
                              claim=False),
    "treelet+stage_partial": dict(order=1, prefetch=3, budget=8192, stage=8192, slot=0, launch="trace",
                                  records=False, claim=False),
    "treelet+stage_latter": dict(order=1, prefetch=5, budget=8192, stage=8192, slot=0, launch="trace",
                                 records=False, claim=False),
    "plain+slots": dict(order=0, prefetch=1, budget=1024, stage=8192, slot=1024, launch="trace", records=False,
                        claim=False),
    "plain+learned": dict(order=0, prefetch=4, budget=1024, stage=8192, slot=1024, launch="trace", records=False,
                          claim=False),
    "plain+batch": dict(order=0, prefetch=0, budget=8192, stage=0, slot=0, launch="batch", records=False,
                        claim=False),
    "treelet+batch": dict(order=1, prefetch=0, budget=8192, stage=0, slot=0, launch="batch", records=False,
                          claim=False),
}
DEFAULTS = {
    "diffuse": ("plain", "plain,plain+input_order,treelet,plain+records,treelet+records,plain+stage,treelet+stage,"
                         "plain+slots,plain+learned"),
    "large": ("plain+batch", "plain+batch,plain,treelet,treelet+batch,plain+records,plain+slots,plain+learned"),
    "c3": ("plain+batch", "plain+batch,treelet+batch"),
    "c1": ("plain", "plain,treelet,plain+records,treelet+records"),
}


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


def host_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    bounded by OMP_NUM_THREADS where the host sets it (the CPU share a GPU's
    process gets on a shared host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(samples, what, target_s=10.0):
    """The oracle (C restatement of traceRay, plain order, hits only) on host cores:
    passes over `samples` (a list of (arena, tlas, rays)) until target_s seconds of work
    have run on all threads, then target_s / 3 seconds on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    def run(threads, budget, items):
        rays = passes = 0
        t0 = time.perf_counter()
        while True:
            for arena, tlas, r in items:
                oracle.trace_parallel(arena, tlas, r, mode=0, threads=threads)
                rays += len(r)
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= budget:
                return rays / dt / 1e6, passes, dt

    threads = host_threads()
    v, passes, dt = run(threads, target_s, samples)
    n = sum(len(r) for _, _, r in samples)
    single = [(a, t, r[:: max(1, threads)]) for a, t, r in samples]
    v1, p1, dt1 = run(1, target_s / 3, single)
    return {"value": round(v, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {n} rays x {passes} passes, plain order, {dt:.1f} s on {threads} host threads",
            "single_thread": {"value": round(v1, 4), "cores": 1,
                              "sample": f"every {max(1, threads)}th of those rays x {p1} passes, {dt1:.1f} s"}}


def training_rays_diffuse(rtx, ctx, scene):
    """Held-out training rays for the learned prefetcher: primary rays of the same
    camera at 768x768 pixels (no pixel centre of that grid lies on the 1024x1024 grid
    of the timed batch) and diffuse bounces of their hits with seeds other than the
    timed batch's."""
    prim = rtx.primary_rays(c2_desc(rtx, 768, 768))
    hits = ctx.trace(prim, mode=0)
    train = [prim[::2]]
    for seed in (1, 2, 3, 4):
        b, _ = scene.secondary_rays(prim[::2], hits[::2], kind=0, seed=seed)
        train.append(b)
    return np.concatenate(train)


def train_learned(rtx, ctx, scene, train, args, budget):
    """Train the next-treelet network and install its table for `scene`.  The
    1M-triangle workload trains one network on traces of both benchmark scenes
    (the 16K-triangle sphere over a quad, 80 % of its pixels, and the held-out
    training rays of this scene); other workloads train on their own scene."""
    from rtx import learned

    t0 = time.perf_counter()
    if args.workload == "diffuse":
        d1 = c1_desc(rtx)
        s1 = rtx.workload_scene(d1)
        p1 = rtx.primary_rays(d1)
        p1 = p1[np.arange(len(p1)) % 5 != 0]
        src = [learned.scene_transitions(ctx, s1.arena, s1.tlas_offset, budget, p1),
               learned.scene_transitions(ctx, scene.arena, scene.tlas_offset, budget, train)]  # leaves this scene set
        lp = learned.JointPrefetcher(ctx, src, n_classes=16384)
        losses = lp.fit(epochs=args.learn_epochs, max_samples=args.learn_samples)
        fit_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        tab = lp.table(1)
        ctx.set_prefetch_table(tab)
        what = "traces of the 16K-triangle scene and of this scene"
    else:
        lp = learned.LearnedPrefetcher(ctx, scene.arena, scene.tlas_offset, n_classes=16384)
        losses = lp.fit(train, epochs=args.learn_epochs, max_samples=args.learn_samples)
        fit_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        tab = lp.install()
        what = "traces of this scene"
    info = {"trained_on": what, "train_rays": len(train), "samples": lp.n_train, "classes": lp.n_classes,
            "epochs": args.learn_epochs, "losses": [round(x, 4) for x in losses], "fit_s": round(fit_s, 2),
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


class Segment:
    """Rays of one scene, resident in device memory with their hit buffer."""

    def __init__(self, rtx, ctx, scene, rays, claim_start=0, scene_id=None):
        self.scene, self.n, self.claim_start, self.scene_id = scene, len(rays), claim_start, scene_id
        self.rays_host = rays
        self.rays = ctx.to_device(rays)
        self.hits = ctx.buffer(len(rays) * rtx.HIT_DTYPE.itemsize)


def setup_single(rtx, ctx, args, info):
    """diffuse / large / c1: one scene, segments of rays, and held-out training rays."""
    if args.workload == "diffuse":
        w = args.width or 1024
        desc = c2_desc(rtx, w, w)
    elif args.workload == "c1":
        w = args.width or 128
        desc = c1_desc(rtx, w, w)
    else:
        w = args.width or 7680
        desc = c5_desc(rtx, w, w * 9 // 16)
    t0 = time.perf_counter()
    builder = rtx.workload_builder(desc)
    mesh_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    scene = builder.finish(ctx=ctx if args.build == "device" else None,
                           quality=rtx.BUILD_SAH_OPTIMAL if args.tree == "sah" else rtx.BUILD_DEFAULT)
    build_s = time.perf_counter() - t0
    log(f"scene: {scene.n_triangles} tris, {scene.n_blas} BLAS, {scene.arena.nbytes / 2**20:.1f} MiB arena, "
        f"meshes in {mesh_s:.2f} s, BVH built ({args.build}) in {build_s:.2f} s")
    ctx.set_scene(scene.arena, scene.tlas_offset)
    primary = rtx.primary_rays(desc)
    hits = ctx.trace(primary, mode=0)
    claim = 0
    if args.workload == "diffuse":
        bounce, _ = scene.secondary_rays(primary, hits, kind=0, seed=99)
        rays = [np.concatenate([primary, bounce])]
        claim = len(primary)  # the bounce rays
        log(f"batch: {len(primary)} primary ({int(hits['hit'].sum())} hits) + {len(bounce)} diffuse = "
            f"{len(rays[0])} rays")
        train_rays = training_rays_diffuse(rtx, ctx, scene)
        workload = (f"1,024 instances x 1,024-tri displaced spheres ({scene.n_triangles} tris), {w}x{w} primary + "
                    f"1 cosine diffuse bounce per hit ({len(rays[0])} rays/step)")
    elif args.workload == "c1":
        rays = [primary]
        train_rays = primary
        log(f"batch: {len(primary)} primary rays ({int(hits['hit'].sum())} hits)")
        workload = (f"displaced UV sphere of 16,384 tris over a 2-tri ground quad ({scene.n_triangles} tris), "
                    f"{w}x{w} primary rays ({len(primary)} rays/step)")
    else:
        # camera rays, then three diffuse bounces from each segment's hits; the
        # training traces come from an earlier frame (other bounce seeds)
        rays = [primary]
        train = [primary[::16]]
        seg, sh = primary, hits
        tseg, tsh = primary[::16], hits[::16]
        for b in range(3):
            nxt, _ = scene.secondary_rays(seg, sh, kind=0, seed=1000 + b)
            rays.append(nxt)
            tn, _ = scene.secondary_rays(tseg, tsh, kind=0, seed=2000 + b)
            train.append(tn)
            if b < 2:
                sh = ctx.trace(nxt, mode=0)
                tsh = ctx.trace(tn, mode=0)
            seg, tseg = nxt, tn
        train_rays = np.concatenate(train)
        log("segments: " + ", ".join(str(len(x)) for x in rays) + " rays")
        workload = (f"20,480 instances x 1,024-tri displaced spheres ({scene.n_triangles} tris), {w // 2}x"
                    f"{w * 9 // 32} pixels x 4 spp (2x2 stratified) x (camera + 3 diffuse bounces, no Russian "
                    f"roulette) ({sum(len(x) for x in rays)} rays/step)")
    del hits
    segs = [Segment(rtx, ctx, scene, r, claim_start=claim if k == 0 else 0) for k, r in enumerate(rays)]
    meta = {"arena_mib": round(scene.arena.nbytes / 2**20, 2), "bvh_build": args.build, "tree": args.tree,
            "bvh_build_s": round(build_s, 3), "mesh_gen_s": round(mesh_s, 3)}
    return workload, segs, train_rays, scene, meta


def setup_c3(rtx, ctx, args):
    from rtx import workloads

    t0 = time.perf_counter()
    items = workloads.multi_scene_batch(ctx, n_scenes=64)
    segs = [Segment(rtx, ctx, scene, rays, scene_id=sid) for sid, scene, rays in items]
    tris = sum(s.scene.n_triangles for s in segs)
    n = sum(s.n for s in segs)
    log(f"c3: 64 scenes, {tris} tris, {n} rays, set up in {time.perf_counter() - t0:.1f} s")
    workload = (f"64 scenes of instanced displaced spheres, log-uniform 1k-2M tris ({tris} tris in all), "
                f"log-uniform 4k-1M rays each (50% primary, 25% shadow, 25% diffuse; {n} rays/step), "
                f"one hipGraph replay of the persistent batch launch")
    meta = {"arena_mib": round(sum(s.scene.arena.nbytes for s in segs) / 2**20, 2), "scenes": len(segs)}
    return workload, segs, meta


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
    ap.add_argument("--mode", default=None, choices=sorted(MODES), help="headline mode")
    ap.add_argument("--modes", default=None, help="modes to measure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--width", type=int, default=0, help="override the image width (height follows the aspect)")
    ap.add_argument("--workload", default="diffuse", choices=["diffuse", "large", "c3", "c1"])
    ap.add_argument("--build", default="host", choices=["host", "device"],
                    help="BLAS builder: host SAH (default) or the device radix-tree builder")
    ap.add_argument("--tree", default="default", choices=["default", "sah"],
                    help="host tree quality of the diffuse, large and c1 scenes: default, or sah "
                         "(three-axis SAH, SAH-optimal 6-wide collapse)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="copy MIB MiB device to device once before tracing: a dispatch of known size "
                         "that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this library traces on one "
              "device (rays are independent, so a multi-GPU run would split the batch with the arena "
              "replicated)", file=sys.stderr)
        return 2
    head_default, modes_default = DEFAULTS[args.workload]
    head_mode = args.mode or head_default
    modes = [m for m in (args.modes or modes_default).split(",") if m in MODES]
    if head_mode not in modes:
        modes.insert(0, head_mode)
    if args.workload == "c3":
        modes = [m for m in modes if MODES[m]["launch"] == "batch"]
        if MODES[head_mode]["launch"] != "batch":
            raise SystemExit("--workload c3 runs batch modes only")

    import rtx

    ctx = rtx.Context(0)
    info = ctx.device_info()
    train_rays = scene = None
    if args.workload == "c3":
        workload, segs, meta = setup_c3(rtx, ctx, args)
    else:
        workload, segs, train_rays, scene, meta = setup_single(rtx, ctx, args, info)
        if any(MODES[m]["launch"] == "batch" for m in modes):
            sid = ctx.add_scene(scene.arena, scene.tlas_offset)
            for s in segs:
                s.scene_id = sid
    if args.copy_probe:
        a, b = ctx.buffer(args.copy_probe << 20), ctx.buffer(args.copy_probe << 20)
        b.copy_from(a, args.copy_probe << 20)
        ctx.synchronize()
        log(f"copy probe: {args.copy_probe} MiB device to device")
        a.free()
        b.free()
    n = sum(s.n for s in segs)

    partitions = {}
    cur_budget = None
    batch_budget = None
    learned_info = None
    results = {}
    for name in modes:
        cfg = dict(MODES[name])
        if args.budget:
            cfg["budget"] = args.budget
        if args.stage and cfg["stage"]:
            cfg["stage"] = args.stage
        if cfg["launch"] == "trace" and (cfg["order"] == 1 or cfg["prefetch"]) and cur_budget != cfg["budget"]:
            t0 = time.perf_counter()
            partitions[cfg["budget"]] = ctx.build_treelets(cfg["budget"])["n_treelets"]
            cur_budget = cfg["budget"]
            log(f"treelets: {partitions[cur_budget]} at {cur_budget} B budget, built in "
                f"{time.perf_counter() - t0:.2f} s")
        if cfg["launch"] == "batch" and cfg["order"] == 1 and batch_budget != cfg["budget"]:
            t0 = time.perf_counter()
            for sid in sorted({s.scene_id for s in segs}):
                ctx.build_scene_treelets(sid, cfg["budget"])
            batch_budget = cfg["budget"]
            log(f"batch scenes: treelets at {batch_budget} B built in {time.perf_counter() - t0:.2f} s")
        if cfg["prefetch"] == 4:
            learned_info = train_learned(rtx, ctx, scene, train_rays, args, cfg["budget"])
        kw = dict(mode=cfg["order"], prefetch=cfg["prefetch"], threshold=args.threshold, stage_bytes=cfg["stage"],
                  slot_bytes=cfg["slot"])
        graph = None
        rec_bufs = {}
        if cfg["launch"] == "batch":
            # the persistent launch claims jobs in order: longest segments first
            order = sorted(range(len(segs)), key=lambda k: -segs[k].n)
            jobs = [(segs[k].scene_id, segs[k].rays, segs[k].n, segs[k].hits) for k in order]
            graph = ctx.batch_graph(jobs, mode=cfg["order"])
        elif cfg["records"]:
            for k, s in enumerate(segs):
                ob = ctx.buffer((s.n + 1) * 8)
                cap = s.n * 48
                rb = ctx.buffer(cap * rtx.RECORD_DTYPE.itemsize)
                st = ctx.trace_device(s.rays, s.n, s.hits, records_buf=rb, record_capacity=cap, offsets_buf=ob,
                                      claim_start=s.claim_start if cfg["claim"] else 0, **kw)
                if st == 5:
                    cap = int(ob.download(np.uint64, s.n + 1)[-1])
                    rb.free()
                    rb = ctx.buffer(max(cap, 1) * rtx.RECORD_DTYPE.itemsize)
                elif st != 0:
                    raise rtx.RtxError(st, f"rtx_trace {name}")
                rec_bufs[k] = (rb, cap, ob)

        def step():
            if graph is not None:
                st, ms = graph.launch(1)
                if st != 0:
                    raise rtx.RtxError(st, f"batch graph {name}")
                return ms
            ms = 0.0
            for k, s in enumerate(segs):
                extra = {}
                if cfg["records"]:
                    rb, cap, ob = rec_bufs[k]
                    extra = dict(records_buf=rb, record_capacity=cap, offsets_buf=ob)
                st = ctx.trace_device(s.rays, s.n, s.hits, claim_start=s.claim_start if cfg["claim"] else 0,
                                      **extra, **kw)
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
        fetch = nodes = recs = 0.0
        pst = {}
        for s in segs:
            if cfg["launch"] == "trace" and cfg["prefetch"]:
                ctx.trace_device(s.rays, s.n, s.hits, **kw)
                for key, v in ctx.prefetch_stats().items():
                    pst[key] = pst.get(key, 0) + v
            h = s.hits.download(rtx.HIT_DTYPE, s.n)
            fetch += float(h["fetch_bytes"].astype(np.float64).sum())
            nodes += float(h["n_nodes"].astype(np.float64).sum())
            recs += float(h["n_records"].astype(np.float64).sum())
        if graph is not None:
            graph.close()
        for rb, _, ob in rec_bufs.values():
            rb.free()
            ob.free()
        kernel_ms = float(np.mean(kms))
        r = {"mrays_s": n / wall / 1e6, "ms_per_step": wall * 1e3, "kernel_ms": kernel_ms,
             "node_fetch_gbs": fetch / (kernel_ms * 1e-3) / 1e9, "fetch_bytes": fetch, "nodes_per_ray": nodes / n,
             "config": cfg}
        if cfg["records"]:
            r["records_per_ray"] = recs / n
            r["record_bytes_per_step"] = recs * rtx.RECORD_DTYPE.itemsize
        if cfg["prefetch"]:
            reads = max(pst["node_reads"], 1)
            r["prefetch"] = pst
            r["lds_hit_rate"] = pst["lds_hits"] / reads
            r["wasted_bytes_per_read"] = pst["wasted_bytes"] / reads
        results[name] = r
        log(f"{name:>22}: {r['mrays_s']:8.1f} Mrays/s  {r['ms_per_step']:.3f} ms/step  kernel {kernel_ms:.3f} ms  "
            f"node fetch {r['node_fetch_gbs']:.0f} GB/s"
            + (f"  {r['records_per_ray']:.1f} records/ray" if cfg["records"] else "")
            + (f"  LDS hit rate {r['lds_hit_rate']:.3f}  wasted {r['wasted_bytes_per_read']:.2f} B/read"
               if cfg["prefetch"] else ""))

    head = results[head_mode]
    pcie_mrays = None
    if args.workload != "c3":
        # host-buffer rate of plain order: upload, trace, download (PCIe-inclusive; not `value`)
        t0 = time.perf_counter()
        for s in segs:
            ctx.trace(s.rays_host, mode=0)
        pcie_mrays = n / (time.perf_counter() - t0) / 1e6
        log(f"host buffers in and out (plain): {pcie_mrays:.1f} Mrays/s")

    single_kernel = "trace_kernel"
    if MODES[head_mode]["launch"] == "trace" and not MODES[head_mode]["stage"] and not MODES[head_mode]["records"]:
        lim = (128 if MODES[head_mode]["order"] == 0 else 256) * info["n_cus"]
        if max(s.n for s in segs) <= lim:
            single_kernel = "trace_wide"
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
        "config": {"workload": workload, "mode": head_mode, "partitions": {str(k): v for k, v in partitions.items()},
                   **meta, "device": info["name"], "n_cus": info["n_cus"]},
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["node_fetch_gbs"], 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["node_fetch_gbs"] / HBM_PEAK_GBS, 4),
            # the committed counters were taken on the default host-built trees
            "traffic": (load_traffic(args.workload, head_mode)
                        if args.build == "host" and args.tree == "default" else None),
            "kernel": single_kernel,
            "kernel_ms": round(head["kernel_ms"], 4),
            "bytes_per_launch": head["fetch_bytes"] / (len(segs) if MODES[head_mode]["launch"] == "trace" else 1),
        },
        "modes": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()
                      if kk not in ("fetch_bytes",)} for k, v in results.items()},
        "learned": learned_info,
        "pcie_inclusive_mrays_s": round(pcie_mrays, 2) if pcie_mrays else None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        if args.workload == "c3":
            samples = [(s.scene.arena, s.scene.tlas_offset, s.rays_host[:: max(1, s.n // 4096)]) for s in segs]
            what = "about 4,096 evenly strided rays of each of the 64 scenes"
        elif args.workload == "large":
            samples = [(scene.arena, scene.tlas_offset,
                        np.concatenate([s.rays_host[:: max(1, s.n // 250000)] for s in segs]))]
            what = "about 250,000 evenly strided rays of each segment"
        else:
            samples = [(scene.arena, scene.tlas_offset, np.concatenate([s.rays_host for s in segs]))]
            what = "the whole batch"
        out["cpu_baseline"] = cpu_baseline(samples, what, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
