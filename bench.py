"""Throughput of the gfx950 wavefront path tracer on the procedural
1M-triangle scene at 1080p.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]

A step renders one full frame: 1920x1080, 64 samples per pixel, max depth 8
(Russian roulette from depth 3), 4096 Disney-shaded instances of a 256-triangle
displaced sphere (1,048,576 triangles), 64 sphere mesh lights and a constant
environment. Scene, BVH and path pool are resident in HBM before timing.

Prints one JSON line: value = Mrays/s (extension + shadow rays per second of
wall time over the timed frames), plus samples/s, the roofline of the
dominant kernel (extension-ray traversal) and the CPU baseline (the
reference's lib/bvh traversal on the same rays, on this host's cores).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s & samples/s @1080p; equal-time relMSE vs FP64 GT, NRC vs plain PT"
HBM_PEAK_GBS = 8000.0
WIDTH, HEIGHT, SPP = 1920, 1080, 64
MIN_DEPTH, MAX_DEPTH = 3, 8
STATE_COUNT = 1 << 21


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_ray(steps, isects):
    """32 B ray in + 20 B hit out + 64 B per traversal step (two 32-byte
    binary nodes) + 36 B per triangle test (three FP32 vertices), with
    lib/bvh's SingleRayTraverser statistics on the same rays."""
    return 32.0 + 20.0 + 64.0 * float(np.mean(steps)) + 36.0 * float(np.mean(isects))


def read_traffic(path):
    """HBM bytes per traced ray from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE of the extension-trace kernel), if present."""
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_ray:\s*([0-9.eE+-]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(rays, world, budget_s, threads):
    """lib/bvh traversal of the captured rays on `threads` host threads,
    repeated over the sample until about budget_s seconds of work"""
    from oracle import oracle
    if not oracle.ref_available():
        return None, None
    ref = oracle.RefBvh(world)
    prim, t, u, v, steps, isects = ref.trace(rays, threads=threads)  # also the work statistics
    traced, dt = 0, 0.0
    while dt < budget_s:
        t0 = time.perf_counter()
        ref.trace(rays, threads=threads)
        dt += time.perf_counter() - t0
        traced += len(rays)
    return dict(value=traced / dt / 1e6, unit="Mrays/s", cores=threads, kind="reference",
                sample=f"{len(rays)} extension rays of the first two bounces of this frame, traced "
                       f"{traced // len(rays)} times by lib/bvh (BinnedSAH16 + SingleRayTraverser, closest "
                       f"hit) on {threads} host threads, {dt:.1f} s"), (steps, isects)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--workload", choices=["field", "media", "nrc", "instanced"], default="field",
                    help="field: the 1M-triangle scene (default); media: the Cornell-style room with a "
                         "128^3 heterogeneous medium, max depth 16; nrc: the Cornell-style room at 256^2, "
                         "1 spp per frame for --frames frames, with the radiance cache against plain path "
                         "tracing at equal time; instanced: the 16.8M-triangle textured two-level field, eight "
                         "views rendered together, 1 spp per frame with the radiance cache for --frames frames "
                         "(64 when not given)")
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--nrc-train-depth", type=int, default=3,
                    help="nrc workload: end training paths at this depth with a cache query (self-training; "
                         "0: trace them out)")
    ap.add_argument("--nrc-steps", type=int, default=4, help="nrc workload: optimizer steps per frame")
    ap.add_argument("--nrc-query-depth", type=int, default=2, help="nrc workload: depth of the cache query")
    ap.add_argument("--nrc-train-stride", type=int, default=16,
                    help="nrc workload: one pixel in this many traces a training path")
    ap.add_argument("--no-bvh-compare", action="store_true",
                    help="skip the frame traced through a GPU-built tree (field workload)")
    ap.add_argument("--gt-spp", type=int, default=1024,
                    help="samples per pixel of the FP64 ground truth for relMSE (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the path tracer runs on one GPU only (asked for --gpus {args.gpus}); a frame would "
              f"shard across GPUs by sample index with one film reduction, which is not built", file=sys.stderr)
        sys.exit(2)

    import btrc_amd
    from btrc_amd import scenes, _abi

    if args.workload == "nrc":
        args.frames = 256 if args.frames is None else args.frames
        print(json.dumps(nrc_workload(args)), flush=True)
        return
    if args.workload == "instanced":
        args.frames = 64 if args.frames is None else args.frames
        print(json.dumps(instanced_workload(args)), flush=True)
        return

    media = args.workload == "media"
    max_depth = 16 if media else MAX_DEPTH
    if media:
        log("bench: building the media scene (128^3 fBm grid)")
        scene, cam = scenes.cornell_media()
    else:
        log("bench: building the 1M-triangle scene")
        scene, cam = scenes.sphere_field()
    ctx = btrc_amd.Context(0)
    t0 = time.perf_counter()
    ctx.upload_scene(scene)
    info = ctx.scene_info()
    log(f"bench: scene uploaded ({info['triangles']} tris, {info['nodes']} BVH4 nodes, depth {info['depth']}, "
        f"stack {info['stack']}) in {time.perf_counter() - t0:.1f} s on {ctx.device_name()}")
    ctx.set_camera(cam, WIDTH, HEIGHT)

    p = _abi.RenderParams()
    p.width, p.height, p.spp = WIDTH, HEIGHT, args.spp
    p.min_depth, p.max_depth = MIN_DEPTH, max_depth
    p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
    p.state_count = STATE_COUNT
    p.time_stages = 0   # timed frames replay the captured bounce loop

    for i in range(args.warmup):
        st = ctx.render(p)
        log(f"bench: warmup {i}: {st.seconds:.3f} s")

    totals = dict(seconds=0.0, ext=0, shadow=0, paths=0)
    for i in range(args.steps):
        st = ctx.render(p)
        totals["seconds"] += st.seconds
        totals["ext"] += st.extension_rays
        totals["shadow"] += st.shadow_rays
        totals["paths"] += st.paths
        log(f"bench: step {i}: {st.seconds:.3f} s, {(st.extension_rays + st.shadow_rays) / st.seconds / 1e6:.1f} "
            f"Mrays/s")
    frame = ctx.film(WIDTH, HEIGHT)[0][..., :3].astype(np.float64)

    # one more frame with HIP events around every stage launch (direct
    # launches, not the graph): per-stage times and the traversal launch time
    ps = _abi.RenderParams.from_buffer_copy(p)
    ps.time_stages = 1
    st = ctx.render(ps)
    staged = dict(seconds=st.seconds, ext=st.extension_rays, stage_ms=np.array(st.stage_ms[:7]),
                  launches=np.array(st.stage_launches[:7]))
    log(f"bench: stage-timed frame: {st.seconds:.3f} s, stage ms {[round(x, 1) for x in st.stage_ms]}")
    quality = None
    if args.gt_spp > 0:
        # FP64 ground truth from the independent sampler, so its samples share
        # nothing with the Halton samples of the timed frame
        g = _abi.RenderParams.from_buffer_copy(p)
        g.spp, g.film_fp64, g.time_stages, g.sampler = args.gt_spp, 1, 0, _abi.SAMPLER_INDEPENDENT
        t0 = time.perf_counter()
        ctx.render(g)
        gt = ctx.film(WIDTH, HEIGHT, fp64=True)[1]
        quality = dict(relmse=float(np.mean((frame - gt) ** 2 / (gt ** 2 + 1e-2))), spp=args.spp,
                       gt_spp=args.gt_spp, gt_sampler="independent", gt_film="fp64",
                       gt_seconds=round(time.perf_counter() - t0, 2))
        log(f"bench: relMSE of the {args.spp} spp frame vs {args.gt_spp} spp FP64 ground truth: "
            f"{quality['relmse']:.5f}")

    secs = totals["seconds"]
    mrays = (totals["ext"] + totals["shadow"]) / secs / 1e6
    samples = totals["paths"] / secs
    trace_ms = staged["stage_ms"][1]

    cpu = None
    bytes_per_ray = None
    if not args.no_cpu and media:
        cpu = media_cpu_baseline(scene, cam, p, args.cpu_seconds)
    elif not args.no_cpu:
        # rays of this frame for the work model and the CPU baseline
        log("bench: capturing rays of the first two bounces")
        pc = _abi.RenderParams.from_buffer_copy(p)
        pc.time_stages = 0
        rays = ctx.capture_rays(pc, 2, 1 << 22)
        rays = rays[np.random.default_rng(0).permutation(len(rays))]
        # world-space triangles in flattened order, as the GPU traverses them
        world = flatten_instances(scene)
        threads = min(16, os.cpu_count() or 1)
        cpu, stats = cpu_baseline(rays, world, args.cpu_seconds, threads)
        if stats is not None:
            bytes_per_ray = algorithmic_bytes_per_ray(*stats)
            cpu["steps_per_ray"] = float(np.mean(stats[0]))
            cpu["isects_per_ray"] = float(np.mean(stats[1]))

    roofline = None
    if media:
        # the medium stage dominates here; its roofline is not modelled yet
        pass
    elif bytes_per_ray is not None and trace_ms > 0:
        launches = max(1, int(staged["launches"][1]))
        rays_per_launch = staged["ext"] / launches
        launch_s = trace_ms * 1e-3 / launches
        achieved = bytes_per_ray * rays_per_launch / launch_s / 1e9
        traffic_per_ray = read_traffic(os.path.join(ROOT, "profiles", "trace_extension_traffic.md"))
        roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4),
                        traffic=None if traffic_per_ray is None else round(traffic_per_ray * rays_per_launch),
                        kernel="k_trace_extension",
                        algorithmic_bytes_per_ray=round(bytes_per_ray, 1),
                        algorithmic_bytes_per_launch=round(bytes_per_ray * rays_per_launch),
                        rays_per_launch=round(rays_per_launch), launch_ms=round(launch_s * 1e3, 4),
                        traffic_bytes_per_ray=traffic_per_ray)

    bvh = None if media or args.no_bvh_compare else gpu_bvh_comparison(ctx, scene, ps, trace_ms, staged["seconds"])

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("Cornell-style room scaled x30 holding a 40^3 invisible box of heterogeneous "
                                "medium (128^3 fBm grid, max sigma_t 2, albedo 0.8, HG g 0.6), 1920x1080, "
                                f"{args.spp} spp, max depth {max_depth}") if media else
                               ("procedural 1M-triangle field (4096 instances x 256-tri displaced sphere, "
                                "Disney per instance, 64 sphere lights + constant env), 1920x1080, "
                                f"{args.spp} spp, max depth {MAX_DEPTH}"),
                   "width": WIDTH, "height": HEIGHT, "spp": args.spp, "max_depth": max_depth,
                   "min_depth": MIN_DEPTH, "triangles": info["triangles"], "state_count": STATE_COUNT},
        "samples_per_s": round(samples, 1),
        "extension_rays": int(totals["ext"]),
        "shadow_rays": int(totals["shadow"]),
        "stage_ms_one_frame": {k: round(float(v), 2) for k, v in
                               zip(("generate", "trace", "shade", "shadow", "medium", "cache", "train"),
                                   staged["stage_ms"])},
        "stage_timed_frame_s": round(staged["seconds"], 4),
        "quality": quality,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bvh_build": bvh,
    }
    print(json.dumps(out), flush=True)


def gpu_bvh_comparison(ctx, scene, staged_params, host_trace_ms, host_frame_s):
    """the same scene with its tree built on the GPU (clustering + capped
    leaf collapse): build wall time against the host's binned-SAH build of
    the upload above, and one stage-timed frame traced through it"""
    from btrc_amd import _abi
    host_ms = ctx.scene_build_ms()
    accel = scene.accel
    scene.accel = _abi.ACCEL_FLAT_GPU
    ctx.upload_scene(scene)           # first build loads the builder's kernels
    ctx.upload_scene(scene)
    gpu_ms = ctx.scene_build_ms()
    info = ctx.scene_info()
    st = ctx.render(staged_params)
    # vertices of the instanced sphere re-sent (all 1M world triangles): records rebuilt, tree refitted
    ctx.update_geometry(scene, 0)
    refit_ms = ctx.update_geometry(scene, 0)
    log(f"bench: GPU-built tree: {gpu_ms:.1f} ms (host SAH build {host_ms:.0f} ms), {info['nodes']} BVH4 nodes, "
        f"trace {st.stage_ms[1]:.1f} ms vs {host_trace_ms:.1f} ms; geometry update + refit {refit_ms:.1f} ms")
    scene.accel = accel
    return dict(host_sah_build_ms=round(host_ms, 1), gpu_build_ms=round(gpu_ms, 2), refit_all_triangles_ms=round(refit_ms, 2),
                gpu_bvh4_nodes=info["nodes"],
                gpu_bvh4_depth=info["depth"], trace_ms_host_sah=round(float(host_trace_ms), 2),
                trace_ms_gpu_built=round(float(st.stage_ms[1]), 2), frame_s_host_sah=round(host_frame_s, 4),
                frame_s_gpu_built=round(st.seconds, 4))


NRC_FLOP_PER_QUERY = 2 * (64 * 64 + 4 * 64 * 64 + 64 * 3)   # 41,344
NRC_FLOP_PER_TRAIN_SAMPLE = 3 * NRC_FLOP_PER_QUERY            # forward, input gradients, weight gradients
MFMA_BF16_PEAK_TFLOPS = 2500.0


def nrc_workload(args):
    """Cornell-style room at 256^2, `frames` frames of 1 spp accumulated:
    plain path tracing, then the radiance cache (queries at depth 2, training
    on 1/16 of the pixels), each timed per frame and scored by relMSE against
    an FP64 ground truth of 4096 spp from the independent sampler."""
    import btrc_amd
    from btrc_amd import _abi, scenes
    W = H = 256
    scene, cam = scenes.cornell()
    ctx = btrc_amd.Context(0)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, W, H)

    def params(spp, fp64=0):
        p = _abi.RenderParams()
        p.width, p.height, p.spp = W, H, spp
        p.min_depth, p.max_depth = 3, 8
        p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
        p.state_count = 1 << 17
        p.film_fp64 = fp64
        return p

    g = params(4096, 1)
    g.state_count = 1 << 20
    g.sampler = _abi.SAMPLER_INDEPENDENT
    ctx.render(g)
    gt = ctx.film(W, H, fp64=True)[1]

    def relmse():
        img = ctx.film(W, H)[0][..., :3].astype(np.float64)
        return float(np.mean((img - gt) ** 2 / (gt ** 2 + 1e-2)))

    checkpoints = sorted({c for c in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024) if c <= args.frames} | {args.frames})

    def run(cache, target=None):
        # with a target, the run goes on past the frame budget (up to 4x,
        # checkpoints every 16 frames) until its relMSE reaches the target
        ctx.nrc_configure(enabled=1 if cache else 0, train=1, train_depth=args.nrc_train_depth,
                          steps=args.nrc_steps, train_stride=args.nrc_train_stride,
                          query_depth=args.nrc_query_depth)
        if cache:
            ctx.nrc_reset(1)
        p = params(1)
        p.time_stages = 1
        curve, seconds, rays, stage = [], 0.0, 0, np.zeros(7)
        queries = records = 0
        losses = []
        f = 0
        while f < args.frames or (target is not None and f < 4 * args.frames and curve[-1][2] > target):
            p.sample_offset, p.accumulate = f, 1 if f else 0
            st = ctx.render(p)
            seconds += st.seconds
            rays += st.extension_rays + st.shadow_rays
            stage += np.array(st.stage_ms[:7])
            if cache:
                ns = ctx.nrc_stats()
                queries += ns.queries
                records += ns.records
                losses.append(float(np.mean(ns.loss[:min(4, max(1, ns.steps_total))])))
            f += 1
            if f in checkpoints or (f > args.frames and f % 16 == 0):
                curve.append((f, round(seconds, 4), relmse()))
        return dict(seconds=seconds, rays=rays, stage_ms=stage, curve=curve, queries=queries, records=records,
                    losses=losses, frames=f)

    cpu = None if args.no_cpu else media_cpu_baseline(scene, cam, params(16), args.cpu_seconds)
    log("bench: plain path tracing")
    plain = run(False)
    log("bench: path tracing with the radiance cache")
    cache = run(True, plain["curve"][-1][2])
    ctx.nrc_configure(enabled=0)
    frames = cache["frames"]

    def time_to(curve, target):
        for frames, t, e in curve:
            if e <= target:
                return t
        return None

    target = plain["curve"][-1][2]
    t_cache = time_to(cache["curve"], target)
    # relMSE of each method at the other's total time, from the checkpoints
    def at_time(curve, t):
        best = None
        for frames, tt, e in curve:
            if tt <= t:
                best = e
        return best

    infer_ms = cache["stage_ms"][5]
    train_ms = cache["stage_ms"][6]
    steps = args.nrc_steps * frames

    # the plain tracer's relMSE at the cache's time for the frame budget,
    # interpolated on log-log axes between its checkpoints
    def plain_at(t):
        pts = plain["curve"]
        for (f0, t0, e0), (f1, t1, e1) in zip(pts, pts[1:]):
            if t0 <= t <= t1 and t1 > t0:
                a = (np.log(t) - np.log(t0)) / (np.log(t1) - np.log(t0))
                return float(np.exp(np.log(e0) + a * (np.log(e1) - np.log(e0))))
        return None

    cache_budget = [c for c in cache["curve"] if c[0] <= args.frames][-1]
    infer_tflops = NRC_FLOP_PER_QUERY * cache["queries"] / (infer_ms * 1e-3) / 1e12 if infer_ms > 0 else None
    train_tflops = NRC_FLOP_PER_TRAIN_SAMPLE * 16384 * steps / (train_ms * 1e-3) / 1e12 if train_ms > 0 else None
    return {
        "metric": METRIC,
        "value": round(cache["rays"] / cache["seconds"] / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": frames,
        "warmup": 0,
        "ms_per_step": round(cache["seconds"] / frames * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 paths, bf16 network (fp32 accumulation)",
        "data": "synthetic",
        "config": {"workload": f"Cornell-style room, 256x256, 1 spp per frame x {args.frames} frames, max depth 8, "
                               f"radiance cache queried at depth {args.nrc_query_depth}, trained on 1/{args.nrc_train_stride} of the "
                               f"pixels ({args.nrc_steps} Adam steps of 16384 records per frame, training paths "
                               f"{'traced out' if args.nrc_train_depth == 0 else f'ended at depth {args.nrc_train_depth} with a cache query'}"
                               "); the cache run goes on (up to 4x) until it reaches the "
                               "plain tracer's final relMSE",
                   "width": W, "height": H, "frames": args.frames, "train_depth": args.nrc_train_depth,
                   "steps_per_frame": args.nrc_steps, "train_stride": args.nrc_train_stride,
                   "query_depth": args.nrc_query_depth},
        "equal_time": {
            "ground_truth": "4096 spp FP64, independent sampler",
            "plain": {"seconds": round(plain["seconds"], 4), "relmse_curve": plain["curve"],
                      "mrays_per_s": round(plain["rays"] / plain["seconds"] / 1e6, 2)},
            "cache": {"seconds": round(cache["seconds"], 4), "relmse_curve": cache["curve"],
                      "frames": frames,
                      "queries_per_frame": round(cache["queries"] / frames),
                      "records_per_frame": round(cache["records"] / frames),
                      "loss_first_last": [round(cache["losses"][0], 4), round(cache["losses"][-1], 4)]},
            "target_relmse": target,
            "cache_time_to_target_s": t_cache,
            "time_ratio": None if t_cache is None else round(t_cache / plain["seconds"], 3),
            "cache_relmse_at_plain_time": at_time(cache["curve"], plain["seconds"]),
            "at_cache_time_for_frame_budget": {"seconds": cache_budget[1], "cache_relmse": cache_budget[2],
                                               "plain_relmse": plain_at(cache_budget[1])},
        },
        "roofline": {"bound": "mfma", "kernel": "k_nrc_infer", "unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS,
                     "achieved": None if infer_tflops is None else round(infer_tflops, 4),
                     "frac": None if infer_tflops is None else round(infer_tflops / MFMA_BF16_PEAK_TFLOPS, 7),
                     "traffic": None, "flop_per_query": NRC_FLOP_PER_QUERY,
                     "train_achieved": None if train_tflops is None else round(train_tflops, 4),
                     "infer_ms_per_frame": round(infer_ms / frames, 4),
                     "train_ms_per_frame": round(train_ms / frames, 4)},
        "cpu_baseline": cpu,
    }


def instanced_workload(args):
    """The 16.8M-triangle textured instanced field (65,536 instances of four
    256-triangle spheres, 1024^2 RGBA8 base-colour textures, two-level
    traversal), eight views from 512x512 to 1920x1080 rendered together in
    one wavefront, one sample per pixel per frame for `frames` frames,
    accumulated, with the radiance cache queried at depth 2 and trained on
    1/16 of the pixels every frame. Frames replay the captured bounce loop."""
    import btrc_amd
    from btrc_amd import _abi, scenes
    log("bench: building the 16.8M-triangle instanced field")
    t0 = time.perf_counter()
    scene, views = scenes.instanced_field()
    log(f"bench: scene built in {time.perf_counter() - t0:.1f} s")
    ctx = btrc_amd.Context(0)
    t0 = time.perf_counter()
    ctx.upload_scene(scene)
    info = ctx.scene_info()
    log(f"bench: scene uploaded ({info['triangles']} tris, {info['nodes']} BVH4 nodes, depth {info['depth']}, "
        f"stack {info['stack']}) in {time.perf_counter() - t0:.1f} s on {ctx.device_name()}")
    pixels = sum(w * h for _, w, h in views)

    p = _abi.RenderParams()
    p.spp = 1
    p.min_depth, p.max_depth = MIN_DEPTH, MAX_DEPTH
    p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
    p.state_count = STATE_COUNT

    ctx.nrc_configure(enabled=1, train=1)
    for f in range(args.warmup):
        p.sample_offset, p.accumulate = f, 1 if f else 0
        st = ctx.render_views(p, views)
        log(f"bench: warmup frame {f}: {st.seconds * 1e3:.1f} ms")
    ctx.nrc_reset(1)
    seconds, ext, shadow, paths = 0.0, 0, 0, 0
    wall0 = time.perf_counter()
    for f in range(args.frames):
        p.sample_offset, p.accumulate = f, 1 if f else 0
        st = ctx.render_views(p, views)
        seconds += st.seconds
        ext += st.extension_rays
        shadow += st.shadow_rays
        paths += st.paths
        if f % 16 == 0 or f == args.frames - 1:
            log(f"bench: frame {f}: {st.seconds * 1e3:.1f} ms, "
                f"{(st.extension_rays + st.shadow_rays) / st.seconds / 1e6:.1f} Mrays/s")
    wall = time.perf_counter() - wall0
    ns = ctx.nrc_stats()
    cam0, w0, h0 = views[0]
    img0 = ctx.view_film(0, w0, h0)[..., :3].astype(np.float64)

    # one more frame with per-stage timing (events between stages: no graph)
    p.sample_offset, p.accumulate, p.time_stages = args.frames, 1, 1
    st = ctx.render_views(p, views)
    stage = np.array(st.stage_ms[:7])
    p.time_stages = 0

    # the same frame over one tree of all 16.8M world-space triangles built
    # on the GPU, against the two-level structure above
    bvh = None if args.no_bvh_compare else instanced_flat_comparison(ctx, scene, p, views, stage)

    quality = None
    if args.gt_spp > 0:
        ctx.nrc_configure(enabled=0)
        ctx.set_camera(cam0, w0, h0)
        g = _abi.RenderParams.from_buffer_copy(p)
        g.width, g.height, g.spp, g.film_fp64, g.sampler = w0, h0, args.gt_spp, 1, _abi.SAMPLER_INDEPENDENT
        g.sample_offset, g.accumulate = 0, 0
        t0 = time.perf_counter()
        ctx.render(g)
        gt = ctx.film(w0, h0, fp64=True)[1]
        quality = dict(view=0, relmse=float(np.mean((img0 - gt) ** 2 / (gt ** 2 + 1e-2))), frames=args.frames,
                       gt_spp=args.gt_spp, gt_sampler="independent", gt_film="fp64", cache=True,
                       gt_seconds=round(time.perf_counter() - t0, 2))
        log(f"bench: view 0 relMSE after {args.frames} frames vs {args.gt_spp} spp FP64 GT: {quality['relmse']:.5f}")

    return {
        "metric": METRIC,
        "value": round((ext + shadow) / seconds / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.frames,
        "warmup": args.warmup,
        "ms_per_step": round(seconds / args.frames * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 paths, bf16 network (fp32 accumulation)",
        "data": "synthetic",
        "config": {"workload": "instanced field: 65,536 instances x 256-tri spheres (16.8M triangles, two-level), "
                               "1024^2 RGBA8 noise base-colour textures, 64 sphere lights + constant env; 8 views "
                               "512x512..1920x1080 in one wavefront; 1 spp per frame with the radiance cache, "
                               f"{args.frames} frames accumulated, max depth {MAX_DEPTH}",
                   "views": [[w, h] for _, w, h in views], "pixels": pixels, "frames": args.frames,
                   "triangles": info["triangles"], "state_count": STATE_COUNT, "launch_graph": True},
        "samples_per_s": round(paths / seconds, 1),
        "extension_rays": int(ext),
        "shadow_rays": int(shadow),
        "wall_seconds": round(wall, 3),
        "stage_ms_one_frame": {k: round(float(v), 3) for k, v in
                               zip(("generate", "trace", "shade", "shadow", "medium", "cache", "train"), stage)},
        "nrc_last_frame": {"queries": int(ns.queries), "records": int(ns.records),
                           "loss": round(float(ns.loss[0]), 5)},
        "quality": quality,
        "roofline": None,
        "cpu_baseline": None,
        "bvh_build": bvh,
    }


def instanced_flat_comparison(ctx, scene, p, views, stage):
    """one stage-timed frame over a single tree of all world-space triangles
    built on the GPU, against the two-level structure's stage times"""
    from btrc_amd import _abi
    two_level_ms = ctx.scene_build_ms()
    scene.accel = _abi.ACCEL_FLAT_GPU
    ctx.upload_scene(scene)
    flat_ms = ctx.scene_build_ms()
    finfo = ctx.scene_info()
    p.time_stages = 1
    stf = ctx.render_views(p, views)
    p.time_stages = 0
    scene.accel = _abi.ACCEL_TWO_LEVEL
    ctx.upload_scene(scene)
    log(f"bench: flat tree of {finfo['triangles']} triangles built on the GPU in {flat_ms:.0f} ms "
        f"({finfo['nodes']} BVH4 nodes): trace {stf.stage_ms[1]:.1f} ms vs two-level {stage[1]:.1f} ms")
    bvh = dict(two_level_host_build_ms=round(two_level_ms, 1), flat_gpu_build_ms=round(flat_ms, 1),
               flat_gpu_bvh4_nodes=finfo["nodes"], flat_gpu_bvh4_depth=finfo["depth"],
               trace_ms_two_level=round(float(stage[1]), 3), trace_ms_flat_gpu_built=round(float(stf.stage_ms[1]), 3),
               shadow_ms_two_level=round(float(stage[3]), 3), shadow_ms_flat_gpu_built=round(float(stf.stage_ms[3]), 3))
    return bvh


def media_cpu_baseline(scene, cam, p, budget_s):
    """the oracle's scalar path tracer (a port of the same algorithm) on 16
    host threads over the first paths of the frame, about budget_s seconds"""
    from oracle import oracle
    os_ = oracle.OracleScene(scene)
    threads = min(16, os.cpu_count() or 1)
    q = type(p).from_buffer_copy(p)
    count, rays, dt = 1 << 14, 0, 0.0
    done = 0
    while dt < budget_s and done + count <= q.width * q.height * q.spp:
        t0 = time.perf_counter()
        _, e, s = os_.render(cam, q, first=done, count=count, threads=threads)
        dt += time.perf_counter() - t0
        rays += e + s
        done += count
        count *= 2
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=threads, kind="port",
                sample=f"first {done} paths of the frame traced by the oracle's scalar path tracer "
                       f"(oracle/oracle.cpp) on {threads} host threads, {dt:.1f} s")


def flatten_instances(scene):
    parts = []
    for inst in scene.instances:
        p = scene.geometries[inst["geometry"]]["positions"].reshape(-1, 3).astype(np.float32)
        m = inst["mat"].astype(np.float32)
        # CTransform3D::apply_to_point operation order
        w = ((m[:, 0] * p[:, :1] + m[:, 1] * p[:, 1:2]) + m[:, 2] * p[:, 2:3]) + m[:, 3]
        parts.append(w[:, :3].reshape(-1, 9))
    return np.concatenate(parts).astype(np.float32)


if __name__ == "__main__":
    main()
