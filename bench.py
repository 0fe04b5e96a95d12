"""Throughput of the gfx950 wavefront path tracer on the procedural
1M-triangle scene at 1080p.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]
    python bench.py --probe {copy,frame,batch,media,nrc}   # short workloads for rocprofv3 counter runs

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
# path states in flight: 33.5 M (about 17 GB of pools, queues and hits on
# the device's 288 GB; the reference's default is 1,000,000). Larger pools
# mean fewer, larger launches per frame: 1080p at 64 spp took 249.6 ms with
# 2^21 states, 229.2 ms with 2^22, 216.3 ms with 2^23, 207.3 ms with 2^24,
# 203.9 ms with 2^25 and 202.1 ms with 2^26.
STATE_COUNT = 1 << 25
CAPTURE_STATE_COUNT = 1 << 21  # the pool of the frame whose first bounces give the baseline's rays


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_ray(steps, isects):
    """32 B ray in + 20 B hit out + 64 B per traversal step (two 32-byte
    binary nodes) + 36 B per triangle test (three FP32 vertices), with
    lib/bvh's SingleRayTraverser statistics on the same rays."""
    return 32.0 + 20.0 + 64.0 * float(np.mean(steps)) + 36.0 * float(np.mean(isects))


def read_traffic(path, key="hbm_bytes_per_ray"):
    """HBM bytes per unit from a committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE of the kernel it covers, per traced ray or per
    path state), if present."""
    if not os.path.exists(path):
        return None
    m = re.search(key + r":\s*([0-9.eE+-]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_threads():
    """threads for the CPU baselines: the job's CPU share (OMP_NUM_THREADS,
    16 per GPU on the benchmark hosts), at most the cores this process may
    run on"""
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(share, len(os.sched_getaffinity(0))))


def host_info(threads):
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(threads_used=threads, nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                cpu_model=model)


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
                       f"hit) on {threads} host threads, {dt:.1f} s", host=host_info(threads)), (steps, isects)


def probe(args):
    """Short fixed workloads for rocprofv3 runs, each printing the count that
    its per-dispatch counters divide by:
      copy  - one device-to-device copy of --mib MiB (torch's copy kernel),
              to scale FETCH_SIZE / WRITE_SIZE readings to bytes;
      frame - one 1080p frame of the 1M-triangle scene at --spp samples per
              pixel (the benchmark frame at 64), with its ray counts, as
              direct launches on one stream (stage timing), so that every
              kernel's trace duration is that of the launch alone;
      batch - the extension rays of that frame's first two bounces (up to
              --rays of them) traced --repeats times as one resident batch;
      media - one 1080p frame of the media workload (the Cornell-style room
              with the 128^3 heterogeneous medium, max depth 16) at --spp
              samples per pixel, as direct launches on one stream;
      nrc   - --frames 1 spp frames of the Cornell-style room at 256^2 with
              the radiance cache (queries at depth 2, training on 1/16 of the
              pixels), with query and record counts."""
    if args.probe == "copy":
        import torch
        n = args.mib * 1024 * 1024 // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"copy: {args.mib} MiB read + {args.mib} MiB written", flush=True)
        return
    import btrc_amd
    from btrc_amd import _abi, scenes
    ctx = btrc_amd.Context(0)
    if args.probe == "nrc":
        scene, cam = scenes.cornell()
        ctx.upload_scene(scene)
        ctx.set_camera(cam, 256, 256)
        p = _abi.RenderParams()
        p.width, p.height, p.spp = 256, 256, 1
        p.min_depth, p.max_depth = MIN_DEPTH, MAX_DEPTH
        p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
        p.state_count = 1 << 17
        ctx.nrc_configure(enabled=1, train=1, train_depth=args.nrc_train_depth, steps=args.nrc_steps,
                          train_stride=args.nrc_train_stride, query_depth=args.nrc_query_depth)
        ctx.nrc_reset(1)
        queries = records = 0
        frames = args.frames or 64
        for f in range(frames):
            p.sample_offset, p.accumulate = f, 1 if f else 0
            ctx.render(p)
            ns = ctx.nrc_stats()
            queries += ns.queries
            records += ns.records
        print(f"nrc: {frames} frames, {queries} queries, {records} training records, "
              f"{frames * args.nrc_steps} optimizer steps", flush=True)
        return
    media = args.probe == "media"
    scene, cam = scenes.cornell_media() if media else scenes.sphere_field()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, WIDTH, HEIGHT)
    p = _abi.RenderParams()
    p.width, p.height, p.spp = WIDTH, HEIGHT, args.spp
    p.min_depth, p.max_depth = MIN_DEPTH, 16 if media else MAX_DEPTH
    p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
    p.state_count = STATE_COUNT
    if args.probe in ("frame", "media"):
        p.time_stages = 1
        st = ctx.render(p)
        print(f"{args.probe}: {args.spp} spp, extension_rays {st.extension_rays} shadow_rays {st.shadow_rays} "
              f"iterations {st.iterations} stage_ms {[round(x, 2) for x in st.stage_ms[:7]]}", flush=True)
        return
    p.state_count = CAPTURE_STATE_COUNT  # the rays the CPU baseline traces
    r = ctx.capture_rays(p, 2, args.rays)
    r = r[np.random.default_rng(0).permutation(len(r))]
    ctx.batch_upload(r)
    for _ in range(args.repeats):
        ms = ctx.batch_trace(_abi.TRI_WATERTIGHT, False, 1)
        print(f"batch: {len(r)} rays in {ms:.3f} ms ({len(r) / ms / 1e3:.1f} Mrays/s)", flush=True)


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
    ap.add_argument("--nrc-train-depth", type=int, default=2,
                    help="nrc workload: end training paths at this depth with a cache query (self-training; "
                         "0: trace them out)")
    ap.add_argument("--nrc-steps", type=int, default=1, help="nrc workload: optimizer steps per frame")
    ap.add_argument("--nrc-warmup", type=int, default=8,
                    help="nrc workload: the cache runs' image restarts after this many frames, which only train "
                         "the cache (their time counts)")
    ap.add_argument("--nrc-query-depth", type=int, default=2, help="nrc workload: depth of the cache query")
    ap.add_argument("--nrc-train-stride", type=int, default=16,
                    help="nrc workload: one pixel in this many traces a training path")
    ap.add_argument("--nrc-ema-decay", type=float, default=0.0,
                    help="radiance cache: frames query a moving average of the weights with this decay (0: off)")
    ap.add_argument("--nrc-seeds", default="1,2,3,4",
                    help="network initialisation seeds of the radiance-cache runs (each trained from scratch)")
    ap.add_argument("--nrc-gt-spp", type=int, default=65536,
                    help="samples per pixel of the radiance-cache comparison's FP64 ground truth")
    ap.add_argument("--no-nrc", action="store_true",
                    help="field workload: skip the radiance-cache equal-time comparison")
    ap.add_argument("--no-bvh-compare", action="store_true",
                    help="skip the frame traced through a GPU-built tree (field workload)")
    ap.add_argument("--gt-spp", type=int, default=None,
                    help="samples per pixel of the FP64 ground truth for relMSE (0: skip; default 16384 for the "
                         "field, 256 for media, 1024 for the instanced views)")
    ap.add_argument("--probe", choices=["copy", "frame", "batch", "media", "nrc"], default=None,
                    help="run a short fixed workload for rocprofv3 instead of the benchmark")
    ap.add_argument("--mib", type=int, default=1024, help="probe copy: MiB copied")
    ap.add_argument("--rays", type=int, default=1 << 22, help="probe batch: rays captured")
    ap.add_argument("--repeats", type=int, default=3, help="probe batch: traversals of the batch")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the path tracer runs on one GPU only (asked for --gpus {args.gpus}); a frame would "
              f"shard across GPUs by sample index with one film reduction, which is not built", file=sys.stderr)
        sys.exit(2)

    if args.probe:
        probe(args)
        return

    import btrc_amd
    from btrc_amd import scenes, _abi

    if args.workload == "nrc":
        args.frames = 256 if args.frames is None else args.frames
        print(json.dumps(nrc_workload(args)), flush=True)
        return
    if args.workload == "instanced":
        args.frames = 64 if args.frames is None else args.frames
        args.gt_spp = 1024 if args.gt_spp is None else args.gt_spp
        print(json.dumps(instanced_workload(args)), flush=True)
        return

    media = args.workload == "media"
    if args.gt_spp is None:
        args.gt_spp = 256 if media else 16384
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
    staged = dict(seconds=st.seconds, ext=st.extension_rays, shadow=st.shadow_rays,
                  stage_ms=np.array(st.stage_ms[:7]), launches=np.array(st.stage_launches[:7]))
    log(f"bench: stage-timed frame: {st.seconds:.3f} s, stage ms {[round(x, 1) for x in st.stage_ms]}")
    quality = None
    if args.gt_spp > 0:
        # FP64 ground truth from the independent sampler, so its samples share
        # nothing with the Halton samples of the timed frame
        # rendered as two halves: their mutual relMSE is about four times the
        # ground truth's own error
        t0 = time.perf_counter()
        halves = []
        for k in range(2):
            g = _abi.RenderParams.from_buffer_copy(p)
            g.spp, g.film_fp64, g.time_stages, g.sampler = args.gt_spp // 2, 1, 0, _abi.SAMPLER_INDEPENDENT
            g.sample_offset = k * (args.gt_spp // 2)
            ctx.render(g)
            halves.append(ctx.film(WIDTH, HEIGHT, fp64=True)[1])
        gt = 0.5 * (halves[0] + halves[1])
        quality = dict(relmse=float(np.mean((frame - gt) ** 2 / (gt ** 2 + 1e-2))), spp=args.spp,
                       gt_spp=args.gt_spp, gt_sampler="independent", gt_film="fp64",
                       gt_split_half_relmse=float(np.mean((halves[0] - halves[1]) ** 2 / (gt ** 2 + 1e-2))),
                       gt_seconds=round(time.perf_counter() - t0, 2))
        log(f"bench: relMSE of the {args.spp} spp frame vs {args.gt_spp} spp FP64 ground truth: "
            f"{quality['relmse']:.5f}")

    secs = totals["seconds"]
    mrays = (totals["ext"] + totals["shadow"]) / secs / 1e6
    samples = totals["paths"] / secs
    trace_ms = staged["stage_ms"][1]

    cpu = None
    bytes_per_ray = None
    medium_stats = None
    if not args.no_cpu and media:
        cpu, medium_stats = media_cpu_baseline(scene, cam, p, args.cpu_seconds)
    elif not args.no_cpu:
        # rays of this frame for the work model and the CPU baseline
        log("bench: capturing rays of the first two bounces")
        pc = _abi.RenderParams.from_buffer_copy(p)
        pc.time_stages = 0
        # the first two bounces of a 2^21-state wave: primary rays and the
        # first bounce's in equal parts, as the traversal meets them over
        # the frame (the first 4 M rays of a larger wave would all be
        # primary rays)
        pc.state_count = CAPTURE_STATE_COUNT
        rays = ctx.capture_rays(pc, 2, 1 << 22)
        rays = rays[np.random.default_rng(0).permutation(len(rays))]
        # world-space triangles in flattened order, as the GPU traverses them
        world = flatten_instances(scene)
        cpu, stats = cpu_baseline(rays, world, args.cpu_seconds, host_threads())
        if stats is not None:
            bytes_per_ray = algorithmic_bytes_per_ray(*stats)
            cpu["steps_per_ray"] = float(np.mean(stats[0]))
            cpu["isects_per_ray"] = float(np.mean(stats[1]))

    copy_gbs = round(ctx.measure_copy(4096, 8), 1)
    log(f"bench: device copy {copy_gbs:.0f} GB/s (read + write)")
    roofline = None
    if media and staged["stage_ms"][4] > 0 and medium_stats is not None:
        roofline = medium_roofline(p, staged, medium_stats)
    elif bytes_per_ray is not None and trace_ms > 0:
        launches = max(1, int(staged["launches"][1]))
        rays_per_launch = staged["ext"] / launches
        launch_s = trace_ms * 1e-3 / launches
        achieved = bytes_per_ray * rays_per_launch / launch_s / 1e9
        traffic_per_ray = read_traffic(os.path.join(ROOT, "profiles", "trace_extension_sphere_field_1080p_64spp.md"))
        roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4),
                        traffic=None if traffic_per_ray is None else round(traffic_per_ray * rays_per_launch),
                        kernel="k_trace_extension",
                        algorithmic_bytes_per_ray=round(bytes_per_ray, 1),
                        algorithmic_bytes_per_launch=round(bytes_per_ray * rays_per_launch),
                        rays_per_launch=round(rays_per_launch), launch_ms=round(launch_s * 1e3, 4),
                        traffic_bytes_per_ray=traffic_per_ray)
    if roofline is not None:
        # the same achieved rate against the copy bandwidth measured above
        roofline["copy_measured"] = copy_gbs
        roofline["frac_of_copy"] = round(roofline["achieved"] / copy_gbs, 4)

    bvh = None if media or args.no_bvh_compare else gpu_bvh_comparison(ctx, scene, ps, trace_ms, staged["seconds"])

    nrc = None
    if not media and not args.no_nrc:
        args.frames = 256 if args.frames is None else args.frames
        nrc = nrc_equal_time(args, sorted({0, args.nrc_train_depth}))

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
        # rays over the traversal kernels' own time (closest hit + shadow
        # stages of the stage-timed frame), beside the wall-time value
        "mrays_per_traversal_s": (round((staged["ext"] + staged["shadow"]) /
                                        ((staged["stage_ms"][1] + staged["stage_ms"][3]) * 1e-3) / 1e6, 1)
                                  if staged["stage_ms"][1] + staged["stage_ms"][3] > 0 else None),
        "hbm_copy": {"value": copy_gbs, "unit": "GB/s", "spec": HBM_PEAK_GBS,
                     "method": "STREAM-style copy kernel (btrc_hip_measure_copy) over 4 GiB, HIP events, median of 8, "
                               "read + write bytes"},
        "quality": quality,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bvh_build": bvh,
        "nrc_equal_time": nrc,
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


def nrc_equal_time(args, train_depths):
    """Cornell-style room at 256^2, `args.frames` frames of 1 spp accumulated:
    plain path tracing, then the radiance cache (queries at depth
    args.nrc_query_depth, training on 1/args.nrc_train_stride of the pixels)
    once per training depth in `train_depths` (0: training paths traced out,
    unbiased targets; d > 0: they end at depth d with a cache query,
    self-training). Timed frames replay the captured bounce loop (no stage
    timing); each cache run goes on past the frame budget (up to 4x) until
    it reaches the plain tracer's final relMSE against an FP64 ground truth
    of args.nrc_gt_spp samples per pixel from the independent sampler (two
    halves, whose mutual relMSE bounds the ground truth's own error). A
    separate stage-timed pass of 32 frames gives the cache kernels' times."""
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
        p.min_depth, p.max_depth = MIN_DEPTH, MAX_DEPTH
        p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
        p.state_count = 1 << 17
        p.film_fp64 = fp64
        return p

    log(f"bench: cache comparison: {args.nrc_gt_spp} spp FP64 ground truth of the 256^2 room")
    t0 = time.perf_counter()
    halves = []
    for k in range(2):
        g = params(args.nrc_gt_spp // 2, 1)
        g.state_count = 1 << 20
        g.sampler = _abi.SAMPLER_INDEPENDENT
        g.sample_offset = k * (args.nrc_gt_spp // 2)
        ctx.render(g)
        halves.append(ctx.film(W, H, fp64=True)[1])
    gt = 0.5 * (halves[0] + halves[1])
    gt_split = float(np.mean((halves[0] - halves[1]) ** 2 / (gt ** 2 + 1e-2)))
    gt_seconds = time.perf_counter() - t0

    def relmse():
        img = ctx.film(W, H)[0][..., :3].astype(np.float64)
        return float(np.mean((img - gt) ** 2 / (gt ** 2 + 1e-2)))

    checkpoints = sorted({c for c in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024) if c <= args.frames} | {args.frames})

    def configure(cache, train_depth, seed=1):
        ctx.nrc_configure(enabled=1 if cache else 0, train=1, train_depth=train_depth, steps=args.nrc_steps,
                          train_stride=args.nrc_train_stride, query_depth=args.nrc_query_depth,
                          ema_decay=args.nrc_ema_decay)
        if cache:
            ctx.nrc_reset(seed)

    def run(cache, train_depth=0, target=None, seed=1):
        configure(cache, train_depth, seed)
        p = params(1)
        p.time_stages = 0
        curve, seconds, rays = [], 0.0, 0
        losses = []
        # the cache's first frames only train it: the image restarts after them
        warm = args.nrc_warmup if cache else 0

        def loss():
            ns = ctx.nrc_stats()
            return float(np.mean(ns.loss[:min(args.nrc_steps, max(1, ns.steps_total))]))
        f = 0
        while f < args.frames or (target is not None and f < 4 * args.frames and curve[-1][2] > target):
            p.sample_offset, p.accumulate = f, 1 if (f and f != warm) else 0
            st = ctx.render(p)
            seconds += st.seconds
            rays += st.extension_rays + st.shadow_rays
            if cache and f == 0:
                # a frame's training overlaps the next frame's bounces; the
                # statistics wait for it, so this wait counts as render time
                t0 = time.perf_counter()
                losses.append(loss())
                seconds += time.perf_counter() - t0
            f += 1
            if f in checkpoints or (f >= 32 and f % 8 == 0):
                curve.append((f, round(seconds, 4), relmse()))
        if cache:
            losses.append(loss())  # the last frame's training is not part of the image
        return dict(seconds=seconds, rays=rays, curve=curve, losses=losses, frames=f)

    def staged(cache, train_depth=0, frames=32):
        configure(cache, train_depth)
        p = params(1)
        p.time_stages = 1
        stage, queries, records = np.zeros(7), 0, 0
        for f in range(frames):
            p.sample_offset, p.accumulate = f, 1 if f else 0
            st = ctx.render(p)
            stage += np.array(st.stage_ms[:7])
            if cache:
                ns = ctx.nrc_stats()
                queries += ns.queries
                records += ns.records
        return stage / frames, queries / frames, records / frames

    def time_to(curve, target):
        for frames, t, e in curve:
            if e <= target:
                return t
        return None

    def at_time(curve, t):
        best = None
        for frames, tt, e in curve:
            if tt <= t:
                best = e
        return best

    log("bench: plain path tracing, 1 spp frames")
    plain = run(False)
    target = plain["curve"][-1][2]

    def plain_at(t):
        """the plain tracer's relMSE at time t, log-log interpolated between checkpoints"""
        pts = plain["curve"]
        for (f0, t0, e0), (f1, t1, e1) in zip(pts, pts[1:]):
            if t0 <= t <= t1 and t1 > t0:
                a = (np.log(t) - np.log(t0)) / (np.log(t1) - np.log(t0))
                return float(np.exp(np.log(e0) + a * (np.log(e1) - np.log(e0))))
        return None

    runs = {}
    seeds = [int(s) for s in args.nrc_seeds.split(",")]
    for d in train_depths:
        log(f"bench: path tracing with the radiance cache, training depth {d}, network seeds {seeds}")
        # each seed trains its own network from scratch; the first seed's run is reported in full
        per_seed = []
        for seed in seeds:
            cs = run(True, d, target, seed)
            ts = time_to(cs["curve"], target)
            per_seed.append(dict(seed=seed, frames=cs["frames"], seconds=round(cs["seconds"], 4),
                                 time_to_target_s=ts,
                                 time_ratio=None if ts is None else round(ts / plain["seconds"], 3),
                                 final_relmse=cs["curve"][-1][2]))
            if seed == seeds[0]:
                c = cs
        ratios = [s["time_ratio"] for s in per_seed if s["time_ratio"] is not None]
        t_cache = time_to(c["curve"], target)
        budget = [x for x in c["curve"] if x[0] <= args.frames][-1]
        stage, qpf, records_pf = staged(True, d)
        infer_ms, train_ms = stage[5], stage[6]
        infer_tflops = NRC_FLOP_PER_QUERY * qpf / (infer_ms * 1e-3) / 1e12 if infer_ms > 0 else None
        train_tflops = (NRC_FLOP_PER_TRAIN_SAMPLE * records_pf / (train_ms * 1e-3) / 1e12
                        if train_ms > 0 else None)
        runs[f"train_depth_{d}"] = {
            "time_ratio_over_seeds": {"mean": None if len(ratios) < len(seeds) else round(float(np.mean(ratios)), 3),
                                      "min": min(ratios) if ratios else None,
                                      "max": max(ratios) if ratios else None, "runs": per_seed},
            "seconds": round(c["seconds"], 4), "frames": c["frames"], "relmse_curve": c["curve"],
            "mrays_per_s": round(c["rays"] / c["seconds"] / 1e6, 2),
            "queries_per_frame": round(qpf),
            "records_per_frame": round(records_pf),
            "loss_first_last": [round(c["losses"][0], 4), round(c["losses"][-1], 4)],
            "time_to_target_s": t_cache,
            "time_ratio": None if t_cache is None else round(t_cache / plain["seconds"], 3),
            "relmse_at_plain_time": at_time(c["curve"], plain["seconds"]),
            "at_frame_budget": {"seconds": budget[1], "cache_relmse": budget[2],
                                "plain_relmse_at_that_time": plain_at(budget[1])},
            "stage_ms_per_frame": {k: round(float(v), 4) for k, v in
                                   zip(("generate", "trace", "shade", "shadow", "medium", "cache", "train"), stage)},
            "infer_tflops": None if infer_tflops is None else round(infer_tflops, 3),
            "train_tflops": None if train_tflops is None else round(train_tflops, 3),
        }
        log(f"bench: cache (training depth {d}): time ratio {runs[f'train_depth_{d}']['time_ratio']}, "
            f"over seeds {runs[f'train_depth_{d}']['time_ratio_over_seeds']['mean']}")
    ctx.nrc_configure(enabled=0)
    plain_stage, _, _ = staged(False)
    return {
        "workload": f"Cornell-style room, 256x256, 1 spp per frame x {args.frames} frames, max depth {MAX_DEPTH}, "
                    f"radiance cache queried at depth {args.nrc_query_depth}, trained on 1/{args.nrc_train_stride} "
                    f"of the pixels ({args.nrc_steps} Adam steps per frame over the frame's training records); the "
                    f"cache runs' image restarts after {args.nrc_warmup} frames that only train the cache (their "
                    "time counts); timed frames replay the captured bounce loop",
        "ground_truth": {"spp": args.nrc_gt_spp, "sampler": "independent", "film": "fp64",
                         "split_half_relmse": gt_split, "seconds": round(gt_seconds, 2)},
        "target_relmse": target,
        "plain": {"seconds": round(plain["seconds"], 4), "relmse_curve": plain["curve"],
                  "mrays_per_s": round(plain["rays"] / plain["seconds"] / 1e6, 2),
                  "stage_ms_per_frame": {k: round(float(v), 4) for k, v in
                                         zip(("generate", "trace", "shade", "shadow"), plain_stage[:4])}},
        "cache": runs,
    }


def nrc_workload(args):
    """the radiance-cache comparison as its own benchmark line: value = the
    cache run's Mrays/s at the chosen training depth"""
    cmp_ = nrc_equal_time(args, sorted({0, args.nrc_train_depth}))
    main_run = cmp_["cache"][f"train_depth_{args.nrc_train_depth}"]
    infer = main_run["infer_tflops"]
    cpu = None
    if not args.no_cpu:
        from btrc_amd import _abi, scenes
        scene, cam = scenes.cornell()
        p = _abi.RenderParams()
        p.width, p.height, p.spp = 256, 256, 16
        p.min_depth, p.max_depth = MIN_DEPTH, MAX_DEPTH
        p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
        cpu, _ = media_cpu_baseline(scene, cam, p, args.cpu_seconds)
    return {
        "metric": METRIC,
        "value": main_run["mrays_per_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": main_run["frames"],
        "warmup": 0,
        "ms_per_step": round(main_run["seconds"] / main_run["frames"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 paths, bf16 network (fp32 accumulation)",
        "data": "synthetic",
        "config": {"workload": cmp_["workload"], "width": 256, "height": 256, "frames": args.frames,
                   "train_depth": args.nrc_train_depth, "steps_per_frame": args.nrc_steps,
                   "train_stride": args.nrc_train_stride, "query_depth": args.nrc_query_depth},
        "equal_time": cmp_,
        "roofline": {"bound": "mfma", "kernel": "k_nrc_infer", "unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS,
                     "achieved": infer, "frac": None if infer is None else round(infer / MFMA_BF16_PEAK_TFLOPS, 6),
                     "traffic": None, "flop_per_query": NRC_FLOP_PER_QUERY,
                     "train_achieved": main_run["train_tflops"]},
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
    staged_ext, trace_launches = st.extension_rays, max(1, int(st.stage_launches[1]))
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

    cpu, roofline = None, None
    if not args.no_cpu:
        cpu, roofline = instanced_baseline(ctx, scene, views, p, stage[1], staged_ext, trace_launches, args)

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
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bvh_build": bvh,
    }


def instanced_baseline(ctx, scene, views, p, trace_ms, ext, launches, args):
    """CPU baseline and traversal work model of the instanced workload: the
    extension rays of the first two bounces of its largest view (1 spp),
    traced by lib/bvh (BinnedSAH16 + SingleRayTraverser) over the 16.8M
    world-space triangles flattened on the host. lib/bvh has no two-level
    traversal, so the work model prices each ray by what a one-level
    traversal of the same triangles reads (32 + 20 + 64 per step + 36 per
    triangle test, its statistics on these rays); the two-level kernel's
    time comes from the stage-timed frame."""
    from btrc_amd import _abi
    cam, w, h = max(views, key=lambda v: v[1] * v[2])
    ctx.nrc_configure(enabled=0)
    ctx.set_camera(cam, w, h)
    q = _abi.RenderParams.from_buffer_copy(p)
    q.width, q.height, q.spp, q.sample_offset, q.accumulate, q.time_stages = w, h, 1, 0, 0, 0
    log("bench: capturing rays of the largest view's first two bounces")
    rays = ctx.capture_rays(q, 2, 1 << 22)
    rays = rays[np.random.default_rng(0).permutation(len(rays))]
    log("bench: flattening the instances for lib/bvh")
    world = flatten_instances(scene)
    cpu, stats = cpu_baseline(rays, world, args.cpu_seconds, host_threads())
    if cpu is None:
        return None, None
    cpu["steps_per_ray"] = float(np.mean(stats[0]))
    cpu["isects_per_ray"] = float(np.mean(stats[1]))
    cpu["sample"] = cpu["sample"].replace("of this frame", f"of the {w}x{h} view") + \
        f", over the {len(world)} flattened world triangles"
    bytes_per_ray = algorithmic_bytes_per_ray(*stats)
    rays_per_launch = ext / launches
    launch_s = trace_ms * 1e-3 / launches
    achieved = bytes_per_ray * rays_per_launch / launch_s / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, kernel="k_trace_extension (two-level)",
                    algorithmic_bytes_per_ray=round(bytes_per_ray, 1),
                    algorithmic_bytes_per_launch=round(bytes_per_ray * rays_per_launch),
                    rays_per_launch=round(rays_per_launch), launch_ms=round(launch_s * 1e3, 4),
                    model="one-level lib/bvh statistics over the flattened triangles, same rays")
    return cpu, roofline


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


MEDIUM_BYTES_PER_STATE = 204.0   # a path state and its hit read by the medium stage (shade-like)
MEDIUM_BYTES_PER_LOOKUP = 32.0   # one trilinear density fetch: eight FP32 voxels


def medium_roofline(p, staged, stats):
    """Work model of the medium stage (k_medium_track + k_medium_scatter):
    204 B per path state it processes (one per extension ray) plus 32 B per
    density lookup of delta tracking. Lookups per path come from the CPU
    baseline's run of the same algorithm on paths spread over the frame
    (media_cpu_baseline); the stage's time is the stage-timed frame's;
    `traffic` is the measured HBM bytes per path state of the committed
    counter summary times the states per launch."""
    lookups_per_path = stats["lookups_per_path"]
    frame_paths = p.width * p.height * p.spp
    launches = max(1, int(staged["launches"][4]))
    total_bytes = MEDIUM_BYTES_PER_STATE * staged["ext"] + MEDIUM_BYTES_PER_LOOKUP * lookups_per_path * frame_paths
    stage_s = staged["stage_ms"][4] * 1e-3
    achieved = total_bytes / stage_s / 1e9
    per_state = read_traffic(os.path.join(ROOT, "profiles", "medium_stage_cornell_media_1080p_64spp.md"),
                             "hbm_bytes_per_state")
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=None if per_state is None else round(per_state * staged["ext"] / launches),
                traffic_bytes_per_state=per_state, kernel="k_medium_track + k_medium_scatter",
                algorithmic_bytes_per_launch=round(total_bytes / launches),
                launch_ms=round(staged["stage_ms"][4] / launches, 4),
                bytes_per_state=MEDIUM_BYTES_PER_STATE, bytes_per_lookup=MEDIUM_BYTES_PER_LOOKUP,
                lookups_per_path=round(lookups_per_path, 3),
                tracking_segments_per_path=round(stats["segments_per_path"], 3),
                sample=stats["sample"])


def oracle_path_baseline(scene, cam, p, budget_s, bands=8):
    """The full-path CPU baseline: the oracle's scalar path tracer (a port of
    the same algorithm, oracle/oracle.cpp) on the host threads over sample-0
    paths of `bands` bands of image rows spread over the frame, a few more
    rows per band each pass until about budget_s seconds. Returns the
    baseline (Mrays/s, with samples/s beside it) and the paths traced."""
    from oracle import oracle
    os_ = oracle.OracleScene(scene)
    threads = host_threads()
    W, H = p.width, p.height
    band_rows = H // bands
    done_rows, rows, rays, dt, paths = 0, 1, 0, 0.0, 0
    while dt < budget_s and done_rows < band_rows:
        rows = min(rows, band_rows - done_rows)
        for k in range(bands):
            r0 = k * band_rows + done_rows
            t0 = time.perf_counter()
            _, e, s = os_.render(cam, p, first=r0 * W, count=rows * W, threads=threads)
            dt += time.perf_counter() - t0
            rays += e + s
            paths += rows * W
        done_rows += rows
        rows *= 2
    sample = (f"{paths} sample-0 paths in {bands} bands of {done_rows} image rows spread over the frame, traced by "
              f"the oracle's scalar path tracer (oracle/oracle.cpp) on {threads} host threads, {dt:.1f} s")
    cpu = dict(value=rays / dt / 1e6, unit="Mrays/s", samples_per_s=round(paths / dt, 1), cores=threads,
               kind="port", sample=sample, host=host_info(threads))
    return cpu, paths


def media_cpu_baseline(scene, cam, p, budget_s, bands=8):
    """CPU baseline of the media workload and the statistics of its work
    model: the full-path baseline above on the media frame, counting the
    oracle's tracking segments and density lookups"""
    from oracle import oracle
    oracle.tracking_stats()
    cpu, paths = oracle_path_baseline(scene, cam, p, budget_s, bands)
    segments, lookups = oracle.tracking_stats()
    return cpu, dict(lookups_per_path=lookups / paths, segments_per_path=segments / paths, sample=cpu["sample"])


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
