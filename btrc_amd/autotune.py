"""Autotuner of the kernels' scheduling knobs and of the path pool size.

The knobs (btrc_tuning in include/btrc_hip.h) change speed only; every
variant returns the same bits (tested in tests/test_gpu_tuning.py):
- traversal: leaf postponing of the closest-hit and of the visibility
  traversal (separately for one- and two-level structures) and the size of
  the persistent traversal grid (wavefronts per CU);
- stages: the grid of the generate / medium / shade kernels (workgroups of
  256 lanes per CU), the tile of path states each workgroup sweeps per pass;
- frame loop: the bounces per captured graph block between read-backs of
  the device counters.
The path pool (`state_count`, the reference's wavefront.h:23 queue size) is
a render parameter; the tuner returns the fastest and the caller passes it
on. Traversal knobs are timed per stage with a short version of the
caller's frame (same scene, camera and film, a few samples per pixel); the
stage grid by the generate + medium + shade time; pool size and graph block
length by the frame's wall time with the captured loop.
"""
import itertools

from . import _abi

WAVES = (16, 20, 24, 28)
STAGE_BLOCKS = (8, 16, 32)
GRAPH_BOUNCES = (2, 4, 8)
# path pools of 2^21 .. 2^25 states (about 1 to 17 GB of device memory;
# larger pools mean fewer, larger launches per frame)
POOLS = (1 << 21, 1 << 23, 1 << 25)


def _staged(ctx, p):
    ctx.render(p)  # the first run of a variant warms caches
    return ctx.render(p)


def autotune(ctx, params, spp=4, waves=WAVES, two_level=False, stage_blocks=STAGE_BLOCKS,
             graph_bounces=GRAPH_BOUNCES, pools=POOLS):
    """Returns (best tuning dict, best state_count, table of the timed
    variants). two_level: the uploaded scene uses BTRC_ACCEL_TWO_LEVEL, whose
    leaf postponing is the two_level_postpone bits (bit 0 closest hit, bit 1
    visibility) instead of closest/any_postpone_leaves. Pass empty sequences
    to skip a group."""
    p = _abi.RenderParams.from_buffer_copy(params)
    p.spp = spp
    p.time_stages = 1
    p.film_fp64 = 0
    table = []
    best = {}
    if waves:
        if two_level:
            variants = [dict(two_level_postpone=b, trace_waves_per_cu=w)
                        for b, w in itertools.product((0, 1, 2, 3), waves)]
        else:
            variants = [dict(closest_postpone_leaves=c, any_postpone_leaves=a, trace_waves_per_cu=w)
                        for c, a, w in itertools.product((1, 0), (0, 1), waves)]
        rows = []
        for v in variants:
            ctx.set_tuning(**v)
            st = _staged(ctx, p)
            rows.append((ctx.get_tuning(), st.stage_ms[1], st.stage_ms[3]))
            table.append(("traversal", v, st.stage_ms[1] + st.stage_ms[3]))
        # the two traversal kernels are independent: pick each by its own time
        best_w = min(waves, key=lambda w: min(t + s for tu, t, s in rows if tu["trace_waves_per_cu"] == w))
        rows = [r for r in rows if r[0]["trace_waves_per_cu"] == best_w]
        if two_level:
            closest_bit = min(rows, key=lambda r: r[1])[0]["two_level_postpone"] & 1
            any_bit = min(rows, key=lambda r: r[2])[0]["two_level_postpone"] & 2
            best.update(two_level_postpone=closest_bit | any_bit, trace_waves_per_cu=best_w)
        else:
            best.update(closest_postpone_leaves=min(rows, key=lambda r: r[1])[0]["closest_postpone_leaves"],
                        any_postpone_leaves=min(rows, key=lambda r: r[2])[0]["any_postpone_leaves"],
                        trace_waves_per_cu=best_w)
        ctx.set_tuning(**best)
    if stage_blocks:
        timed = []
        for b in stage_blocks:
            ctx.set_tuning(stage_blocks_per_cu=b)
            st = _staged(ctx, p)
            ms = st.stage_ms[0] + st.stage_ms[2] + st.stage_ms[4]
            timed.append((ms, b))
            table.append(("stage_blocks_per_cu", b, ms))
        best["stage_blocks_per_cu"] = min(timed)[1]
        ctx.set_tuning(**best)
    q = _abi.RenderParams.from_buffer_copy(p)
    q.time_stages = 0
    best_pool = params.state_count
    if pools or graph_bounces:
        ctx.set_tuning(launch_graph=1)
        timed = []
        for pool, g in itertools.product(pools or (params.state_count,), graph_bounces or (4,)):
            q.state_count = pool
            ctx.set_tuning(graph_bounces=g)
            s = _staged(ctx, q).seconds
            timed.append((s, pool, g))
            table.append(("pool_and_graph", (pool, g), s * 1e3))
        _, best_pool, best_g = min(timed)
        best["graph_bounces"] = best_g
        ctx.set_tuning(**best)
    return best, best_pool, table
