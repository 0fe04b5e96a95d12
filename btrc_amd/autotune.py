"""Autotuner of the kernels' scheduling knobs (btrc_tuning in
include/btrc_hip.h): leaf postponing of the closest-hit and of the
visibility traversal, and the size of the persistent traversal grid.

The knobs change speed only; every variant returns the same bits (tested in
tests/test_gpu_tuning.py). The tuner renders a short version of the caller's
frame (same scene, camera and film, a few samples per pixel) for each
candidate and keeps the fastest traversal settings.
"""
import itertools

from . import _abi

WAVES = (16, 20, 24, 28)


def autotune(ctx, params, spp=4, waves=WAVES):
    """Returns (best tuning dict, list of (tuning, trace ms, shadow ms))."""
    p = _abi.RenderParams.from_buffer_copy(params)
    p.spp = spp
    p.time_stages = 1
    p.film_fp64 = 0
    table = []
    for closest, anyp, w in itertools.product((1, 0), (0, 1), waves):
        ctx.set_tuning(closest_postpone_leaves=closest, any_postpone_leaves=anyp, trace_waves_per_cu=w)
        ctx.render(p)  # first run of a variant warms caches
        st = ctx.render(p)
        table.append((ctx.get_tuning(), st.stage_ms[1], st.stage_ms[3]))
    # the two traversal kernels are independent: pick each by its own time
    best_w = min(waves, key=lambda w: min(t + s for tu, t, s in table if tu["trace_waves_per_cu"] == w))
    rows = [r for r in table if r[0]["trace_waves_per_cu"] == best_w]
    best_closest = min(rows, key=lambda r: r[1])[0]["closest_postpone_leaves"]
    best_any = min(rows, key=lambda r: r[2])[0]["any_postpone_leaves"]
    best = dict(closest_postpone_leaves=best_closest, any_postpone_leaves=best_any, trace_waves_per_cu=best_w)
    ctx.set_tuning(**best)
    return best, table
