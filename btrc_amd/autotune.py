"""Autotuner of the kernels' scheduling knobs (btrc_tuning in
include/btrc_hip.h): leaf postponing of the closest-hit and of the
visibility traversal (separately for one- and two-level structures), and the
size of the persistent traversal grid.

The knobs change speed only; every variant returns the same bits (tested in
tests/test_gpu_tuning.py). The tuner renders a short version of the caller's
frame (same scene, camera and film, a few samples per pixel) for each
candidate and keeps the fastest traversal settings.
"""
import itertools

from . import _abi

WAVES = (16, 20, 24, 28)


def autotune(ctx, params, spp=4, waves=WAVES, two_level=False):
    """Returns (best tuning dict, list of (tuning, trace ms, shadow ms)).
    two_level: the uploaded scene uses BTRC_ACCEL_TWO_LEVEL, whose leaf
    postponing is the two_level_postpone bits (bit 0 closest hit, bit 1
    visibility) instead of closest/any_postpone_leaves."""
    p = _abi.RenderParams.from_buffer_copy(params)
    p.spp = spp
    p.time_stages = 1
    p.film_fp64 = 0
    table = []
    if two_level:
        variants = [dict(two_level_postpone=b, trace_waves_per_cu=w) for b, w in itertools.product((0, 1, 2, 3), waves)]
    else:
        variants = [dict(closest_postpone_leaves=c, any_postpone_leaves=a, trace_waves_per_cu=w)
                    for c, a, w in itertools.product((1, 0), (0, 1), waves)]
    for v in variants:
        ctx.set_tuning(**v)
        ctx.render(p)  # first run of a variant warms caches
        st = ctx.render(p)
        table.append((ctx.get_tuning(), st.stage_ms[1], st.stage_ms[3]))
    # the two traversal kernels are independent: pick each by its own time
    best_w = min(waves, key=lambda w: min(t + s for tu, t, s in table if tu["trace_waves_per_cu"] == w))
    rows = [r for r in table if r[0]["trace_waves_per_cu"] == best_w]
    if two_level:
        closest_bit = min(rows, key=lambda r: r[1])[0]["two_level_postpone"] & 1
        any_bit = min(rows, key=lambda r: r[2])[0]["two_level_postpone"] & 2
        best = dict(two_level_postpone=closest_bit | any_bit, trace_waves_per_cu=best_w)
    else:
        best_closest = min(rows, key=lambda r: r[1])[0]["closest_postpone_leaves"]
        best_any = min(rows, key=lambda r: r[2])[0]["any_postpone_leaves"]
        best = dict(closest_postpone_leaves=best_closest, any_postpone_leaves=best_any, trace_waves_per_cu=best_w)
    ctx.set_tuning(**best)
    return best, table
