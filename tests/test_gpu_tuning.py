"""Every scheduling variant of the traversal returns the same bits."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes

from util import assert_hits_equal, random_rays, render_params

pytestmark = pytest.mark.gpu

VARIANTS = [dict(closest_postpone_leaves=c, any_postpone_leaves=a, trace_waves_per_cu=w)
            for c in (0, 1) for a in (0, 1) for w in (1, 7, 24, 32)]


def test_tuning_roundtrip_and_validation(ctx):
    ctx.set_tuning(trace_waves_per_cu=7)
    assert ctx.get_tuning()["trace_waves_per_cu"] == 7
    with pytest.raises(Exception):
        ctx.set_tuning(trace_waves_per_cu=0)
    ctx.set_tuning(closest_postpone_leaves=1, any_postpone_leaves=0, trace_waves_per_cu=20)


def test_traversal_variants_bitwise_equal(ctx):
    scene = scenes.sphere_field(instances=1024, lights=8, seed=6, extent=50.0)[0]
    ctx.upload_scene(scene)
    rays = random_rays(200000, [0, 0, 0], [50, 50, 50], 31)
    short = rays.copy()
    short["tmax"] = np.float32(5.0)
    ref_h = ref_o = None
    for v in VARIANTS:
        ctx.set_tuning(**v)
        for mode in (_abi.TRI_WATERTIGHT, _abi.TRI_MOLLER_TRUMBORE):
            h = ctx.trace_closest(rays, mode)
            o = ctx.trace_any(short, mode)
            if ref_h is None:
                ref_h, ref_o = {}, {}
            if mode not in ref_h:
                ref_h[mode], ref_o[mode] = h, o
            else:
                assert_hits_equal(h, ref_h[mode])
                assert np.array_equal(o, ref_o[mode])
    ctx.set_tuning(closest_postpone_leaves=1, any_postpone_leaves=0, trace_waves_per_cu=20)


def test_render_variants_equal(ctx):
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    p = render_params(96, 96, 4, state_count=5000)
    ctx.set_camera(cam, 96, 96)
    imgs = []
    for v in VARIANTS[::3]:
        ctx.set_tuning(**v)
        ctx.render(p)
        imgs.append(ctx.film(96, 96)[0])
    for im in imgs[1:]:
        # only the FP32 atomic accumulation order may differ
        np.testing.assert_allclose(im, imgs[0], rtol=1e-5, atol=1e-6)
    ctx.set_tuning(closest_postpone_leaves=1, any_postpone_leaves=0, trace_waves_per_cu=20)


def test_graph_loop_equals_direct_launches(ctx):
    # the bounce loop replayed from a captured graph, across progressive
    # frames (new sample offsets, same launches) and a changed pool size
    scene, cam = scenes.sphere_field(instances=128, lights=8, seed=6, extent=20.0)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 64, 48)
    films = {}
    for graph in (1, 0):
        ctx.set_tuning(launch_graph=graph)
        out = []
        for frame, pool in ((0, 3000), (1, 3000), (2, 3000), (3, 2000)):
            p = render_params(64, 48, 1, state_count=pool, film_fp64=1)
            p.sample_offset = frame
            st = ctx.render(p)
            out.append((ctx.film(64, 48, fp64=True)[1], st.extension_rays, st.shadow_rays))
        films[graph] = out
    for (a, ea, sa), (b, eb, sb) in zip(films[1], films[0]):
        assert ea == eb and sa == sb
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())
    ctx.set_tuning(launch_graph=1)


@pytest.mark.parametrize("kind", ["field", "media", "cache"])
def test_shadow_stream_overlap_equals_direct_launches(ctx, kind):
    """bounces of a million paths run their shadow rays on a second stream
    beside the next bounce's first stages; the film, the ray counts and the
    cache's training statistics equal the direct launches' (one stream)"""
    if kind == "media":
        scene, cam = scenes.cornell_media(grid_n=24)
    elif kind == "cache":
        scene, cam = scenes.cornell()
    else:
        scene, cam = scenes.sphere_field(instances=256, lights=8, seed=6, extent=30.0)
    ctx.upload_scene(scene)
    w, h, spp = (1024, 1024, 1) if kind == "cache" else (256, 256, 16)
    ctx.set_camera(cam, w, h)
    got = {}
    for graph in (1, 0):
        ctx.set_tuning(launch_graph=graph)
        if kind == "cache":
            ctx.nrc_configure(enabled=1, train=1, train_depth=0)
            ctx.nrc_reset(3)
        out = []
        for frame in range(2):
            p = render_params(w, h, spp, max_depth=10 if kind == "media" else 8, state_count=1 << 20, film_fp64=1)
            p.sample_offset, p.accumulate = frame, 1 if frame else 0
            st = ctx.render(p)
            stats = ctx.nrc_stats() if kind == "cache" else None
            out.append((ctx.film(w, h, fp64=True)[1], st.extension_rays, st.shadow_rays,
                        None if stats is None else (stats.records, stats.queries, tuple(stats.loss[:4]))))
        got[graph] = out
        if kind == "cache":
            ctx.nrc_configure(enabled=0)
    for (a, ea, sa, na), (b, eb, sb, nb) in zip(got[1], got[0]):
        assert ea == eb and sa == sb and na == nb
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())
    ctx.set_tuning(launch_graph=1)


def test_stage_grid_and_graph_block_keep_results(ctx):
    """the stage kernels' grid and the bounces per captured block change
    speed only"""
    scene, cam = scenes.sphere_field(instances=128, lights=8, seed=7, extent=20.0)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 64, 48)
    ref = None
    for blocks, bounces in ((0, 0), (1, 2), (64, 16), (3, 6)):
        ctx.set_tuning(stage_blocks_per_cu=blocks, graph_bounces=bounces, launch_graph=1)
        st = ctx.render(render_params(64, 48, 2, state_count=2500, film_fp64=1))
        got = (ctx.film(64, 48, fp64=True)[1], st.extension_rays, st.shadow_rays)
        if ref is None:
            ref = got
        assert got[1:] == ref[1:]
        assert np.abs(got[0] - ref[0]).max() <= 1e-12 * max(1.0, np.abs(ref[0]).max())
    import btrc_amd
    for bad in (dict(graph_bounces=3), dict(graph_bounces=18), dict(stage_blocks_per_cu=65)):
        with pytest.raises(btrc_amd.BtrcException):
            ctx.set_tuning(**bad)
    ctx.set_tuning(stage_blocks_per_cu=0, graph_bounces=0)


def test_coplanar_hits_every_schedule(ctx):
    # the floor and coplanar tiles: box culling must keep equal-distance
    # candidates whatever the visit order, in both structures
    from oracle import oracle
    from test_oracle_traversal import coplanar_rays, coplanar_scene
    rays = coplanar_rays(50000, 5)
    for accel in (_abi.ACCEL_FLAT, _abi.ACCEL_TWO_LEVEL):
        scene = coplanar_scene()
        scene.accel = accel
        ctx.upload_scene(scene)
        want = oracle.OracleScene(scene).trace(rays)
        for c in (0, 1):
            for w in (1, 20):
                ctx.set_tuning(closest_postpone_leaves=c, two_level_postpone=c, trace_waves_per_cu=w)
                assert_hits_equal(ctx.trace_closest(rays), want)
    ctx.set_tuning(closest_postpone_leaves=1, two_level_postpone=0, trace_waves_per_cu=20)


@pytest.mark.parametrize("two_level", [False, True], ids=["one_level", "two_level"])
def test_autotune_picks_a_candidate_and_keeps_results(ctx, two_level):
    """the tuner times every candidate, leaves the best one set, and the
    frame it tunes renders the same paths before and after"""
    from btrc_amd.autotune import autotune
    scene, cam = scenes.sphere_field(instances=256, lights=8, seed=4, extent=24.0)
    if two_level:
        scene.accel = _abi.ACCEL_TWO_LEVEL
    ctx.upload_scene(scene)
    p = render_params(128, 96, 4, state_count=1 << 14)
    ctx.set_camera(cam, p.width, p.height)
    before = ctx.render(p)
    img0, _ = ctx.film(p.width, p.height)
    best, pool, table = autotune(ctx, p, spp=1, waves=(16, 20), two_level=two_level, stage_blocks=(8, 16),
                                 graph_bounces=(2, 4), pools=(1 << 13, 1 << 14))
    assert len(table) == 8 + 2 + 4
    assert all(ms > 0 for _, _, ms in table)
    assert pool in (1 << 13, 1 << 14)
    t = ctx.get_tuning()
    for k, v in best.items():
        assert t[k] == v
    assert best["stage_blocks_per_cu"] in (8, 16) and best["graph_bounces"] in (2, 4)
    after = ctx.render(p)
    img1, _ = ctx.film(p.width, p.height)
    assert (before.extension_rays, before.shadow_rays) == (after.extension_rays, after.shadow_rays)
    # the pool size the tuner picked renders the same paths too
    q = render_params(128, 96, 4, state_count=pool)
    ctx.render(q)
    img2, _ = ctx.film(p.width, p.height)
    assert np.abs(img0 - img2).max() / max(1.0, np.abs(img0).max()) <= 1e-5
    after = ctx.render(p)
    img1, _ = ctx.film(p.width, p.height)
    assert (before.extension_rays, before.shadow_rays) == (after.extension_rays, after.shadow_rays)
    rel = np.abs(img0 - img1) / np.maximum(np.abs(img0), 1e-2)
    assert rel.max() <= 1e-5


def test_large_pool_traces_the_same_paths(ctx):
    """a frame of 33 M paths in one 2^25-state wave (dense indices up to
    33 M, pool slots beyond 2^25) and in 2^21-state waves: equal ray counts
    and FP64 films"""
    scene, cam = scenes.sphere_field(instances=256, lights=8, seed=6, extent=30.0)
    ctx.upload_scene(scene)
    w, h, spp = 1920, 1080, 16
    ctx.set_camera(cam, w, h)
    got = []
    for pool in (1 << 25, 1 << 21):
        st = ctx.render(render_params(w, h, spp, max_depth=8, state_count=pool, film_fp64=1))
        got.append((st.extension_rays, st.shadow_rays, ctx.film(w, h, fp64=True)[1]))
    (ea, sa, a), (eb, sb, b) = got
    assert ea == eb and sa == sb
    assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())
