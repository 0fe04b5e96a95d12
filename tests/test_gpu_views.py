"""Batched multi-view rendering: each view of one btrc_hip_render_views
call gets the samples a render of that view alone gets."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

pytestmark = pytest.mark.gpu


def field_views():
    scene, cam = scenes.sphere_field(instances=128, lights=8, seed=12, extent=20.0)
    c = 10.0
    views = [(cam, 64, 48),
             (scenes.Camera((c, 35.0, c), (c, c, c + 0.5), (0.0, 0.0, 1.0), 50.0), 40, 40),
             (scenes.Camera((-8.0, c, c), (c, c, c), (0.0, 1.0, 0.0), 70.0), 96, 32)]
    return scene, views


@pytest.mark.parametrize("sampler,tile", [(_abi.SAMPLER_HALTON, 0), (_abi.SAMPLER_INDEPENDENT, 0),
                                          (_abi.SAMPLER_HALTON, 1)])
def test_views_equal_separate_renders(ctx, sampler, tile):
    scene, views = field_views()
    ctx.upload_scene(scene)
    p = render_params(0, 0, 4, state_count=3000, film_fp64=1)
    p.sampler, p.tile = sampler, tile
    stats = ctx.render_views(p, views)
    total = sum(w * h for _, w, h in views)
    rgba_all, rgb64_all = ctx.film(total, 1, fp64=True)
    rgb64_all = rgb64_all.reshape(-1, 3)
    batches = [ctx.view_film(k, w, h) for k, (_, w, h) in enumerate(views)]
    base, ext, shadow = 0, 0, 0
    for k, (cam, w, h) in enumerate(views):
        assert np.array_equal(batches[k], rgba_all.reshape(-1, 4)[base:base + w * h].reshape(h, w, 4))
        batch64 = rgb64_all[base:base + w * h].reshape(h, w, 3)
        q = render_params(w, h, 4, state_count=3000, film_fp64=1)
        q.sampler, q.tile = sampler, tile
        ctx.set_camera(cam, w, h)
        s1 = ctx.render(q)
        _, alone64 = ctx.film(w, h, fp64=True)
        assert np.abs(batch64 - alone64).max() <= 1e-12 * max(1.0, np.abs(alone64).max())
        ext += s1.extension_rays
        shadow += s1.shadow_rays
        base += w * h
    assert stats.extension_rays == ext and stats.shadow_rays == shadow
    assert stats.paths == total * 4


def test_views_match_oracle(ctx):
    scene, views = field_views()
    ctx.upload_scene(scene)
    p = render_params(0, 0, 4, state_count=4096)
    ctx.render_views(p, views)
    for k, (cam, w, h) in enumerate(views):
        got = ctx.view_film(k, w, h)
        ref = oracle.OracleScene(scene).render(cam, render_params(w, h, 4, state_count=4096))[0] / 4
        rel = np.abs(got[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
        assert rel.max() <= 1e-4, f"view {k}: max rel err {rel.max()}"


def test_view_arguments(ctx):
    import btrc_amd
    scene, views = field_views()
    ctx.upload_scene(scene)
    p = render_params(0, 0, 1)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.render_views(p, [])
    with pytest.raises(btrc_amd.BtrcException):
        ctx.render_views(p, [(views[0][0], 0, 10)])
    with pytest.raises(btrc_amd.BtrcException):
        ctx.render_views(p, [views[0]] * 65)
    ctx.render_views(p, views)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.view_film(3, 8, 8)
