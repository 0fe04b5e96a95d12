"""The image-based environment light on the GPU against the oracle: the
deterministic atan2 / acos bit for bit, the importance table and the light's
eval_le / pdf_li / sample_li bit for bit, renders per pixel (both samplers,
with a medium), and the diffuse sphere under a two-level sky against its
closed form."""
import numpy as np
import pytest

import btrc_amd
from btrc_amd import _abi, scenes
from oracle import oracle

from test_oracle_envir import check_sphere_under_sky, ibl_scene, index_texture, sphere_under_sky, two_level_texture
from test_oracle_texture import rgba8
from util import render_params

pytestmark = pytest.mark.gpu


def test_atan2_acos_bitwise_equal_to_host(ctx):
    rng = np.random.default_rng(21)
    n = 200000
    y = (rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))).astype(np.float32)
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))).astype(np.float32)
    y[:8] = [0.0, -0.0, 0.0, -0.0, np.inf, -np.inf, np.inf, 1.0]
    x[:8] = [0.0, 0.0, -0.0, -0.0, np.inf, np.inf, -np.inf, -np.inf]
    assert np.array_equal(ctx.eval_detmath(7, y, x).view(np.uint32), oracle.detmath(7, y, x).view(np.uint32))
    c = np.concatenate([rng.uniform(-1, 1, n), 1 - np.exp(rng.uniform(-17, 0, 1000)), [1.0, -1.0, 0.0, 1.5]])
    c = c.astype(np.float32)
    got, want = ctx.eval_detmath(8, c, c), oracle.detmath(8, c, c)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32))


TEXTURES = {
    "sky": lambda: (scenes.sky_texture(64, 32, seed=3), dict()),
    "index_point": lambda: (index_texture(8, 4), dict(filter="point")),
    "srgb8_wrap": lambda: (rgba8(37, 19, 4), dict(address_u="wrap", srgb_to_linear=True)),
    "two_level": lambda: (two_level_texture(16, 16, 3.0, 1.0), dict(filter="point")),
}


def texture_scene(kind, up, lw=40, lh=20):
    texels, kw = TEXTURES[kind]()
    s = scenes.Scene()
    s.set_ibl(s.add_texture(texels, **kw), up=up, lut_width=lw, lut_height=lh)
    return s


@pytest.mark.parametrize("kind", list(TEXTURES))
@pytest.mark.parametrize("up", [(0.0, 0.0, 1.0), (0.3, -1.0, 0.2), (0.99, 0.1, 0.0)])
def test_table_and_queries_bitwise_vs_oracle(ctx, kind, up):
    scene = texture_scene(kind, up)
    ctx.upload_scene(scene)
    o = oracle.OracleScene(scene)
    assert np.array_equal(ctx.envir_table().view(np.uint32), o.envir_table().view(np.uint32))
    rng = np.random.default_rng(31)
    d = rng.standard_normal((60000, 3)).astype(np.float32)
    d[:6] = [[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]]
    sm = rng.uniform(0, 1, (len(d), 3)).astype(np.float32)
    sm[:4] = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.999999, 0.999999, 0.999999], [0.25, 0.0, 1.0]]
    got, want = ctx.eval_envir(d, sm), o.envir_eval(d, sm)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_default_table_size_bitwise(ctx):
    s = scenes.Scene()
    s.set_ibl(s.add_texture(scenes.sky_texture(256, 128, seed=9)), up=(0.0, 1.0, 0.0))
    ctx.upload_scene(s)
    t = ctx.envir_table()
    assert t.shape == (400 * 200,)
    assert np.array_equal(t.view(np.uint32), oracle.OracleScene(s).envir_table().view(np.uint32))


def test_bad_ibl_arguments(ctx):
    s = texture_scene("sky", (0.0, 0.0, 1.0))
    s.envir = ("ibl", 3, (0.0, 0.0, 1.0), 40, 20)      # no such texture
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(s)
    s.envir = ("ibl", 0, (0.0, 0.0, 1.0), 0, 20)       # empty table
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(s)
    s.envir = ("ibl", 0, (0.0, 0.0, 0.0), 40, 20)      # no up direction
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(s)
    s.envir = None
    ctx.upload_scene(s)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.eval_envir(np.zeros((1, 3), np.float32), np.zeros((1, 3), np.float32))
    assert ctx.envir_table().size == 0


def render_pair(ctx, scene, cam, p):
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    stats = ctx.render(p)
    rgba, _ = ctx.film(p.width, p.height)
    ref, e, s = oracle.OracleScene(scene).render(cam, p)
    return rgba[..., :3], ref / p.spp, stats, e, s


@pytest.mark.parametrize("sampler", [_abi.SAMPLER_HALTON, _abi.SAMPLER_INDEPENDENT], ids=["halton", "indep"])
def test_ibl_render_per_pixel_parity(ctx, sampler):
    scene, cam = scenes.sphere_field(instances=64, lights=8, seed=11, extent=16.0)
    scene.set_ibl(scene.add_texture(scenes.sky_texture(128, 64, seed=5)), up=(0.0, 1.0, 0.0), lut_width=64,
                  lut_height=32)
    p = render_params(96, 64, 8, state_count=4096)
    p.sampler = sampler
    img, ref, stats, e, s = render_pair(ctx, scene, cam, p)
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s


def test_ibl_with_media_per_pixel_parity(ctx):
    scene, cam = scenes.cornell_media(grid=scenes.fbm_grid(24, octaves=3, seed=3, max_sigma=0.3))
    scene.set_ibl(scene.add_texture(scenes.sky_texture(64, 32, seed=6)), up=(0.0, 1.0, 0.0), lut_width=32,
                  lut_height=16)
    p = render_params(48, 40, 4, min_depth=3, max_depth=10, state_count=2500)
    img, ref, stats, e, s = render_pair(ctx, scene, cam, p)
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s


def test_diffuse_sphere_under_a_two_level_sky_gpu(ctx):
    albedo, lt, lb = 0.7, 2.0, 0.5
    s, cam = sphere_under_sky(albedo, lt, lb)
    p = render_params(48, 48, 256, min_depth=5, max_depth=6)
    ctx.upload_scene(s)
    ctx.set_camera(cam, p.width, p.height)
    ctx.render(p)
    img, _ = ctx.film(p.width, p.height)
    check_sphere_under_sky(img[..., :3].astype(np.float64), albedo, lt, lb)
