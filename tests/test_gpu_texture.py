"""Array2D textures on the GPU against the oracle, bit for bit, and a
textured render per pixel."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from test_oracle_texture import MODES, rgba8, tex_scene
from util import render_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("filt", ["point", "linear"])
@pytest.mark.parametrize("srgb", [False, True])
def test_sample_bitwise_vs_oracle(ctx, mode, filt, srgb):
    scene = tex_scene(rgba8(37, 53, 7), address_u=mode, address_v="clamp" if mode == "wrap" else mode, filter=filt,
                      srgb_to_linear=srgb, border=(0.5, 0.25, 0.125, 1.0))
    ctx.upload_scene(scene)
    rng = np.random.default_rng(8)
    uv = rng.uniform(-3.0, 4.0, (50000, 2)).astype(np.float32)
    uv[:4] = [[np.nan, 0.5], [0.5, np.inf], [-1e30, 1e30], [1e9, -7e8]]  # non-finite and huge coordinates
    got = ctx.eval_texture(0, uv)
    want = oracle.OracleScene(scene).texture_eval(0, uv)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_float_texture_bitwise_vs_oracle(ctx):
    t = np.random.default_rng(9).normal(size=(64, 32, 4)).astype(np.float32)
    scene = tex_scene(t, address_u="wrap", address_v="mirror")
    ctx.upload_scene(scene)
    uv = np.random.default_rng(10).uniform(-2, 2, (20000, 2)).astype(np.float32)
    assert np.array_equal(ctx.eval_texture(0, uv).view(np.uint32),
                          oracle.OracleScene(scene).texture_eval(0, uv).view(np.uint32))


def test_bad_texture_arguments(ctx):
    import btrc_amd
    s = tex_scene(rgba8(4, 4, 1))
    ctx.upload_scene(s)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.eval_texture(1, np.zeros((1, 2), np.float32))
    s.materials[0]["color_texture"] = 5   # no such texture
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(s)


def test_textured_render_per_pixel_parity(ctx):
    scene, cam = scenes.sphere_field(instances=64, lights=8, seed=11, extent=16.0)
    uvs = scenes.uv_sphere_uvs()
    scene.geometries[0]["uvs"] = uvs
    texs = [scene.add_texture(scenes.noise_texture(64, seed=k), address_u="wrap", srgb_to_linear=(k == 1))
            for k in range(3)]
    for k, m in enumerate(scene.materials):
        if m["type"] == _abi.MATERIAL_DISNEY:
            m["color_texture"] = texs[k % 3]
            if k % 2:
                m["roughness_texture"] = texs[(k + 1) % 3]
    scene.accel = _abi.ACCEL_TWO_LEVEL
    p = render_params(96, 64, 8, state_count=4096)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    stats = ctx.render(p)
    rgba, _ = ctx.film(p.width, p.height)
    ref, e, s = oracle.OracleScene(scene).render(cam, p)
    ref /= p.spp
    rel = np.abs(rgba[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s


def test_every_parameter_textured_and_normal_maps_per_pixel(ctx):
    """noise textures on all thirteen Disney parameters and normal maps on
    Disney and diffuse materials: image and first-hit g-buffer per pixel"""
    scene, cam = scenes.sphere_field(instances=48, lights=6, seed=21, extent=12.0)
    scene.geometries[0]["uvs"] = scenes.uv_sphere_uvs()
    rng = np.random.default_rng(4)

    def float_tex(lo, hi, size=32):
        t = rng.uniform(lo, hi, (size, size, 4)).astype(np.float32)
        return scene.add_texture(t, address_u="wrap")
    unit = [float_tex(0.0, 1.0) for _ in range(4)]
    ior = float_tex(1.2, 1.8)
    nm = rng.uniform(0.3, 0.7, (32, 32, 4)).astype(np.float32)
    nm[..., 2] = rng.uniform(0.85, 1.0, (32, 32))
    nmap = scene.add_texture(nm, address_u="wrap")
    names = [n for n in _abi.MATERIAL_TEXTURE_SLOTS if n not in ("color", "normal", "ior")]
    k = 0
    for m in scene.materials:
        if m["type"] != _abi.MATERIAL_DISNEY:
            continue
        m["color_texture"] = unit[k % 4]
        for j, name in enumerate(names):
            if (k + j) % 3:
                m[name + "_texture"] = unit[(k + j) % 4]
        m["ior_texture"] = ior
        if k % 2:
            m["normal_texture"] = nmap
        k += 1
    for i in range(0, len(scene.instances), 5):
        if scene.instances[i].get("light") is None:
            scene.instances[i]["material"] = scene.add_diffuse([0.7, 0.6, 0.5], normal_texture=nmap)
    p = render_params(96, 64, 8, state_count=4096)
    p.albedo, p.normal = 1, 1
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    stats = ctx.render(p)
    rgba, _ = ctx.film(p.width, p.height)
    alb, nrm = ctx.aov(p.width, p.height)
    ref, e, s, ralb, rnrm = oracle.OracleScene(scene).render(cam, p, aov=True)
    ref /= p.spp
    rel = np.abs(rgba[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s
    assert np.abs(alb[..., :3] - ralb / p.spp).max() <= 1e-5
    n = rnrm / p.spp
    ln = np.linalg.norm(n, axis=2, keepdims=True)
    n = np.where(ln > 1e-3, n / np.maximum(ln, 1e-30), 0.0)
    assert np.abs(nrm[..., :3] - n).max() <= 1e-4
