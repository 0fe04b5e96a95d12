"""Pins the oracle's Array2D texture sampling (texture2d/array2d.cpp over
utils/cuda/texture.cpp:63-105): exact texel values at texel centres, an
independent float64 bilinear/address-mode restatement in numpy, wrap
periodicity, border colour, the sRGB decode table against IEC 61966-2-1
values, and a constant texture rendering exactly like the constant colour."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

MODES = ["clamp", "wrap", "mirror", "border"]


def tex_scene(texels, **kw):
    s = scenes.Scene()
    t = s.add_texture(texels, **kw)
    s.add_instance(s.add_geometry(scenes.uv_sphere(), uvs=scenes.uv_sphere_uvs()), s.add_diffuse([1, 1, 1], color_texture=t))
    return s


def rgba8(h, w, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4), dtype=np.uint8)


def decode(texels, srgb):
    x = texels.astype(np.float64) / 255.0
    if srgb:
        rgb = x[..., :3]
        x[..., :3] = np.where(rgb <= 0.04045, rgb / 12.92, ((rgb + 0.055) / 1.055) ** 2.4)
    return x


def address(i, n, mode):
    if mode == "wrap":
        return np.mod(i, n)
    if mode == "mirror":
        m = np.mod(i, 2 * n)
        return np.where(m < n, m, 2 * n - 1 - m)
    if mode == "border":
        return np.where((i < 0) | (i >= n), -1, i)
    return np.clip(i, 0, n - 1)


def ref_sample(values, u, v, mode_u, mode_v, filt, border):
    """float64 restatement: texel i centred at (i + 0.5) / n"""
    h, w = values.shape[:2]

    def texel(x, y):
        out = values[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)].copy()
        out[(x < 0) | (y < 0)] = border
        return out

    if filt == "point":
        return texel(address(np.floor(u * w).astype(np.int64), w, mode_u),
                     address(np.floor(v * h).astype(np.int64), h, mode_v))
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    a, b = (x - x0)[:, None], (y - y0)[:, None]
    xa, xb = address(x0, w, mode_u), address(x0 + 1, w, mode_u)
    ya, yb = address(y0, h, mode_v), address(y0 + 1, h, mode_v)
    return ((1 - a) * (1 - b) * texel(xa, ya) + a * (1 - b) * texel(xb, ya) + (1 - a) * b * texel(xa, yb)
            + a * b * texel(xb, yb))


@pytest.mark.parametrize("filt", ["point", "linear"])
@pytest.mark.parametrize("srgb", [False, True])
def test_texel_centres_exact(filt, srgb):
    t = rgba8(8, 16, 1)
    os_ = oracle.OracleScene(tex_scene(t, filter=filt, srgb_to_linear=srgb))
    y, x = np.mgrid[0:8, 0:16]
    uv = np.stack([(x.ravel() + 0.5) / 16, (y.ravel() + 0.5) / 8], axis=1).astype(np.float32)
    got = os_.texture_eval(0, uv)
    want = decode(t, srgb).reshape(-1, 4).astype(np.float32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("filt", ["point", "linear"])
def test_matches_float64_restatement(mode, filt):
    t = rgba8(12, 20, 2)
    border = (0.25, 0.5, 0.75, 1.0)
    os_ = oracle.OracleScene(tex_scene(t, address_u=mode, address_v=mode, filter=filt, border=border))
    uv = np.random.default_rng(3).uniform(-2.3, 3.1, (20000, 2)).astype(np.float32)
    got = os_.texture_eval(0, uv).astype(np.float64)
    want = ref_sample(decode(t, False), uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64), mode, mode, filt,
                      np.array(border))
    # point sampling is exact except where u * W rounds across an integer
    err = np.abs(got - want).max(axis=1)
    assert (err < 1e-5).mean() > (0.999 if filt == "point" else 1.0 - 1e-9)


def test_float_texels_and_mixed_modes():
    t = np.random.default_rng(4).normal(size=(6, 9, 4)).astype(np.float32)
    os_ = oracle.OracleScene(tex_scene(t, address_u="wrap", address_v="mirror"))
    uv = np.random.default_rng(5).uniform(-1.5, 2.5, (5000, 2)).astype(np.float32)
    got = os_.texture_eval(0, uv).astype(np.float64)
    want = ref_sample(t.astype(np.float64), uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64), "wrap",
                      "mirror", "linear", np.zeros(4))
    assert np.abs(got - want).max() < 1e-5


def test_wrap_is_periodic_and_border_is_constant():
    t = rgba8(16, 16, 6)
    os_ = oracle.OracleScene(tex_scene(t, address_u="wrap", address_v="wrap"))
    u = (np.arange(1024) / 1024.0).astype(np.float32)
    uv = np.stack([u, u[::-1]], axis=1)
    base = os_.texture_eval(0, uv)
    for k in (-3, 1, 7):
        assert np.array_equal(os_.texture_eval(0, uv + np.float32(k)), base)
    b = oracle.OracleScene(tex_scene(t, address_u="border", address_v="border", border=(0.1, 0.2, 0.3, 0.4)))
    far = np.array([[-5.0, 0.5], [0.5, 9.0], [3.0, -3.0]], np.float32)
    assert np.array_equal(b.texture_eval(0, far), np.tile(np.float32([0.1, 0.2, 0.3, 0.4]), (3, 1)))


def test_srgb_decode_values():
    t = np.array([[[0, 128, 255, 128]]], np.uint8)
    got = oracle.OracleScene(tex_scene(t, srgb_to_linear=True, filter="point")).texture_eval(0, [[0.5, 0.5]])[0]
    # IEC 61966-2-1: 0 -> 0, 255 -> 1, 128/255 -> 0.2158605; alpha stays linear
    assert got[0] == 0.0 and got[2] == 1.0
    assert abs(got[1] - 0.21586050011389926) < 1e-7
    assert got[3] == np.float32(128 / 255)


def test_constant_texture_renders_like_constant_colour():
    c = np.array([51, 153, 204, 255], np.uint8)
    tex = np.tile(c, (32, 32, 1))
    scene, cam = scenes.sphere_field(instances=16, lights=4, seed=3, extent=8.0)
    plain, _ = scenes.sphere_field(instances=16, lights=4, seed=3, extent=8.0)
    t = scene.add_texture(tex)
    for m in scene.materials:
        if m["type"] == _abi.MATERIAL_DISNEY:
            m["color_texture"] = t
    for m in plain.materials:
        if m["type"] == _abi.MATERIAL_DISNEY:
            m["color"] = [float(np.float32(v / 255.0)) for v in c[:3]]
    p = render_params(24, 16, 2)
    a = oracle.OracleScene(scene).render(cam, p)[0]
    b = oracle.OracleScene(plain).render(cam, p)[0]
    assert np.array_equal(a, b)
