"""Pins the oracle's textured material parameters and normal maps
(DisneyMaterial::create_shader, material/disney.cpp:813-842; DiffuseMaterial,
material/diffuse.cpp:21-33; NormalMap, material/utils/normal_map.cpp:13-21):

- every Disney parameter read from a one-texel texture renders exactly like
  the same constant, and the constant matters (a different value changes
  the image), so a texture wired to the wrong parameter fails;
- a flat normal map (RGB 0.5, 0.5, 1: local +z) leaves the image unchanged up
  to the rounding of re-normalising the frame;
- a constant tilted normal map (RGB = (1 + sin a, 1, 1 + cos a) / 2, local
  (sin a, 0, cos a)) turns the first-hit shading normal by exactly a from
  the surface normal, for Disney and diffuse materials.
"""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

VALUES = dict(metallic=0.3, roughness=0.35, specular=0.7, specular_tint=0.4, anisotropic=0.5, sheen=0.6,
              sheen_tint=0.3, clearcoat=0.8, clearcoat_gloss=0.4, transmission=0.3, transmission_roughness=0.2,
              ior=1.33)
OTHER = dict(metallic=0.9, roughness=0.8, specular=0.1, specular_tint=0.9, anisotropic=0.0, sheen=0.0,
             sheen_tint=0.9, clearcoat=0.1, clearcoat_gloss=0.95, transmission=0.9, transmission_roughness=0.7,
             ior=2.0)


def disney_field(values):
    scene, cam = scenes.sphere_field(instances=24, lights=4, seed=6, extent=8.0)
    for m in scene.materials:
        if m["type"] == _abi.MATERIAL_DISNEY:
            m.update(values)
    return scene, cam


def one_texel(scene, value):
    return scene.add_texture(np.full((1, 1, 4), value, np.float32), filter="point")


@pytest.mark.parametrize("name", sorted(VALUES))
def test_each_parameter_texture_renders_like_its_constant(name):
    p = render_params(24, 16, 2)
    plain, cam = disney_field(VALUES)
    textured, _ = disney_field(dict(VALUES, **{name: OTHER[name]}))   # the texture must win over the constant
    t = one_texel(textured, VALUES[name])
    for m in textured.materials:
        if m["type"] == _abi.MATERIAL_DISNEY:
            m[name + "_texture"] = t
    other, _ = disney_field(dict(VALUES, **{name: OTHER[name]}))
    a = oracle.OracleScene(plain).render(cam, p)[0]
    b = oracle.OracleScene(textured).render(cam, p)[0]
    c = oracle.OracleScene(other).render(cam, p)[0]
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


# the quad's plane is turned away from the world axes, so that a normal map
# read in world instead of shading coordinates changes the angle
TILT = scenes.rotate_y(0.5) @ np.array([[1, 0, 0, 0], [0, np.cos(0.4), -np.sin(0.4), 0],
                                        [0, np.sin(0.4), np.cos(0.4), 0], [0, 0, 0, 1]])
SURFACE_NORMAL = TILT[:3, :3] @ np.array([0.0, 0.0, 1.0])


def flat_quad_scene(material):
    s = scenes.Scene()
    q = scenes.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))
    uvs = np.array([[0, 0, 1, 0, 1, 1], [0, 0, 1, 1, 0, 1]], np.float32)
    mat = material(s)
    s.add_instance(s.add_geometry(q, uvs=uvs), mat, TILT)
    s.set_gradient_sky((0, 1, 0), (0.3, 0.3, 0.3), (1.0, 1.0, 1.0))
    eye = 3.0 * SURFACE_NORMAL
    cam = scenes.Camera(tuple(eye), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 20.0)
    return s, cam


@pytest.mark.parametrize("kind", ["disney", "diffuse"])
def test_flat_normal_map_changes_nothing(kind):
    def mat(s, nmap):
        t = one_texel(s, [0.5, 0.5, 1.0, 1.0]) if nmap else None
        if kind == "disney":
            return s.add_disney([0.6, 0.5, 0.4], 0.2, 0.4, normal_texture=t)
        return s.add_diffuse([0.6, 0.5, 0.4], normal_texture=t)
    p = render_params(16, 16, 8)
    s0, cam = flat_quad_scene(lambda s: mat(s, False))
    s1, _ = flat_quad_scene(lambda s: mat(s, True))
    a = oracle.OracleScene(s0).render(cam, p)[0]
    b = oracle.OracleScene(s1).render(cam, p)[0]
    assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


@pytest.mark.parametrize("kind", ["disney", "diffuse"])
@pytest.mark.parametrize("deg", [25.0, 60.0])
def test_tilted_normal_map_turns_the_shading_normal(kind, deg):
    a = np.radians(deg)
    rgb = [(1 + np.sin(a)) / 2, 0.5, (1 + np.cos(a)) / 2, 1.0]

    def mat(s):
        t = one_texel(s, rgb)
        if kind == "disney":
            return s.add_disney([0.6, 0.5, 0.4], 0.2, 0.4, normal_texture=t)
        return s.add_diffuse([0.6, 0.5, 0.4], normal_texture=t)
    s, cam = flat_quad_scene(mat)
    p = render_params(16, 16, 4)
    _, _, _, _, nsum = oracle.OracleScene(s).render(cam, p, aov=True)
    c = 8   # the quad fills the centre of the image
    n = nsum[c - 2:c + 2, c - 2:c + 2].reshape(-1, 3)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    cosines = n @ SURFACE_NORMAL
    assert np.abs(cosines - np.cos(a)).max() < 2e-5, (cosines, np.cos(a))


def test_normal_map_changes_the_image():
    def mat(s, rgb):
        return s.add_disney([0.6, 0.5, 0.4], 0.0, 0.3, normal_texture=one_texel(s, rgb))
    p = render_params(16, 16, 8)
    s0, cam = flat_quad_scene(lambda s: mat(s, [0.5, 0.5, 1.0, 1.0]))
    s1, _ = flat_quad_scene(lambda s: mat(s, [0.8, 0.5, 0.9, 1.0]))
    a = oracle.OracleScene(s0).render(cam, p)[0]
    b = oracle.OracleScene(s1).render(cam, p)[0]
    assert np.abs(a - b).max() > 0.01
