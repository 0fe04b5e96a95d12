"""Closed-form scenes that pin the estimator (generate -> shade -> shadow, MIS,
light sampling, Russian roulette) independently of any restatement: the CPU
oracle and the GPU path tracer must both reproduce them."""
import numpy as np
import pytest

from btrc_amd import scenes
from oracle import oracle

from util import render_params


def furnace_scene(albedo, L):
    """a convex Lambert object under a constant environment: every reflected
    direction escapes, so radiance leaving it is albedo * L"""
    s = scenes.Scene()
    g = s.add_geometry(scenes.uv_sphere())
    s.add_instance(g, s.add_diffuse([albedo] * 3, shadow_terminator=False), scenes.scale(1.0))
    s.set_gradient_sky((0, 1, 0), (L, L, L), (L, L, L))
    cam = scenes.Camera((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 30.0)
    return s, cam


def closed_box_scene(albedo, Le):
    """inside a closed box whose six walls emit Le and reflect with albedo a:
    with depth capped at D and no roulette the pixel value is
    Le * (1 + a + ... + a^D)"""
    s = scenes.Scene()
    for f in scenes.box_faces(2.0, 2.0, 2.0):
        inward = f[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]]  # swap b and c: normals point inside
        s.add_instance(s.add_geometry(inward), s.add_diffuse([albedo] * 3), light=(Le, Le, Le))
    cam = scenes.Camera((0.1, -0.2, 0.3), (0.5, 0.3, -1.0), (0.0, 1.0, 0.0), 70.0)
    return s, cam


def check_furnace(img, albedo, L, spp):
    c = img.shape[0] // 2
    inner = img[c - 3:c + 3, c - 3:c + 3]     # well inside the silhouette
    corner = img[:2, :2]                        # environment only
    assert np.all(corner == L)                  # camera rays that miss see L exactly
    m, s = inner.mean(), inner.std() / np.sqrt(inner.size)
    assert abs(m - albedo * L) < 5 * s + 1e-3, (m, albedo * L, s)


def check_box(img, albedo, Le, D):
    want = Le * (1 - albedo ** (D + 1)) / (1 - albedo)
    m, s = img.mean(), img.std() / np.sqrt(img.size)
    assert abs(m - want) < 5 * s + 1e-4 * want, (m, want, s)


def oracle_image(scene, cam, p):
    film, _, _ = oracle.OracleScene(scene).render(cam, p)
    return film / p.spp


def test_furnace_oracle():
    scene, cam = furnace_scene(0.6, 1.5)
    p = render_params(32, 32, 64, min_depth=5, max_depth=8)
    check_furnace(oracle_image(scene, cam, p), 0.6, 1.5, 64)


def test_closed_box_oracle():
    scene, cam = closed_box_scene(0.5, 2.0)
    D = 4
    p = render_params(16, 16, 32, min_depth=D, max_depth=D)
    check_box(oracle_image(scene, cam, p), 0.5, 2.0, D)


@pytest.mark.gpu
def test_furnace_gpu(ctx):
    scene, cam = furnace_scene(0.6, 1.5)
    p = render_params(128, 128, 64, min_depth=5, max_depth=8)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    ctx.render(p)
    img, _ = ctx.film(p.width, p.height)
    check_furnace(img[..., :3].astype(np.float64), 0.6, 1.5, 64)


@pytest.mark.gpu
def test_closed_box_gpu(ctx):
    scene, cam = closed_box_scene(0.5, 2.0)
    D = 4
    p = render_params(128, 128, 64, min_depth=D, max_depth=D)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    ctx.render(p)
    img, _ = ctx.film(p.width, p.height)
    check_box(img[..., :3].astype(np.float64), 0.5, 2.0, D)
