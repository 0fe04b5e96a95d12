"""Closed-form scenes that pin the estimator (generate -> shade -> shadow, MIS,
light sampling, Russian roulette) independently of any restatement: the CPU
oracle and the GPU path tracer must both reproduce them.

Russian roulette (simple_russian_roulette, renderer/wavefront/helper.cpp:67-97)
is pinned by the closed box with roulette on every bounce from depth 1:
the pixel mean stays Le (1 - a^(D+1)) / (1 - a) only if survivors are divided
by the continuation probability, and the number of extension rays per path
follows the truncated geometric law of the survival draws (u <= p continues,
luminance(beta) < threshold selects the roulette), which a swapped comparison
or a missing threshold test changes."""
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


def rr_ray_law(m, D, p):
    """mean and variance of extension rays per path in a closed box when
    every vertex at depth m..D-1 plays roulette with continuation p: the
    camera ray and depths 1..m always exist, then K survivals with
    P(K = k) = p^k (1 - p) for k < D - m and P(K = D - m) = p^(D - m)"""
    ks = np.arange(D - m + 1)
    prob = p ** ks * (1 - p)
    prob[-1] = p ** (D - m)
    rays = m + 1 + ks
    mean = float((prob * rays).sum())
    return mean, float((prob * (rays - mean) ** 2).sum())


def check_rr_rays(ext_rays, paths, m, D, p, rr_on):
    if not rr_on:
        assert ext_rays == paths * (D + 1)   # no roulette: every path reaches depth D
        return
    mean, var = rr_ray_law(m, D, p)
    got = ext_rays / paths
    assert abs(got - mean) < 5 * np.sqrt(var / paths), (got, mean)


RR_CASES = [(0.5, 10.0), (0.8, 10.0), (0.8, 0.0)]  # (continuation probability, luminance threshold)


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


@pytest.mark.parametrize("cont,thr", RR_CASES)
def test_closed_box_rr_oracle(cont, thr):
    """roulette at every bounce from depth 1 (threshold above any throughput
    of the box) or never (threshold 0)"""
    a, Le, m, D = 0.5, 2.0, 1, 6
    scene, cam = closed_box_scene(a, Le)
    p = render_params(16, 16, 32, min_depth=m, max_depth=D, rr_threshold=thr, rr_cont_prob=cont)
    film, ext, _ = oracle.OracleScene(scene).render(cam, p)
    check_box(film / p.spp, a, Le, D)
    check_rr_rays(ext, 16 * 16 * 32, m, D, cont, thr > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("cont,thr", RR_CASES)
def test_closed_box_rr_gpu(ctx, cont, thr):
    a, Le, m, D = 0.5, 2.0, 1, 6
    scene, cam = closed_box_scene(a, Le)
    p = render_params(128, 128, 64, min_depth=m, max_depth=D, rr_threshold=thr, rr_cont_prob=cont)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    stats = ctx.render(p)
    img, _ = ctx.film(p.width, p.height)
    check_box(img[..., :3].astype(np.float64), a, Le, D)
    check_rr_rays(stats.extension_rays, 128 * 128 * 64, m, D, cont, thr > 0)


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
