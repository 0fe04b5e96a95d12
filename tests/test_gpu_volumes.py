"""The volume-primitive medium on the GPU against the oracle: tracking bit
for bit on segments, and renders per pixel with volumes alone, volumes under
an environment light, and volumes next to instance media."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

pytestmark = pytest.mark.gpu


def volume_scene():
    s = scenes.Scene()
    grid = scenes.fbm_grid(16, octaves=3, seed=7, max_sigma=3.0)
    alb = np.random.default_rng(3).uniform(0.3, 0.9, grid.shape + (3,)).astype(np.float32)
    s.add_volume((-0.6, 0.1, -0.6), (1.2, 0, 0), (0, 1.0, 0), (0, 0, 1.2), grid, alb)
    s.add_volume((0.0, 0.3, -0.2), (0.8, 0, 0), (0, 0.8, 0), (0, 0, 0.8), 1.5, (0.8, 0.6, 0.4))
    c, sn = np.cos(0.7), np.sin(0.7)
    s.add_volume((-0.8, 1.2, 0.2), (0.5 * c, 0.5 * sn, 0), (-0.5 * sn, 0.5 * c, 0), (0, 0, 0.5), 2.0, (1, 1, 1))
    return s


def test_tracking_bitwise_vs_oracle(ctx):
    scene = volume_scene()
    ctx.upload_scene(scene)
    rng = np.random.default_rng(11)
    n = 20000
    seg = np.zeros(n, dtype=_abi.ray_dtype())
    a = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32) + np.float32([0, 0.8, 0])
    b = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32) + np.float32([0, 0.8, 0])
    seg["o"], seg["d"], seg["tmax"] = a, b - a, 1.0
    seeds = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    got = ctx.eval_volumes(seg, seeds)
    want = oracle.OracleScene(scene).volume_eval(seg, seeds)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert 0.1 < got[:, 0].mean() < 0.9


def cornell_with_volumes():
    base, cam = scenes.cornell()
    v = volume_scene()
    base.volumes = v.volumes
    return base, cam


def open_scene_with_volumes():
    s = volume_scene()
    m = s.add_diffuse([0.6, 0.6, 0.6])
    s.add_instance(s.add_geometry(scenes.quad((-3, 0, 3), (3, 0, 3), (3, 0, -3), (-3, 0, -3))), m)
    s.set_gradient_sky((0, 1, 0), (0.2, 0.3, 0.5), (1.0, 0.9, 0.8))
    cam = scenes.Camera((0.0, 1.2, 3.5), (0.0, 0.6, 0.0), (0.0, 1.0, 0.0), 50.0)
    return s, cam


def media_and_volumes():
    s, cam = scenes.cornell_media(grid=scenes.fbm_grid(16, octaves=3, seed=3, max_sigma=0.3))
    s.add_volume((-20.0, 5.0, -20.0), (10, 0, 0), (0, 10, 0), (0, 0, 10), 0.05, (0.9, 0.9, 0.9))
    return s, cam


SCENES = {"cornell": cornell_with_volumes, "open-sky": open_scene_with_volumes, "with-media": media_and_volumes}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("sampler", [_abi.SAMPLER_HALTON, _abi.SAMPLER_INDEPENDENT], ids=["halton", "indep"])
def test_render_parity(ctx, name, sampler):
    scene, cam = SCENES[name]()
    w, h, spp = 40, 32, 4
    p = render_params(w, h, spp, min_depth=3, max_depth=12, state_count=2000)
    p.sampler = sampler
    ctx.upload_scene(scene)
    ctx.set_camera(cam, w, h)
    stats = ctx.render(p)
    rgba, _ = ctx.film(w, h)
    ref, e, s = oracle.OracleScene(scene).render(cam, p)
    ref /= spp
    assert stats.extension_rays == e and stats.shadow_rays == s
    rel = np.abs(rgba[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()} at {np.unravel_index(rel.argmax(), rel.shape)}"
    assert ref.sum() > 0


def test_bad_volumes_rejected(ctx):
    import btrc_amd
    s = volume_scene()
    s.volumes[0]["x"] = (0.0, 0.0, 0.0)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(s)
    s = volume_scene()
    s.volumes[1]["sigma_t"] = -1.0
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(s)
