"""GPU parity of the participating-media path (medium stage, shadow-ray
transmittance, invisible boundaries) against the oracle: the tracking
estimators bit for bit on the same streams, and whole renders per pixel."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

pytestmark = pytest.mark.gpu


def medium_only(add):
    s = scenes.Scene()
    # one tiny triangle far away: the context needs a scene to hold the medium
    s.add_instance(s.add_geometry(scenes.quad((90, 90, 90), (91, 90, 90), (91, 91, 90), (90, 91, 90))),
                   s.add_diffuse([0.5, 0.5, 0.5]))
    add(s)
    return s


MEDIA = {
    "homogeneous": lambda s: s.add_homogeneous(0.8, (0.3, 0.6, 0.9), g=-0.4),
    "grid": lambda s: s.add_heterogeneous(scenes.fbm_grid(20, octaves=3, seed=4, max_sigma=2.0), (0.8, 0.8, 0.8),
                                          g=0.6, world_to_local=np.diag([0.25, 0.25, 0.25, 1.0])),
    "grid-albedo": lambda s: s.add_heterogeneous(
        scenes.fbm_grid(12, octaves=2, seed=5, max_sigma=3.0), (0.5, 0.5, 0.5), g=0.2,
        albedo_grid=np.random.default_rng(1).random((12, 12, 12, 3)).astype(np.float32)),
}


@pytest.mark.parametrize("kind", list(MEDIA))
def test_tracking_bitwise(ctx, kind):
    scene = medium_only(MEDIA[kind])
    ctx.upload_scene(scene)
    rng = np.random.default_rng(8)
    n = 20000
    seg = np.zeros(n, dtype=_abi.ray_dtype())
    seg["o"] = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    seg["d"] = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    seg["tmax"] = 1.0
    seeds = rng.integers(1, 2 ** 63, n, dtype=np.uint64)
    gpu = ctx.eval_medium(0, seg, seeds)
    ref = oracle.OracleScene(scene).medium_eval(0, seg, seeds)
    diff = np.nonzero((gpu.view(np.uint32) != ref.view(np.uint32)).any(axis=1))[0]
    assert len(diff) == 0, f"{len(diff)} segments differ, first {diff[:3]}: {gpu[diff[:2]]} vs {ref[diff[:2]]}"
    assert 0.05 < gpu[:, 0].mean() < 0.95 and 0.05 < gpu[:, 7].mean() < 0.95


def homogeneous_box():
    """Cornell-style room with a homogeneous medium filling an invisible box
    that also encloses one of the inner boxes"""
    base, cam = scenes.cornell()
    s = scenes.Scene()
    s.geometries, s.materials = base.geometries, base.materials
    for inst in base.instances:
        s.add_instance(inst["geometry"], inst["material"], inst["mat"], light=inst["light"])
    med = s.add_homogeneous(1.2, (0.7, 0.8, 0.9), g=-0.3)
    box = np.concatenate(scenes.box_faces(1.0, 1.0, 1.0))
    s.add_instance(s.add_geometry(box), s.add_invisible(), scenes.translate(0.3, 0.5, 0.3), inner_medium=med)
    return s, cam


SCENES = {
    "homogeneous": homogeneous_box,
    "grid": lambda: scenes.cornell_media(grid=scenes.fbm_grid(24, octaves=3, seed=3, max_sigma=0.3)),
}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("sampler", [_abi.SAMPLER_HALTON, _abi.SAMPLER_INDEPENDENT], ids=["halton", "indep"])
def test_media_render_parity(ctx, name, sampler):
    scene, cam = SCENES[name]()
    w, h, spp = 48, 40, 4
    p = render_params(w, h, spp, min_depth=3, max_depth=10, state_count=2500)
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
    assert stats.stage_launches[4] == 0 or stats.stage_ms[4] >= 0
    assert ref.sum() > 0


@pytest.mark.parametrize("name", list(SCENES))
def test_media_pool_sizes_give_one_image(ctx, name):
    """the medium stage hands states to its lanes in batches and queues;
    pools smaller than one batch, of odd sizes and large ones trace the same
    paths: equal ray counts and FP64 films"""
    scene, cam = SCENES[name]()
    w, h, spp = 40, 32, 3
    ctx.upload_scene(scene)
    ctx.set_camera(cam, w, h)
    got = []
    for pool in (37, 1000, 4097, 1 << 16):
        st = ctx.render(render_params(w, h, spp, min_depth=3, max_depth=10, state_count=pool, film_fp64=1))
        got.append((st.extension_rays, st.shadow_rays, ctx.film(w, h, fp64=True)[1]))
    for e, s, img in got[1:]:
        assert (e, s) == got[0][:2]
        assert np.abs(img - got[0][2]).max() <= 1e-12 * max(1.0, np.abs(got[0][2]).max())
