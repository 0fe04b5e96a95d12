"""GPU parity of the generate / film side: the independent sampler, tile
path order, the Gaussian film filter and the albedo / normal outputs, each
against the oracle on the same inputs."""
import numpy as np
import pytest

import btrc_amd
from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

pytestmark = pytest.mark.gpu


def test_independent_sampler_bitwise(ctx):
    rng = np.random.default_rng(12)
    n = 4096
    px = rng.integers(0, 4096, n)
    py = rng.integers(0, 4096, n)
    si = rng.integers(0, 5000, n)
    si[:4] = [0, 1, 2 ** 15, 31103]
    gpu = ctx.eval_independent(px, py, si, 40)
    ref = oracle.independent(px, py, si, 40)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


VARIANTS = [
    dict(sampler=_abi.SAMPLER_INDEPENDENT, tile=0, filt=None),
    dict(sampler=_abi.SAMPLER_HALTON, tile=1, filt=("gaussian", 1.3, 1.6)),
    dict(sampler=_abi.SAMPLER_INDEPENDENT, tile=1, filt=("gaussian", 0.8, 3.0)),
]


@pytest.mark.parametrize("v", VARIANTS, ids=["indep", "halton-tile-gauss", "indep-tile-gauss"])
def test_render_variant_parity(ctx, v):
    scene, cam = scenes.cornell()
    w, h, spp = 80, 64, 6
    p = render_params(w, h, spp, state_count=3000)
    p.sampler, p.tile, p.albedo, p.normal = v["sampler"], v["tile"], 1, 1
    ctx.upload_scene(scene)
    ctx.set_camera(cam, w, h)
    try:
        if v["filt"]:
            ctx.set_filter(*v["filt"])
        stats = ctx.render(p)
        rgba, _ = ctx.film(w, h)
        alb, nrm = ctx.aov(w, h)
    finally:
        ctx.set_filter("box")
    ref, e, s, ralb, rnrm = oracle.OracleScene(scene).render(cam, p, filt=v["filt"], aov=True)
    ref /= spp
    rel = np.abs(rgba[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s
    # albedo: sums of material colours over first hits / spp
    assert np.abs(alb[..., :3] - ralb / spp).max() <= 1e-5
    # normal: the normalised mean shading normal, zero when it nearly cancels
    m = rnrm / spp
    ln = np.linalg.norm(m, axis=-1, keepdims=True)
    want = np.where(ln > 1e-3, m / np.maximum(ln, 1e-30), 0.0)
    assert np.abs(nrm[..., :3] - want).max() <= 1e-4
    assert (alb[..., 3] == 1).all() and (nrm[..., 3] == 1).all()


def test_renderer_outputs_and_errors(ctx):
    scene, cam = scenes.cornell()
    r = btrc_amd.WavefrontPathTracer(context=ctx, spp=2, state_count=1 << 12, albedo=True, normal=True,
                                     sampler="independent", tile=True)
    r.set_scene(scene)
    r.set_camera(cam)
    r.set_film(32, 24)
    r.set_film_filter("gaussian")
    try:
        res = r.render()
    finally:
        ctx.set_filter("box")
    assert res.albedo.shape == (24, 32, 4) and res.normal.shape == (24, 32, 4)
    # unit normals where camera rays hit the room, zero where all of a
    # pixel's rays leave through the open front (no albedo there either)
    ln = np.linalg.norm(res.normal[..., :3], axis=-1)
    unit = np.abs(ln - 1) < 1e-5
    assert (unit | (ln == 0)).all()
    assert (res.albedo[~unit, :3] == 0).all()
    assert unit.mean() > 0.8
    # a render without AOVs cannot read them back
    ctx.render(render_params(8, 8, 1))
    with pytest.raises(btrc_amd.BtrcException) as e:
        ctx.aov(8, 8)
    assert e.value.status == _abi.BTRC_ERR_STATE
    with pytest.raises(btrc_amd.BtrcException):
        ctx.set_filter("gaussian", radius=-1.0)
    bad = render_params(8, 8, 1)
    bad.sampler = 7
    with pytest.raises(btrc_amd.BtrcException):
        ctx.render(bad)
