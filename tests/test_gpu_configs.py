"""Each benchmark configuration's own workload on the GPU, at its full size:

- the 1M-triangle Disney field at 1920x1080 (depth 8) and the media room
  (128^3 fBm grid, sigma_t 2, HG g 0.6, depth 16) at 1920x1080: a band of
  rows of the full-size frame, the same paths as the benchmark frame's
  (same Halton streams), per pixel against the oracle;
- the radiance cache on the 256^2 room at its settings (2^16 record slots,
  four Adam steps of 16,384 records per frame): records, queries, a finite
  and falling loss, and the gain over ending the same paths without it;
- the 16.8M-triangle textured two-level field with its eight views in one
  wavefront: each view equals its own render, a band of view 0 matches the
  oracle per pixel, and with the cache and the captured bounce loop on,
  several frames stay finite and count their rays.

(The 256^2 Cornell-style room without media or cache is rendered at its
full size, 16 spp, in test_gpu_parity.test_cornell_per_pixel_parity.)
"""
import math

import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params

pytestmark = pytest.mark.gpu


def oracle_band(os_, cam, p, r0, r1):
    """oracle radiance of image rows [r0, r1) of the frame `p` describes:
    every sample index, the frame's own path indices (uniform order)"""
    W, H = p.width, p.height
    film = np.zeros((H, W, 3))
    for s in range(p.spp):
        film += os_.render(cam, p, first=s * W * H + r0 * W, count=(r1 - r0) * W)[0]
    return film[r0:r1] / p.spp


def assert_band(gpu_rgba, ref, r0, r1, what):
    got = gpu_rgba[r0:r1, :, :3].astype(np.float64)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"{what}: max rel err {rel.max()} at {np.unravel_index(rel.argmax(), rel.shape)}"
    assert ref.max() > 0


def test_field_1080p_band_matches_oracle(ctx):
    scene, cam = scenes.sphere_field()
    ctx.upload_scene(scene)
    p = render_params(1920, 1080, 2, min_depth=3, max_depth=8, state_count=1 << 21)
    ctx.set_camera(cam, p.width, p.height)
    st = ctx.render(p)
    assert st.paths == 1920 * 1080 * 2
    rgba, _ = ctx.film(p.width, p.height)
    assert np.isfinite(rgba).all()
    assert_band(rgba, oracle_band(oracle.OracleScene(scene), cam, p, 536, 544), 536, 544, "1M-triangle field")


def test_media_1080p_band_matches_oracle(ctx):
    scene, cam = scenes.cornell_media()
    ctx.upload_scene(scene)
    p = render_params(1920, 1080, 1, min_depth=3, max_depth=16, state_count=1 << 21)
    ctx.set_camera(cam, p.width, p.height)
    ctx.render(p)
    rgba, _ = ctx.film(p.width, p.height)
    assert np.isfinite(rgba).all()
    # rows through the lit upper part of the medium box (the centre of the frame is nearly opaque)
    assert_band(rgba, oracle_band(oracle.OracleScene(scene), cam, p, 298, 302), 298, 302, "media room")


def relmse(img, gt):
    return float(np.mean((img - gt) ** 2 / (gt ** 2 + 1e-2)))


def test_cache_at_its_configuration(ctx):
    W = H = 256
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, W, H)
    g = render_params(W, H, 512, min_depth=3, max_depth=8, state_count=1 << 20, film_fp64=1)
    g.sampler = _abi.SAMPLER_INDEPENDENT
    ctx.nrc_configure(enabled=0)
    ctx.render(g)
    gt = ctx.film(W, H, fp64=True)[1]

    def frames(n, first, train=True, reset=False):
        ctx.nrc_configure(enabled=1, train=1 if train else 0, train_depth=0, steps=4, train_stride=16,
                          query_depth=2)
        if reset:
            ctx.nrc_reset(3)
        stats = []
        for f in range(n):
            p = render_params(W, H, 1, min_depth=3, max_depth=8, state_count=1 << 17)
            p.sample_offset, p.accumulate = first + f, 1 if f else 0
            ctx.render(p)
            stats.append(ctx.nrc_stats())
        return ctx.film(W, H)[0][..., :3].astype(np.float64), stats

    _, stats = frames(32, 0, reset=True)
    for s in stats:
        # 4096 training paths of up to 9 vertices in 2^16 slots; one query per other path at depth 2
        assert 0 < s.records <= 4096 * 9 and s.records <= 1 << 16
        assert 0 < s.queries <= W * H
        assert s.skipped == 0 and all(math.isfinite(x) for x in s.loss)
    early = np.mean([np.mean(s.loss) for s in stats[:8]])
    late = np.mean([np.mean(s.loss) for s in stats[-8:]])
    assert late < early, (early, late)
    assert stats[-1].steps_total == 32 * 4
    img, _ = frames(16, 32)
    ctx.nrc_set_weights(np.zeros(_abi.NRC_PARAMS, dtype=np.float32))
    trunc, _ = frames(16, 32, train=False)
    ctx.nrc_configure(enabled=0)
    e_cache, e_trunc = relmse(img, gt), relmse(trunc, gt)
    print(f"256^2 cache: relMSE at 16 frames: cache {e_cache:.5f}, truncated {e_trunc:.5f}")
    assert e_cache < 0.5 * e_trunc


@pytest.fixture(scope="module")
def instanced():
    return scenes.instanced_field()


def test_instanced_views_equal_separate_renders(ctx, instanced):
    scene, views = instanced
    ctx.upload_scene(scene)
    assert ctx.scene_info()["triangles"] >= 65536 * 256
    p = render_params(0, 0, 1, min_depth=3, max_depth=8, state_count=1 << 21, film_fp64=1)
    st = ctx.render_views(p, views)
    total = sum(w * h for _, w, h in views)
    assert st.paths == total
    _, all64 = ctx.film(total, 1, fp64=True)
    all64 = all64.reshape(-1, 3)
    base, ext, shadow = 0, 0, 0
    for k, (cam, w, h) in enumerate(views):
        ctx.set_camera(cam, w, h)
        s1 = ctx.render(render_params(w, h, 1, min_depth=3, max_depth=8, state_count=1 << 21, film_fp64=1))
        alone = ctx.film(w, h, fp64=True)[1]
        batch = all64[base:base + w * h].reshape(h, w, 3)
        assert np.isfinite(batch).all()
        assert np.abs(batch - alone).max() <= 1e-12 * max(1.0, np.abs(alone).max()), k
        ext += s1.extension_rays
        shadow += s1.shadow_rays
        base += w * h
    assert st.extension_rays == ext and st.shadow_rays == shadow


def test_instanced_view_band_matches_oracle(ctx, instanced):
    scene, views = instanced
    ctx.upload_scene(scene)
    p = render_params(0, 0, 1, min_depth=3, max_depth=8, state_count=1 << 21)
    ctx.render_views(p, views)
    cam, w, h = views[0]
    got = ctx.view_film(0, w, h)
    q = render_params(w, h, 1, min_depth=3, max_depth=8)
    assert_band(got, oracle_band(oracle.OracleScene(scene), cam, q, h // 2, h // 2 + 2), h // 2, h // 2 + 2,
                "instanced field, view 0")


def test_instanced_views_with_cache_and_graph(ctx, instanced):
    scene, views = instanced
    ctx.upload_scene(scene)
    ctx.set_tuning(launch_graph=1)
    ctx.nrc_configure(enabled=1, train=1, train_depth=0)
    ctx.nrc_reset(5)
    total = sum(w * h for _, w, h in views)
    p = render_params(0, 0, 1, min_depth=3, max_depth=8, state_count=1 << 21)
    for f in range(4):
        p.sample_offset, p.accumulate = f, 1 if f else 0
        st = ctx.render_views(p, views)
        ns = ctx.nrc_stats()
        assert st.paths == total and st.extension_rays >= total and st.shadow_rays > 0
        assert 0 < ns.queries <= total and ns.records > 0
        assert all(math.isfinite(x) for x in ns.loss)
    ctx.nrc_configure(enabled=0)
    for k, (_, w, h) in enumerate(views):
        img = ctx.view_film(k, w, h)
        assert np.isfinite(img).all() and img[..., :3].mean() > 0
