"""The radiance cache on the GPU: the MFMA network against the FP64 oracle
(outputs and gradients, with the device's BF16 rounding points), Adam's
first step in closed form, a regression fit, and rendering with the cache:
queries, training records, and what the trained cache gains over ending
the paths at the same depth without it."""
import math

import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from test_oracle_nrc import random_vertices

from util import render_params

pytestmark = pytest.mark.gpu


@pytest.fixture
def cornell_ctx(ctx):
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 64, 64)
    yield ctx, scene, cam
    ctx.nrc_configure(enabled=0)


def scene_vertices(n, box, seed):
    rng = np.random.default_rng(seed)
    v = random_vertices(n, rng)
    lo, inv = box[:3], box[3:]
    v[:, 0:3] = lo + rng.uniform(0, 1, (n, 3)) / inv
    return v


def test_forward_matches_oracle(cornell_ctx):
    ctx, _, _ = cornell_ctx
    ctx.nrc_reset(5)
    w = ctx.nrc_weights()
    box = ctx.scene_box()
    v = scene_vertices(3000, box, 1)
    gpu = ctx.nrc_eval(v).astype(np.float64)
    ref = oracle.nrc_forward(w, box, v, bf16=True)
    scale = np.abs(ref).mean()
    err = np.abs(gpu - ref)
    assert err.max() <= 0.05 * scale and err.mean() <= 0.005 * scale, (err.max(), err.mean(), scale)


def test_gradients_match_oracle(cornell_ctx):
    ctx, _, _ = cornell_ctx
    ctx.nrc_reset(6)
    w = ctx.nrc_weights()
    box = ctx.scene_box()
    v = scene_vertices(5000, box, 2)
    t = np.random.default_rng(3).uniform(0, 2, (5000, 3)).astype(np.float32)
    g, loss = ctx.nrc_gradients(v, t)
    gr, lr_ = oracle.nrc_gradients(w, box, v, t, bf16=True)
    assert abs(loss - lr_) <= 1e-2 * lr_
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert rel < 0.05 and cos > 0.998, (rel, cos)
    assert (g[5 * 4096 + 3 * 64:] == 0).all()


def test_first_adam_step_closed_form(cornell_ctx):
    # from zero moments the first Adam step moves each weight by -lr sign(g)
    ctx, _, _ = cornell_ctx
    ctx.nrc_configure(learning_rate=1e-2)
    ctx.nrc_reset(7)
    w0 = ctx.nrc_weights()
    v = scene_vertices(4096, ctx.scene_box(), 4)
    t = np.random.default_rng(5).uniform(0, 1, (4096, 3)).astype(np.float32)
    g, _ = ctx.nrc_gradients(v, t)
    ctx.nrc_fit(v, t, 1)
    dw = ctx.nrc_weights().astype(np.float64) - w0
    live = np.abs(g) > 1e-12
    assert np.allclose(dw[live], -1e-2 * np.sign(g[live]), rtol=1e-4, atol=1e-6)
    assert (dw[~live] == 0).all()


def test_fit_reduces_loss(cornell_ctx):
    ctx, _, _ = cornell_ctx
    ctx.nrc_configure(learning_rate=1e-2)
    ctx.nrc_reset(8)
    box = ctx.scene_box()
    v = scene_vertices(16384, box, 6)
    q = (v[:, :3] - box[:3]) * box[3:]
    t = np.stack([0.5 + 0.5 * np.sin(3 * q[:, 0]), 0.5 + 0.5 * np.cos(2 * q[:, 1]), 0.3 + 0.2 * q[:, 2]],
                 axis=1).astype(np.float32)
    losses = np.concatenate([ctx.nrc_fit(v, t, 50) for _ in range(6)])
    assert np.isfinite(losses).all()
    assert losses[-10:].mean() < 0.1 * losses[:5].mean(), (losses[:5], losses[-10:])


def relmse(img, gt):
    return float(np.mean((img - gt) ** 2 / (gt ** 2 + 1e-2)))


def cache_frames(ctx, frames, first, w, h, depth=8, train=True, reset=True):
    ctx.nrc_configure(enabled=1, train=1 if train else 0)
    if reset:
        ctx.nrc_reset(9)
    stats = []
    for f in range(frames):
        p = render_params(w, h, 1, min_depth=3, max_depth=depth, state_count=1 << 14)
        p.sample_offset = first + f
        p.accumulate = 1 if f > 0 else 0
        ctx.render(p)
        stats.append(ctx.nrc_stats())
    return ctx.film(w, h)[0][..., :3].astype(np.float64), stats


def test_cache_render_queries_and_training(cornell_ctx):
    ctx, _, _ = cornell_ctx
    img, stats = cache_frames(ctx, 24, 0, 64, 64)
    for st in stats:
        assert 0 < st.queries <= 64 * 64 and 0 < st.records <= 64 * 64 // 16 * 9
        assert st.skipped == 0  # every batch mixes the frame's records
        assert all(math.isfinite(x) for x in st.loss)
    early = np.mean([np.mean(s.loss) for s in stats[:6]])
    late = np.mean([np.mean(s.loss) for s in stats[-6:]])
    assert late < early, (early, late)
    assert stats[-1].steps_total == 24 * 4
    assert np.isfinite(img).all() and img.mean() > 0


def test_trained_cache_beats_truncated_paths(cornell_ctx):
    ctx, _, _ = cornell_ctx
    w, h = 64, 64
    g = render_params(w, h, 2048, min_depth=3, max_depth=8, state_count=1 << 18, film_fp64=1)
    g.sampler = _abi.SAMPLER_INDEPENDENT
    ctx.nrc_configure(enabled=0)
    ctx.render(g)
    gt = ctx.film(w, h, fp64=True)[1]
    # warm the cache, then measure 32 frames while it keeps learning
    cache_frames(ctx, 96, 0, w, h)
    img, _ = cache_frames(ctx, 32, 96, w, h, reset=False)
    # the same prefix without the cache: zero weights, no training
    ctx.nrc_set_weights(np.zeros(_abi.NRC_PARAMS, dtype=np.float32))
    trunc, _ = cache_frames(ctx, 32, 96, w, h, train=False, reset=False)
    # plain path tracing at the same sample count
    ctx.nrc_configure(enabled=0)
    p = render_params(w, h, 32, min_depth=3, max_depth=8, state_count=1 << 16)
    p.sample_offset = 96
    ctx.render(p)
    plain = ctx.film(w, h)[0][..., :3].astype(np.float64)
    e_cache, e_trunc, e_plain = relmse(img, gt), relmse(trunc, gt), relmse(plain, gt)
    print(f"relMSE at 32 spp: cache {e_cache:.5f}, truncated {e_trunc:.5f}, plain {e_plain:.5f}")
    assert e_cache < 0.5 * e_trunc


def test_self_training_ends_training_paths_early(cornell_ctx):
    """train_depth k: training paths stop at depth k with a cache query, so
    the frame's bounce loop is shorter; the cache still learns and still
    beats the truncated paths"""
    ctx, _, _ = cornell_ctx
    w, h = 64, 64
    ctx.nrc_configure(enabled=1, train=1, train_depth=0)
    ctx.nrc_reset(9)
    p = render_params(w, h, 1, min_depth=3, max_depth=8, state_count=1 << 14)
    full = ctx.render(p)
    ctx.nrc_configure(enabled=1, train=1, train_depth=4)
    ctx.nrc_reset(9)
    short = ctx.render(p)
    # the loop checks for completion every few bounces, so compare, not count
    assert short.iterations < full.iterations
    import btrc_amd
    with pytest.raises(btrc_amd.BtrcException):
        ctx.nrc_configure(train_depth=1)        # before the query depth
    # at the query depth: training paths end with the rendered paths, no
    # bounce beyond theirs
    ctx.nrc_configure(enabled=1, train=1, train_depth=2)
    ctx.nrc_reset(9)
    same = ctx.render(p)
    assert same.iterations <= short.iterations
    ns = ctx.nrc_stats()
    assert ns.records > 0 and all(math.isfinite(x) for x in ns.loss[:4])
    ctx.nrc_configure(train_depth=4)
    img, stats = cache_frames(ctx, 96, 0, w, h)
    assert all(math.isfinite(x) for s in stats for x in s.loss)
    early = np.mean([np.mean(s.loss) for s in stats[:12]])
    late = np.mean([np.mean(s.loss) for s in stats[-12:]])
    assert late < early, (early, late)
    g = render_params(w, h, 1024, min_depth=3, max_depth=8, state_count=1 << 18, film_fp64=1)
    g.sampler = _abi.SAMPLER_INDEPENDENT
    ctx.nrc_configure(enabled=0)
    ctx.render(g)
    gt = ctx.film(w, h, fp64=True)[1]
    ctx.nrc_configure(enabled=1, train_depth=4)
    img, _ = cache_frames(ctx, 32, 96, w, h, reset=False)
    ctx.nrc_set_weights(np.zeros(_abi.NRC_PARAMS, dtype=np.float32))
    trunc, _ = cache_frames(ctx, 32, 96, w, h, train=False, reset=False)
    ctx.nrc_configure(train_depth=0)
    e_cache, e_trunc = relmse(img, gt), relmse(trunc, gt)
    print(f"self-training: relMSE at 32 spp: cache {e_cache:.5f}, truncated {e_trunc:.5f}")
    assert e_cache < 0.5 * e_trunc


def test_weight_moving_average(cornell_ctx):
    """with ema_decay d every optimizer step sets e <- d e + (1 - d) w, starting
    at the weights of the reset; without it the average stays there"""
    ctx, _, _ = cornell_ctx
    v = scene_vertices(4096, ctx.scene_box(), 9)
    t = np.random.default_rng(10).uniform(0, 1, (4096, 3)).astype(np.float32)
    d = 0.9
    ctx.nrc_configure(learning_rate=1e-2, ema_decay=d)
    ctx.nrc_reset(11)
    e = ctx.nrc_weights().astype(np.float64)
    assert np.array_equal(ctx.nrc_ema_weights(), ctx.nrc_weights())
    for _ in range(3):
        ctx.nrc_fit(v, t, 1)
        e = d * e + (1 - d) * ctx.nrc_weights().astype(np.float64)
        assert np.allclose(ctx.nrc_ema_weights(), e, rtol=1e-5, atol=1e-7)
    ctx.nrc_configure(ema_decay=0.0)
    before = ctx.nrc_ema_weights()
    ctx.nrc_fit(v, t, 1)
    assert np.array_equal(ctx.nrc_ema_weights(), before)
    import btrc_amd
    with pytest.raises(btrc_amd.BtrcException):
        ctx.nrc_configure(ema_decay=1.0)
