"""Two-level traversal (instance tree over per-geometry trees) against the
oracle's two-level restatement, bit for bit, through the C ABI."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from test_oracle_two_level import baked_field, two_level
from util import assert_hits_equal, random_rays, render_params

pytestmark = pytest.mark.gpu


def upload(ctx, scene):
    ctx.upload_scene(scene)
    return oracle.OracleScene(scene)


@pytest.mark.parametrize("mode", [_abi.TRI_WATERTIGHT, _abi.TRI_MOLLER_TRUMBORE])
def test_closest_hit_bitwise_vs_oracle(ctx, mode):
    scene = two_level(scenes.sphere_field(instances=512, lights=8, seed=5, extent=40.0)[0])
    os_ = upload(ctx, scene)
    rays = random_rays(100000, [0, 0, 0], [40, 40, 40], 17)
    gpu = ctx.trace_closest(rays, mode)
    assert_hits_equal(gpu, os_.trace(rays, tri_test=mode))
    assert (gpu["instance"] != 0xFFFFFFFF).mean() > 0.3


def test_any_hit_vs_oracle(ctx):
    scene = two_level(scenes.sphere_field(instances=512, lights=8, seed=5, extent=40.0)[0])
    os_ = upload(ctx, scene)
    rays = random_rays(100000, [0, 0, 0], [40, 40, 40], 19, tmax=np.float32(6.0))
    assert np.array_equal(ctx.trace_any(rays), os_.trace(rays, any_hit=True))


def test_identity_transforms_equal_one_level(ctx):
    rays = random_rays(50000, [0, 0, 0], [12, 12, 12], 31)
    ctx.upload_scene(baked_field(instances=64))
    flat = ctx.trace_closest(rays)
    ctx.upload_scene(two_level(baked_field(instances=64)))
    two = ctx.trace_closest(rays)
    assert_hits_equal(flat, two)
    # and back: buffers of the previous structure are replaced
    ctx.upload_scene(baked_field(instances=64))
    assert_hits_equal(flat, ctx.trace_closest(rays))


def test_dense_overlap_deep_stack(ctx):
    # 1024 overlapping instances in a 6-unit cube: instance-tree stacks run
    # past the LDS entries into the global overflow
    scene = two_level(scenes.sphere_field(instances=1024, lights=2, seed=9, extent=6.0)[0])
    os_ = upload(ctx, scene)
    info = ctx.scene_info()
    assert info["stack"] > 16
    rays = random_rays(3000, [-2, -2, -2], [8, 8, 8], 41)
    assert_hits_equal(ctx.trace_closest(rays), os_.trace(rays))


def test_degenerate_instance_sets(ctx):
    rays = random_rays(4096, [-3, -3, -3], [3, 3, 3], 43)
    # one instance (the instance tree is a single leaf)
    s = scenes.Scene()
    s.add_instance(s.add_geometry(scenes.uv_sphere()), s.add_diffuse([0.5, 0.5, 0.5]),
                   scenes.translate(0.5, 0.0, 0.0) @ scenes.scale(2.0))
    os_ = upload(ctx, two_level(s))
    assert_hits_equal(ctx.trace_closest(rays), os_.trace(rays))
    # an instance of an empty geometry next to a real one
    s = scenes.Scene()
    m = s.add_diffuse([0.5, 0.5, 0.5])
    s.add_instance(s.add_geometry(np.zeros((0, 9), np.float32)), m)
    s.add_instance(s.add_geometry(scenes.uv_sphere()), m, scenes.scale(1.5))
    os_ = upload(ctx, two_level(s))
    hits = ctx.trace_closest(rays)
    assert_hits_equal(hits, os_.trace(rays))
    assert (hits["instance"][hits["instance"] != 0xFFFFFFFF] == 1).all()


def test_render_per_pixel_parity(ctx):
    scene, cam = scenes.sphere_field(instances=256, lights=16, seed=4, extent=30.0)
    two_level(scene)
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


def test_schedules_bitwise_equal(ctx):
    # leaves tested when reached or postponed (inside the current instance)
    scene = two_level(scenes.sphere_field(instances=768, lights=8, seed=13, extent=30.0)[0])
    os_ = upload(ctx, scene)
    rays = random_rays(60000, [0, 0, 0], [30, 30, 30], 47)
    short = rays.copy()
    short["tmax"] = np.float32(4.0)
    want_h, want_o = os_.trace(rays), os_.trace(short, any_hit=True)
    for bits in range(4):
        ctx.set_tuning(two_level_postpone=bits)
        assert_hits_equal(ctx.trace_closest(rays), want_h)
        assert np.array_equal(ctx.trace_any(short), want_o)
    ctx.set_tuning(two_level_postpone=0)
