"""Parity of the gfx950 kernels with the CPU oracle and with the reference's
own CPU traversal (lib/bvh goldens). Called through the C ABI."""
import os

import numpy as np
import pytest

import btrc_amd
from btrc_amd import _abi, scenes
from oracle import oracle

from util import assert_hits_equal, random_rays, render_params, single_mesh_scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def upload(ctx, scene):
    ctx.upload_scene(scene)
    return oracle.OracleScene(scene)


# ------------------------------------------------------------ building blocks

@pytest.mark.parametrize("fn,lo,hi", [(0, -40.0, 40.0), (1, -40.0, 40.0), (2, 1e-30, 1e30), (3, -100.0, 88.0),
                                      (5, 0.0, 1e6), (6, -10.0, 10.0)])
def test_detmath_bitwise_equal_to_host(ctx, fn, lo, hi):
    rng = np.random.default_rng(fn)
    if fn == 2:
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 200000)).astype(np.float32)
    else:
        x = rng.uniform(lo, hi, 200000).astype(np.float32)
    y = rng.uniform(0.1, 3.0, len(x)).astype(np.float32)
    assert np.array_equal(ctx.eval_detmath(fn, x, y).view(np.uint32), oracle.detmath(fn, x, y).view(np.uint32))


def test_detmath_pow_bitwise(ctx):
    rng = np.random.default_rng(4)
    x = rng.uniform(1e-4, 1.0, 100000).astype(np.float32)
    y = rng.uniform(0.0, 2.0, 100000).astype(np.float32)
    assert np.array_equal(ctx.eval_detmath(4, x, y).view(np.uint32), oracle.detmath(4, x, y).view(np.uint32))


@pytest.mark.parametrize("w,h,dims", [(256, 256, 48), (1920, 1080, 48), (7, 3, 48), (64, 64, 300)])
def test_halton_bitwise_equal_to_host(ctx, w, h, dims):
    """even samples take the first-digit tables (dimensions below 256), odd
    ones the plain digit loop; 300 dimensions cross the tables' end"""
    rng = np.random.default_rng(w)
    n = 4096
    px = rng.integers(0, w, n)
    py = rng.integers(0, h, n)
    si = rng.integers(0, 64, n)
    gpu = ctx.eval_halton(w, h, px, py, si, dims)
    cpu = oracle.halton(w, h, px, py, si, dims)
    assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32))


# ------------------------------------------------------------ traversal

@pytest.mark.parametrize("mode", [_abi.TRI_MOLLER_TRUMBORE, _abi.TRI_WATERTIGHT])
def test_simple_example(ctx, mode):
    tris = np.array([[1, -1, 1, 1, 1, 1, -1, 1, 1], [1, -1, 1, -1, -1, 1, -1, 1, 1]], np.float32)
    ctx.upload_scene(single_mesh_scene(tris))
    r = np.zeros(1, _abi.ray_dtype())
    r["o"], r["d"], r["tmin"], r["tmax"] = (0, 0, 0), (0, 0, 1), 0.0, 100.0
    h = ctx.trace_closest(r, mode)[0]
    assert h["prim"] == 0 and h["t"] == 1.0 and h["u"] == 0.0 and h["v"] == 0.5


def test_mt_matches_lib_bvh_goldens(ctx):
    for name in ("bvhref_cornell", "bvhref_field"):
        g = np.load(os.path.join(GOLDEN, name + ".npz"))
        if name == "bvhref_cornell":
            scene = single_mesh_scene(scenes.load_obj_triangles(os.path.join(GOLDEN, "cornell_box.obj")))
        else:
            import make_golden_shim as mg
            scene = mg.field_scene()
        ctx.upload_scene(scene)
        hits = ctx.trace_closest(g["rays"], _abi.TRI_MOLLER_TRUMBORE)
        flat = hits["prim"].copy()
        ok = flat != 0xFFFFFFFF
        if name == "bvhref_field":
            flat[ok] = hits["instance"][ok] * 256 + hits["prim"][ok]
        assert np.array_equal(ok, g["prim"] != 0xFFFFFFFF)
        same = flat == g["prim"]
        tie = ok & ~same
        assert np.array_equal(hits["t"][tie].view(np.uint32), g["t"][tie].view(np.uint32))
        m = ok & same
        for k in ("t", "u", "v"):
            assert np.array_equal(hits[k][m].view(np.uint32), g[k][m].view(np.uint32)), (name, k)


def test_cornell_image_known_answer(ctx):
    from PIL import Image
    tris = scenes.load_obj_triangles(os.path.join(GOLDEN, "cornell_box.obj"))
    ctx.upload_scene(single_mesh_scene(tris))
    rays = oracle.kat_rays(1080, 720, (0, 0.9, 2.5), (0, 0.001, -1), (0, 1, 0), 60.0)
    for mode in (_abi.TRI_MOLLER_TRUMBORE, _abi.TRI_WATERTIGHT):
        img = oracle.kat_pixels(1080, 720, tris, ctx.trace_closest(rays, mode))
        ref = np.asarray(Image.open(os.path.join(GOLDEN, "cornell_box_reference.png")))
        assert (img != ref).any(axis=2).sum() == 0


@pytest.mark.parametrize("mode", [_abi.TRI_WATERTIGHT, _abi.TRI_MOLLER_TRUMBORE])
def test_closest_hit_bitwise_vs_oracle_field(ctx, mode):
    scene = scenes.sphere_field(instances=512, lights=8, seed=5, extent=40.0)[0]
    os_ = upload(ctx, scene)
    rays = random_rays(100000, [0, 0, 0], [40, 40, 40], 17)
    assert_hits_equal(ctx.trace_closest(rays, mode), os_.trace(rays, tri_test=mode))


def test_closest_hit_bitwise_vs_oracle_cornell(ctx):
    scene, _ = scenes.cornell()
    os_ = upload(ctx, scene)
    rays = random_rays(100000, [-0.99, 0.01, -0.99], [0.99, 1.99, 0.99], 18)
    assert_hits_equal(ctx.trace_closest(rays), os_.trace(rays))


def test_any_hit_vs_oracle(ctx):
    scene = scenes.sphere_field(instances=512, lights=8, seed=5, extent=40.0)[0]
    os_ = upload(ctx, scene)
    rays = random_rays(100000, [0, 0, 0], [40, 40, 40], 19, tmax=np.float32(6.0))
    assert np.array_equal(ctx.trace_any(rays), os_.trace(rays, any_hit=True))


def test_empty_and_degenerate_batches(ctx):
    scene, _ = scenes.cornell()
    ctx.upload_scene(scene)
    assert len(ctx.trace_closest(np.zeros(0, _abi.ray_dtype()))) == 0
    r = random_rays(64, [-0.5, 0.5, -0.5], [0.5, 1.5, 0.5], 3)
    r["d"][:8] = 0.0          # zero direction: no hit, no fault
    r["tmax"][8:16] = 0.0     # empty interval
    r["d"][16:24] = [[0, 1, 0]] * 8   # axis-aligned
    os_ = oracle.OracleScene(scene)
    assert_hits_equal(ctx.trace_closest(r), os_.trace(r))


def test_large_batch_matches_oracle_sample(ctx):
    scene = scenes.sphere_field(instances=4096, lights=64, seed=2)[0]
    os_ = upload(ctx, scene)
    rays = random_rays(1 << 21, [0, 0, 0], [100, 100, 100], 23)
    hits = ctx.trace_closest(rays)
    idx = np.random.default_rng(1).choice(len(rays), 20000, replace=False)
    assert_hits_equal(hits[idx], os_.trace(rays[idx]))


# ------------------------------------------------------------ rendering

def render_pair(ctx, scene, cam, p):
    ctx.upload_scene(scene)
    ctx.set_camera(cam, p.width, p.height)
    stats = ctx.render(p)
    rgba, rgb64 = ctx.film(p.width, p.height, fp64=bool(p.film_fp64))
    ref, e, s = oracle.OracleScene(scene).render(cam, p)
    ref /= p.spp
    return rgba, rgb64, ref, stats, (e, s)


def test_cornell_per_pixel_parity(ctx):
    # the Cornell-style scene at its full size: 256x256, 16 spp, depth 8, RR after 3
    scene, cam = scenes.cornell()
    p = render_params(256, 256, 16, film_fp64=1)
    rgba, rgb64, ref, stats, (e, s) = render_pair(ctx, scene, cam, p)
    gpu = rgba[..., :3].astype(np.float64)
    rel = np.abs(gpu - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()} at {np.unravel_index(rel.argmax(), rel.shape)}"
    assert np.abs(rgb64 - ref).max() <= 1e-9 * max(1.0, ref.max())
    assert stats.paths == 256 * 256 * 16
    assert stats.extension_rays == e and stats.shadow_rays == s
    assert (rgba[..., 3] == 1.0).all()


def test_field_per_pixel_parity(ctx):
    scene, cam = scenes.sphere_field(instances=256, lights=16, seed=4, extent=30.0)
    p = render_params(96, 64, 8, state_count=4096)
    rgba, _, ref, stats, (e, s) = render_pair(ctx, scene, cam, p)
    rel = np.abs(rgba[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s


def test_small_pool_refill_matches(ctx):
    # a pool much smaller than the image exercises the refill rule
    scene, cam = scenes.cornell()
    p = render_params(64, 48, 3, state_count=1000)
    rgba, _, ref, stats, _ = render_pair(ctx, scene, cam, p)
    assert np.abs(rgba[..., :3] - ref).max() / max(1.0, ref.max()) <= 1e-5
    assert stats.paths == 64 * 48 * 3


def test_render_argument_errors(ctx):
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 16, 16)
    with pytest.raises(btrc_amd.BtrcException) as e:
        ctx.render(render_params(0, 16, 1))
    assert e.value.status == _abi.BTRC_ERR_INVALID
    bad = scenes.Scene()
    bad.add_instance(5, 0)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.upload_scene(bad)


def test_renderer_mirror_defaults(ctx):
    scene, cam = scenes.cornell()
    r = btrc_amd.WavefrontPathTracer(context=ctx, spp=2, state_count=1 << 12)
    r.set_scene(scene)
    r.set_camera(cam)
    r.set_film(32, 24)
    res = r.render()
    assert res.color.shape == (24, 32, 4)
    assert np.isfinite(res.color).all()
    assert r.params["min_depth"] == 5 and r.params["max_depth"] == 10  # wavefront.h:18-19
