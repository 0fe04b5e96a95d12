"""The GPU BVH builder against lib/bvh (the reference's vendored CPU BVH
library, compiled from its sources): the clustering tree, the leaf collapse
and the refit node for node and bit for bit; the capped collapse the trace
kernels use by its invariants; and scenes built on the GPU tracing and
rendering exactly like scenes built on the host."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import random_rays, render_params

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (lib/bvh) not built")


def random_boxes(n, seed, kind):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        c = rng.uniform(-50, 50, (n, 3))
        e = rng.uniform(0.01, 2.0, (n, 3))
    elif kind == "clustered":
        centres = rng.uniform(-50, 50, (max(1, n // 50), 3))
        c = centres[rng.integers(0, len(centres), n)] + rng.normal(0, 0.5, (n, 3))
        e = rng.uniform(0.01, 0.3, (n, 3))
    else:  # duplicates: many identical boxes, ties everywhere
        base = rng.uniform(-5, 5, (max(1, n // 8), 3))
        c = base[rng.integers(0, len(base), n)]
        e = np.full((n, 3), 0.25)
    c = c.astype(np.float32)
    e = e.astype(np.float32)
    return np.concatenate([c - e, c + e], axis=1)


def as_words(nodes):
    return np.ascontiguousarray(nodes).view(np.uint32)


@needs_ref
@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 100000])
@pytest.mark.parametrize("kind", ["uniform", "clustered", "duplicates"])
@pytest.mark.parametrize("collapse", [False, True], ids=["ploc", "collapsed"])
def test_build_matches_lib_bvh(ctx, n, kind, collapse):
    boxes = random_boxes(n, n + len(kind), kind)
    nodes, prims, ms = ctx.build_bvh(boxes, max_leaf_size=0 if collapse else -1)
    ref_nodes, ref_prims = oracle.ref_ploc(boxes, collapse=collapse)
    assert len(nodes) == len(ref_nodes)
    np.testing.assert_array_equal(prims, ref_prims)
    np.testing.assert_array_equal(as_words(nodes), as_words(ref_nodes))
    assert ms >= 0


@needs_ref
@pytest.mark.parametrize("collapse", [False, True], ids=["ploc", "collapsed"])
def test_refit_matches_lib_bvh(ctx, collapse):
    boxes = random_boxes(20000, 5, "uniform")
    moved = boxes + np.random.default_rng(6).normal(0, 1.0, (len(boxes), 1)).astype(np.float32)
    nodes, prims, _ = ctx.build_bvh(boxes, max_leaf_size=0 if collapse else -1, refit_boxes=moved)
    ref_nodes, ref_prims = oracle.ref_ploc(boxes, collapse=collapse, refit_boxes=moved)
    np.testing.assert_array_equal(prims, ref_prims)
    np.testing.assert_array_equal(as_words(nodes), as_words(ref_nodes))


def check_tree(nodes, prims, boxes, max_leaf):
    """every primitive in exactly one leaf of at most max_leaf; every node
    box is exactly the union of what it holds"""
    n = len(boxes)
    assert sorted(prims.tolist()) == list(range(n))
    seen = np.zeros(n, np.int32)
    lo = nodes["bounds"][:, 0::2]
    hi = nodes["bounds"][:, 1::2]
    stack = [0]
    while stack:
        i = stack.pop()
        c, f = int(nodes["primitive_count"][i]), int(nodes["first"][i])
        if c:
            assert c <= max_leaf
            p = prims[f:f + c]
            seen[p] += 1
            np.testing.assert_array_equal(lo[i], boxes[p, :3].min(axis=0))
            np.testing.assert_array_equal(hi[i], boxes[p, 3:].max(axis=0))
        else:
            np.testing.assert_array_equal(lo[i], np.minimum(lo[f], lo[f + 1]))
            np.testing.assert_array_equal(hi[i], np.maximum(hi[f], hi[f + 1]))
            stack += [f, f + 1]
    assert (seen == 1).all()


@pytest.mark.parametrize("kind", ["uniform", "duplicates"])
def test_capped_collapse_invariants(ctx, kind):
    boxes = random_boxes(50000, 9, kind)
    nodes, prims, _ = ctx.build_bvh(boxes, max_leaf_size=4)
    check_tree(nodes, prims, boxes, 4)


@pytest.mark.parametrize("instances", [64, 4096])
def test_gpu_built_scene_traces_like_host_built(ctx, instances):
    scene, cam = scenes.sphere_field(instances=instances, lights=8, seed=3, extent=16.0 if instances == 64 else 100.0)
    rays = random_rays(200000, -20.0 if instances == 64 else -10.0, 20.0 if instances == 64 else 110.0, 4)
    hits = {}
    for accel in (_abi.ACCEL_FLAT, _abi.ACCEL_FLAT_GPU):
        scene.accel = accel
        ctx.upload_scene(scene)
        hits[accel] = (ctx.trace_closest(rays), ctx.trace_any(rays), ctx.scene_info())
    a, b = hits[_abi.ACCEL_FLAT], hits[_abi.ACCEL_FLAT_GPU]
    assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
    assert np.array_equal(a[1], b[1])
    assert b[2]["triangles"] == a[2]["triangles"]


def test_gpu_built_scene_renders_like_host_built(ctx):
    scene, cam = scenes.sphere_field(instances=64, lights=8, seed=11, extent=16.0)
    p = render_params(96, 64, 8, state_count=4096)
    out = []
    for accel in (_abi.ACCEL_FLAT, _abi.ACCEL_FLAT_GPU):
        scene.accel = accel
        ctx.upload_scene(scene)
        ctx.set_camera(cam, p.width, p.height)
        stats = ctx.render(p)
        img, _ = ctx.film(p.width, p.height)
        out.append((img, stats.extension_rays, stats.shadow_rays))
    # identical paths; the film's FP32 atomics add them in another order
    rel = np.abs(out[0][0] - out[1][0]) / np.maximum(np.abs(out[0][0]), 1e-2)
    assert rel.max() <= 1e-5
    assert out[0][1:] == out[1][1:]


def test_single_triangle_and_build_time(ctx):
    s = scenes.Scene()
    g = s.add_geometry(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32))
    s.add_instance(g, s.add_diffuse([0.5] * 3))
    s.accel = _abi.ACCEL_FLAT_GPU
    ctx.upload_scene(s)
    r = np.zeros(1, dtype=_abi.ray_dtype())
    r["o"] = (0.2, 0.2, 1.0)
    r["d"] = (0.0, 0.0, -1.0)
    r["tmax"] = 10.0
    h = ctx.trace_closest(r)
    assert h["t"][0] == 1.0 and h["prim"][0] == 0
    assert ctx.scene_build_ms() > 0


def displace(scene, geometry, seed):
    g = scene.geometries[geometry]
    rng = np.random.default_rng(seed)
    p = g["positions"].reshape(-1, 3, 3)
    # move every vertex shared by several triangles consistently: by a
    # function of its position
    off = 0.15 * np.sin(p * 3.0 + rng.uniform(0, 6.28, 3)).astype(np.float32)
    g["positions"] = np.ascontiguousarray((p + off).reshape(-1, 9), dtype=np.float32)


@pytest.mark.parametrize("accel", [_abi.ACCEL_FLAT, _abi.ACCEL_FLAT_GPU], ids=["host", "gpu"])
def test_refit_after_update_traces_like_a_fresh_upload(ctx, accel):
    """moving one geometry's vertices and refitting gives the hits of a
    fresh upload of the moved scene, bit for bit"""
    scene, cam = scenes.sphere_field(instances=256, lights=8, seed=5, extent=24.0)
    scene.accel = accel
    ctx.upload_scene(scene)
    displace(scene, 0, 1)
    ms = ctx.update_geometry(scene, 0)
    assert ms > 0
    rays = random_rays(200000, -2.0, 26.0, 8)
    got = (ctx.trace_closest(rays), ctx.trace_any(rays))
    ctx.upload_scene(scene)
    want = (ctx.trace_closest(rays), ctx.trace_any(rays))
    assert np.array_equal(got[0].view(np.uint8), want[0].view(np.uint8))
    assert np.array_equal(got[1], want[1])


def test_refit_after_update_renders_like_the_oracle(ctx):
    """a moved emissive geometry: shading records, light sampling tables and
    the tree all follow"""
    scene, cam = scenes.sphere_field(instances=64, lights=8, seed=11, extent=16.0)
    scene.accel = _abi.ACCEL_FLAT_GPU
    ctx.upload_scene(scene)
    for gi in range(len(scene.geometries)):
        displace(scene, gi, 10 + gi)
        ctx.update_geometry(scene, gi)
    p = render_params(96, 64, 8, state_count=4096)
    ctx.set_camera(cam, p.width, p.height)
    stats = ctx.render(p)
    img, _ = ctx.film(p.width, p.height)
    ref, e, s = oracle.OracleScene(scene).render(cam, p)
    ref /= p.spp
    rel = np.abs(img[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert stats.extension_rays == e and stats.shadow_rays == s


def test_update_rejects_bad_requests(ctx):
    import btrc_amd
    scene, cam = scenes.sphere_field(instances=16, lights=2, seed=1, extent=8.0)
    scene.accel = _abi.ACCEL_TWO_LEVEL
    ctx.upload_scene(scene)
    with pytest.raises(btrc_amd.BtrcException):
        ctx.update_geometry(scene, 0)                  # two-level structure
    scene.accel = _abi.ACCEL_FLAT
    ctx.upload_scene(scene)
    g = scene.geometries[0]
    keep = g["positions"]
    g["positions"] = keep[:-1].copy()
    with pytest.raises(btrc_amd.BtrcException):
        ctx.update_geometry(scene, 0)                  # triangle count changed
    g["positions"] = keep
    assert ctx.lib.btrc_hip_update_geometry(ctx.h, len(scene.geometries), None, None) == _abi.BTRC_ERR_INVALID
    assert ctx.lib.btrc_hip_update_geometry(ctx.h, 0, None, None) == _abi.BTRC_ERR_INVALID
