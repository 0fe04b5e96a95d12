"""Pins the oracle's two-level traversal (instance tree over per-geometry
trees, Scene::commit's TLAS/BLAS split, core/scene.cpp:92-239): with identity
transforms it must equal the one-level oracle bit for bit (already pinned on
lib/bvh's goldens); with general transforms it must find the same
(instance, primitive) as one level up to the rounding of the ray transform;
its instance culling must not change any hit."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import random_rays


def baked_field(instances=24, seed=3, extent=12.0):
    """the sphere field with every instance's world vertices baked into its
    own geometry and an identity transform"""
    src = scenes.sphere_field(instances=instances, lights=2, seed=seed, extent=extent)[0]
    s = scenes.Scene()
    mat = s.add_diffuse([0.5, 0.5, 0.5])
    for inst in src.instances:
        p = src.geometries[inst["geometry"]]["positions"].reshape(-1, 3).astype(np.float32)
        m = inst["mat"].astype(np.float32)
        w = ((m[:, 0] * p[:, :1] + m[:, 1] * p[:, 1:2]) + m[:, 2] * p[:, 2:3]) + m[:, 3]
        s.add_instance(s.add_geometry(w[:, :3].reshape(-1, 9)), mat)
    return s


def two_level(scene):
    scene.accel = _abi.ACCEL_TWO_LEVEL
    return scene


@pytest.mark.parametrize("mode", [_abi.TRI_WATERTIGHT, _abi.TRI_MOLLER_TRUMBORE])
def test_identity_transforms_equal_one_level_bitwise(mode):
    flat = oracle.OracleScene(baked_field())
    two = oracle.OracleScene(two_level(baked_field()))
    rays = random_rays(20000, [0, 0, 0], [12, 12, 12], 31)
    a = flat.trace(rays, tri_test=mode)
    b = two.trace(rays, tri_test=mode)
    assert a.tobytes() == b.tobytes()
    assert (a["instance"] != 0xFFFFFFFF).mean() > 0.2


@pytest.mark.parametrize("mode", [_abi.TRI_WATERTIGHT, _abi.TRI_MOLLER_TRUMBORE])
def test_instance_culling_equals_brute_force(mode):
    scene = two_level(scenes.sphere_field(instances=32, lights=4, seed=3, extent=12.0)[0])
    os_ = oracle.OracleScene(scene)
    rays = random_rays(20000, [0, 0, 0], [12, 12, 12], 5)
    assert os_.trace(rays, tri_test=mode, use_bvh=True).tobytes() == \
        os_.trace(rays, tri_test=mode, use_bvh=False).tobytes()
    rays["tmax"] = np.float32(3.0)
    assert np.array_equal(os_.trace(rays, tri_test=mode, any_hit=True, use_bvh=True),
                          os_.trace(rays, tri_test=mode, any_hit=True, use_bvh=False))


def test_general_transforms_agree_with_one_level():
    scene = scenes.sphere_field(instances=48, lights=4, seed=7, extent=14.0)[0]
    flat = oracle.OracleScene(scene)
    two = oracle.OracleScene(two_level(scenes.sphere_field(instances=48, lights=4, seed=7, extent=14.0)[0]))
    rays = random_rays(20000, [0, 0, 0], [14, 14, 14], 8)
    a = flat.trace(rays)
    b = two.trace(rays)
    same = (a["instance"] == b["instance"]) & (a["prim"] == b["prim"])
    assert same.mean() > 0.999
    hit = same & (a["instance"] != 0xFFFFFFFF)
    assert hit.mean() > 0.2
    # the two place rounding differently (world vertices vs. the moved ray):
    # distances agree to about 2^-17 of the scene's 14-unit extent
    assert np.abs(a["t"][hit] - b["t"][hit]).max() < 1e-4
    # barycentrics of the same triangle agree to the same order
    assert np.abs(a["u"][hit] - b["u"][hit]).max() < 1e-3


def test_any_hit_is_closest_hit_within_tmax():
    scene = two_level(scenes.sphere_field(instances=32, lights=4, seed=3, extent=12.0)[0])
    os_ = oracle.OracleScene(scene)
    rays = random_rays(20000, [0, 0, 0], [12, 12, 12], 9, tmax=np.float32(2.0))
    assert np.array_equal(os_.trace(rays, any_hit=True), os_.trace(rays)["instance"] != 0xFFFFFFFF)
