"""Shared helpers of the test suite."""
import numpy as np

from btrc_amd import _abi, scenes

FMAX = np.float32(3.4028234663852886e38)


def random_rays(n, lo, hi, seed, tmax=FMAX):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, dtype=_abi.ray_dtype())
    rays["o"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = tmax
    return rays


def single_mesh_scene(tris):
    s = scenes.Scene()
    s.add_instance(s.add_geometry(tris), s.add_diffuse([0.5, 0.5, 0.5]))
    return s


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_hits_equal(a, b):
    """bitwise equality of two btrc_hit arrays"""
    assert a.dtype == b.dtype
    ab = np.ascontiguousarray(a).view(np.uint32).reshape(len(a), 5)
    bb = np.ascontiguousarray(b).view(np.uint32).reshape(len(b), 5)
    diff = np.nonzero((ab != bb).any(axis=1))[0]
    assert len(diff) == 0, f"{len(diff)} of {len(a)} hits differ, first {diff[:5]}: {a[diff[:3]]} vs {b[diff[:3]]}"


def render_params(w, h, spp, min_depth=3, max_depth=8, rr_threshold=0.1, rr_cont_prob=0.5, state_count=1 << 20,
                  film_fp64=0, time_stages=0):
    p = _abi.RenderParams()
    p.width, p.height, p.spp = w, h, spp
    p.min_depth, p.max_depth = min_depth, max_depth
    p.rr_threshold, p.rr_cont_prob = rr_threshold, rr_cont_prob
    p.state_count = state_count
    p.film_fp64 = film_fp64
    p.time_stages = time_stages
    return p
