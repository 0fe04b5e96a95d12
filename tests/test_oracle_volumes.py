"""Pins the oracle's volume-primitive medium (VolumePrimitiveMedium,
core/volume.cpp:108-197; volume::Aggregate tracking,
core/volume/aggregate.cpp:136-254) on closed forms: ratio-tracking mean and
delta-tracking scattering probability against exp(-optical depth) for
constant, overlapping, rotated and linearly varying densities; segments that
miss every volume; and a white furnace (albedo 1, constant environment)."""
import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from util import render_params


def segments(n, a, b):
    seg = np.zeros(n, dtype=_abi.ray_dtype())
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    seg["o"] = a
    seg["d"] = b - a
    seg["tmax"] = 1.0
    return seg


def eval_volumes(scene, a, b, n=40000, seed=1):
    os_ = oracle.OracleScene(scene)
    seeds = np.random.default_rng(seed).integers(0, 2 ** 63, n, dtype=np.uint64)
    return os_.volume_eval(segments(n, a, b), seeds)


def unit_box(s, o, size, sigma, albedo=(0.6, 0.7, 0.8)):
    return s.add_volume(o, (size, 0, 0), (0, size, 0), (0, 0, size), sigma, albedo)


def check_depth(out, tau):
    tr = out[:, 7].astype(np.float64)
    n = len(tr)
    assert abs(tr.mean() - np.exp(-tau)) < 5 * tr.std() / np.sqrt(n) + 1e-4
    p = out[:, 0].mean()
    assert abs(p - (1 - np.exp(-tau))) < 5 * np.sqrt(p * (1 - p) / n) + 1e-4


def test_constant_box_closed_form():
    s = scenes.Scene()
    unit_box(s, (0, 0, 0), 1.0, 1.3)
    out = eval_volumes(s, (-1, 0.4, 0.5), (2, 0.4, 0.5))   # 1 unit inside
    check_depth(out, 1.3)
    sc = out[out[:, 0] == 1]
    assert (sc[:, 1] >= 0).all() and (sc[:, 1] <= 1).all()          # scatter points inside
    assert np.allclose(sc[:, 4:7], [0.6, 0.7, 0.8])


def test_overlapping_boxes_add_densities():
    s = scenes.Scene()
    unit_box(s, (0, 0, 0), 1.0, 0.7)
    unit_box(s, (0.5, 0, 0), 1.0, 0.4, albedo=(0.2, 0.2, 0.2))
    out = eval_volumes(s, (-1, 0.5, 0.5), (3, 0.5, 0.5))
    check_depth(out, 0.7 * 1.0 + 0.4 * 1.0)
    # albedo of a scatter point in the overlap is the density-weighted mix
    sc = out[(out[:, 0] == 1) & (out[:, 1] > 0.55) & (out[:, 1] < 0.95)]
    assert np.allclose(sc[:, 4], (0.7 * 0.6 + 0.4 * 0.2) / 1.1, atol=1e-5)


def test_rotated_box_and_misses():
    s = scenes.Scene()
    c, sn = np.cos(0.6), np.sin(0.6)
    s.add_volume((0, 0, 0), (c, sn, 0), (-sn, c, 0), (0, 0, 1), 0.9, (1, 1, 1))
    # along the rotated x edge direction through the middle: 1 unit inside
    mid = np.array([0.5 * c - 0.5 * sn, 0.5 * sn + 0.5 * c, 0.5])
    e = np.array([c, sn, 0.0])
    out = eval_volumes(s, mid - 2 * e, mid + 2 * e)
    check_depth(out, 0.9)
    miss = eval_volumes(s, (5, 5, 5), (6, 7, 8), n=1000)
    assert (miss[:, 0] == 0).all() and (miss[:, 7:10] == 1).all()


def test_linear_density_grid():
    # sigma_t(u) = 2u sampled on a 64-cell grid: the linear filter is exact
    # between the outer texel centres; along u from 0.1 to 0.9 the depth is
    # the integral of 2u = 0.8
    n = 64
    u = (np.arange(n) + 0.5) / n
    grid = np.tile((2 * u).astype(np.float32), (4, 4, 1))
    s = scenes.Scene()
    unit_box(s, (0, 0, 0), 1.0, grid)
    out = eval_volumes(s, (0.1, 0.5, 0.5), (0.9, 0.5, 0.5))
    check_depth(out, 0.8)


def test_white_furnace():
    # a non-absorbing volume under a constant environment: every pixel sees
    # the environment's radiance (up to paths cut at the depth limit)
    s = scenes.Scene()
    s.add_volume((-1, -1, -1), (2, 0, 0), (0, 2, 0), (0, 0, 2), 0.8, (1, 1, 1))
    s.set_gradient_sky((0, 1, 0), (0.5, 0.5, 0.5), (0.5, 0.5, 0.5))
    cam = scenes.Camera((0, 0, -4), (0, 0, 0), (0, 1, 0), 40.0)
    p = render_params(16, 16, 64, min_depth=64, max_depth=64)
    img, e, sh = oracle.OracleScene(s).render(cam, p)
    img /= p.spp
    assert abs(img.mean() - 0.5) < 0.01
    assert np.abs(img - 0.5).max() < 0.25   # 64 samples per pixel
