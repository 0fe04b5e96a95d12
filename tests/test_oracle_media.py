"""Pins the oracle's participating media (medium/homogeneous.cpp,
medium/hetergeneous.cpp, core/medium.cpp, utils/henyey_greenstein.h) by
closed forms and the statistics tracking estimators must have: the
Henyey-Greenstein density and its mean cosine, the linear grid filter on
linear fields, free-flight probabilities exp(-tau) and the unbiasedness of
ratio tracking against quadrature of the optical depth."""
import math

import numpy as np
import pytest
from scipy import stats

from btrc_amd import _abi, scenes
from oracle import oracle


def hg_density(g, u):
    """density of u = -cos(wi, wo) on [-1, 1]"""
    return 0.5 * (1 - g * g) / (1 + g * g - 2 * g * u) ** 1.5


@pytest.mark.parametrize("g", [-0.7, 0.0, 0.3, 0.6, 0.9])
def test_hg_sampling_matches_density(g):
    rng = np.random.default_rng(5)
    out = oracle.hg_sample(g, rng.random((200000, 2), dtype=np.float32))
    d = out[:, :3]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    # the pdf returned with a sample is the density of that sample
    assert np.allclose(out[:, 3], out[:, 4], rtol=2e-3, atol=1e-6)
    u = -d[:, 2].astype(np.float64)
    # mean of u is g (mean cosine between wi and -wo)
    assert abs(u.mean() - g) < 5 * u.std() / math.sqrt(len(u)) + 1e-4
    edges = np.linspace(-1, 1, 41)
    h, _ = np.histogram(u, bins=edges)
    fine = np.linspace(-1, 1, 40 * 64 + 1)
    cdf = np.concatenate([[0], np.cumsum((hg_density(g, fine[1:]) + hg_density(g, fine[:-1])) / 2 * np.diff(fine))])
    p = np.diff(np.interp(edges, fine, cdf / cdf[-1]))
    keep = p * len(u) > 20
    chi2 = ((h - len(u) * p) ** 2 / (len(u) * p))[keep].sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4
    # azimuth uniform
    phi = np.arctan2(d[:, 1], d[:, 0])
    hp, _ = np.histogram(phi, bins=16, range=(-math.pi, math.pi))
    assert stats.chisquare(hp).pvalue > 1e-4


def test_hg_pdf_normalised():
    # 2 pi * integral of the phase function over u = 1
    for g in (-0.5, 0.0, 0.6, 0.95):
        u = np.linspace(-1, 1, 200001)
        f = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * u) ** 1.5)
        assert abs(2 * math.pi * np.trapezoid(f, u) - 1) < 1e-4


def test_grid_filter_reproduces_multilinear_fields():
    n = (5, 7, 9)  # nz, ny, nx
    z, y, x = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    # trilinear filtering reproduces multilinear fields exactly
    grid = ((0.5 * x + 1) * (0.25 * y + 2) * (3 - 0.125 * z)).astype(np.float32)
    rng = np.random.default_rng(2)
    # interior points: between the first and last texel centres
    uvw = np.stack([(0.5 + rng.random(200) * (n[2] - 1)) / n[2], (0.5 + rng.random(200) * (n[1] - 1)) / n[1],
                    (0.5 + rng.random(200) * (n[0] - 1)) / n[0]], axis=1)
    got = oracle.tex3(grid, uvw)
    tx, ty, tz = uvw[:, 0] * n[2] - 0.5, uvw[:, 1] * n[1] - 0.5, uvw[:, 2] * n[0] - 0.5
    want = (0.5 * tx + 1) * (0.25 * ty + 2) * (3 - 0.125 * tz)
    assert np.abs(got - want).max() < 1e-4
    # clamp addressing: beyond the outer texel centres the edge value holds
    edge = oracle.tex3(grid, [(-0.3, 0.5, 0.5), (1.7, 0.5, 0.5)])
    inner = oracle.tex3(grid, [(0.5 / n[2], 0.5, 0.5), (1 - 0.5 / n[2], 0.5, 0.5)])
    assert np.array_equal(edge, inner)


def medium_scene(add):
    s = scenes.Scene()
    add(s)
    return oracle.OracleScene(s)


def segments(n, a, b):
    seg = np.zeros(n, dtype=_abi.ray_dtype())
    seg["o"] = a
    seg["d"] = np.subtract(b, a)
    seg["tmax"] = 1.0
    return seg


def test_homogeneous_free_flight_and_tr():
    sigma_t, albedo = 0.7, (0.2, 0.5, 0.9)
    os_ = medium_scene(lambda s: s.add_homogeneous(sigma_t, albedo, g=0.3))
    n, L = 100000, 2.0
    out = os_.medium_eval(0, segments(n, (0, 0, 0), (L, 0, 0)), np.arange(n, dtype=np.uint64) * 7919 + 1)
    sc = out[:, 0] == 1
    p = math.exp(-sigma_t * L)
    assert abs((~sc).mean() - p) < 5 * math.sqrt(p * (1 - p) / n)
    # scattering distance: exponential with rate sigma_t, truncated at L
    x = out[sc, 1].astype(np.float64)
    ks = stats.kstest(x, lambda t: (1 - np.exp(-sigma_t * t)) / (1 - p))
    assert ks.pvalue > 1e-4
    assert np.allclose(out[sc, 4:7], albedo)
    # the reference's transmittance uses sigma_a = sigma_t (1 - albedo) only
    want = np.exp(-L * sigma_t * (1 - np.array(albedo)))
    assert np.allclose(out[:, 7:10], want, rtol=1e-5)


def optical_depth(grid, w2l, a, b, steps=20000):
    t = (np.arange(steps) + 0.5) / steps
    p = a[None, :] + t[:, None] * (b - a)[None, :]
    hp = np.concatenate([p, np.ones((steps, 1))], axis=1) @ w2l.T
    uvw = 0.5 + 0.5 * hp[:, :3]
    return oracle.tex3(grid, uvw).astype(np.float64).mean() * np.linalg.norm(b - a)


def test_heterogeneous_tracking_statistics():
    grid = scenes.fbm_grid(24, octaves=3, seed=9, max_sigma=1.5)
    w2l = np.diag([1 / 4, 1 / 4, 1 / 4, 1.0]) @ scenes.translate(-1, 0, 0)
    os_ = medium_scene(lambda s: s.add_heterogeneous(grid, (0.8, 0.8, 0.8), g=0.6, world_to_local=w2l))
    rng = np.random.default_rng(4)
    for _ in range(3):
        a = rng.uniform(-3, 5, 3)
        b = rng.uniform(-3, 5, 3)
        tau = optical_depth(grid, w2l, a, b)
        n = 40000
        out = os_.medium_eval(0, segments(n, a, b), rng.integers(1, 2 ** 63, n, dtype=np.uint64))
        sc = out[:, 0] == 1
        p = math.exp(-tau)
        # delta tracking: no scattering with probability exp(-tau)
        assert abs((~sc).mean() - p) < 5 * math.sqrt(p * (1 - p) / n) + 2e-3
        # ratio tracking: unbiased for exp(-tau)
        tr = out[:, 7].astype(np.float64)
        assert abs(tr.mean() - p) < 5 * tr.std() / math.sqrt(n) + 2e-3
        assert (out[:, 7] == out[:, 8]).all() and (out[:, 8] == out[:, 9]).all()
        # scattering points lie on the segment and carry the medium albedo
        if sc.any():
            v = out[sc, 1:4] - a
            s_ = v @ (b - a) / np.dot(b - a, b - a)
            assert ((s_ >= -1e-5) & (s_ <= 1 + 1e-5)).all()
            assert np.allclose(out[sc, 4:7], 0.8)


def test_heterogeneous_albedo_grid_is_filtered():
    grid = np.full((4, 4, 4), 5.0, dtype=np.float32)
    alb = np.zeros((4, 4, 4, 3), dtype=np.float32)
    alb[..., 0] = np.arange(4)[None, None, :] / 4.0  # varies along x
    alb[..., 1] = 0.5
    alb[..., 2] = 0.25
    os_ = medium_scene(lambda s: s.add_heterogeneous(grid, (0.9, 0.9, 0.9), albedo_grid=alb))
    n = 2000
    out = os_.medium_eval(0, segments(n, (-1, 0, 0), (1, 0, 0)), np.arange(1, n + 1, dtype=np.uint64))
    sc = out[:, 0] == 1
    assert sc.mean() > 0.99
    x = out[sc, 1]
    u = 0.5 + 0.5 * x
    want = np.clip(u * 4 - 0.5, 0, 3) / 4.0
    assert np.abs(out[sc, 4] - want).max() < 1e-5
    assert np.allclose(out[sc, 5], 0.5) and np.allclose(out[sc, 6], 0.25)


def test_grid_text_round_trip(tmp_path):
    g = scenes.fbm_grid(8, octaves=2, seed=1)
    p = tmp_path / "grid.txt"
    scenes.save_grid_text(p, g)
    head = open(p).readline().split()
    assert head == ["float", "1", "8", "8", "8"]
    assert np.array_equal(scenes.load_grid_text(p), g)
