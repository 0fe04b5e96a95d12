"""Pins the oracle's film filters against their densities: the box filter's
offsets are uniform on [-0.5, 0.5)^2 and GaussianFilter (film_filter/
gaussian.cpp:27-66) draws offsets from exp(-a x^2) - exp(-a r^2) times the
same in y, truncated to the disc of radius r."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle

R, ALPHA = 1.3, 1.6


def density_cells(r, alpha, bins):
    """probability of each of bins x bins square cells of [-r, r]^2 under
    the truncated separable density, by midpoint quadrature on a fine grid"""
    sub = 8
    n = bins * sub
    c = (np.arange(n) + 0.5) / n * 2 * r - r
    x, y = np.meshgrid(c, c)
    g = lambda t: np.maximum(0.0, np.exp(-alpha * t * t) - math.exp(-alpha * r * r))
    f = g(x) * g(y) * (x * x + y * y < r * r)
    f = f.reshape(bins, sub, bins, sub).sum(axis=(1, 3))
    return f / f.sum()


def test_gaussian_filter_matches_density():
    rng = np.random.default_rng(3)
    n = 400000
    u = rng.random((n, 3), dtype=np.float32)
    off = oracle.gaussian_filter(R, ALPHA, u)
    assert (np.abs(off) <= R + 1e-6).all()
    # support: the cells whose centres lie inside the disc. CAliasTable::sample
    # accepts cell i when n*u - i <= accept_prob (calias.cpp:19), so a draw
    # with n*u an exact integer takes its own cell even at zero weight; those
    # are the only draws allowed outside
    outside = np.hypot(off[:, 0], off[:, 1]) > R + 2 * R / 256 * math.sqrt(2)
    nu = u[:, 0].astype(np.float64) * 65536
    assert (nu[outside] == np.floor(nu[outside])).all()
    assert outside.mean() < 2e-3
    bins = 16
    h, _, _ = np.histogram2d(off[:, 1], off[:, 0], bins=bins, range=[[-R, R], [-R, R]])
    p = density_cells(R, ALPHA, bins)
    keep = p * n > 20
    chi2 = (((h - n * p) ** 2) / np.maximum(n * p, 1e-12))[keep].sum()
    dof = keep.sum() - 1
    assert stats.chi2.sf(chi2, dof) > 1e-4, (chi2, dof)
    # symmetric, zero mean
    assert abs(off[:, 0].mean()) < 5 * off[:, 0].std() / math.sqrt(n)
    assert abs(off[:, 1].mean()) < 5 * off[:, 1].std() / math.sqrt(n)


@pytest.mark.parametrize("alpha", [0.5, 4.0])
def test_gaussian_filter_variance_tracks_alpha(alpha):
    rng = np.random.default_rng(4)
    off = oracle.gaussian_filter(R, alpha, rng.random((200000, 3), dtype=np.float32))
    p = density_cells(R, alpha, 64)
    c = (np.arange(64) + 0.5) / 64 * 2 * R - R
    var_x = (p.sum(axis=0) * c * c).sum() + (2 * R / 64) ** 2 / 12
    assert abs(off[:, 0].var() / var_x - 1) < 0.02
