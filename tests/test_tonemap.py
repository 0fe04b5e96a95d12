"""ACES tone mapping (ACESToneMap, postprocess/tonemap.cpp:14-24,35-51): the
Narkowicz fit f(x) = x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14) of the
exposure-scaled colour, saturated to [0, 1], alpha unchanged. The oracle is
pinned at points where the fit has known values; the device kernel must
equal it bit for bit on the same film."""
import numpy as np
import pytest

from oracle import oracle


def fit64(x):
    return x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14)


def one_channel(xs, exposure=1.0):
    a = np.zeros((len(xs), 4), np.float32)
    a[:, 0] = xs
    a[:, 3] = 0.25
    out = oracle.aces(a, exposure)
    assert np.all(out[:, 3] == 0.25)
    return out[:, 0].astype(np.float64)


def test_aces_fit_known_points():
    knee = (0.56 + np.sqrt(0.56 ** 2 + 4 * 0.08 * 0.14)) / (2 * 0.08)   # f(knee) = 1
    xs = np.array([0.0, 1.0, 0.18, knee - 0.01, knee + 0.01, 100.0, -0.005, -1.0])
    got = one_channel(xs)
    assert got[0] == 0.0
    assert abs(got[1] - 2.54 / 3.16) < 1e-6
    assert abs(got[2] - fit64(0.18)) < 1e-6 and abs(got[2] - 0.26690) < 1e-4
    assert got[3] < 1.0 and got[4] == 1.0 and got[5] == 1.0
    assert got[6] == 0.0                       # small negative values: the numerator is negative
    assert got[7] == 1.0                       # large negative values: the rational is positive, saturated
    xs = np.linspace(0.0, 7.0, 1001)
    assert np.abs(one_channel(xs) - fit64(xs)).max() < 2e-6
    assert np.all(np.diff(one_channel(xs)) >= 0)


def test_aces_exposure_is_a_scale():
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 4, (64, 4)).astype(np.float32)
    e = np.float32(0.7)
    scaled = a.copy()
    scaled[:, :3] *= e
    assert np.array_equal(oracle.aces(a, e), oracle.aces(scaled, 1.0))


@pytest.mark.gpu
@pytest.mark.parametrize("exposure", [1.0, 0.37, 4.0])
def test_device_tonemap_equals_oracle(ctx, exposure):
    from btrc_amd import scenes
    from util import render_params
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 64, 48)
    ctx.render(render_params(64, 48, 4))
    film = ctx.film(64, 48)[0]
    got = ctx.tonemapped(64, 48, exposure)
    assert np.array_equal(got.view(np.uint32), oracle.aces(film, exposure).view(np.uint32))
    assert (got[..., :3] >= 0).all() and (got[..., :3] <= 1).all() and (got[..., 3] == 1).all()
    if exposure >= 1.0:
        assert 0 < (got[..., :3] == 1).sum() < got[..., :3].size     # the light saturates, the rest does not
