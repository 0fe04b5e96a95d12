"""Pinhole camera orientation, field of view and aspect ratio
(PinholeCamera::commit / generate_ray_inline, camera/pinhole.cpp:55-99;
pixel to film coordinate, renderer/wavefront/generate.cpp:64-75).

Camera rays that miss every triangle see a gradient sky, whose radiance is
linear in the cosine between the ray and the sky's up axis
(light/gradient_sky.cpp:20-25). The expected image is computed here from the
textbook pinhole model, independently of the renderer and the oracle: image
row 0 is the side `up` points to, column 0 the side opposite to
right = forward x up, the vertical field of view spans the image height and
the horizontal extent is `width / height` times the vertical one. The sky's
up axis is tilted in both image directions, so an upside-down or mirrored
film, a transposed film, a wrong field of view or a wrong aspect ratio each
change the pixels well beyond the tolerance.
"""
import numpy as np
import pytest

from btrc_amd import scenes
from oracle import oracle

from util import render_params

EYE = np.array([0.3, -0.2, 1.0])
DST = EYE + np.array([0.25, 0.1, -1.0])
UP = np.array([0.05, 1.0, 0.1])            # not orthogonal to the view direction
FOV = 40.0
SKY_UP = np.array([1.0, 2.0, 0.5])
LOWER = np.array([0.2, 0.5, 1.0])
UPPER = np.array([1.0, 0.4, 0.1])
W, H = 48, 32


def sky_scene():
    s = scenes.Scene()
    fwd = (DST - EYE) / np.linalg.norm(DST - EYE)
    back = EYE - 5.0 * fwd    # one small triangle behind the camera: no camera ray reaches it
    tri = np.array([[*back, *(back + [0.1, 0, 0]), *(back + [0, 0.1, 0])]], np.float32)
    s.add_instance(s.add_geometry(tri), s.add_diffuse([0.5, 0.5, 0.5]))
    s.set_gradient_sky(tuple(SKY_UP), tuple(LOWER), tuple(UPPER))
    return s, scenes.Camera(tuple(EYE), tuple(DST), tuple(UP), FOV)


def expected_image(w, h, sub=32):
    """pixel means of the sky over each pixel's footprint (box filter)"""
    fwd = (DST - EYE) / np.linalg.norm(DST - EYE)
    right = np.cross(fwd, UP)
    right /= np.linalg.norm(right)
    true_up = np.cross(right, fwd)
    t = np.tan(np.radians(FOV) / 2)
    a = w / h
    sky_up = SKY_UP / np.linalg.norm(SKY_UP)
    ss = (np.arange(sub) + 0.5) / sub
    img = np.zeros((h, w, 3))
    for py in range(h):
        v = 1.0 - 2.0 * (py + ss) / h                 # +1 at the top edge
        for px in range(w):
            u = 2.0 * (px + ss) / w - 1.0             # -1 at the left edge
            d = (fwd[None, None, :] + (t * a * u)[None, :, None] * right[None, None, :]
                 + (t * v)[:, None, None] * true_up[None, None, :])
            d /= np.linalg.norm(d, axis=2, keepdims=True)
            s = (0.5 * (d @ sky_up + 1.0)).mean()
            img[py, px] = LOWER * (1 - s) + UPPER * s
    return img


def check(img):
    want = expected_image(img.shape[1], img.shape[0])
    err = np.abs(img - want).max()
    assert err < 2e-3, err
    # the test discriminates: the mirrored and the flipped image differ by far more
    assert np.abs(want[:, ::-1] - want).max() > 0.05 and np.abs(want[::-1] - want).max() > 0.05


def test_camera_orientation_oracle():
    scene, cam = sky_scene()
    p = render_params(W, H, 64, min_depth=0, max_depth=1)
    film, ext, _ = oracle.OracleScene(scene).render(cam, p)
    assert ext == W * H * 64    # camera rays only
    check(film / p.spp)


@pytest.mark.gpu
def test_camera_orientation_gpu(ctx):
    scene, cam = sky_scene()
    p = render_params(W, H, 64, min_depth=0, max_depth=1)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, W, H)
    ctx.render(p)
    img, _ = ctx.film(W, H)
    check(img[..., :3].astype(np.float64))
