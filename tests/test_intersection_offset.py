"""The self-intersection offset of new ray origins (intersection_offset,
utils/intersection.cpp:5-31; Waechter and Binder, "A Fast and Robust Method
for Avoiding Self-Intersection", Ray Tracing Gems ch. 6).

The method acts on each axis alone. Away from the origin (|p| >= 1/32) it
moves the coordinate by trunc(512 |n|) representable floats towards the sign
of the normal coordinate n; within 1/32 of the origin it adds 2^-16 n in
float arithmetic. The checks below state those properties without the bit
trick: ulp steps are taken with nextafter, one at a time.
"""
import numpy as np
import pytest

from oracle import oracle

FN_OFFSET = 9


def step_ulps(p, k, direction):
    """p moved by k representable float32 values towards +inf (direction > 0) or -inf"""
    r = np.float32(p)
    target = np.float32(np.inf if direction > 0 else -np.inf)
    for _ in range(k):
        r = np.nextafter(r, target)
    return r


def cases():
    rng = np.random.default_rng(7)
    mags = np.concatenate([10.0 ** rng.uniform(-6, 4, 300), [1 / 32, np.nextafter(np.float32(1 / 32), 0), 1.0, 2.0,
                                                              np.nextafter(np.float32(2.0), 0), 0.0]])
    p = (mags * rng.choice([-1.0, 1.0], len(mags))).astype(np.float32)
    n = rng.uniform(-1, 1, len(p)).astype(np.float32)
    n[:20] = rng.uniform(-1 / 512, 1 / 512, 20)       # below one ulp step
    n[20:24] = [1.0, -1.0, 0.0, -0.0]
    return p, n


def expected(p, n):
    out = np.empty_like(p)
    for i, (a, b) in enumerate(zip(p, n)):
        if abs(a) < np.float32(1 / 32):
            out[i] = np.float32(a) + np.float32(2.0 ** -16) * np.float32(b)
        else:
            out[i] = step_ulps(a, int(abs(np.float32(512.0) * b)), 1 if b > 0 else -1)
    return out


def test_offset_properties_oracle():
    p, n = cases()
    got = oracle.detmath(FN_OFFSET, p, n)
    want = expected(p, n)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), np.nonzero(got != want)[0][:5]
    # the new origin lies on the normal's side of the surface point, or on it
    assert np.all((got - p) * n >= 0)
    far = np.abs(p) >= np.float32(1 / 32)
    moved = far & (np.abs(n) >= 1 / 512)
    assert np.all(got[moved] != p[moved]) and np.all(got[far & (np.abs(n) < 1 / 512)] == p[far & (np.abs(n) < 1 / 512)])


@pytest.mark.gpu
def test_offset_bitwise_gpu(ctx):
    p, n = cases()
    assert np.array_equal(ctx.eval_detmath(FN_OFFSET, p, n).view(np.uint32),
                          oracle.detmath(FN_OFFSET, p, n).view(np.uint32))
