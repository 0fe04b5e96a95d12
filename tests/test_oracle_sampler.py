"""Pins the oracle's Halton sampler (sampler/halton.cpp, utils/low_discrepancy.cpp)
and the shared FP32 transcendentals by mathematics, not by restating them."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]


def radical_inverse(base, a):
    """textbook radical inverse (digit reversal) as an exact rational"""
    r, f = Fraction(0), Fraction(1, base)
    while a:
        r += (a % base) * f
        a //= base
        f /= base
    return r


def stratified(x, m, eps=2e-7):
    """the k-th smallest of m points lies in [k/m, (k+1)/m) up to float32
    rounding of the scrambled value"""
    xs = np.sort(x.astype(np.float64))
    k = np.arange(m)
    return bool(((xs >= k / m - eps) & (xs < (k + 1) / m + eps)).all())


def test_dimensions_are_stratified():
    # A 1x1 film gives halton index = sample index. Owen scrambling by nested
    # digit permutations keeps the radical inverse a (0,1)-sequence: the first
    # b^k indices of base-b dimension d fall one into each of the b^k
    # intervals [j/b^k, (j+1)/b^k), and so does every later aligned block.
    dims = 12
    n = 5 ** 4
    out = oracle.halton(1, 1, np.zeros(n), np.zeros(n), np.arange(n), dims)
    for k in range(dims):
        base = PRIMES[k + 2]  # get1d starts at dimension 2 (halton.cpp:64-65)
        m = 1
        while m * base <= n:
            m *= base
        assert stratified(out[:m, k], m), f"dimension {k + 2} (base {base})"
        if 2 * m <= n:
            assert stratified(out[m:2 * m, k], m), f"dimension {k + 2} (base {base}), second block"
        # and coarser strata of every prefix length that is a power of the base
        mm = base
        while mm <= m:
            assert stratified(out[:mm, k], mm)
            mm *= base


def test_scrambling_differs_per_dimension_and_stays_in_unit_interval():
    n = 4096
    out = oracle.halton(1, 1, np.zeros(n), np.zeros(n), np.arange(n), 40)
    assert (out >= 0).all() and (out < 1).all()
    # distinct scrambles per dimension: no two columns equal
    for a in range(40):
        for b in range(a + 1, 40):
            assert not np.array_equal(out[:, a], out[:, b])
    # roughly uniform
    assert abs(out.mean() - 0.5) < 0.01


@pytest.mark.parametrize("w,h", [(256, 256), (1920, 1080), (5, 3), (128, 243)])
def test_pixel_index_mapping(w, h):
    # The index of (pixel, sample 0) is the unique i < 2^a 3^b whose base-2
    # and base-3 radical inverses put the point inside the pixel's cell of
    # the 2^a x 3^b grid (pbrt-v4's Halton pixel stratification), and sample
    # s adds s * 2^a 3^b.
    sa = 1
    while sa < min(w, 128):
        sa *= 2
    sb = 1
    while sb < min(h, 128):
        sb *= 3
    stride = sa * sb
    seen = set()
    ny, nx = min(h, sb, 128), min(w, sa, 128)
    for py in range(ny):
        for px in range(nx):
            i = oracle.halton_index(w, h, px, py, 0)
            assert i < stride
            # pixels enter modulo 128 (halton.cpp:55-56)
            assert math.floor(radical_inverse(2, i) * sa) == (px % 128) % sa
            assert math.floor(radical_inverse(3, i) * sb) == (py % 128) % sb
            seen.add(i)
            assert oracle.halton_index(w, h, px, py, 7) == i + 7 * stride
    assert len(seen) == nx * ny
    # pixels repeat with period 128
    if w > 130:
        assert oracle.halton_index(w, h, 3, 1, 2) == oracle.halton_index(w, h, 131, 1, 2)
    if h > 130:
        assert oracle.halton_index(w, h, 3, 1, 2) == oracle.halton_index(w, h, 3, 129, 2)


def ulp(x):
    x = np.abs(np.asarray(x, np.float64)).astype(np.float32)
    return np.spacing(x).astype(np.float64)


@pytest.mark.parametrize("fn,ref,lo,hi", [(0, np.sin, -30.0, 30.0), (1, np.cos, -30.0, 30.0),
                                          (3, np.exp, -87.0, 88.0)])
def test_detmath_accuracy(fn, ref, lo, hi):
    x = np.random.default_rng(fn).uniform(lo, hi, 400000).astype(np.float32)
    got = oracle.detmath(fn, x).astype(np.float64)
    want = ref(x.astype(np.float64))
    err = np.abs(got - want) / np.maximum(ulp(want), 2.0 ** -149)
    assert err.max() <= 2.0, (fn, err.max(), x[err.argmax()])


def test_detmath_log_and_pow_accuracy():
    rng = np.random.default_rng(9)
    x = np.exp(rng.uniform(-80, 80, 400000)).astype(np.float32)
    got = oracle.detmath(2, x).astype(np.float64)
    want = np.log(x.astype(np.float64))
    assert (np.abs(got - want) / np.maximum(ulp(want), 2.0 ** -149)).max() <= 2.0
    a = rng.uniform(1e-4, 1.0, 200000).astype(np.float32)
    b = rng.uniform(0.0, 2.0, 200000).astype(np.float32)
    got = oracle.detmath(4, a, b).astype(np.float64)
    want = np.power(a.astype(np.float64), b.astype(np.float64))
    # exp(y log x): error grows with |y log x| <= 18.5 here
    assert (np.abs(got - want) / want).max() < 4e-6


def test_detmath_special_values():
    x = np.array([0.0, -0.0, 1.0, np.inf, -1.0, np.nan], np.float32)
    lg = oracle.detmath(2, x)
    assert lg[0] == -np.inf and lg[1] == -np.inf and lg[2] == 0.0 and lg[3] == np.inf
    assert np.isnan(lg[4]) and np.isnan(lg[5])
    ex = oracle.detmath(3, np.array([0.0, -200.0, 200.0, 1.0], np.float32))
    assert ex[0] == 1.0 and ex[1] == 0.0 and ex[2] == np.inf
    assert oracle.detmath(0, np.array([0.0], np.float32))[0] == 0.0
    assert oracle.detmath(1, np.array([0.0], np.float32))[0] == 1.0
