"""Pins the oracle's radiance-cache network: its input encoding and forward
pass by weights that make the network compute known functions, and its
backward pass by central finite differences (FP64, no BF16 rounding)."""
import math

import numpy as np

from oracle import oracle

P = oracle.NRC_PARAMS
BOX = np.array([-1, -1, -1, 0.5, 0.5, 0.5], dtype=np.float32)  # [-1, 1]^3


def layer(w, l):
    return w[l * 4096:(l + 1) * 4096].reshape(64, 64) if l < 5 else w[5 * 4096:].reshape(32, 64)


def random_vertices(n, rng):
    v = np.zeros((n, 16), dtype=np.float32)
    v[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    v[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    v[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    v[:, 9] = rng.uniform(0.05, 1, n)
    v[:, 10:16] = rng.uniform(0, 1, (n, 6))
    return v


def random_weights(rng, scale=1.0):
    w = np.zeros(P, dtype=np.float32)
    for l in range(6):
        rows = 3 if l == 5 else 64
        layer(w, l)[:rows] = rng.uniform(-1, 1, (rows, 64)) * scale * math.sqrt(6 / 64)
    return w


def test_forward_reads_the_encoding():
    rng = np.random.default_rng(1)
    v = random_vertices(50, rng)
    w = np.zeros(P, dtype=np.float32)
    # unit 0 <- constant-one input (61); unit 1 <- direction x (48);
    # unit 2 <- sin(pi * x_normalised) (0); unit 3 <- diffuse red (55)
    for u, k in ((0, 61), (1, 48), (2, 0), (3, 55)):
        layer(w, 0)[u, k] = 1.0
    for l in range(1, 5):
        for u in range(4):
            layer(w, l)[u, u] = 1.0
    layer(w, 5)[0, 0] = 2.5
    layer(w, 5)[1, 1] = 1.0
    layer(w, 5)[1, 2] = 1.0
    layer(w, 5)[2, 3] = -1.0
    y = oracle.nrc_forward(w, BOX, v, bf16=False)
    q = (v[:, 0] - BOX[0]) * BOX[3]
    assert np.allclose(y[:, 0], 2.5)
    assert np.allclose(y[:, 1], np.maximum(v[:, 3], 0) + np.maximum(np.sin(np.pi * q), 0), atol=1e-6)
    assert np.allclose(y[:, 2], -v[:, 10], atol=1e-7)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    v = random_vertices(40, rng)
    t = rng.uniform(0, 2, (40, 3)).astype(np.float32)
    w = random_weights(rng)
    g, loss = oracle.nrc_gradients(w, BOX, v, t, bf16=False)
    y0 = oracle.nrc_forward(w, BOX, v, bf16=False)
    y0f = y0.astype(np.float32)
    den = (y0f * y0f + np.float32(0.01)).astype(np.float64)  # held constant, as in the loss
    assert abs(loss - np.mean((y0 - t) ** 2 / den)) < 1e-12

    def frozen_loss(wp):
        y = oracle.nrc_forward(wp, BOX, v, bf16=False)
        return np.mean((y - t) ** 2 / den)

    idx = np.concatenate([rng.choice(5 * 4096, 40, replace=False), 5 * 4096 + rng.choice(3 * 64, 10, replace=False)])
    for k in idx:
        h = 1e-4 * max(1.0, abs(float(w[k])))
        wp, wm = w.copy(), w.copy()
        wp[k] += np.float32(h)
        wm[k] -= np.float32(h)
        hk = (float(wp[k]) - float(wm[k])) / 2
        fd = (frozen_loss(wp) - frozen_loss(wm)) / (2 * hk)
        assert abs(fd - g[k]) <= 2e-4 * max(1e-3, abs(g[k])) + 1e-7, (k, fd, g[k])
    # padding rows of the output layer get no gradient
    assert (g[5 * 4096 + 3 * 64:] == 0).all()


def test_bf16_rounding_is_a_small_perturbation():
    rng = np.random.default_rng(3)
    v = random_vertices(200, rng)
    w = random_weights(rng)
    exact = oracle.nrc_forward(w, BOX, v, bf16=False)
    rounded = oracle.nrc_forward(w, BOX, v, bf16=True)
    rel = np.abs(rounded - exact) / (np.abs(exact).mean() + 1e-9)
    assert rel.max() < 0.1 and rel.mean() < 0.02


def frequency_input(f, a, c):
    """the input index of sin (c = 0) or cos (c = 1) of 2^f pi q_a: element
    e = 6 (f % 4) + 2 a + c of half f // 4, which is input 16 (e // 8) +
    8 half + e % 8"""
    h, e = f // 4, 6 * (f % 4) + 2 * a + c
    return 16 * (e // 8) + 8 * h + e % 8


def test_frequency_encoding_order_and_values():
    """each frequency input is sin or cos of 2^f pi q_a, q the box-normalised
    position, read through two ReLU units (+x and -x); the other inputs are
    the vertex's features and three constant ones"""
    rng = np.random.default_rng(4)
    v = random_vertices(64, rng)
    q = ((v[:, :3] - BOX[:3]) * BOX[3:]).astype(np.float64)
    for f in range(8):
        for a in range(3):
            for c in range(2):
                k = frequency_input(f, a, c)
                w = np.zeros(P, dtype=np.float32)
                layer(w, 0)[0, k], layer(w, 0)[1, k] = 1.0, -1.0
                for l in range(1, 5):
                    layer(w, l)[0, 0] = layer(w, l)[1, 1] = 1.0
                layer(w, 5)[0, 0], layer(w, 5)[0, 1] = 1.0, -1.0
                got = oracle.nrc_forward(w, BOX, v, bf16=False)[:, 0]
                want = (np.cos if c else np.sin)(2.0 ** f * np.pi * q[:, a])
                assert np.abs(got - want).max() < 2e-4, (f, a, c, np.abs(got - want).max())


def test_feature_inputs():
    """inputs 48..55: direction, normal, roughness, diffuse red; 56..60:
    diffuse green, blue, specular; 61..63: constant ones"""
    rng = np.random.default_rng(5)
    v = random_vertices(16, rng)
    want = {48: v[:, 3], 49: v[:, 4], 50: v[:, 5], 51: v[:, 6], 52: v[:, 7], 53: v[:, 8], 54: v[:, 9],
            55: v[:, 10], 56: v[:, 11], 57: v[:, 12], 58: v[:, 13], 59: v[:, 14], 60: v[:, 15],
            61: 1.0, 62: 1.0, 63: 1.0}
    freq = {frequency_input(f, a, c) for f in range(8) for a in range(3) for c in range(2)}
    assert freq == set(range(48))
    for k, x in want.items():
        w = np.zeros(P, dtype=np.float32)
        layer(w, 0)[0, k], layer(w, 0)[1, k] = 1.0, -1.0
        for l in range(1, 5):
            layer(w, l)[0, 0] = layer(w, l)[1, 1] = 1.0
        layer(w, 5)[0, 0], layer(w, 5)[0, 1] = 1.0, -1.0
        got = oracle.nrc_forward(w, BOX, v, bf16=False)[:, 0]
        assert np.allclose(got, x, atol=1e-7), k
