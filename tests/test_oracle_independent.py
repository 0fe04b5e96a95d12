"""Pins the oracle's independent sampler (sampler/independent.cpp:6-25): the
MurmurHash64A pixel hash (utils/hash.h) and the PCG32 stream, by published
known-answer values and by properties of the generator. How CUJ's cstd::PCG
seeds its stream and turns bits into floats is not in the reference tree;
the oracle follows pbrt-v4's RNG (which utils/hash.h itself follows), so
bit parity with the reference's own runs is unpinned for that step."""
import struct

import numpy as np

from oracle import oracle

MASK = (1 << 64) - 1


def test_murmur64a_smhasher_verification():
    # SMHasher's VerificationTest: hash keys 0..i-1 (i = 0..255) with seed
    # 256 - i, hash the 256 concatenated hashes with seed 0; MurmurHash64A's
    # published verification value is the low 32 bits of that, 0x1F0D3804
    key = bytes(range(256))
    hashes = b"".join(oracle.murmur64a(key[:i], 256 - i).to_bytes(8, "little") for i in range(256))
    assert oracle.murmur64a(hashes, 0) & 0xFFFFFFFF == 0x1F0D3804


def test_murmur64a_tail_bytes_matter():
    # every tail length changes the hash (a dropped tail case would collide)
    base = bytes(range(1, 16))
    seen = {oracle.murmur64a(base[:n], 0) for n in range(16)}
    assert len(seen) == 16


def test_pcg32_reference_outputs():
    # pcg32-demo (pcg-random.org reference C code): pcg32_srandom_r(rng, 42, 54)
    out = oracle.pcg32(54, 42, 6)
    assert [int(x) for x in out] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_jump_ahead_equals_stepping():
    seq, off = 0x123456789ABCDEF, 77
    stream = oracle.pcg32(seq, off, 5000)
    for k in (0, 1, 2, 3, 63, 64, 65, 1000, 4095):
        assert np.array_equal(oracle.pcg32(seq, off, 10, skip=k), stream[k:k + 10])


def test_pcg32_period_wraps():
    # the LCG has period 2^64: jumping 2^64 - k after k numbers returns to the start
    seq, off = 99, 5
    start = oracle.pcg32(seq, off, 8)
    for k in (1, 3, 7):
        assert np.array_equal(oracle.pcg32(seq, off, 8, skip=(1 << 64) - k)[k:], start[:8 - k])


def mix_bits(v):
    v ^= v >> 31
    v = (v * 0x7FB5D329728EA185) & MASK
    v ^= v >> 27
    v = (v * 0x81DADEF4BC2DD44D) & MASK
    v ^= v >> 33
    return v


def test_independent_sampler_streams():
    # IndependentSampler(pixel, s): stream murmur(x, y), offset mix_bits(seq)
    # (pbrt-v4 RNG(seq)), advanced by 65536 numbers per sample index
    for x, y in ((0, 0), (17, 3), (1919, 1079)):
        seq = oracle.murmur64a(struct.pack("<II", x, y), 0)
        raw = oracle.pcg32(seq, mix_bits(seq), 16, skip=0)
        for s in (0, 1, 5):
            got = oracle.independent([x], [y], [s], 16)[0]
            bits = oracle.pcg32(seq, mix_bits(seq), 16, skip=65536 * s)
            want = np.minimum((bits.astype(np.float64) / 2.0 ** 32).astype(np.float32),
                              np.float32(float.fromhex("0x1.fffffep-1")))
            assert np.array_equal(got, want)
        assert not np.array_equal(raw, oracle.pcg32(seq, mix_bits(seq), 16, skip=65536))


def test_independent_sampler_uniformity():
    n = 20000
    px = np.arange(n) % 256
    py = np.arange(n) // 256
    u = oracle.independent(px, py, np.zeros(n), 4)
    assert ((u >= 0) & (u < 1)).all()
    # mean and variance of U[0, 1) within 5 sigma
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 5 * 0.0745 / np.sqrt(u.size)
    # neighbouring pixels are decorrelated
    assert abs(np.corrcoef(u[:-1, 0], u[1:, 0])[0, 1]) < 0.05
