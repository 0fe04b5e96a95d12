// Halton sampler with per-digit Owen scrambling and pixel stratification,
// restating HaltonSampler (builtin/sampler/halton.cpp:39-133) and the
// low-discrepancy helpers it calls (utils/low_discrepancy.cpp:10-110, after
// pbrt-v4). The wavefront tracer of the reference aliases its global sampler
// to an independent PCG stream (renderer/wavefront/common.h:37); this build
// wires the Halton sampler in its place, so every random number of a path is
// a pure function of (pixel, sample index, dimension).
#pragma once

#include "vecmath.h"

namespace btrc {

constexpr int PRIME_TABLE_SIZE = 1000; // utils/prime.h:9
constexpr int MAX_HALTON_RESOLUTION = 128; // halton.cpp:9

// The first PRIME_TABLE_SIZE primes (utils/prime.cpp), produced by trial
// division at compile time.
struct PrimeTable
{
    int p[PRIME_TABLE_SIZE];
};

constexpr PrimeTable make_prime_table()
{
    PrimeTable t{};
    int n = 0;
    for(int c = 2; n < PRIME_TABLE_SIZE; ++c)
    {
        bool is_prime = true;
        for(int d = 2; d * d <= c; ++d)
        {
            if(c % d == 0)
            {
                is_prime = false;
                break;
            }
        }
        if(is_prime)
            t.p[n++] = c;
    }
    return t;
}

// pbrt-v4 MixBits, low_discrepancy.cpp:49-57
BTRC_HD uint64_t mix_bits(uint64_t v)
{
    v = v ^ (v >> 31);
    v = v * 0x7fb5d329728ea185ull;
    v = v ^ (v >> 27);
    v = v * 0x81dadef4bc2dd44dull;
    v = v ^ (v >> 33);
    return v;
}

// Kensler's hashed permutation, low_discrepancy.cpp:12-45
BTRC_HD uint32_t permutation_element(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w = w | (w >> 1);
    w = w | (w >> 2);
    w = w | (w >> 4);
    w = w | (w >> 8);
    w = w | (w >> 16);
    for(;;)
    {
        i = i ^ p;
        i = i * 0xe170893du;
        i = i ^ (p >> 16);
        i = i ^ ((i & w) >> 4);
        i = i ^ (p >> 8);
        i = i * 0x0929eb3fu;
        i = i ^ (p >> 23);
        i = i ^ ((i & w) >> 1);
        i = i * (1u | p >> 27);
        i = i * 0x6935fa69u;
        i = i ^ ((i & w) >> 11);
        i = i * 0x74dcb303u;
        i = i ^ ((i & w) >> 2);
        i = i * 0x9e501cc3u;
        i = i ^ ((i & w) >> 2);
        i = i * 0xc860a3dfu;
        i = i & w;
        i = i ^ (i >> 5);
        if(i < l)
            break;
    }
    return (i + p) % l;
}

// Quotient and remainder of a by a small base; a stays below 2^32 for every
// image this tracer renders, which keeps the division in 32 bits.
BTRC_HD void divmod_base(uint64_t a, uint32_t base, uint64_t *q, uint64_t *r)
{
    if(a < 0x100000000ull)
    {
        const uint32_t a32 = (uint32_t)a;
        const uint32_t q32 = a32 / base;
        *q = q32;
        *r = a32 - q32 * base;
    }
    else
    {
        const uint64_t q64 = a / base;
        *q = q64;
        *r = a - q64 * base;
    }
}

// owen_scrambled_radical_inverse, low_discrepancy.cpp:83-101
BTRC_HD float owen_scrambled_radical_inverse(uint32_t base, uint64_t a, uint64_t hash)
{
    const float inv_base = 1.0f / (float)base;
    float inv_base_m = 1.0f;
    uint64_t reversed_digits = 0;
    while(1.0f - inv_base_m < 1.0f)
    {
        uint64_t next, digit_value;
        divmod_base(a, base, &next, &digit_value);
        const uint64_t digit_hash = mix_bits(hash ^ reversed_digits);
        digit_value = permutation_element((uint32_t)digit_value, base, (uint32_t)digit_hash);
        reversed_digits = reversed_digits * base + digit_value;
        inv_base_m = inv_base_m * inv_base;
        a = next;
    }
    return fmin_(inv_base_m * (float)reversed_digits, 0x1.fffffep-1f);
}

// inverse_radical_inverse, low_discrepancy.cpp:103-114
BTRC_HD uint64_t inverse_radical_inverse(uint64_t inverse, uint32_t base, int digits)
{
    uint64_t index = 0;
    for(int i = 0; i < digits; ++i)
    {
        const uint64_t digit = inverse % base;
        inverse = inverse / base;
        index = index * base + digit;
    }
    return index;
}

// Per-film constants of the Halton sampler (initialize_consts, halton.cpp:118-133)
struct HaltonConsts
{
    int32_t base_scales[2];
    int32_t base_exponents[2];
    int32_t mult_inverse[2];
};

struct HaltonState
{
    uint32_t dimension;
    uint64_t index;
};

// halton.cpp:39-44 (sample_halton)
BTRC_HD float sample_halton(const int *primes, uint32_t dim, uint64_t halton_index)
{
    return owen_scrambled_radical_inverse(
        (uint32_t)primes[dim], halton_index, mix_bits((uint64_t)(1u + (dim << 4))));
}

// HaltonSampler(res, pixel, sample_index), halton.cpp:48-66
BTRC_HD HaltonState halton_start(const HaltonConsts &hc, uint32_t px, uint32_t py, int32_t sample_index)
{
    HaltonState st;
    st.index = 0;
    const int32_t sample_stride = hc.base_scales[0] * hc.base_scales[1];
    if(sample_stride > 1)
    {
        const uint32_t pm[2] = { px % MAX_HALTON_RESOLUTION, py % MAX_HALTON_RESOLUTION };
        for(int i = 0; i < 2; ++i)
        {
            const uint64_t dim_offset = inverse_radical_inverse(
                (uint64_t)pm[i], i == 0 ? 2u : 3u, hc.base_exponents[i]);
            st.index = st.index + dim_offset * (uint64_t)(sample_stride / hc.base_scales[i])
                                             * (uint64_t)hc.mult_inverse[i];
        }
        st.index = st.index % (uint64_t)sample_stride;
    }
    st.index = st.index + (uint64_t)(sample_index * sample_stride);
    st.dimension = 2; // halton.cpp:64-65
    return st;
}

// get1d / get2d / get3d, halton.cpp:80-116
BTRC_HD float halton_get1d(const int *primes, HaltonState &st)
{
    if(st.dimension >= (uint32_t)PRIME_TABLE_SIZE)
        st.dimension = 2;
    const float r = sample_halton(primes, st.dimension, st.index);
    st.dimension = st.dimension + 1;
    return r;
}

BTRC_HD void halton_get2d(const int *primes, HaltonState &st, float *a, float *b)
{
    if(st.dimension + 1 >= (uint32_t)PRIME_TABLE_SIZE)
        st.dimension = 2;
    *a = sample_halton(primes, st.dimension, st.index);
    *b = sample_halton(primes, st.dimension + 1, st.index);
    st.dimension = st.dimension + 2;
}

BTRC_HD void halton_get3d(const int *primes, HaltonState &st, float *a, float *b, float *c)
{
    if(st.dimension + 2 >= (uint32_t)PRIME_TABLE_SIZE)
        st.dimension = 2;
    *a = sample_halton(primes, st.dimension, st.index);
    *b = sample_halton(primes, st.dimension + 1, st.index);
    *c = sample_halton(primes, st.dimension + 2, st.index);
    st.dimension = st.dimension + 3;
}

} // namespace btrc
