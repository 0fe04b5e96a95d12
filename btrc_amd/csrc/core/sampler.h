// The two samplers of the wavefront path:
//  * Halton with per-digit Owen scrambling and pixel stratification,
//    restating HaltonSampler (builtin/sampler/halton.cpp:39-133) and the
//    low-discrepancy helpers it calls (utils/low_discrepancy.cpp:10-110,
//    after pbrt-v4); every number of a path is a pure function of (pixel,
//    sample index, dimension);
//  * the independent sampler the reference's wavefront tracer uses by
//    default (renderer/wavefront/common.h:37, sampler/independent.cpp:6-25):
//    one PCG32 stream per pixel seeded by MurmurHash64A of the pixel
//    coordinates, advanced by 65536 numbers per sample index.
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
// the mask Kensler's permutation works in: l - 1 with every lower bit set
BTRC_HD uint32_t permutation_mask(uint32_t l)
{
    uint32_t w = l - 1;
    w = w | (w >> 1);
    w = w | (w >> 2);
    w = w | (w >> 4);
    w = w | (w >> 8);
    w = w | (w >> 16);
    return w;
}

BTRC_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

// x mod l from m = floor((2^32 - 1) / l): mulhi(x, m) > x / l - 1, so one
// correction step gives the exact remainder
BTRC_HD uint32_t mod_recip(uint32_t x, uint32_t l, uint32_t m)
{
    uint32_t r = x - mulhi32(x, m) * l;
    if(r >= l)
        r = r - l;
    return r;
}

// Kensler's permutation of i over [0, l) with w = permutation_mask(l) and
// m = floor((2^32 - 1) / l)
BTRC_HD uint32_t permutation_element_masked(uint32_t i, uint32_t l, uint32_t w, uint32_t m, uint32_t p)
{
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
    return mod_recip(i + p, l, m);
}

BTRC_HD uint32_t permutation_element(uint32_t i, uint32_t l, uint32_t p)
{
    return permutation_element_masked(i, l, permutation_mask(l), 0xffffffffu / l, p);
}

// Quotient and remainder of a by a small base; a stays below 2^32 for every
// image this tracer renders, which keeps the division in 32 bits. There it
// multiplies by m = floor((2^32 - 1) / base): mulhi(a, m) > a / base - 1, so
// one correction step gives the exact quotient.
BTRC_HD void divmod_base(uint64_t a, uint32_t base, uint32_t m, uint64_t *q, uint64_t *r)
{
    if(a < 0x100000000ull)
    {
        const uint32_t a32 = (uint32_t)a;
        uint32_t q32 = mulhi32(a32, m);
        uint32_t r32 = a32 - q32 * base;
        if(r32 >= base)
        {
            q32 = q32 + 1;
            r32 = r32 - base;
        }
        *q = q32;
        *r = r32;
    }
    else
    {
        const uint64_t q64 = a / base;
        *q = q64;
        *r = a - q64 * base;
    }
}

// One digit of owen_scrambled_radical_inverse (low_discrepancy.cpp:83-101):
// the lowest base-`base` digit of a, permuted by the hash of the digits
// reversed so far, appended to them; a loses the digit. m and w are the
// base's reciprocal (divmod_base) and permutation mask.
BTRC_HD uint64_t owen_digit(uint32_t base, uint32_t m, uint32_t w, uint64_t hash, uint64_t &a, uint64_t reversed)
{
    uint64_t next, digit_value;
    divmod_base(a, base, m, &next, &digit_value);
    const uint64_t digit_hash = mix_bits(hash ^ reversed);
    digit_value = permutation_element_masked((uint32_t)digit_value, base, w, m, (uint32_t)digit_hash);
    a = next;
    return reversed * base + digit_value;
}

// owen_scrambled_radical_inverse, low_discrepancy.cpp:83-101 (the digit
// split's reciprocal and the permutation's mask computed once per call)
BTRC_HD float owen_scrambled_radical_inverse(uint32_t base, uint64_t a, uint64_t hash)
{
    const float inv_base = 1.0f / (float)base;
    const uint32_t m = 0xffffffffu / base;
    const uint32_t w = permutation_mask(base);
    float inv_base_m = 1.0f;
    uint64_t reversed_digits = 0;
    while(1.0f - inv_base_m < 1.0f)
    {
        uint64_t next, digit_value;
        divmod_base(a, base, m, &next, &digit_value);
        const uint64_t digit_hash = mix_bits(hash ^ reversed_digits);
        digit_value = permutation_element_masked((uint32_t)digit_value, base, w, m, (uint32_t)digit_hash);
        reversed_digits = reversed_digits * base + digit_value;
        inv_base_m = inv_base_m * inv_base;
        a = next;
    }
    return fmin_(inv_base_m * (float)reversed_digits, 0x1.fffffep-1f);
}

// Tables of the first digits. The first K scrambled digits of a depend only
// on a mod base^K, so for the first HALTON_TABLE_DIMS dimensions a table
// holds, for every r < base^K (K the largest with base^K <= 4096), the
// reversed digits after K steps of the loop above on a = r. The rest of the
// digits continue the loop from there; results are identical.
constexpr int HALTON_TABLE_DIMS = 256;
constexpr uint32_t HALTON_TABLE_SPAN = 4096;

// a dimension's table: where its values start, base^K, K, and
// floor((2^32 - 1) / base^K) for divmod_base
struct HaltonDim
{
    uint32_t first, span, digits, recip;
};

struct HaltonTables
{
    const uint16_t *values; // all dimensions' tables
    const HaltonDim *dims;  // HALTON_TABLE_DIMS entries
};

// the same as owen_scrambled_radical_inverse(base, a, hash) with the first
// info.digits digits from table t
BTRC_HD float owen_scrambled_radical_inverse_tabled(uint32_t base, uint64_t a, uint64_t hash, const uint16_t *t,
                                                    HaltonDim info)
{
    const float inv_base = 1.0f / (float)base;
    const uint32_t m = 0xffffffffu / base;
    const uint32_t w = permutation_mask(base);
    uint64_t high, low;
    divmod_base(a, info.span, info.recip, &high, &low); // a / base^K, a mod base^K
    uint64_t reversed_digits = t[info.first + low];
    a = high;
    float inv_base_m = 1.0f;
    for(uint32_t k = 0; k < info.digits; ++k)
        inv_base_m = inv_base_m * inv_base;
    while(1.0f - inv_base_m < 1.0f)
    {
        reversed_digits = owen_digit(base, m, w, hash, a, reversed_digits);
        inv_base_m = inv_base_m * inv_base;
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

// halton.cpp:39-44 (sample_halton); with tables (non-null values) the
// first digits of the first HALTON_TABLE_DIMS dimensions come from them
BTRC_HD float sample_halton(const int *primes, uint32_t dim, uint64_t halton_index,
                            HaltonTables tab = HaltonTables{ nullptr, nullptr })
{
    const uint64_t hash = mix_bits((uint64_t)(1u + (dim << 4)));
    if(tab.values && dim < (uint32_t)HALTON_TABLE_DIMS)
        return owen_scrambled_radical_inverse_tabled((uint32_t)primes[dim], halton_index, hash, tab.values,
                                                     tab.dims[dim]);
    return owen_scrambled_radical_inverse((uint32_t)primes[dim], halton_index, hash);
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
BTRC_HD float halton_get1d(const int *primes, HaltonState &st, HaltonTables tab = HaltonTables{ nullptr, nullptr })
{
    if(st.dimension >= (uint32_t)PRIME_TABLE_SIZE)
        st.dimension = 2;
    const float r = sample_halton(primes, st.dimension, st.index, tab);
    st.dimension = st.dimension + 1;
    return r;
}

BTRC_HD void halton_get2d(const int *primes, HaltonState &st, float *a, float *b,
                          HaltonTables tab = HaltonTables{ nullptr, nullptr })
{
    if(st.dimension + 1 >= (uint32_t)PRIME_TABLE_SIZE)
        st.dimension = 2;
    *a = sample_halton(primes, st.dimension, st.index, tab);
    *b = sample_halton(primes, st.dimension + 1, st.index, tab);
    st.dimension = st.dimension + 2;
}

BTRC_HD void halton_get3d(const int *primes, HaltonState &st, float *a, float *b, float *c,
                          HaltonTables tab = HaltonTables{ nullptr, nullptr })
{
    if(st.dimension + 2 >= (uint32_t)PRIME_TABLE_SIZE)
        st.dimension = 2;
    *a = sample_halton(primes, st.dimension, st.index, tab);
    *b = sample_halton(primes, st.dimension + 1, st.index, tab);
    *c = sample_halton(primes, st.dimension + 2, st.index, tab);
    st.dimension = st.dimension + 3;
}

// one dimension's table (host side): the reversed digits after the first
// info.digits steps for every a < info.span
BTRC_HD void fill_halton_table(uint32_t base, uint32_t dim, HaltonDim info, uint16_t *out)
{
    const uint64_t hash = mix_bits((uint64_t)(1u + (dim << 4)));
    const uint32_t m = 0xffffffffu / base, w = permutation_mask(base);
    for(uint32_t r = 0; r < info.span; ++r)
    {
        uint64_t a = r, reversed = 0;
        for(uint32_t j = 0; j < info.digits; ++j)
            reversed = owen_digit(base, m, w, hash, a, reversed);
        out[r] = (uint16_t)reversed;
    }
}

// K and base^K of a base's table: the most digits with base^K <= HALTON_TABLE_SPAN
BTRC_HD HaltonDim halton_dim_shape(uint32_t base, uint32_t first)
{
    uint32_t span = 1, k = 0;
    while(span * base <= HALTON_TABLE_SPAN)
    {
        span *= base;
        ++k;
    }
    return HaltonDim{ first, span, k, 0xffffffffu / span };
}

// ---- independent sampler

// MurmurHash64A (Austin Appleby, public domain; the form of utils/hash.h:12-71)
// of the 8-byte key (x, y) as hash::hash(pixel.x, pixel.y) lays it out
// (two little-endian uint32, seed 0): one whole block, no tail bytes.
BTRC_HD uint64_t murmur_hash_pixel(uint32_t x, uint32_t y)
{
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    uint64_t h = 8ull * m;
    uint64_t k = (uint64_t)x | ((uint64_t)y << 32);
    k = k * m;
    k = k ^ (k >> 47);
    k = k * m;
    h = (h ^ k) * m;
    h = h ^ (h >> 47);
    h = h * m;
    h = h ^ (h >> 47);
    return h;
}

// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically
// Good Algorithms for Random Number Generation", 2014): 64-bit LCG state,
// XSH-RR output. Seeding by stream index as pbrt-v4's RNG(seq_index).
struct Pcg32
{
    uint64_t state;
    uint64_t inc;
};

constexpr uint64_t PCG32_MULT = 0x5851f42d4c957f2dull;

BTRC_HD uint32_t pcg32_next(Pcg32 &r)
{
    const uint64_t old = r.state;
    r.state = old * PCG32_MULT + r.inc;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

// stream `seq`, starting offset `offset` (pcg32_srandom_r)
BTRC_HD Pcg32 pcg32_seed(uint64_t seq, uint64_t offset)
{
    Pcg32 r;
    r.state = 0;
    r.inc = (seq << 1) | 1ull;
    (void)pcg32_next(r);
    r.state = r.state + offset;
    (void)pcg32_next(r);
    return r;
}

// jump ahead by delta numbers in O(log delta) (Brown, "Random Number
// Generation with Arbitrary Strides", 1994)
BTRC_HD void pcg32_advance(Pcg32 &r, uint64_t delta)
{
    uint64_t cur_mult = PCG32_MULT, cur_plus = r.inc;
    uint64_t acc_mult = 1, acc_plus = 0;
    while(delta > 0)
    {
        if(delta & 1ull)
        {
            acc_mult = acc_mult * cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1ull) * cur_plus;
        cur_mult = cur_mult * cur_mult;
        delta >>= 1;
    }
    r.state = acc_mult * r.state + acc_plus;
}

// uniform float in [0, 1): 32 random bits times 2^-32, capped below 1
BTRC_HD float pcg32_uniform_float(Pcg32 &r)
{
    const float f = (float)pcg32_next(r) * 2.3283064365386963e-10f;
    return f < 0x1.fffffep-1f ? f : 0x1.fffffep-1f;
}

// IndependentSampler(res, pixel, sample_index), independent.cpp:11-15
BTRC_HD Pcg32 independent_start(uint32_t px, uint32_t py, int32_t sample_index)
{
    const uint64_t seq = murmur_hash_pixel(px, py);
    Pcg32 r = pcg32_seed(seq, mix_bits(seq));
    pcg32_advance(r, (uint64_t)((int64_t)sample_index * 65536));
    return r;
}

} // namespace btrc
