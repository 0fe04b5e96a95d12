// Deterministic single-precision transcendental functions.
//
// Every function here is a fixed sequence of IEEE-754 binary32 operations
// (add, mul, div, sqrt, fma, integer bit manipulation). Each of those is
// correctly rounded both on the x86-64 host and on gfx950, so one source
// compiled by g++ for the host and by hipcc for the GPU yields the same bits
// for the same input. The path tracer uses these functions wherever a
// transcendental value feeds a sampling decision, so that a GPU path and a
// host path driven by the same sample stream follow the same branch at every
// vertex. Builds must use -ffp-contract=off; fused operations are spelled out
// with fmaf() where they are intended.
//
// Methods (textbook, no vendor library constants):
//   sin/cos : Cody-Waite reduction by pi/2 (three-part split of pi/2, each
//             part the float nearest the remaining residual), then Taylor
//             polynomials on [-pi/4, pi/4] (sin to x^9, cos to x^10).
//   log     : x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
//             s = (m-1)/(m+1), atanh series to s^11.
//   exp     : x = k ln2 + r, |r| <= ln2/2, Taylor to r^8, scale by 2^k.
//   pow     : exp(y * log(x)) for x > 0.
//   atan2   : octant reduction to t = min/max in [0, 1], then t > tan(pi/8)
//             maps to pi/4 + atan((t-1)/(t+1)); Taylor series of atan to
//             s^21 on |s| <= tan(pi/8); pi/4, pi/2, pi added as two-part
//             splits.
//   acos    : 2 atan2(sqrt(1-x), sqrt(1+x)).
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BTRC_DM_FN __host__ __device__ __forceinline__
#else
#include <cmath>
#define BTRC_DM_FN static inline
#endif

#pragma STDC FP_CONTRACT OFF

namespace btrc_detmath {

BTRC_DM_FN uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
BTRC_DM_FN float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// round half away from zero, result as float (|x| < 2^31 assumed)
BTRC_DM_FN float round_away(float x)
{
    const int32_t k = (int32_t)(x + (x >= 0.0f ? 0.5f : -0.5f));
    return (float)k;
}

// 2^k for integer k in [-126, 127]
BTRC_DM_FN float exp2i(int32_t k) { return u2f((uint32_t)(k + 127) << 23); }

// sin and cos of x. Accurate to about 1 ulp for |x| < 2^16.
BTRC_DM_FN void sincos(float x, float *s, float *c)
{
    const float TWO_OVER_PI = 0x1.45f306p-1f;
    const float P1 = 0x1.921fb6p+0f;   // float nearest pi/2
    const float P2 = -0x1.777a5cp-25f; // float nearest pi/2 - P1
    const float P3 = -0x1.ee59dap-50f; // float nearest pi/2 - P1 - P2

    const float kf = round_away(x * TWO_OVER_PI);
    float r = fmaf(-kf, P1, x);
    r = fmaf(-kf, P2, r);
    r = fmaf(-kf, P3, r);
    const int32_t q = ((int32_t)kf) & 3;

    const float r2 = r * r;
    // sin(r) = r + r^3 * (-1/3! + r^2 (1/5! + r^2 (-1/7! + r^2 / 9!)))
    float ps = fmaf(r2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    ps = fmaf(r2, ps, 0x1.111112p-7f);
    ps = fmaf(r2, ps, -0x1.555556p-3f);
    const float sr = fmaf(r * r2, ps, r);
    // cos(r) = 1 + r^2 * (-1/2 + r^2 (1/4! + r^2 (-1/6! + r^2 (1/8! - r^2/10!))))
    float pc = fmaf(r2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    pc = fmaf(r2, pc, -0x1.6c16c2p-10f);
    pc = fmaf(r2, pc, 0x1.555556p-5f);
    pc = fmaf(r2, pc, -0.5f);
    const float cr = fmaf(r2, pc, 1.0f);

    // quadrant q: (sin, cos) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr),
    // chosen without branches (selects and sign flips only)
    const bool odd = (q & 1) != 0;
    const float so = odd ? cr : sr;
    const float co = odd ? sr : cr;
    *s = (q & 2) ? -so : so;
    *c = ((q + 1) & 2) ? -co : co;
}

BTRC_DM_FN float sin(float x) { float s, c; sincos(x, &s, &c); return s; }
BTRC_DM_FN float cos(float x) { float s, c; sincos(x, &s, &c); return c; }

// natural logarithm; log(0) = -inf, log(x<0) = NaN, log(inf) = inf
BTRC_DM_FN float log(float x)
{
    if(!(x > 0.0f))
        return x == 0.0f ? u2f(0xff800000u) : u2f(0x7fc00000u);
    if(!(x < u2f(0x7f800000u)))
        return x;
    int32_t eadj = 0;
    if(x < 0x1p-126f)
    {
        x = x * 0x1p23f;
        eadj = -23;
    }
    const uint32_t bits = f2u(x);
    int32_t e = (int32_t)((bits >> 23) & 0xffu) - 127 + eadj;
    float m = u2f((bits & 0x007fffffu) | 0x3f800000u);
    if(m > 0x1.6a09e6p+0f) // sqrt(2)
    {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    // 2 atanh(s) = 2s + s^3 (2/3 + s^2 (2/5 + s^2 (2/7 + s^2 (2/9 + s^2 2/11))))
    float p = fmaf(s2, 0x1.745d18p-3f, 0x1.c71c72p-3f);
    p = fmaf(s2, p, 0x1.24924ap-2f);
    p = fmaf(s2, p, 0x1.99999ap-2f);
    p = fmaf(s2, p, 0x1.555556p-1f);
    const float lnm = fmaf(s * s2, p, s + s);
    const float LN2_HI = 0x1.62e430p-1f;
    const float LN2_LO = -0x1.05c610p-29f;
    const float ef = (float)e;
    return fmaf(ef, LN2_HI, fmaf(ef, LN2_LO, lnm));
}

// e^x
BTRC_DM_FN float exp(float x)
{
    if(x != x)
        return x;
    if(x > 88.72284f)
        return u2f(0x7f800000u);
    if(x < -103.97208f)
        return 0.0f;
    const float LOG2E = 0x1.715476p+0f;
    const float LN2_HI = 0x1.62e430p-1f;
    const float LN2_LO = -0x1.05c610p-29f;
    const float kf = round_away(x * LOG2E);
    float r = fmaf(-kf, LN2_HI, x);
    r = fmaf(-kf, LN2_LO, r);
    // Taylor: 1 + r + r^2/2! + ... + r^8/8!
    float p = fmaf(r, 0x1.a01a02p-16f, 0x1.a01a02p-13f);
    p = fmaf(r, p, 0x1.6c16c2p-10f);
    p = fmaf(r, p, 0x1.111112p-7f);
    p = fmaf(r, p, 0x1.555556p-5f);
    p = fmaf(r, p, 0x1.555556p-3f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    int32_t k = (int32_t)kf;
    if(k > 127)
    {
        p = p * 0x1p127f;
        k = k - 127;
    }
    else if(k < -126)
    {
        p = p * 0x1p-126f;
        k = k + 126;
    }
    return p * exp2i(k);
}

// x^y for x >= 0 (x = 0 gives 0 for y > 0 and 1 for y = 0)
BTRC_DM_FN float pow(float x, float y)
{
    if(y == 0.0f)
        return 1.0f;
    if(x == 0.0f)
        return y > 0.0f ? 0.0f : u2f(0x7f800000u);
    return exp(y * log(x));
}

// arctangent of s for |s| <= tan(pi/8):
// s + s^3 (-1/3 + s^2 (1/5 + s^2 (-1/7 + ... + s^2 (-1/19 + s^2 / 21))))
BTRC_DM_FN float atan_small(float s)
{
    const float s2 = s * s;
    float p = fmaf(s2, 1.0f / 21.0f, -1.0f / 19.0f);
    p = fmaf(s2, p, 1.0f / 17.0f);
    p = fmaf(s2, p, -1.0f / 15.0f);
    p = fmaf(s2, p, 1.0f / 13.0f);
    p = fmaf(s2, p, -1.0f / 11.0f);
    p = fmaf(s2, p, 1.0f / 9.0f);
    p = fmaf(s2, p, -1.0f / 7.0f);
    p = fmaf(s2, p, 1.0f / 5.0f);
    p = fmaf(s2, p, -1.0f / 3.0f);
    return fmaf(s * s2, p, s);
}

// atan2(y, x) in [-pi, pi], with the C library's signed-zero and infinity
// cases; NaN in gives NaN out
BTRC_DM_FN float atan2(float y, float x)
{
    if(x != x || y != y)
        return x + y;
    const float INF = u2f(0x7f800000u);
    const float PI4_HI = 0x1.921fb6p-1f, PI4_LO = -0x1.777a5cp-26f;
    const float PI2_HI = 0x1.921fb6p+0f, PI2_LO = -0x1.777a5cp-25f;
    const float PI_HI = 0x1.921fb6p+1f, PI_LO = -0x1.777a5cp-24f;
    const bool xneg = (f2u(x) >> 31) != 0, yneg = (f2u(y) >> 31) != 0;
    const float ax = u2f(f2u(x) & 0x7fffffffu), ay = u2f(f2u(y) & 0x7fffffffu);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    float a;
    if(mx == 0.0f)
        a = 0.0f;
    else
    {
        float t;
        if(mx == INF)
            t = mn == INF ? 1.0f : 0.0f;
        else
            t = mn / mx;
        if(t > 0x1.a8279ap-2f) // tan(pi/8)
            a = (atan_small((t - 1.0f) / (t + 1.0f)) + PI4_LO) + PI4_HI;
        else
            a = atan_small(t);
        if(ay > ax)
            a = (PI2_LO - a) + PI2_HI;
    }
    if(xneg)
        a = (PI_LO - a) + PI_HI;
    return yneg ? -a : a;
}

// arccosine in [0, pi]; NaN outside [-1, 1]
BTRC_DM_FN float acos(float x)
{
    return 2.0f * atan2(sqrtf(1.0f - x), sqrtf(1.0f + x));
}

} // namespace btrc_detmath
