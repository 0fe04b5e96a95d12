// Small vector/frame/transform toolkit shared by host setup code and gfx950
// kernels. Every operation has a fixed evaluation order (builds use
// -ffp-contract=off) so that host and device produce identical bits.
#pragma once

#include <cstdint>
#include <cstring>

#include "detmath.h"

#if defined(__HIPCC__)
#define BTRC_HD __host__ __device__ __forceinline__
#else
#define BTRC_HD static inline
#endif

#pragma STDC FP_CONTRACT OFF

namespace btrc {

constexpr float PI = 3.1415926535f;          // btrc_pi, utils/math/scalar.h:8
constexpr float MAX_FLOAT = 3.40282347e+38f; // btrc_max_float, utils/math/scalar.h:7

BTRC_HD uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
BTRC_HD float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
BTRC_HD int32_t f2i_bits(float f) { int32_t u; std::memcpy(&u, &f, 4); return u; }
BTRC_HD float i2f_bits(int32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

BTRC_HD float fmin_(float a, float b) { return a < b ? a : b; }
BTRC_HD float fmax_(float a, float b) { return a > b ? a : b; }
BTRC_HD float fabs_(float a) { return u2f(f2u(a) & 0x7fffffffu); }
BTRC_HD float saturate(float x) { return fmin_(fmax_(x, 0.0f), 1.0f); }
BTRC_HD float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
BTRC_HD float sqr(float x) { return x * x; }
BTRC_HD bool isfinite_(float x) { return (f2u(x) & 0x7f800000u) != 0x7f800000u; }

struct V3
{
    float x, y, z;
};

BTRC_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
BTRC_HD V3 v3s(float s) { return v3(s, s, s); }
BTRC_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
BTRC_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
BTRC_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
BTRC_HD V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
BTRC_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
BTRC_HD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
BTRC_HD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
BTRC_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
BTRC_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BTRC_HD V3 cross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
BTRC_HD float length_square(V3 a) { return dot(a, a); }
BTRC_HD float length(V3 a) { return sqrtf(dot(a, a)); }
BTRC_HD V3 normalize(V3 a) { const float inv = 1.0f / sqrtf(dot(a, a)); return a * inv; }
BTRC_HD bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
BTRC_HD bool isfinite3(V3 a) { return isfinite_(a.x) && isfinite_(a.y) && isfinite_(a.z); }
BTRC_HD float max3(V3 a) { return fmax_(fmax_(a.x, a.y), a.z); }
BTRC_HD V3 lerp3(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
BTRC_HD float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
BTRC_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Spectrum::get_lum, core/spectrum.cpp:84-87
BTRC_HD float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// Orthonormal frame; CFrame, utils/cmath/cframe.cpp
struct Frame
{
    V3 x, y, z;
};

// CFrame(x, y, z) normalizes each axis (cframe.cpp:18-23)
BTRC_HD Frame make_frame(V3 x, V3 y, V3 z)
{
    Frame f;
    f.x = normalize(x);
    f.y = normalize(y);
    f.z = normalize(z);
    return f;
}

BTRC_HD V3 local_to_global(const Frame &f, V3 l) { return f.x * l.x + f.y * l.y + f.z * l.z; }
BTRC_HD V3 global_to_local(const Frame &f, V3 g) { return v3(dot(g, f.x), dot(g, f.y), dot(g, f.z)); }
BTRC_HD bool in_positive_z_hemisphere(const Frame &f, V3 v) { return dot(v, f.z) > 0.0f; }

// CFrame::rotate_to_new_z, cframe.cpp:79-84
BTRC_HD Frame rotate_to_new_z(const Frame &f, V3 new_z)
{
    const V3 new_x = cross(f.y, new_z);
    const V3 new_y = cross(new_z, new_x);
    return make_frame(new_x, new_y, new_z);
}

// CFrame::from_z, cframe.cpp:56-68
BTRC_HD Frame frame_from_z(V3 z)
{
    const V3 new_z = normalize(z);
    V3 new_y;
    if(1.0f - fabs_(new_z.x) < 0.1f)
        new_y = cross(new_z, v3(0, 1, 0));
    else
        new_y = cross(new_z, v3(1, 0, 0));
    return make_frame(cross(new_z, new_y), new_y, new_z);
}

// Affine transform rows of local_to_world (mat) and world_to_local (inv);
// CTransform3D, utils/cmath/ctransform.cpp:31-55
struct Xform
{
    float m[12]; // row-major 3x4 of mat
    float n[12]; // row-major 3x4 of inv
};

BTRC_HD V3 xf_point(const Xform &t, V3 p)
{
    return v3(t.m[0] * p.x + t.m[1] * p.y + t.m[2] * p.z + t.m[3],
              t.m[4] * p.x + t.m[5] * p.y + t.m[6] * p.z + t.m[7],
              t.m[8] * p.x + t.m[9] * p.y + t.m[10] * p.z + t.m[11]);
}

BTRC_HD V3 xf_vector(const Xform &t, V3 v)
{
    return v3(t.m[0] * v.x + t.m[1] * v.y + t.m[2] * v.z,
              t.m[4] * v.x + t.m[5] * v.y + t.m[6] * v.z,
              t.m[8] * v.x + t.m[9] * v.y + t.m[10] * v.z);
}

// normal transform by the transposed inverse, ctransform.cpp:49-55
BTRC_HD V3 xf_normal(const Xform &t, V3 n)
{
    return v3(t.n[0] * n.x + t.n[4] * n.y + t.n[8] * n.z,
              t.n[1] * n.x + t.n[5] * n.y + t.n[9] * n.z,
              t.n[2] * n.x + t.n[6] * n.y + t.n[10] * n.z);
}

// Self-intersection-safe ray origin (Waechter and Binder, Ray Tracing Gems
// ch. 6), restating intersection_offset, utils/intersection.cpp:5-31.
BTRC_HD V3 intersection_offset(V3 pos, V3 nor)
{
    const float ORIGIN = 1.0f / 32.0f;
    const float FLOAT_SCALE = 1.0f / 65536.0f;
    const float INT_SCALE = 256.0f;
    const int32_t oix = (int32_t)(2.0f * INT_SCALE * nor.x);
    const int32_t oiy = (int32_t)(2.0f * INT_SCALE * nor.y);
    const int32_t oiz = (int32_t)(2.0f * INT_SCALE * nor.z);
    const float pix = i2f_bits(f2i_bits(pos.x) + (pos.x < 0.0f ? -oix : oix));
    const float piy = i2f_bits(f2i_bits(pos.y) + (pos.y < 0.0f ? -oiy : oiy));
    const float piz = i2f_bits(f2i_bits(pos.z) + (pos.z < 0.0f ? -oiz : oiz));
    return v3(fabs_(pos.x) < ORIGIN ? pos.x + FLOAT_SCALE * nor.x : pix,
              fabs_(pos.y) < ORIGIN ? pos.y + FLOAT_SCALE * nor.y : piy,
              fabs_(pos.z) < ORIGIN ? pos.z + FLOAT_SCALE * nor.z : piz);
}

// intersection.cpp:33-37: offset to the side the next direction leaves from
BTRC_HD V3 intersection_offset(V3 pos, V3 nor, V3 next_dir)
{
    return intersection_offset(pos, dot(nor, next_dir) > 0.0f ? nor : -nor);
}

// ---- distributions, utils/cmath/cdistribution.cpp ----

// cdistribution.cpp:5-13
BTRC_HD V3 sample_sphere_uniform(float u1, float u2)
{
    const float z = 1.0f - (u1 + u1);
    const float phi = 2.0f * PI * u2;
    const float r = sqrtf(fmax_(0.0f, 1.0f - z * z));
    float s, c;
    btrc_detmath::sincos(phi, &s, &c);
    return v3(r * c, r * s, z);
}

BTRC_HD float pdf_sample_sphere_uniform() { return 1.0f / (4.0f * PI); }

// concentric-disk cosine hemisphere, cdistribution.cpp:30-52
BTRC_HD V3 sample_hemisphere_zweighted(float u1, float u2)
{
    float sx = 0.0f, sy = 0.0f;
    u1 = 2.0f * u1 - 1.0f;
    u2 = 2.0f * u2 - 1.0f;
    if(u1 != 0.0f || u2 != 0.0f)
    {
        float theta, r;
        if(fabs_(u1) > fabs_(u2))
        {
            r = u1;
            theta = 0.25f * PI * (u2 / u1);
        }
        else
        {
            r = u2;
            theta = 0.5f * PI - 0.25f * PI * (u1 / u2);
        }
        float s, c;
        btrc_detmath::sincos(theta, &s, &c);
        sx = r * c;
        sy = r * s;
    }
    const float z = sqrtf(fmax_(0.0f, 1.0f - (sx * sx + sy * sy)));
    return v3(sx, sy, z);
}

// cdistribution.cpp:64-67
BTRC_HD float pdf_sample_hemisphere_zweighted(V3 v) { return v.z >= 0.0f ? v.z / PI : 0.0f; }

// cdistribution.cpp:69-73: returns (b1, b2)
BTRC_HD void sample_triangle_uniform(float u1, float u2, float *b1, float *b2)
{
    const float t = sqrtf(u1);
    *b1 = 1.0f - t;
    *b2 = t * u2;
}

} // namespace btrc
