// 2D image textures: the Array2D texture plugin (texture2d/array2d.cpp:19-30)
// over a CUDA texture object (utils/cuda/texture.cpp:63-105) with its
// description (texture2d/description.cpp:32-93): normalised coordinates,
// per-axis address mode (clamp by default, wrap, mirror, border), point or
// linear filtering (linear by default), UNorm8 texels read as value / 255,
// optional sRGB decoding of 8-bit colour channels.
//
// Here the filtering is done in software: texel x of a W-wide image is
// centred at (x + 0.5) / W; linear filtering blends the four texels around
// (u W - 0.5, v H - 0.5) with FP32 weights; the address mode maps each texel
// index (clamp to the edge, wrap modulo the size, mirror with period 2 W,
// border: the border colour for indices outside the image). Row 0 is the
// image's first row (v = 0). sRGB decoding applies to red, green and blue of
// 8-bit texels before filtering, from a 256-entry table.
#pragma once

#include "vecmath.h"

namespace btrc {

enum : int32_t
{
    TEXTURE_RGBA8 = 0,    // UNorm8x4
    TEXTURE_RGBA32F = 1,  // F32x4
};

enum : int32_t
{
    ADDRESS_CLAMP = 0,
    ADDRESS_WRAP = 1,
    ADDRESS_MIRROR = 2,
    ADDRESS_BORDER = 3,
};

enum : int32_t
{
    TEX_FILTER_LINEAR = 0,
    TEX_FILTER_POINT = 1,
};

struct TextureRec
{
    int32_t width, height;
    uint32_t offset;      // first texel in the format's texel array
    int32_t format;
    int32_t address_u, address_v;
    int32_t filter;
    int32_t srgb;
    float border[4];
};

// texel arrays and the 8-bit decode table: [0, 256) value / 255,
// [256, 512) sRGB to linear
struct TextureView
{
    const TextureRec *textures;
    const uint32_t *texels8;
    const float4 *texels32;
    const float *decode;
};

BTRC_HD float4 tex_float4(float x, float y, float z, float w)
{
    float4 r;
    r.x = x; r.y = y; r.z = z; r.w = w;
    return r;
}

// texel index under an address mode; -1 for a border texel
BTRC_HD int32_t tex_address(int32_t i, int32_t n, int32_t mode)
{
    if(mode == ADDRESS_WRAP)
    {
        int32_t m = i % n;
        return m < 0 ? m + n : m;
    }
    if(mode == ADDRESS_MIRROR)
    {
        int32_t m = i % (2 * n);
        m = m < 0 ? m + 2 * n : m;
        return m < n ? m : 2 * n - 1 - m;
    }
    if(mode == ADDRESS_BORDER)
        return (i < 0 || i >= n) ? -1 : i;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

BTRC_HD float4 tex_texel(const TextureView &tv, const TextureRec &t, int32_t x, int32_t y)
{
    if(x < 0 || y < 0)
        return tex_float4(t.border[0], t.border[1], t.border[2], t.border[3]);
    const uint32_t i = t.offset + (uint32_t)y * (uint32_t)t.width + (uint32_t)x;
    if(t.format == TEXTURE_RGBA32F)
        return tv.texels32[i];
    const uint32_t p = tv.texels8[i];
    const float *lut = tv.decode + (t.srgb ? 256 : 0);
    return tex_float4(lut[p & 0xffu], lut[(p >> 8) & 0xffu], lut[(p >> 16) & 0xffu], tv.decode[p >> 24]);
}

// integer part of a texel coordinate, kept inside the int32 range
BTRC_HD int32_t tex_floor(float x, float *frac)
{
    float f = floorf(x);
    *frac = x - f;
    f = fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
    return (int32_t)f;
}

// sample_texture2d_4f at normalised (u, v)
BTRC_HD float4 tex_sample(const TextureView &tv, int32_t id, float u, float v)
{
    const TextureRec &t = tv.textures[id];
    const int32_t W = t.width, H = t.height;
    if(t.filter == TEX_FILTER_POINT)
    {
        float fa, fb;
        const int32_t x = tex_address(tex_floor(u * (float)W, &fa), W, t.address_u);
        const int32_t y = tex_address(tex_floor(v * (float)H, &fb), H, t.address_v);
        return tex_texel(tv, t, x, y);
    }
    float a, b;
    const int32_t x0 = tex_floor(u * (float)W - 0.5f, &a);
    const int32_t y0 = tex_floor(v * (float)H - 0.5f, &b);
    if(!(a >= 0.0f && a < 1.0f))
        a = 0.0f; // non-finite coordinate
    if(!(b >= 0.0f && b < 1.0f))
        b = 0.0f;
    const int32_t xa = tex_address(x0, W, t.address_u), xb = tex_address(x0 + 1, W, t.address_u);
    const int32_t ya = tex_address(y0, H, t.address_v), yb = tex_address(y0 + 1, H, t.address_v);
    const float4 t00 = tex_texel(tv, t, xa, ya), t10 = tex_texel(tv, t, xb, ya);
    const float4 t01 = tex_texel(tv, t, xa, yb), t11 = tex_texel(tv, t, xb, yb);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return tex_float4(((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
                      ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
                      ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z,
                      ((w00 * t00.w + w10 * t10.w) + w01 * t01.w) + w11 * t11.w);
}

} // namespace btrc
