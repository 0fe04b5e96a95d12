// Neural radiance cache: network layout and input encoding.
//
// The cache (Mueller et al., "Real-time Neural Radiance Caching for Path
// Tracing", SIGGRAPH 2021) maps a path vertex to the radiance that leaves it
// toward the previous vertex, excluding the vertex's own emission. The
// reference has no cache (README.md:17 lists it as future work); its
// interface here is this build's own.
//
// Network: 64 inputs, five hidden layers of 64 ReLU units, 3 linear outputs,
// no biases (a constant-one input plays that role): 5 * 64 * 64 + 3 * 64 =
// 20,672 weights. Weights and activations are BF16, products accumulate in
// FP32 (MFMA). The output layer is stored padded to 32 rows; rows 3..31 stay
// zero.
//
// Input encoding (64 values) of a vertex record (NrcVertex), laid out for
// the MFMA operand it feeds: input k = 16 s + 8 h + j (slice s = 0..3, half
// h = 0..1, j = 0..7) is element 8 s + j of half h's 32 values:
//   e = 0..23  sin (e even) or cos (e odd) of 2^f pi x_a, f = 4 h + e / 6,
//              axis a = (e % 6) / 2, x the position normalised to the scene
//              box; each half starts its chain at 2^(4 h) pi x_a directly
//              and reaches the next three frequencies by the double-angle
//              identities sin 2t = 2 sin t cos t,
//              cos 2t = (cos t - sin t)(cos t + sin t)
//   e = 24..31 half 0: direction toward the previous vertex (3), shading
//              normal (3), roughness, diffuse red; half 1: diffuse green and
//              blue, specular reflectance (3), three constant ones
#pragma once

#include "detmath.h"
#include "material.h"
#include "vecmath.h"

namespace btrc {

constexpr int NRC_IN = 64;
constexpr int NRC_WIDTH = 64;
constexpr int NRC_LAYERS = 6;       // 5 hidden + output
constexpr int NRC_OUT = 3;
constexpr int NRC_OUT_ROWS = 32;    // output layer rows as stored
constexpr int NRC_FREQS = 8;
constexpr int NRC_PARAMS = 5 * NRC_WIDTH * NRC_IN + NRC_OUT_ROWS * NRC_WIDTH; // 22,528 stored
constexpr int NRC_REAL_PARAMS = 5 * NRC_WIDTH * NRC_IN + NRC_OUT * NRC_WIDTH; // 20,672 trained

// first weight of layer l; layer l is [rows][64] row-major (out x in)
BTRC_HD constexpr int nrc_layer_offset(int l) { return l * NRC_WIDTH * NRC_IN; }
BTRC_HD constexpr int nrc_layer_rows(int l) { return l == NRC_LAYERS - 1 ? NRC_OUT_ROWS : NRC_WIDTH; }

// what the shading stage records at a vertex (16 floats)
struct NrcVertex
{
    float p[3];      // world position
    float wo[3];     // unit direction toward the previous vertex
    float n[3];      // shading normal
    float roughness;
    float diffuse[3];
    float specular[3];
};

struct NrcBox
{
    float lo[3];
    float inv_extent[3]; // 1 / (hi - lo), per axis
};

// Diffuse and specular reflectance and roughness of a material, the inputs
// the cache sees: Disney's base colour split by metallic and transmission,
// its specular colour at normal incidence (0.08 * specular, tinted toward the
// base colour by metallic); Lambert: diffuse = albedo, specular 0,
// roughness 1.
BTRC_HD void nrc_material_inputs(const MaterialRec &m, float *roughness, V3 *diffuse, V3 *specular)
{
    const V3 c = v3(m.color[0], m.color[1], m.color[2]);
    if(m.type == MATERIAL_DISNEY)
    {
        *roughness = m.roughness;
        *diffuse = c * ((1.0f - m.metallic) * (1.0f - m.transmission));
        *specular = lerp3(v3s(0.08f * m.specular), c, m.metallic);
    }
    else
    {
        *roughness = 1.0f;
        *diffuse = c;
        *specular = v3s(0.0f);
    }
}

// the 32 network inputs of half h of a vertex record, in the order above
BTRC_HD void nrc_encode_half(const NrcVertex &v, const NrcBox &box, int h, float *x)
{
    const float base = h ? 16.0f * PI : PI;
    float s[3], c[3];
    for(int a = 0; a < 3; ++a)
    {
        const float q = (v.p[a] - box.lo[a]) * box.inv_extent[a];
        btrc_detmath::sincos(base * q, &s[a], &c[a]);
    }
    int k = 0;
    for(int f = 0; f < 4; ++f)
    {
        for(int a = 0; a < 3; ++a)
        {
            x[k++] = s[a];
            x[k++] = c[a];
            const float s2 = 2.0f * s[a] * c[a];
            const float c2 = (c[a] - s[a]) * (c[a] + s[a]);
            s[a] = s2;
            c[a] = c2;
        }
    }
    const float tail0[8] = { v.wo[0], v.wo[1], v.wo[2], v.n[0], v.n[1], v.n[2], v.roughness, v.diffuse[0] };
    const float tail1[8] = { v.diffuse[1], v.diffuse[2], v.specular[0], v.specular[1], v.specular[2],
                             1.0f, 1.0f, 1.0f };
    for(int j = 0; j < 8; ++j)
        x[k++] = h ? tail1[j] : tail0[j];
}

// relative L2 loss of one output channel (Mueller et al. eq. 3): the
// prediction in the denominator is a constant for the gradient
BTRC_HD float nrc_rel_l2(float y, float t) { return (y - t) * (y - t) / (y * y + 0.01f); }
BTRC_HD float nrc_rel_l2_grad(float y, float t) { return 2.0f * (y - t) / (y * y + 0.01f); }

} // namespace btrc
