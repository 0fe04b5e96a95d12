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
// Input encoding (64 values) of a vertex record (NrcVertex):
//   [0, 48)  sin(2^f pi x_a), cos(2^f pi x_a) for f = 0..7 and the three
//            axes of the position normalised to the scene box, ordered
//            (f, axis, sin/cos); f = 0 evaluated directly, each higher
//            frequency from the one below by the double-angle identities
//            sin 2t = 2 sin t cos t, cos 2t = (cos t - sin t)(cos t + sin t)
//   [48, 51) direction toward the previous vertex (unit)
//   [51, 54) shading normal
//   [54]     roughness
//   [55, 58) diffuse reflectance
//   [58, 61) specular reflectance
//   [61, 64) constant 1
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

// the 64 network inputs of a vertex record
BTRC_HD void nrc_encode(const NrcVertex &v, const NrcBox &box, float *x)
{
    float q[3];
    for(int a = 0; a < 3; ++a)
        q[a] = (v.p[a] - box.lo[a]) * box.inv_extent[a];
    int k = 0;
    float s[3], c[3];
    for(int a = 0; a < 3; ++a)
    {
        const float t = PI * q[a];
        s[a] = btrc_detmath::sin(t);
        c[a] = btrc_detmath::cos(t);
    }
    for(int f = 0; f < NRC_FREQS; ++f)
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
    for(int a = 0; a < 3; ++a)
        x[k++] = v.wo[a];
    for(int a = 0; a < 3; ++a)
        x[k++] = v.n[a];
    x[k++] = v.roughness;
    for(int a = 0; a < 3; ++a)
        x[k++] = v.diffuse[a];
    for(int a = 0; a < 3; ++a)
        x[k++] = v.specular[a];
    while(k < NRC_IN)
        x[k++] = 1.0f;
}

// relative L2 loss of one output channel (Mueller et al. eq. 3): the
// prediction in the denominator is a constant for the gradient
BTRC_HD float nrc_rel_l2(float y, float t) { return (y - t) * (y - t) / (y * y + 0.01f); }
BTRC_HD float nrc_rel_l2_grad(float y, float t) { return 2.0f * (y - t) / (y * y + 0.01f); }

} // namespace btrc
