// Participating media of the wavefront path: homogeneous and grid-based
// heterogeneous media with a Henyey-Greenstein phase function.
//
//  * HomogeneousMedium (builtin/medium/homogeneous.cpp:25-73): free-flight
//    distance with sigma_t; its transmittance uses sigma_a = sigma_t (1 -
//    albedo) only, as the reference does.
//  * HetergeneousMedium (builtin/medium/hetergeneous.cpp:25-106): delta
//    tracking for scattering, ratio tracking for transmittance, majorant =
//    max voxel (at least 0.001); grid coordinates 0.5 + 0.5 * uvw.
//  * TransformMedium (core/medium.cpp:98-129): uvw = world_to_local(p).
//  * HenyeyGreensteinPhaseShader (core/medium.cpp:32-76) and
//    henyey_greenstein (utils/henyey_greenstein.h:8-16).
//
// Grid lookups filter trilinearly with clamp addressing in the convention
// of normalised texture coordinates (texel i centred at (i + 0.5) / n), with
// full FP32 weights.
#pragma once

#include "detmath.h"
#include "vecmath.h"

namespace btrc {

enum : int32_t
{
    MEDIUM_HOMOGENEOUS = 0,
    MEDIUM_HETEROGENEOUS = 1,
};

// Tracking loops stop after this many steps. A loop reaches it only when a
// segment length or density is not finite; the path then sees no
// scattering and the transmittance reached so far.
constexpr int32_t MAX_TRACKING_STEPS = 1 << 20;

struct MediumRec
{
    int32_t type;
    int32_t has_xform;
    float sigma_t;          // homogeneous extinction
    float inv_max_density;  // heterogeneous: 1 / max(max voxel, 0.001)
    float albedo[3];        // constant albedo (grid albedo overrides)
    float g;                // Henyey-Greenstein asymmetry
    int32_t nx, ny, nz;     // voxel grid, x fastest
    int32_t albedo_grid;    // 1: 3 floats per voxel at albedo_offset
    uint32_t density_offset; // first voxel in SceneView::voxels
    uint32_t albedo_offset;
    float world_to_local[12]; // row-major 3x4
    float pad[2];
};

BTRC_HD float henyey_greenstein(float g, float u)
{
    const float g2 = g * g;
    const float dem = 1.0f + g2 - 2.0f * (g * u);
    return (1.0f - g2) / (4.0f * PI * dem * sqrtf(dem));
}

struct PhaseSample
{
    V3 phase;
    V3 dir;
    float pdf;
};

// HenyeyGreensteinPhaseShader::sample, core/medium.cpp:36-66
BTRC_HD PhaseSample hg_sample(float g, V3 color, V3 wo, float s0, float s1)
{
    const float s = s0 + s0 - 1.0f;
    float u;
    if(fabs_(g) < 0.001f)
        u = s;
    else
    {
        const float g2 = g * g;
        const float q = (1.0f - g2) / (1.0f + g * s);
        u = (1.0f + g2 - q * q) / (g + g);
    }
    const float cos_theta = -u;
    const float sin_theta = sqrtf(fmax_(0.0f, 1.0f - cos_theta * cos_theta));
    const float phi = 2.0f * PI * s1;
    const V3 local = v3(sin_theta * btrc_detmath::sin(phi), sin_theta * btrc_detmath::cos(phi), cos_theta);
    const float val = henyey_greenstein(g, u);
    const Frame f = frame_from_z(wo);
    PhaseSample r;
    r.phase = color * val;
    r.dir = f.x * local.x + f.y * local.y + f.z * local.z;
    r.pdf = val;
    return r;
}

// eval / pdf, core/medium.cpp:68-77: u = -cos(wi, wo)
BTRC_HD float hg_pdf(float g, V3 wi, V3 wo)
{
    return henyey_greenstein(g, -dot(normalize(wi), normalize(wo)));
}

BTRC_HD float clamp_texel_coord(float x, int32_t n)
{
    if(!(x >= -1.0f))
        x = -1.0f;
    if(x > (float)n)
        x = (float)n;
    return x;
}

// trilinear lookup of channel c of a grid with `ch` channels per voxel
BTRC_HD float grid_lookup(const float *g, int32_t nx, int32_t ny, int32_t nz, int32_t ch, int32_t c, V3 uvw)
{
    const float x = clamp_texel_coord(uvw.x * (float)nx - 0.5f, nx);
    const float y = clamp_texel_coord(uvw.y * (float)ny - 0.5f, ny);
    const float z = clamp_texel_coord(uvw.z * (float)nz - 0.5f, nz);
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const float tx = x - fx, ty = y - fy, tz = z - fz;
    const int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    const int32_t x0 = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix), x1 = ix + 1 < 0 ? 0 : (ix + 1 >= nx ? nx - 1 : ix + 1);
    const int32_t y0 = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy), y1 = iy + 1 < 0 ? 0 : (iy + 1 >= ny ? ny - 1 : iy + 1);
    const int32_t z0 = iz < 0 ? 0 : (iz >= nz ? nz - 1 : iz), z1 = iz + 1 < 0 ? 0 : (iz + 1 >= nz ? nz - 1 : iz + 1);
    auto at = [&](int32_t i, int32_t j, int32_t k) {
        return g[(((size_t)k * ny + j) * nx + i) * ch + c];
    };
    const float c00 = lerpf(at(x0, y0, z0), at(x1, y0, z0), tx);
    const float c10 = lerpf(at(x0, y1, z0), at(x1, y1, z0), tx);
    const float c01 = lerpf(at(x0, y0, z1), at(x1, y0, z1), tx);
    const float c11 = lerpf(at(x0, y1, z1), at(x1, y1, z1), tx);
    const float c0 = lerpf(c00, c10, ty);
    const float c1 = lerpf(c01, c11, ty);
    return lerpf(c0, c1, tz);
}

BTRC_HD V3 medium_uvw(const MediumRec &m, V3 p)
{
    if(!m.has_xform)
        return p;
    const float *w = m.world_to_local;
    return v3(w[0] * p.x + w[1] * p.y + w[2] * p.z + w[3],
              w[4] * p.x + w[5] * p.y + w[6] * p.z + w[7],
              w[8] * p.x + w[9] * p.y + w[10] * p.z + w[11]);
}

struct MediumSample
{
    bool scattered;
    V3 position;
    V3 color;  // phase colour (albedo) at the scattering point
    float g;
};

// Delta and ratio tracking through a grid medium, one step at a time: the
// loops of medium_sample and medium_tr below are made of these steps, and
// the wavefront's medium stage runs them one step per pass so that a lane
// whose segment is done takes the next path state.
struct GridTrack
{
    V3 a;           // segment start (world space)
    V3 local_a;     // its grid coordinate
    V3 local_step;  // grid coordinate per unit of t
    V3 world_step;  // world position per unit of t
    float t, t_max; // distance tracked so far and the segment's length
    int32_t it;     // steps taken (at most MAX_TRACKING_STEPS)
};

BTRC_HD GridTrack grid_track_start(const MediumRec &m, V3 a, V3 b)
{
    GridTrack g;
    const V3 uvw_a = medium_uvw(m, a), uvw_b = medium_uvw(m, b);
    g.t_max = length(b - a);
    g.local_a = v3s(0.5f) + 0.5f * uvw_a;
    const V3 local_b = v3s(0.5f) + 0.5f * uvw_b;
    g.local_step = (local_b - g.local_a) / g.t_max;
    g.world_step = (b - a) / g.t_max;
    g.a = a;
    g.t = 0.0f;
    g.it = 0;
    return g;
}

// One free-flight step against the majorant: false once the segment is left
// (or the step limit reached), else the grid point and its density. Smp
// provides float get1d().
template<class Smp>
BTRC_HD bool grid_track_step(const MediumRec &m, const float *voxels, GridTrack &g, Smp &smp, V3 *uvw,
                             float *density)
{
    if(g.it >= MAX_TRACKING_STEPS)
        return false;
    const float dt = -btrc_detmath::log(1.0f - smp.get1d()) * m.inv_max_density;
    g.t = g.t + dt;
    if(g.t >= g.t_max)
        return false;
    ++g.it;
    *uvw = g.local_a + g.t * g.local_step;
    *density = grid_lookup(voxels + m.density_offset, m.nx, m.ny, m.nz, 1, 0, *uvw);
    return true;
}

// the phase colour at grid point uvw
BTRC_HD V3 grid_albedo(const MediumRec &m, const float *voxels, V3 uvw)
{
    if(!m.albedo_grid)
        return v3(m.albedo[0], m.albedo[1], m.albedo[2]);
    const float *alb = voxels + m.albedo_offset;
    return v3(grid_lookup(alb, m.nx, m.ny, m.nz, 3, 0, uvw), grid_lookup(alb, m.nx, m.ny, m.nz, 3, 1, uvw),
              grid_lookup(alb, m.nx, m.ny, m.nz, 3, 2, uvw));
}

// Medium::sample along the segment a -> b (a, b in world space). Smp
// provides float get1d().
template<class Smp>
BTRC_HD MediumSample medium_sample(const MediumRec &m, const float *voxels, V3 a, V3 b, Smp &smp)
{
    MediumSample r;
    r.scattered = false;
    r.position = v3s(0.0f);
    r.color = v3s(0.0f);
    r.g = m.g;
    if(m.type == MEDIUM_HOMOGENEOUS)
    {
        const float t_max = length(b - a);
        const float st = -btrc_detmath::log(1.0f - smp.get1d()) / m.sigma_t;
        if(st < t_max)
        {
            const float t = st / t_max;
            r.scattered = true;
            r.position = a * (1.0f - t) + b * t;
            r.color = v3(m.albedo[0], m.albedo[1], m.albedo[2]);
        }
        return r;
    }
    GridTrack g = grid_track_start(m, a, b);
    V3 uvw;
    float density;
    while(grid_track_step(m, voxels, g, smp, &uvw, &density))
    {
        if(smp.get1d() < density * m.inv_max_density)
        {
            r.scattered = true;
            r.position = g.a + g.t * g.world_step;
            r.color = grid_albedo(m, voxels, uvw);
            break;
        }
    }
    return r;
}

// Medium::tr along a -> b
template<class Smp>
BTRC_HD V3 medium_tr(const MediumRec &m, const float *voxels, V3 a, V3 b, Smp &smp)
{
    if(m.type == MEDIUM_HOMOGENEOUS)
    {
        const V3 sigma_a = m.sigma_t * (v3s(1.0f) - v3(m.albedo[0], m.albedo[1], m.albedo[2]));
        const V3 e = -length(a - b) * sigma_a;
        return v3(btrc_detmath::exp(e.x), btrc_detmath::exp(e.y), btrc_detmath::exp(e.z));
    }
    GridTrack g = grid_track_start(m, a, b);
    float result = 1.0f;
    V3 uvw;
    float density;
    while(grid_track_step(m, voxels, g, smp, &uvw, &density))
        result = result * (1.0f - density * m.inv_max_density);
    return v3s(result);
}

} // namespace btrc
