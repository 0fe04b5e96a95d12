// The volume-primitive medium (VolumePrimitiveMedium, core/volume.cpp:108-197;
// volume::Aggregate, core/volume/aggregate.cpp:128-254; volume::BVH queries,
// core/volume/bvh.cpp:127-324).
//
// Each volume primitive is the parallelepiped o + u x + v y + w z with u, v,
// w in [0, 1], a density sigma_t(u, v, w) and an albedo(u, v, w) (grids
// filtered like the 3D textures of the media, or constants). Along a
// segment a -> b the medium walks from boundary to boundary: the next point
// where the segment enters or leaves any primitive splits it, the primitives
// holding the piece's midpoint (at most four) form its overlap set, and
// delta / ratio tracking runs over the piece with the summed densities and
// the summed maximum densities as majorant; scattering is isotropic with
// the density-weighted albedo. The reference finds boundaries and overlaps
// with a BVH over the primitives; a linear scan gives the same sets (the
// closest boundary is a minimum; the overlap set is every primitive that
// holds the point), summed in primitive order.
#pragma once

#include "detmath.h"
#include "medium.h"
#include "vecmath.h"

namespace btrc {

constexpr int32_t MAX_VOLUME_OVERLAP = 4; // MAX_OVERLAP_COUNT, core/volume/bvh.h:10

struct VolumeRec
{
    float o[3];
    float x[3], y[3], z[3];             // edges
    float inv_x2, inv_y2, inv_z2;       // 1 / |edge|^2 (VolumePrimitive::world_pos_to_uvw)
    float xd[3], yd[3], zd[3];          // edge / |edge|^2 (the BVH's primitive record)
    float max_density;                  // grid maximum or the constant
    float sigma_t;                      // constant density (nx == 0)
    int32_t nx, ny, nz;                 // density grid, x fastest; 0: constant
    uint32_t density_offset;            // into the voxel array
    int32_t ax, ay, az;                 // albedo grid (3 floats per voxel); 0: constant
    uint32_t albedo_offset;
    float albedo[3];
    float pad;
};

struct VolumeView
{
    const VolumeRec *vols;
    int32_t count;
    const float *voxels;
};

BTRC_HD float dot3(const float *a, V3 b) { return a[0] * b.x + a[1] * b.y + a[2] * b.z; }

// VolumePrimitive::world_pos_to_uvw, core/volume.cpp:44-51
BTRC_HD V3 volume_uvw(const VolumeRec &v, V3 p)
{
    const V3 op = p - v3(v.o[0], v.o[1], v.o[2]);
    return v3(dot3(v.x, op) * v.inv_x2, dot3(v.y, op) * v.inv_y2, dot3(v.z, op) * v.inv_z2);
}

// volume::BVH::is_in_prim, core/volume/bvh.cpp:275-285
BTRC_HD bool volume_contains(const VolumeRec &v, V3 p)
{
    const V3 op = p - v3(v.o[0], v.o[1], v.o[2]);
    const float u = dot3(v.xd, op), w2 = dot3(v.yd, op), w3 = dot3(v.zd, op);
    return 0.0f <= u && u <= 1.0f && 0.0f <= w2 && w2 <= 1.0f && 0.0f <= w3 && w3 <= 1.0f;
}

// volume::BVH::find_closest_intersection for one primitive, bvh.cpp:294-324:
// the first boundary crossing of p + t d in (0, t_max), or -1
BTRC_HD float volume_crossing(const VolumeRec &v, V3 p, V3 d, float t_max)
{
    const V3 op = p - v3(v.o[0], v.o[1], v.o[2]);
    float t0 = 0.0f, t1 = t_max;
    const float *axes[3] = { v.xd, v.yd, v.zd };
    for(int k = 0; k < 3; ++k)
    {
        const float opd = dot3(axes[k], op);
        const float dd = dot3(axes[k], d);
        const float nd = (0.0f - opd) / dd;
        const float fd = (1.0f - opd) / dd;
        t0 = fmaxf(t0, fminf(nd, fd));
        t1 = fminf(t1, fmaxf(nd, fd));
    }
    if(t0 > t1)
        return -1.0f;
    if(t0 > 0.0f)
        return t0;
    if(t1 < t_max)
        return t1;
    return -1.0f;
}

// closest crossing along a -> b over all primitives; false when none
BTRC_HD bool volume_closest_crossing(const VolumeView &vv, V3 a, V3 b, V3 *out)
{
    const V3 dir = b - a;
    float final_t = MAX_FLOAT, t_max = 1.0f;
    for(int32_t i = 0; i < vv.count; ++i)
    {
        const float t = volume_crossing(vv.vols[i], a, dir, t_max);
        if(t >= 0.0f)
        {
            final_t = fminf(final_t, t);
            t_max = final_t;
        }
    }
    if(final_t > 10.0f)
        return false;
    *out = a * (1.0f - final_t) + b * final_t;
    return true;
}

struct VolumeOverlap
{
    int32_t count;
    int32_t id[MAX_VOLUME_OVERLAP];
};

// volume::BVH::get_overlap, bvh.cpp:197-245: primitives holding p (first four)
BTRC_HD VolumeOverlap volume_overlap(const VolumeView &vv, V3 p)
{
    VolumeOverlap r;
    r.count = 0;
    for(int32_t i = 0; i < vv.count && r.count < MAX_VOLUME_OVERLAP; ++i)
        if(volume_contains(vv.vols[i], p))
            r.id[r.count++] = i;
    return r;
}

BTRC_HD float volume_density(const VolumeView &vv, const VolumeRec &v, V3 uvw)
{
    if(v.nx == 0)
        return v.sigma_t;
    return grid_lookup(vv.voxels + v.density_offset, v.nx, v.ny, v.nz, 1, 0, uvw);
}

BTRC_HD V3 volume_albedo(const VolumeView &vv, const VolumeRec &v, V3 uvw)
{
    if(v.ax == 0)
        return v3(v.albedo[0], v.albedo[1], v.albedo[2]);
    const float *g = vv.voxels + v.albedo_offset;
    return v3(grid_lookup(g, v.ax, v.ay, v.az, 3, 0, uvw), grid_lookup(g, v.ax, v.ay, v.az, 3, 1, uvw),
              grid_lookup(g, v.ax, v.ay, v.az, 3, 2, uvw));
}

// volume::Aggregate::get_max_density, aggregate.cpp:128-134
BTRC_HD float volume_majorant(const VolumeView &vv, const VolumeOverlap &ov)
{
    float m = 0.0f;
    for(int32_t k = 0; k < ov.count; ++k)
        m = m + vv.vols[ov.id[k]].max_density;
    return fmaxf(m, 0.01f);
}

// VolumePrimitiveMedium::sample, core/volume.cpp:119-159, with
// sample_scattering_in_overlap (aggregate.cpp:136-215) on each piece
template<class Smp>
BTRC_HD MediumSample volume_sample(const VolumeView &vv, V3 a, V3 b, Smp &smp)
{
    MediumSample r;
    r.scattered = false;
    r.position = v3s(0.0f);
    r.color = v3s(0.0f);
    r.g = 0.0f;
    if(vv.count == 0)
        return r;
    V3 o = a;
    const V3 d = normalize(b - a);
    for(int32_t it = 0; it < MAX_TRACKING_STEPS; ++it)
    {
        if(dot(b - o, d) <= 0.0f || length_square(b - o) < 0.00001f)
            break;
        V3 inct;
        if(!volume_closest_crossing(vv, o, b, &inct))
            inct = b;
        const VolumeOverlap ov = volume_overlap(vv, 0.5f * (o + inct));
        if(ov.count != 0)
        {
            const float inv_max = 1.0f / volume_majorant(vv, ov);
            V3 ua[MAX_VOLUME_OVERLAP], ub[MAX_VOLUME_OVERLAP];
            for(int32_t k = 0; k < ov.count; ++k)
            {
                ua[k] = volume_uvw(vv.vols[ov.id[k]], o);
                ub[k] = volume_uvw(vv.vols[ov.id[k]], inct);
            }
            const float t_max = length(inct - o);
            float t = 0.0f;
            for(int32_t step = 0; step < MAX_TRACKING_STEPS; ++step)
            {
                const float dt = -btrc_detmath::log(1.0f - smp.get1d()) * inv_max;
                t = t + dt;
                if(t >= t_max)
                    break;
                const float tf = t / t_max;
                float dens[MAX_VOLUME_OVERLAP];
                float sum = 0.0f;
                for(int32_t k = 0; k < ov.count; ++k)
                {
                    const V3 uvw = ua[k] * (1.0f - tf) + ub[k] * tf;
                    dens[k] = volume_density(vv, vv.vols[ov.id[k]], uvw);
                    sum = sum + dens[k];
                }
                if(smp.get1d() < sum * inv_max)
                {
                    r.scattered = true;
                    r.position = o * (1.0f - tf) + inct * tf;
                    V3 alb = v3s(0.0f);
                    for(int32_t k = 0; k < ov.count; ++k)
                    {
                        const V3 uvw = ua[k] * (1.0f - tf) + ub[k] * tf;
                        alb = alb + (dens[k] / sum) * volume_albedo(vv, vv.vols[ov.id[k]], uvw);
                    }
                    r.color = alb;
                    return r;
                }
            }
        }
        o = inct + 0.001f * d;
    }
    return r;
}

// VolumePrimitiveMedium::tr, core/volume.cpp:161-191, with tr_in_overlap
// (aggregate.cpp:217-254) on each piece
template<class Smp>
BTRC_HD V3 volume_tr(const VolumeView &vv, V3 a, V3 b, Smp &smp)
{
    float result = 1.0f;
    if(vv.count == 0)
        return v3s(result);
    V3 o = a;
    const V3 d = normalize(b - a);
    for(int32_t it = 0; it < MAX_TRACKING_STEPS; ++it)
    {
        if(dot(b - o, d) <= 0.0f || length_square(b - o) < 0.0001f)
            break;
        V3 inct;
        if(!volume_closest_crossing(vv, o, b, &inct))
            inct = b;
        const VolumeOverlap ov = volume_overlap(vv, 0.5f * (o + inct));
        if(ov.count != 0)
        {
            const float inv_max = 1.0f / volume_majorant(vv, ov);
            V3 ua[MAX_VOLUME_OVERLAP], ub[MAX_VOLUME_OVERLAP];
            for(int32_t k = 0; k < ov.count; ++k)
            {
                ua[k] = volume_uvw(vv.vols[ov.id[k]], o);
                ub[k] = volume_uvw(vv.vols[ov.id[k]], inct);
            }
            const float t_max = length(inct - o);
            float seg = 1.0f, t = 0.0f;
            for(int32_t step = 0; step < MAX_TRACKING_STEPS; ++step)
            {
                const float dt = -btrc_detmath::log(1.0f - smp.get1d()) * inv_max;
                t = t + dt;
                if(t >= t_max)
                    break;
                const float tf = t / t_max;
                float sum = 0.0f;
                for(int32_t k = 0; k < ov.count; ++k)
                    sum = sum + volume_density(vv, vv.vols[ov.id[k]], ua[k] * (1.0f - tf) + ub[k] * tf);
                seg = seg * (1.0f - sum * inv_max);
            }
            result = result * seg;
        }
        o = inct + 0.001f * d;
    }
    return v3s(result);
}

} // namespace btrc
