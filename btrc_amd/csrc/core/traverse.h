// Ray/box and ray/triangle tests of the software traversal that replaces the
// OptiX trace and shadow programs (renderer/wavefront/trace.cpp:38-78,
// shadow.cpp:24-76). Host and device share this header so that a closest hit
// computed anywhere has the same (t, b1, b2) bits.
//
// Two triangle tests are provided:
//  * watertight (Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection",
//    JCGT 2013), single precision only, used by the renderer;
//  * Moeller-Trumbore in the exact operation order of lib/bvh's Triangle
//    (lib/bvh/include/bvh/triangle.hpp:26-115), so that hits can be compared
//    bit for bit with the reference's CPU traversal on the same rays.
// The closest hit is the lexicographic minimum of (t, flattened triangle
// index), which makes it independent of the BVH layout and visit order.
#pragma once

#include "vecmath.h"

namespace btrc {

enum : int32_t
{
    TRI_WATERTIGHT = 0,
    TRI_MOLLER_TRUMBORE = 1,
};

// 4-wide BVH node, 128 bytes: one 128-byte cache line of gfx950's L2.
// Slot j holds the box of child j. child[j] >= 0 names an inner node;
// child[j] < 0 is a leaf whose triangles start at ~child[j] and number
// count[j]. Unused slots carry an empty box (lo = +inf, hi = -inf).
struct alignas(16) BvhNode4
{
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    // as built: inner child index, or ~first triangle of a leaf with
    // count[j] triangles (count 0: empty slot); the device copy holds the
    // traversal codes instead (inner index, leaf code, or all ones)
    int32_t child[4];
    int32_t count[4];
};
static_assert(sizeof(BvhNode4) == 128, "BvhNode4 must be one 128-byte line");

struct RayBoxPre
{
    V3 o;
    V3 inv;      // 1 / d
    V3 inv_pad;  // 1 / d enlarged by 2 ulps (Ize, "Robust BVH Ray Traversal", JCGT 2013)
    bool neg_x, neg_y, neg_z;
};

BTRC_HD float add_ulp_magnitude(float x, uint32_t ulps)
{
    return isfinite_(x) ? u2f(f2u(x) + ulps) : x;
}

BTRC_HD RayBoxPre raybox_pre(V3 o, V3 d)
{
    RayBoxPre p;
    p.o = o;
    p.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    p.inv_pad = v3(add_ulp_magnitude(p.inv.x, 2), add_ulp_magnitude(p.inv.y, 2), add_ulp_magnitude(p.inv.z, 2));
    p.neg_x = (f2u(d.x) >> 31) != 0;
    p.neg_y = (f2u(d.y) >> 31) != 0;
    p.neg_z = (f2u(d.z) >> 31) != 0;
    return p;
}

// robust_max / robust_min (lib/bvh/include/bvh/utilities.hpp:62-71): a NaN on
// the left is dropped in favour of the right operand
BTRC_HD float rmax(float a, float b) { return a > b ? a : b; }
BTRC_HD float rmin(float a, float b) { return a < b ? a : b; }

// The slab test with the near and far planes already chosen by the ray's
// direction signs; same arithmetic as raybox.
BTRC_HD bool raybox_nf(const RayBoxPre &p, float nx, float fx, float ny, float fy, float nz, float fz,
                       float tmin, float tmax, float *tentry)
{
    const float t0x = (nx - p.o.x) * p.inv.x;
    const float t0y = (ny - p.o.y) * p.inv.y;
    const float t0z = (nz - p.o.z) * p.inv.z;
    const float t1x = (fx - p.o.x) * p.inv_pad.x;
    const float t1y = (fy - p.o.y) * p.inv_pad.y;
    const float t1z = (fz - p.o.z) * p.inv_pad.z;
    const float te = fmaxf(t0x, fmaxf(t0y, fmaxf(t0z, tmin)));
    const float tx = fminf(t1x, fminf(t1y, fminf(t1z, tmax)));
    *tentry = te;
    return te <= tx;
}

// Conservative slab test of one box; returns the entry distance through *tentry.
BTRC_HD bool raybox(const RayBoxPre &p, float lox, float hix, float loy, float hiy, float loz, float hiz,
                    float tmin, float tmax, float *tentry)
{
    const float nx = p.neg_x ? hix : lox, fx = p.neg_x ? lox : hix;
    const float ny = p.neg_y ? hiy : loy, fy = p.neg_y ? loy : hiy;
    const float nz = p.neg_z ? hiz : loz, fz = p.neg_z ? loz : hiz;
    const float t0x = (nx - p.o.x) * p.inv.x;
    const float t0y = (ny - p.o.y) * p.inv.y;
    const float t0z = (nz - p.o.z) * p.inv.z;
    const float t1x = (fx - p.o.x) * p.inv_pad.x;
    const float t1y = (fy - p.o.y) * p.inv_pad.y;
    const float t1z = (fz - p.o.z) * p.inv_pad.z;
    // robust max / min: tmin and tmax are never NaN, so every right operand
    // of the chains is a number and IEEE maxNum / minNum (which drop a NaN
    // operand) give exactly rmax / rmin, in fewer instructions
    const float te = fmaxf(t0x, fmaxf(t0y, fmaxf(t0z, tmin)));
    const float tx = fminf(t1x, fminf(t1y, fminf(t1z, tmax)));
    *tentry = te;
    return te <= tx;
}

// Box culling distance for the current closest hit. A box's entry distance
// and a triangle's t come from different arithmetic: the slab entry carries
// ~2 ulps of rounding, the watertight t an error of order 2^-24 times the
// distance from the ray origin to the triangle's vertices (it works on
// v - o), which for a ray starting close to a large triangle is many ulps of
// t (11 ulps were seen for a hit at t = 0.037 on a 2-unit floor). Culling
// therefore keeps every box whose entry lies within best_t (1 + 2^-20) +
// margin, with margin = 2^-18 (scene extent + |o|) from cull_margin: a hit
// no farther than the current one (a tie included) is never culled, so the
// (t, index) minimum does not depend on the visit order.
BTRC_HD float cull_margin(float scene_extent, V3 o)
{
    const float m = fmaxf(fabs_(o.x), fmaxf(fabs_(o.y), fabs_(o.z)));
    return (scene_extent + m) * (1.0f / 262144.0f);
}

BTRC_HD float cull_distance(float best_t, float margin)
{
    return best_t < MAX_FLOAT ? best_t * 1.00000095f + margin : best_t;
}

// ---- watertight test ----

struct WatertightPre
{
    V3 o;
    int32_t kx, ky, kz;
    float Sx, Sy, Sz;
};

BTRC_HD WatertightPre watertight_pre(V3 o, V3 d)
{
    WatertightPre w;
    w.o = o;
    const float ax = fabs_(d.x), ay = fabs_(d.y), az = fabs_(d.z);
    int32_t kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int32_t kx = kz + 1 == 3 ? 0 : kz + 1;
    int32_t ky = kx + 1 == 3 ? 0 : kx + 1;
    if(comp(d, kz) < 0.0f)
    {
        const int32_t t = kx;
        kx = ky;
        ky = t;
    }
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    const float dz = comp(d, kz);
    w.Sx = comp(d, kx) / dz;
    w.Sy = comp(d, ky) / dz;
    w.Sz = 1.0f / dz;
    return w;
}

// Returns true on a hit with t in [tmin, tmax]; (b1, b2) are the barycentric
// weights of v1 and v2.
BTRC_HD bool watertight_hit(const WatertightPre &w, V3 v0, V3 v1, V3 v2, float tmin, float tmax,
                            float *t_out, float *b1_out, float *b2_out)
{
    const V3 A = v0 - w.o;
    const V3 B = v1 - w.o;
    const V3 C = v2 - w.o;
    const float Akz = comp(A, w.kz), Bkz = comp(B, w.kz), Ckz = comp(C, w.kz);
    const float Ax = comp(A, w.kx) - w.Sx * Akz;
    const float Ay = comp(A, w.ky) - w.Sy * Akz;
    const float Bx = comp(B, w.kx) - w.Sx * Bkz;
    const float By = comp(B, w.ky) - w.Sy * Bkz;
    const float Cx = comp(C, w.kx) - w.Sx * Ckz;
    const float Cy = comp(C, w.ky) - w.Sy * Ckz;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    const float det = U + V + W;
    if(det == 0.0f)
        return false;
    const float Az = w.Sz * Akz;
    const float Bz = w.Sz * Bkz;
    const float Cz = w.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if(!(t >= tmin && t <= tmax))
        return false;
    *t_out = t;
    *b1_out = V * rcp;
    *b2_out = W * rcp;
    return true;
}

// ---- lib/bvh Moeller-Trumbore (triangle.hpp:95-115) ----

struct MTPre
{
    V3 o, d;
};

BTRC_HD MTPre mt_pre(V3 o, V3 d)
{
    MTPre m;
    m.o = o;
    m.d = d;
    return m;
}

// Returns lib/bvh's (t, u, v) for the triangle (p0, p1, p2).
BTRC_HD bool mt_hit(const MTPre &m, V3 v0, V3 v1, V3 v2, float tmin, float tmax,
                    float *t_out, float *u_out, float *v_out)
{
    const V3 e1 = v0 - v1;
    const V3 e2 = v2 - v0;
    const V3 n = cross(e1, e2);
    const V3 c = v0 - m.o;
    const V3 r = cross(m.d, c);
    const float inv_det = 1.0f / dot(n, m.d);
    const float u = dot(r, e2) * inv_det;
    const float v = dot(r, e1) * inv_det;
    const float w = 1.0f - u - v;
    if(u >= 0.0f && v >= 0.0f && w >= 0.0f)
    {
        const float t = dot(n, c) * inv_det;
        if(t >= tmin && t <= tmax)
        {
            *t_out = t;
            *u_out = u;
            *v_out = v;
            return true;
        }
    }
    return false;
}

} // namespace btrc
