// Software BVH traversal for gfx950 (CDNA4 has no ray-tracing instructions).
//
// One lane owns one ray. The traversal stack lives in LDS laid out
// [entry][lane], so pushes and pops of a wavefront touch 64 consecutive
// dwords and never conflict on banks. Leaves met in a node are intersected
// at once (eager leaves, as lib/bvh's SingleRayTraverser does); inner
// children that the ray enters are visited nearest first. Boxes are tested
// with Ize's conservative slab test so no triangle is culled by rounding,
// which makes the closest hit the BVH-independent minimum of
// (t, flattened triangle index).
#pragma once

#include <hip/hip_runtime.h>

#include "../core/traverse.h"

namespace btrc {

struct TraceScene
{
    const BvhNode4 *nodes;
    // 3 float4 per triangle in leaf order:
    //   (v0.xyz, flattened index), (v1.xyz, instance), (v2.xyz, prim in geometry)
    const float4 *tris;
};

struct HitRecord
{
    float t, b1, b2;
    uint32_t gid;   // flattened triangle index, 0xffffffff on a miss
    uint32_t inst;
    uint32_t prim;
};

// Walk the BVH for one ray. ANY stops at the first hit in [tmin, tmax].
template<int TRI, bool ANY>
__device__ __forceinline__ HitRecord trace_ray(
    const TraceScene &sc, int32_t *stack, int32_t stride, V3 o, V3 d, float tmin, float tmax)
{
    HitRecord h;
    h.t = tmax;
    h.b1 = h.b2 = 0.0f;
    h.gid = 0xffffffffu;
    h.inst = h.prim = 0xffffffffu;

    const RayBoxPre rb = raybox_pre(o, d);
    WatertightPre wt;
    MTPre mt;
    if(TRI == TRI_WATERTIGHT)
        wt = watertight_pre(o, d);
    else
        mt = mt_pre(o, d);

    int32_t sp = 0;
    int32_t node = 0;
    for(;;)
    {
        const float4 *np = reinterpret_cast<const float4 *>(sc.nodes + node);
        const float4 lox = np[0], hix = np[1], loy = np[2], hiy = np[3], loz = np[4], hiz = np[5];
        const int4 child = reinterpret_cast<const int4 *>(np)[6];
        const int4 count = reinterpret_cast<const int4 *>(np)[7];
        const float blox[4] = { lox.x, lox.y, lox.z, lox.w };
        const float bhix[4] = { hix.x, hix.y, hix.z, hix.w };
        const float bloy[4] = { loy.x, loy.y, loy.z, loy.w };
        const float bhiy[4] = { hiy.x, hiy.y, hiy.z, hiy.w };
        const float bloz[4] = { loz.x, loz.y, loz.z, loz.w };
        const float bhiz[4] = { hiz.x, hiz.y, hiz.z, hiz.w };
        const int32_t ch[4] = { child.x, child.y, child.z, child.w };
        const int32_t cn[4] = { count.x, count.y, count.z, count.w };

        float te[4];
        bool hit[4];
        {
            const float cull = cull_distance(h.t);
#pragma unroll
            for(int j = 0; j < 4; ++j)
                hit[j] = raybox(rb, blox[j], bhix[j], bloy[j], bhiy[j], bloz[j], bhiz[j], tmin, cull, &te[j]);
        }

        // leaves of this node, at once
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            if(hit[j] && ch[j] < 0)
            {
                const int32_t first = ~ch[j];
                const int32_t n = cn[j];
                for(int32_t k = 0; k < n; ++k)
                {
                    const float4 a = sc.tris[3 * (first + k) + 0];
                    const float4 b = sc.tris[3 * (first + k) + 1];
                    const float4 c = sc.tris[3 * (first + k) + 2];
                    const uint32_t gid = __float_as_uint(a.w);
                    float t, u, v;
                    bool ok;
                    if(TRI == TRI_WATERTIGHT)
                        ok = watertight_hit(wt, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), tmin, h.t, &t, &u, &v);
                    else
                        ok = mt_hit(mt, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), tmin, h.t, &t, &u, &v);
                    if(ok && (t < h.t || gid < h.gid))
                    {
                        h.t = t;
                        h.b1 = u;
                        h.b2 = v;
                        h.gid = gid;
                        h.inst = __float_as_uint(b.w);
                        h.prim = __float_as_uint(c.w);
                        if(ANY)
                            return h;
                    }
                }
            }
        }

        // inner children, nearest first
        const float cull = cull_distance(h.t);
        float k[4];
        int32_t c[4];
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const bool in = hit[j] && ch[j] >= 0 && te[j] <= cull;
            k[j] = in ? te[j] : INFINITY;
            c[j] = in ? ch[j] : -1;
        }
#define BTRC_CSWAP(a, b)                                  \
        if(k[b] < k[a])                                   \
        {                                                 \
            const float tk = k[a]; k[a] = k[b]; k[b] = tk; \
            const int32_t tc = c[a]; c[a] = c[b]; c[b] = tc; \
        }
        BTRC_CSWAP(0, 1)
        BTRC_CSWAP(2, 3)
        BTRC_CSWAP(0, 2)
        BTRC_CSWAP(1, 3)
        BTRC_CSWAP(1, 2)
#undef BTRC_CSWAP
        if(c[0] >= 0)
        {
            if(c[3] >= 0) stack[(sp++) * stride] = c[3];
            if(c[2] >= 0) stack[(sp++) * stride] = c[2];
            if(c[1] >= 0) stack[(sp++) * stride] = c[1];
            node = c[0];
            continue;
        }
        if(sp == 0)
            break;
        node = stack[(--sp) * stride];
    }
    return h;
}

} // namespace btrc
