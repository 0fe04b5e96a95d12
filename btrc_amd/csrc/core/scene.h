// Device-resident scene records and the surface/light queries the shading
// stages make on them: get_hitinfo (core/scene.cpp:9-64), the uniform light
// sampler (core/light_sampler.cpp:41-63), MeshLight (light/mesh_light.cpp),
// GradientSky (light/gradient_sky.cpp), the alias table
// (utils/cmath/calias.cpp:10-23) and TriangleMesh::sample_inline
// (geometry/triangle_mesh.cpp:188-243).
#pragma once

#include "material.h"
#include "medium.h"
#include "texture.h"
#include "volume.h"

#if !defined(__HIPCC__)
struct float4 { float x, y, z, w; };
#endif

namespace btrc {

enum : int32_t
{
    LIGHT_MESH = 0,
    LIGHT_GRADIENT_SKY = 1,
};

// InstanceInfo, core/scene.h:21-30, with the transform as two 3x4 blocks
struct InstanceRec
{
    int32_t geometry_id;
    int32_t material_id;
    int32_t light_id;     // index into the light list, -1 if not emissive
    uint32_t inner_medium;
    uint32_t outer_medium;
    uint32_t flag;
    uint32_t prim_base;   // index of this instance's first triangle in the flattened list
    uint32_t pad;
    Xform xf;
};

// Per-geometry ranges into the shared per-triangle arrays
struct GeometryRec
{
    uint32_t prim_offset;
    uint32_t prim_count;
    float total_area;
    float pad;
};

struct LightRec
{
    int32_t type;
    int32_t instance;
    int32_t geometry;
    float scale;        // MeshLight::scale_ (mesh_light.cpp:17-22)
    float intensity[3];
    float up[3];        // GradientSky, normalized (gradient_sky.cpp:15-18)
    float lower[3];
    float upper[3];
    float pad[2];
};

struct AliasUnit
{
    float accept_prob;
    uint32_t another_idx;
};

// Everything the shading kernels read, as device pointers.
struct SceneView
{
    const InstanceRec *instances;
    const GeometryRec *geometries;
    const MaterialRec *materials;
    const LightRec *lights;
    // TriangleMesh::commit per-triangle frame records (triangle_mesh.cpp:40-64),
    // indexed by geometry prim_offset + prim
    const float4 *gx_u_a;
    const float4 *gy_u_ba;
    const float4 *gz_u_ca;
    const float4 *sn_v_a;
    const float4 *sn_v_ba;
    const float4 *sn_v_ca;
    // local-space (a, b - a, c - a) per triangle for light sampling
    const float4 *light_pos;
    const AliasUnit *alias;
    int32_t light_count;
    int32_t envir_light_index; // -1 without an environment light
    int32_t instance_count;
    // media, by descending priority (core/scene.cpp:144-161): ids 0..n-1;
    // id n is the volume-primitive medium (no volume primitives here, so
    // vacuum) and id n + 1 stands for "no medium"
    int32_t medium_count;
    const MediumRec *media;
    const float *voxels;
    TextureView tex;
    VolumeView vol;           // the volume-primitive medium (id medium_count)
    float world_diagonal;     // 1.2 |bbox of the scene and the cameras| (wavefront.cpp:138-139)
};

struct SurfacePoint
{
    V3 position;
    Frame frame;     // geometric frame, world space
    V3 interp_z;     // interpolated shading normal, world space
    float u, v;      // barycentrics of the hit
    float tu, tv;    // texture coordinate
};

BTRC_HD V3 f4xyz(float4 a) { return v3(a.x, a.y, a.z); }

// get_hitinfo, core/scene.cpp:9-64
BTRC_HD SurfacePoint get_hitinfo(
    const SceneView &sc, V3 o, V3 d, const InstanceRec &inst, uint32_t prim, float t, float u, float v)
{
    const GeometryRec &g = sc.geometries[inst.geometry_id];
    const uint32_t i = g.prim_offset + prim;
    SurfacePoint sp;
    sp.position = o + t * d;
    const Frame gf = make_frame(f4xyz(sc.gx_u_a[i]), f4xyz(sc.gy_u_ba[i]), f4xyz(sc.gz_u_ca[i]));
    sp.frame.x = normalize(xf_vector(inst.xf, gf.x));
    sp.frame.y = normalize(xf_vector(inst.xf, gf.y));
    sp.frame.z = normalize(xf_normal(inst.xf, gf.z));
    const V3 n = f4xyz(sc.sn_v_a[i]) + f4xyz(sc.sn_v_ba[i]) * u + f4xyz(sc.sn_v_ca[i]) * v;
    sp.interp_z = normalize(xf_normal(inst.xf, n));
    sp.u = u;
    sp.v = v;
    sp.tu = sc.gx_u_a[i].w + sc.gy_u_ba[i].w * u + sc.gz_u_ca[i].w * v;
    sp.tv = sc.sn_v_a[i].w + sc.sn_v_ba[i].w * u + sc.sn_v_ca[i].w * v;
    return sp;
}

// The material at a surface point: textured parameters sampled at its
// texture coordinate (DisneyMaterial::create_shader, disney.cpp:813-842:
// Texture2D::sample_spectrum / sample_float; sample_float reads the first
// channel).
BTRC_HD MaterialRec material_at(const SceneView &sc, int32_t id, const SurfacePoint &sp)
{
    MaterialRec m = sc.materials[id];
    if(m.color_tex >= 0)
    {
        const float4 c = tex_sample(sc.tex, m.color_tex, sp.tu, sp.tv);
        m.color[0] = c.x;
        m.color[1] = c.y;
        m.color[2] = c.z;
    }
    if(m.metallic_tex >= 0)
        m.metallic = tex_sample(sc.tex, m.metallic_tex, sp.tu, sp.tv).x;
    if(m.roughness_tex >= 0)
        m.roughness = tex_sample(sc.tex, m.roughness_tex, sp.tu, sp.tv).x;
    return m;
}

// geometric normal only, as the medium stage reads it (wavefront/medium.cpp:230-234)
BTRC_HD V3 hit_geometric_normal(const SceneView &sc, const InstanceRec &inst, uint32_t prim)
{
    const GeometryRec &g = sc.geometries[inst.geometry_id];
    return xf_normal(inst.xf, f4xyz(sc.gz_u_ca[g.prim_offset + prim]));
}

// CAliasTable::sample, utils/cmath/calias.cpp:15-21
BTRC_HD uint32_t alias_sample(const AliasUnit *table, uint32_t n, float u)
{
    const float nu = (float)n * u;
    uint32_t i = (uint32_t)nu;
    i = i < n - 1 ? i : n - 1;
    const float s = nu - (float)i;
    return s <= table[i].accept_prob ? i : table[i].another_idx;
}

// UniformLightSampler::sample, core/light_sampler.cpp:41-58
BTRC_HD int32_t light_select(int32_t n, float u, float *pdf)
{
    if(n == 0)
    {
        *pdf = 0.0f;
        return -1;
    }
    int32_t idx = (int32_t)(u * (float)n);
    idx = idx < n - 1 ? idx : n - 1;
    *pdf = 1.0f / (float)n;
    return idx;
}

BTRC_HD float light_select_pdf(int32_t n) { return 1.0f / (float)n; }

// GradientSky::eval_le, gradient_sky.cpp:20-25
BTRC_HD V3 sky_eval_le(const LightRec &l, V3 to_light)
{
    const float cos_theta = dot(v3(l.up[0], l.up[1], l.up[2]), normalize(to_light));
    const float s = saturate(0.5f * (cos_theta + 1.0f));
    return v3(l.lower[0], l.lower[1], l.lower[2]) * (1.0f - s) + v3(l.upper[0], l.upper[1], l.upper[2]) * s;
}

// MeshLight::eval_le, mesh_light.cpp:24-36
BTRC_HD V3 mesh_light_eval_le(const LightRec &l, V3 frame_z, V3 wr)
{
    if(dot(frame_z, wr) >= 0.0f)
        return v3(l.intensity[0], l.intensity[1], l.intensity[2]);
    return v3s(0.0f);
}

// MeshLight::pdf_li, mesh_light.cpp:60-76, with TriangleMesh::pdf = 1/area
BTRC_HD float mesh_light_pdf_li(const SceneView &sc, const LightRec &l, V3 ref_pos, V3 pos, V3 nor)
{
    if(dot(nor, ref_pos - pos) <= 0.0f)
        return 0.0f;
    const float pdf_area = 1.0f / sc.geometries[l.geometry].total_area;
    const V3 spt_to_ref = ref_pos - pos;
    const float dist2 = length_square(spt_to_ref);
    const float dist3 = sqrtf(dist2) * dist2;
    return 1.0f / (l.scale * l.scale) * pdf_area * dist3 / fabs_(dot(nor, spt_to_ref));
}

struct LightSample
{
    V3 position;
    V3 normal;
    float pdf;
    V3 radiance;
};

// MeshLight::sample_li, mesh_light.cpp:38-58 over TriangleMesh::sample_inline
BTRC_HD LightSample mesh_light_sample_li(
    const SceneView &sc, const LightRec &l, V3 ref_pos, float s0, float s1, float s2)
{
    const GeometryRec &g = sc.geometries[l.geometry];
    const uint32_t prim = alias_sample(sc.alias + g.prim_offset, g.prim_count, s0);
    float b1, b2;
    sample_triangle_uniform(s1, s2, &b1, &b2);
    const uint32_t i = g.prim_offset + prim;
    const V3 a = f4xyz(sc.light_pos[3 * i + 0]);
    const V3 ba = f4xyz(sc.light_pos[3 * i + 1]);
    const V3 ca = f4xyz(sc.light_pos[3 * i + 2]);
    const V3 pos = a + ba * b1 + ca * b2;
    const V3 ez = normalize(f4xyz(sc.gz_u_ca[i]));
    const float pdf_area = 1.0f / g.total_area;

    const InstanceRec &inst = sc.instances[l.instance];
    LightSample r;
    r.position = xf_point(inst.xf, pos);
    r.normal = normalize(xf_normal(inst.xf, ez));
    const V3 pos_to_ref = ref_pos - r.position;
    const float dist2 = length_square(pos_to_ref);
    r.pdf = 1.0f / (l.scale * l.scale) * pdf_area * dist2 / fabs_(dot(r.normal, normalize(pos_to_ref)));
    r.radiance = dot(pos_to_ref, r.normal) > 0.0f ? v3(l.intensity[0], l.intensity[1], l.intensity[2]) : v3s(0.0f);
    return r;
}

// A light sample seen from a surface point, sample_light_li_impl
// (renderer/wavefront/helper.cpp:10-65) with dir_to_o of
// sample_surface_light_li (helper.cpp:127-142).
struct ShadowSample
{
    bool success;
    V3 o, d;
    float t1;
    float light_pdf;
    V3 li;
};

// offset_origin: surface points move the origin off the surface
// (sample_surface_light_li, helper.cpp:127-142); medium points keep it
// (sample_medium_light_li, helper.cpp:115-125)
BTRC_HD ShadowSample sample_light_li(
    const SceneView &sc, V3 ref_pos, V3 ref_nor, bool offset_origin, float l0, float l1, float l2, float lsel)
{
    ShadowSample r;
    float select_pdf;
    const int32_t idx = light_select(sc.light_count, lsel, &select_pdf);
    r.success = idx >= 0;
    if(!r.success)
    {
        r.o = r.d = r.li = v3s(0.0f);
        r.t1 = 0.0f;
        r.light_pdf = 0.0f;
        return r;
    }
    const LightRec &l = sc.lights[idx];
    if(l.type == LIGHT_MESH)
    {
        const LightSample s = mesh_light_sample_li(sc, l, ref_pos, l0, l1, l2);
        const V3 diff = s.position - ref_pos;
        const V3 shadow_dst = intersection_offset(s.position, s.normal, -diff);
        r.o = offset_origin ? intersection_offset(ref_pos, ref_nor, diff) : ref_pos;
        r.d = shadow_dst - r.o;
        r.t1 = 1.0f;
        r.li = s.radiance;
        r.light_pdf = select_pdf * s.pdf;
    }
    else
    {
        // GradientSky::sample_li, gradient_sky.cpp:27-35
        const V3 dir = sample_sphere_uniform(l0, l1);
        r.o = offset_origin ? intersection_offset(ref_pos, ref_nor, dir) : ref_pos;
        r.d = dir;
        r.t1 = MAX_FLOAT;
        r.li = sky_eval_le(l, dir);
        r.light_pdf = select_pdf * pdf_sample_sphere_uniform();
    }
    return r;
}

BTRC_HD ShadowSample sample_surface_light_li(
    const SceneView &sc, V3 ref_pos, V3 ref_nor, float l0, float l1, float l2, float lsel)
{
    return sample_light_li(sc, ref_pos, ref_nor, true, l0, l1, l2, lsel);
}

// eval_miss_le, helper.cpp:144-172
BTRC_HD V3 eval_miss_le(const SceneView &sc, V3 d, V3 beta_le, float bsdf_pdf)
{
    if(sc.envir_light_index < 0)
        return v3s(0.0f);
    const LightRec &l = sc.lights[sc.envir_light_index];
    const V3 le = sky_eval_le(l, d);
    if(bsdf_pdf < 0.0f)
        return beta_le * le / -bsdf_pdf;
    const float select_light_pdf = light_select_pdf(sc.light_count);
    const float light_pdf = select_light_pdf * pdf_sample_sphere_uniform();
    return beta_le * le / (bsdf_pdf + light_pdf);
}

// handle_intersected_light, helper.cpp:174-209
BTRC_HD V3 handle_intersected_light(
    const SceneView &sc, V3 o, V3 d, const SurfacePoint &sp, V3 beta_le, float bsdf_pdf, int32_t light_index)
{
    if(light_index < 0)
        return v3s(0.0f);
    const float select_light_pdf = light_select_pdf(sc.light_count);
    const LightRec &l = sc.lights[light_index];
    if(l.type != LIGHT_MESH)
        return v3s(0.0f);
    const V3 le = mesh_light_eval_le(l, sp.frame.z, -d);
    if(bsdf_pdf < 0.0f)
        return beta_le * le / -bsdf_pdf;
    const float light_dir_pdf = mesh_light_pdf_li(sc, l, o, sp.position, sp.frame.z);
    const float light_pdf = select_light_pdf * light_dir_pdf;
    return beta_le * le / (bsdf_pdf + light_pdf);
}

} // namespace btrc
