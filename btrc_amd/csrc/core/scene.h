// Device-resident scene records and the surface/light queries the shading
// stages make on them: get_hitinfo (core/scene.cpp:9-64), the uniform light
// sampler (core/light_sampler.cpp:41-63), MeshLight (light/mesh_light.cpp),
// GradientSky (light/gradient_sky.cpp), IBL (light/ibl.cpp) with its
// importance table (light/env_sampler.cpp), the alias table
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
    LIGHT_IBL = 2,
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
    float up[3];        // GradientSky: normalized (gradient_sky.cpp:15-18); IBL: frame z
    float lower[3];
    float upper[3];
    float frame_x[3];   // IBL: Frame::from_z(up) (ibl.cpp:11-14), axes normalized
    float frame_y[3];   // once more on the device (CFrame(Frame), cframe.cpp:11-22)
    int32_t texture;    // IBL radiance texture
    int32_t lut_w, lut_h; // IBL importance table resolution
    float pad[3];
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
    // the IBL's tile probabilities and their alias table (lut_w * lut_h each)
    const float *env_probs;
    const AliasUnit *env_alias;
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
// channel), and the normal map's RGB (normal_map.cpp:13-21).
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
    auto scalar = [&](int32_t tex, float &value) {
        if(tex >= 0)
            value = tex_sample(sc.tex, tex, sp.tu, sp.tv).x;
    };
    scalar(m.metallic_tex, m.metallic);
    scalar(m.roughness_tex, m.roughness);
    scalar(m.specular_tex, m.specular);
    scalar(m.specular_tint_tex, m.specular_tint);
    scalar(m.anisotropic_tex, m.anisotropic);
    scalar(m.sheen_tex, m.sheen);
    scalar(m.sheen_tint_tex, m.sheen_tint);
    scalar(m.clearcoat_tex, m.clearcoat);
    scalar(m.clearcoat_gloss_tex, m.clearcoat_gloss);
    scalar(m.transmission_tex, m.transmission);
    scalar(m.transmission_roughness_tex, m.transmission_roughness);
    scalar(m.ior_tex, m.ior);
    if(m.normal_tex >= 0)
    {
        const float4 c = tex_sample(sc.tex, m.normal_tex, sp.tu, sp.tv);
        m.normal_rgb[0] = c.x;
        m.normal_rgb[1] = c.y;
        m.normal_rgb[2] = c.z;
    }
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

// local_angle::phi, utils/local_angle.h:63-72: the azimuth in [0, 2 pi)
BTRC_HD float local_phi(V3 w)
{
    float r = 0.0f;
    if(w.x != 0.0f || w.y != 0.0f)
    {
        r = btrc_detmath::atan2(w.y, w.x);
        r = r < 0.0f ? r + 2.0f * PI : r;
    }
    return r;
}

// local_angle::theta, local_angle.h:45-49
BTRC_HD float local_theta(V3 w) { return btrc_detmath::acos(fmin_(fmax_(w.z, -1.0f), 1.0f)); }

// hammersley2d, utils/math/hammersley.h:7-21: (i / n, bit-reversed i)
BTRC_HD void hammersley2d(uint32_t i, uint32_t n, float *x, float *y)
{
    uint32_t b = (i << 16u) | (i >> 16u);
    b = ((b & 0x55555555u) << 1u) | ((b & 0xAAAAAAAAu) >> 1u);
    b = ((b & 0x33333333u) << 2u) | ((b & 0xCCCCCCCCu) >> 2u);
    b = ((b & 0x0F0F0F0Fu) << 4u) | ((b & 0xF0F0F0F0u) >> 4u);
    b = ((b & 0x00FF00FFu) << 8u) | ((b & 0xFF00FF00u) >> 8u);
    *x = (float)i / (float)n;
    *y = (float)b * 2.3283064365386963e-10f;
}

// One tile of EnvirLightSampler's table (env_sampler.cpp:22-57): the mean
// luminance of the texture at n Hammersley points spread over the tile
// widened by a tenth of its size on every side, and the tile's solid angle
// 2 pi du |cos(pi v1) - cos(pi v0)|.
BTRC_HD void envir_tile_stats(const TextureView &tv, int32_t tex, int32_t w, int32_t h, int32_t xi, int32_t yi,
                              int32_t n, float *lum, float *area)
{
    const float x0 = (float)xi / (float)w;
    const float x1 = (float)(xi + 1) / (float)w;
    const float y0 = (float)yi / (float)h;
    const float y1 = (float)(yi + 1) / (float)h;
    float sum = 0.0f;
    for(int32_t i = 0; i < n; ++i)
    {
        float hx, hy;
        hammersley2d((uint32_t)i, (uint32_t)n, &hx, &hy);
        const float sx = 1.2f * hx - 0.1f;
        const float sy = 1.2f * hy - 0.1f;
        const float4 c = tex_sample(tv, tex, lerpf(x0, x1, sx), lerpf(y0, y1, sy));
        sum = sum + luminance(v3(c.x, c.y, c.z));
    }
    *lum = sum / (float)n;
    *area = fabs_(2.0f * PI * (x1 - x0) * (btrc_detmath::cos(PI * y1) - btrc_detmath::cos(PI * y0)));
}

// i32(floor(x)) clamped to [0, n - 1]; NaN gives 0, as the device's
// saturating float-to-int conversion does
BTRC_HD int32_t envir_tile_index(float x, int32_t n)
{
    const float f = floorf(x);
    if(!(f >= 0.0f))
        return 0;
    if(f >= (float)(n - 1))
        return n - 1;
    return (int32_t)f;
}

BTRC_HD V3 ibl_to_local(const LightRec &l, V3 g)
{
    return v3(dot(g, v3(l.frame_x[0], l.frame_x[1], l.frame_x[2])), dot(g, v3(l.frame_y[0], l.frame_y[1], l.frame_y[2])),
              dot(g, v3(l.up[0], l.up[1], l.up[2])));
}

BTRC_HD V3 ibl_to_global(const LightRec &l, V3 w)
{
    return v3(l.frame_x[0], l.frame_x[1], l.frame_x[2]) * w.x + v3(l.frame_y[0], l.frame_y[1], l.frame_y[2]) * w.y +
           v3(l.up[0], l.up[1], l.up[2]) * w.z;
}

// IBL::eval_local, ibl.cpp:71-78: the texture at (phi / 2 pi, theta / pi)
BTRC_HD V3 ibl_eval_local(const SceneView &sc, const LightRec &l, V3 w)
{
    const float phi = local_phi(w);
    const float theta = local_theta(w);
    const float u = saturate((1.0f / (2.0f * PI)) * phi);
    const float v = saturate((1.0f / PI) * theta);
    const float4 c = tex_sample(sc.tex, l.texture, u, v);
    return v3(c.x, c.y, c.z);
}

// EnvirLightSampler::sample, env_sampler.cpp:124-159: a tile from the alias
// table, then a direction uniform in the tile's solid angle
BTRC_HD void ibl_sample_local(const SceneView &sc, const LightRec &l, float s0, float s1, float s2, V3 *dir, float *pdf)
{
    const uint32_t tile = alias_sample(sc.env_alias, (uint32_t)(l.lut_w * l.lut_h), s0);
    const int32_t ty = (int32_t)tile / l.lut_w;
    const int32_t tx = (int32_t)tile % l.lut_w;
    const float tile_pdf = sc.env_probs[tile];
    const float u0 = (float)tx / (float)l.lut_w;
    const float u1 = (float)(tx + 1) / (float)l.lut_w;
    const float v0 = (float)ty / (float)l.lut_h;
    const float v1 = (float)(ty + 1) / (float)l.lut_h;
    float cv0 = btrc_detmath::cos(PI * v0);
    float cv1 = btrc_detmath::cos(PI * v1);
    if(cv0 > cv1)
    {
        const float t = cv0;
        cv0 = cv1;
        cv1 = t;
    }
    const float cos_theta = cv0 + s2 * (cv1 - cv0);
    const float sin_theta = sqrtf(fmax_(0.0f, 1.0f - cos_theta * cos_theta));
    const float phi = 2.0f * PI * lerpf(u0, u1, s1);
    float sp, cp;
    btrc_detmath::sincos(phi, &sp, &cp);
    *dir = v3(sin_theta * cp, sin_theta * sp, cos_theta);
    const float in_tile_pdf = 1.0f / (2.0f * PI * (u1 - u0) * (cv1 - cv0));
    *pdf = tile_pdf * in_tile_pdf;
}

// EnvirLightSampler::pdf, env_sampler.cpp:161-187
BTRC_HD float ibl_pdf_local(const SceneView &sc, const LightRec &l, V3 to_light)
{
    const V3 dir = normalize(to_light);
    const float theta = btrc_detmath::acos(dir.z);
    const float phi = local_phi(dir);
    const float u = phi / (2.0f * PI);
    const float v = theta / PI;
    const int32_t tx = envir_tile_index(u * (float)l.lut_w, l.lut_w);
    const int32_t ty = envir_tile_index(v * (float)l.lut_h, l.lut_h);
    const float tile_pdf = sc.env_probs[ty * l.lut_w + tx];
    const float u0 = (float)tx / (float)l.lut_w;
    const float u1 = (float)(tx + 1) / (float)l.lut_w;
    const float v0 = (float)ty / (float)l.lut_h;
    const float v1 = (float)(ty + 1) / (float)l.lut_h;
    const float c0 = btrc_detmath::cos(PI * v0);
    const float c1 = btrc_detmath::cos(PI * v1);
    const float in_tile_pdf = 1.0f / fabs_(2.0f * PI * (u1 - u0) * (c1 - c0));
    return tile_pdf * in_tile_pdf;
}

// EnvirLight::eval_le: GradientSky::eval_le or IBL::eval_le_inline
// (ibl.cpp:27-31, the direction taken into the light's frame and normalized)
BTRC_HD V3 envir_eval_le(const SceneView &sc, const LightRec &l, V3 to_light)
{
    if(l.type == LIGHT_IBL)
        return ibl_eval_local(sc, l, normalize(ibl_to_local(l, to_light)));
    return sky_eval_le(l, to_light);
}

// EnvirLight::pdf_li: GradientSky's uniform sphere (gradient_sky.cpp:37-40)
// or IBL::pdf_li_inline (ibl.cpp:45-49)
BTRC_HD float envir_pdf_li(const SceneView &sc, const LightRec &l, V3 to_light)
{
    if(l.type == LIGHT_IBL)
        return ibl_pdf_local(sc, l, ibl_to_local(l, to_light));
    return pdf_sample_sphere_uniform();
}

// EnvirLight::sample_li: GradientSky::sample_li (gradient_sky.cpp:27-35) or
// IBL::sample_li_inline (ibl.cpp:33-43; the radiance is looked up at the
// local direction as sampled)
BTRC_HD void envir_sample_li(const SceneView &sc, const LightRec &l, float s0, float s1, float s2, V3 *dir, V3 *li,
                             float *pdf)
{
    if(l.type == LIGHT_IBL)
    {
        V3 local;
        ibl_sample_local(sc, l, s0, s1, s2, &local, pdf);
        *dir = ibl_to_global(l, local);
        *li = ibl_eval_local(sc, l, local);
        return;
    }
    *dir = sample_sphere_uniform(s0, s1);
    *li = sky_eval_le(l, *dir);
    *pdf = pdf_sample_sphere_uniform();
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
        V3 dir, li;
        float pdf;
        envir_sample_li(sc, l, l0, l1, l2, &dir, &li, &pdf);
        r.o = offset_origin ? intersection_offset(ref_pos, ref_nor, dir) : ref_pos;
        r.d = dir;
        r.t1 = MAX_FLOAT;
        r.li = li;
        r.light_pdf = select_pdf * pdf;
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
    const V3 le = envir_eval_le(sc, l, d);
    if(bsdf_pdf < 0.0f)
        return beta_le * le / -bsdf_pdf;
    const float select_light_pdf = light_select_pdf(sc.light_count);
    const float light_pdf = select_light_pdf * envir_pdf_li(sc, l, d);
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
