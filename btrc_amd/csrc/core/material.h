// Surface scattering: the Lambert and Disney BSDF components and the
// single-component BSDF aggregate with its shading-frame fixes.
//
// Restates builtin/material/{diffuse.cpp, component/diffuse.cpp, disney.cpp}
// and material/utils/{aggregate.cpp, shader_frame.cpp, microfacet.cpp,
// fresnel.cpp}. Both materials register exactly one component of sample
// weight 1 (diffuse.cpp:30-31, disney.cpp:840-841), so the aggregate's
// component selection reduces to that component with the sample unchanged.
//
// Where the reference evaluates sin(phi) and cos(phi) of phi = atan2(y, x)
// (utils/local_angle.h:43-76), this code uses the identities
// cos(phi) = x / r, sin(phi) = y / r with r = sqrt(x^2 + y^2) (and phi = 0
// when x = y = 0), which avoids two transcendentals per angle.
#pragma once

#include "vecmath.h"

namespace btrc {

enum : int32_t
{
    MATERIAL_DIFFUSE = 0,
    MATERIAL_DISNEY = 1,
    MATERIAL_INVISIBLE = 2, // InvisibleSurface, builtin/material/invisible.cpp:6-57
};

// Per-material constants. Textures are constant per material in this path
// (Constant2D); every field is read by the shading stage.
struct MaterialRec
{
    int32_t type;
    int32_t shadow_terminator;
    float color[3];
    float metallic;
    float roughness;
    float specular;
    float specular_tint;
    float anisotropic;
    float sheen;
    float sheen_tint;
    float clearcoat;
    float clearcoat_gloss;
    float transmission;
    float transmission_roughness;
    float ior;
    // textures in place of the parameters above, -1: the constant
    int32_t color_tex, metallic_tex, roughness_tex, specular_tex, specular_tint_tex, anisotropic_tex;
    int32_t sheen_tex, sheen_tint_tex, clearcoat_tex, clearcoat_gloss_tex, transmission_tex;
    int32_t transmission_roughness_tex, ior_tex;
    // normal map texture (-1: none) and, once sampled at a surface point by
    // material_at, its RGB
    int32_t normal_tex;
    float normal_rgb[3];
};

struct BSDFSample
{
    V3 bsdf;
    V3 dir;
    float pdf;
    bool is_delta;
};

BTRC_HD void bsdf_sample_clear(BSDFSample &s)
{
    s.bsdf = v3s(0.0f);
    s.dir = v3s(0.0f);
    s.pdf = 0.0f;
    s.is_delta = false;
}

// ---- local angles, utils/local_angle.h ----

BTRC_HD float cos2sin(float c) { return sqrtf(fmax_(0.0f, 1.0f - c * c)); }

BTRC_HD float tan_theta(V3 w)
{
    const float t = 1.0f - w.z * w.z;
    float r = 0.0f;
    if(t > 0.0f)
        r = sqrtf(t) / w.z;
    return r;
}

// sin(phi(w)) and cos(phi(w))
BTRC_HD void sincos_phi(V3 w, float *s, float *c)
{
    if(w.x != 0.0f || w.y != 0.0f)
    {
        const float r = sqrtf(w.x * w.x + w.y * w.y);
        *s = w.y / r;
        *c = w.x / r;
    }
    else
    {
        *s = 0.0f;
        *c = 1.0f;
    }
}

// ---- fresnel.cpp ----

BTRC_HD float dielectric_fresnel(float eta_i, float eta_o, float cos_theta_i)
{
    if(cos_theta_i < 0.0f)
    {
        const float teta = eta_i;
        eta_i = eta_o;
        eta_o = teta;
        cos_theta_i = -cos_theta_i;
    }
    const float sin_theta_i = sqrtf(fmax_(0.0f, 1.0f - cos_theta_i * cos_theta_i));
    const float sin_theta_t = eta_o / eta_i * sin_theta_i;
    if(sin_theta_t >= 1.0f)
        return 1.0f;
    const float cos_theta_t = sqrtf(fmax_(0.0f, 1.0f - sin_theta_t * sin_theta_t));
    const float para = (eta_i * cos_theta_i - eta_o * cos_theta_t) / (eta_i * cos_theta_i + eta_o * cos_theta_t);
    const float perp = (eta_o * cos_theta_i - eta_i * cos_theta_t) / (eta_o * cos_theta_i + eta_i * cos_theta_t);
    return 0.5f * (para * para + perp * perp);
}

BTRC_HD float schlick_approx(float r0, float cos_theta_i)
{
    const float t = 1.0f - cos_theta_i;
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    return r0 + (1.0f - r0) * t5;
}

BTRC_HD V3 schlick_approx3(V3 r0, float cos_theta_i)
{
    const float t = 1.0f - cos_theta_i;
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    return r0 + (v3s(1.0f) - r0) * t5;
}

// ---- microfacet.cpp ----

BTRC_HD float mf_compute_a(float sin_phi_h, float cos_phi_h, float ax, float ay)
{
    return sqr(cos_phi_h / ax) + sqr(sin_phi_h / ay);
}

BTRC_HD float mf_smith_gtr2(float tan_theta_, float alpha)
{
    float r = 1.0f;
    if(tan_theta_ != 0.0f)
    {
        const float root = alpha * tan_theta_;
        r = 2.0f / (1.0f + sqrtf(1.0f + root * root));
    }
    return r;
}

BTRC_HD float mf_anisotropic_gtr2(float sin_phi_h, float cos_phi_h, float sin_theta_h, float cos_theta_h, float ax, float ay)
{
    const float A = mf_compute_a(sin_phi_h, cos_phi_h, ax, ay);
    const float RD = sqr(sin_theta_h) * A + sqr(cos_theta_h);
    return 1.0f / (PI * ax * ay * sqr(RD));
}

BTRC_HD float mf_smith_anisotropic_gtr2(float cos_phi, float sin_phi, float ax, float ay, float tan_theta_)
{
    const float t = sqr(ax * cos_phi) + sqr(ay * sin_phi);
    const float sqr_val = 1.0f + t * sqr(tan_theta_);
    const float lambda = -0.5f + 0.5f * sqrtf(sqr_val);
    return 1.0f / (1.0f + lambda);
}

BTRC_HD V3 mf_sample_anisotropic_gtr2(float ax, float ay, float s0, float s1)
{
    float sn, cs;
    btrc_detmath::sincos(2.0f * PI * s0, &sn, &cs);
    float sin_phi_h = ay * sn;
    float cos_phi_h = ax * cs;
    const float nor = 1.0f / sqrtf(sqr(sin_phi_h) + sqr(cos_phi_h));
    sin_phi_h = sin_phi_h * nor;
    cos_phi_h = cos_phi_h * nor;
    const float A = mf_compute_a(sin_phi_h, cos_phi_h, ax, ay);
    const float cos_theta_h = sqrtf(A * (1.0f - s1) / ((1.0f - A) * s1 + A));
    const float sin_theta_h = cos2sin(cos_theta_h);
    return normalize(v3(sin_theta_h * cos_phi_h, sin_theta_h * sin_phi_h, cos_theta_h));
}

BTRC_HD float mf_gtr1(float sin_theta_h, float cos_theta_h, float alpha)
{
    const float U = sqr(alpha) - 1.0f;
    const float LD = 2.0f * PI * btrc_detmath::log(alpha);
    const float RD = sqr(alpha * cos_theta_h) + sqr(sin_theta_h);
    return U / (LD * RD);
}

BTRC_HD V3 mf_sample_gtr1(float alpha, float s0, float s1)
{
    const float phi = 2.0f * PI * s0;
    const float cos_theta = sqrtf((btrc_detmath::pow(alpha, 2.0f - 2.0f * s1) - 1.0f) / (sqr(alpha) - 1.0f));
    const float sin_theta = cos2sin(cos_theta);
    float sn, cs;
    btrc_detmath::sincos(phi, &sn, &cs);
    return normalize(v3(sin_theta * cs, sin_theta * sn, cos_theta));
}

// visible-normal sampling, microfacet.cpp:87-110
BTRC_HD V3 mf_sample_anisotropic_gtr2_vnor(V3 ve, float ax, float ay, float s0, float s1)
{
    const V3 vh = normalize(v3(ax * ve.x, ay * ve.y, ve.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = v3(1.0f, 0.0f, 0.0f);
    if(lensq > 3e-4f)
        t1 = v3(-vh.y, vh.x, 0.0f) / sqrtf(lensq);
    const V3 t2 = cross(vh, t1);
    const float r = sqrtf(s0);
    float sn, cs;
    btrc_detmath::sincos(2.0f * PI * s1, &sn, &cs);
    const float t_1 = r * cs;
    const float _t_2 = r * sn;
    const float s = 0.5f * (1.0f + vh.z);
    const float t_2 = (1.0f - s) * sqrtf(1.0f - t_1 * t_1) + s * _t_2;
    const V3 nh = t_1 * t1 + t_2 * t2 + sqrtf(fmax_(0.0f, 1.0f - t_1 * t_1 - t_2 * t_2)) * vh;
    return normalize(v3(ax * nh.x, ay * nh.y, fmax_(0.0f, nh.z)));
}

// ---- Lambert component, component/diffuse.cpp ----

BTRC_HD void lambert_sample(V3 albedo, V3 lwo, float s0, float s1, BSDFSample &out)
{
    bsdf_sample_clear(out);
    if(lwo.z <= 0.0f)
        return;
    const V3 lwi = sample_hemisphere_zweighted(s0, s1);
    out.bsdf = albedo / PI;
    out.dir = lwi;
    out.pdf = pdf_sample_hemisphere_zweighted(lwi);
}

BTRC_HD V3 lambert_eval(V3 albedo, V3 lwi, V3 lwo)
{
    if(lwi.z <= 0.0f || lwo.z <= 0.0f)
        return v3s(0.0f);
    return albedo / PI;
}

BTRC_HD float lambert_pdf(V3 lwi, V3 lwo)
{
    if(lwi.z <= 0.0f || lwo.z <= 0.0f)
        return 0.0f;
    return pdf_sample_hemisphere_zweighted(lwi);
}

// ---- Disney component, disney.cpp:9-736 ----

struct Disney
{
    V3 C, C_tint;
    float metallic, roughness, specular, specular_tint, anisotropic;
    float sheen, sheen_tint, clearcoat, clearcoat_roughness, transmission;
    float ior, trans_ax, trans_ay, ax, ay;
    float diffuse_weight, specular_weight, clearcoat_weight, transmission_weight;
};

BTRC_HD float disney_one_minus_5(float x)
{
    const float t = 1.0f - x;
    const float t2 = t * t;
    return t2 * t2 * t;
}

// DisneyBSDF constructor, disney.cpp:434-475
BTRC_HD Disney disney_make(const MaterialRec &m)
{
    Disney d;
    d.C = v3(m.color[0], m.color[1], m.color[2]);
    const float lum = luminance(d.C);
    d.C_tint = lum > 0.0f ? d.C / lum : v3s(1.0f);
    d.metallic = m.metallic;
    d.roughness = m.roughness;
    d.specular = m.specular;
    d.specular_tint = m.specular_tint;
    d.anisotropic = m.anisotropic;
    d.sheen = m.sheen;
    d.sheen_tint = m.sheen_tint;
    d.transmission = m.transmission;
    d.ior = fmax_(1.01f, m.ior);

    float aspect = 1.0f;
    if(m.anisotropic > 0.0f)
        aspect = sqrtf(1.0f - 0.9f * m.anisotropic);
    d.ax = fmax_(0.001f, sqr(m.roughness) / aspect);
    d.ay = fmax_(0.001f, sqr(m.roughness) * aspect);
    d.trans_ax = fmax_(0.001f, sqr(m.transmission_roughness) / aspect);
    d.trans_ay = fmax_(0.001f, sqr(m.transmission_roughness) * aspect);

    d.clearcoat = m.clearcoat;
    d.clearcoat_roughness = fmax_(0.0001f, sqr(lerpf(0.1f, 0.0f, m.clearcoat_gloss)));

    const float A = clampf(lum * (1.0f - m.metallic), 0.3f, 0.7f);
    const float B = 1.0f - A;
    d.diffuse_weight = A * (1.0f - m.transmission);
    d.transmission_weight = A * m.transmission;
    d.specular_weight = B * 2.0f / (2.0f + m.clearcoat);
    d.clearcoat_weight = B * m.clearcoat / (2.0f + m.clearcoat);
    return d;
}

BTRC_HD bool disney_refract(V3 w, V3 n, float eta, V3 &out)
{
    const float cos_theta_i = dot(w, n);
    const float sin2_theta_t = sqr(eta) * (1.0f - sqr(cos_theta_i));
    const float cos2_theta_t = 1.0f - sin2_theta_t;
    if(cos2_theta_t <= 0.0f)
        return false;
    const float cos_theta_t = sqrtf(cos2_theta_t);
    out = normalize((eta * cos_theta_i - cos_theta_t) * n - eta * w);
    return true;
}

BTRC_HD V3 disney_eval_diffuse(const Disney &d, float cos_theta_i, float cos_theta_o, float cos_theta_d)
{
    const V3 lambert = d.C / PI;
    const float fl = disney_one_minus_5(cos_theta_i);
    const float fv = disney_one_minus_5(cos_theta_o);
    const float rr = 2.0f * d.roughness * cos_theta_d * cos_theta_d;
    const V3 retro_refl = d.C / PI * rr * (fl + fv + fl * fv * (rr - 1.0f));
    return lambert * (1.0f - 0.5f * fl) * (1.0f - 0.5f * fv) + retro_refl;
}

BTRC_HD V3 disney_eval_sheen(const Disney &d, float cos_theta_d)
{
    return 4.0f * d.sheen * lerp3(v3s(1.0f), d.C_tint, d.sheen_tint) * disney_one_minus_5(cos_theta_d);
}

BTRC_HD V3 disney_eval_clearcoat(
    const Disney &d, float cos_theta_i, float cos_theta_o, float tan_theta_i, float tan_theta_o,
    float sin_theta_h, float cos_theta_h, float cos_theta_d)
{
    const float D = mf_gtr1(sin_theta_h, cos_theta_h, d.clearcoat_roughness);
    const float F = schlick_approx(0.04f, cos_theta_d);
    const float G = mf_smith_gtr2(tan_theta_i, 0.25f) * mf_smith_gtr2(tan_theta_o, 0.25f);
    const float val = d.clearcoat * D * F * G / fabs_(4.0f * cos_theta_i * cos_theta_o);
    return v3s(val);
}

// D * G of an anisotropic GTR2 lobe for half vector lwh and directions lwi, lwo
BTRC_HD float disney_dg(V3 lwh, V3 lwi, V3 lwo, float ax, float ay, float *D_out)
{
    float sin_phi_h, cos_phi_h;
    sincos_phi(lwh, &sin_phi_h, &cos_phi_h);
    const float cos_theta_h = lwh.z;
    const float sin_theta_h = cos2sin(cos_theta_h);
    const float D = mf_anisotropic_gtr2(sin_phi_h, cos_phi_h, sin_theta_h, cos_theta_h, ax, ay);
    float sin_phi_i, cos_phi_i, sin_phi_o, cos_phi_o;
    sincos_phi(lwi, &sin_phi_i, &cos_phi_i);
    sincos_phi(lwo, &sin_phi_o, &cos_phi_o);
    const float tan_theta_i = tan_theta(lwi);
    const float tan_theta_o = tan_theta(lwo);
    const float G = mf_smith_anisotropic_gtr2(cos_phi_i, sin_phi_i, ax, ay, tan_theta_i)
                  * mf_smith_anisotropic_gtr2(cos_phi_o, sin_phi_o, ax, ay, tan_theta_o);
    *D_out = D;
    return G;
}

// disney.cpp:100-145 (Radiance transport)
BTRC_HD V3 disney_eval_trans(const Disney &d, V3 lwi, V3 lwo)
{
    const float cos_theta_i = lwi.z;
    const float cos_theta_o = lwo.z;
    const float eta = cos_theta_o > 0.0f ? d.ior : 1.0f / d.ior;
    V3 lwh = normalize(lwo + eta * lwi);
    if(lwh.z < 0.0f)
        lwh = -lwh;
    const float cos_theta_d = dot(lwo, lwh);
    const float F = dielectric_fresnel(d.ior, 1.0f, cos_theta_d);
    float D;
    const float G = disney_dg(lwh, lwi, lwo, d.trans_ax, d.trans_ay, &D);
    const float sdem = cos_theta_d + eta * dot(lwi, lwh);
    const float corr_factor = 1.0f / eta;
    const V3 sqrtC = v3(sqrtf(d.C.x), sqrtf(d.C.y), sqrtf(d.C.z));
    const float val = (1.0f - F) * D * G * eta * eta
                    * dot(lwi, lwh) * dot(lwo, lwh)
                    * corr_factor * corr_factor
                    / (cos_theta_i * cos_theta_o * sdem * sdem);
    const float trans_factor = cos_theta_o > 0.0f ? d.transmission : 1.0f;
    return (1.0f - d.metallic) * trans_factor * sqrtC * fabs_(val);
}

// disney.cpp:147-180
BTRC_HD V3 disney_eval_inner_refl(const Disney &d, V3 lwi, V3 lwo)
{
    const V3 lwh = -normalize(lwi + lwo);
    const float cos_theta_d = dot(lwo, lwh);
    const float F = dielectric_fresnel(d.ior, 1.0f, cos_theta_d);
    float D;
    const float G = disney_dg(lwh, lwi, lwo, d.trans_ax, d.trans_ay, &D);
    return d.transmission * d.C * fabs_(F * D * G / (4.0f * lwi.z * lwo.z));
}

// disney.cpp:182-220
BTRC_HD V3 disney_eval_specular(const Disney &d, V3 lwi, V3 lwo)
{
    const float cos_theta_i = lwi.z;
    const float cos_theta_o = lwo.z;
    const V3 lwh = normalize(lwi + lwo);
    const float cos_theta_d = dot(lwi, lwh);
    const V3 Cspec = lerp3(lerp3(v3s(1.0f), d.C_tint, d.specular_tint), d.C, d.metallic);
    const V3 dielectric = Cspec * dielectric_fresnel(d.ior, 1.0f, cos_theta_d);
    const V3 conductor = schlick_approx3(Cspec, cos_theta_d);
    const V3 F = lerp3(d.specular * dielectric, conductor, d.metallic);
    float D;
    const float G = disney_dg(lwh, lwi, lwo, d.ax, d.ay, &D);
    return F * D * G / fabs_(4.0f * cos_theta_i * cos_theta_o);
}

// disney.cpp:227-331
BTRC_HD V3 disney_sample_specular(const Disney &d, V3 lwo, float s0, float s1)
{
    const V3 lwh = mf_sample_anisotropic_gtr2_vnor(lwo, d.ax, d.ay, s0, s1);
    if(lwh.z <= 0.0f)
        return v3s(0.0f);
    V3 lwi = 2.0f * dot(lwo, lwh) * lwh - lwo;
    if(lwi.z <= 0.0f)
        return v3s(0.0f);
    return normalize(lwi);
}

BTRC_HD V3 disney_sample_clearcoat(const Disney &d, V3 lwo, float s0, float s1)
{
    const V3 lwh = mf_sample_gtr1(d.clearcoat_roughness, s0, s1);
    if(lwh.z <= 0.0f)
        return v3s(0.0f);
    V3 lwi = 2.0f * dot(lwo, lwh) * lwh - lwo;
    if(lwi.z <= 0.0f)
        return v3s(0.0f);
    return normalize(lwi);
}

BTRC_HD V3 disney_sample_transmission(const Disney &d, V3 lwo, float s0, float s1)
{
    const V3 lwh = mf_sample_anisotropic_gtr2(d.trans_ax, d.trans_ay, s0, s1);
    if(lwh.z <= 0.0f)
        return v3s(0.0f);
    if((lwo.z > 0.0f) != (dot(lwh, lwo) > 0.0f))
        return v3s(0.0f);
    const float eta = lwo.z > 0.0f ? 1.0f / d.ior : d.ior;
    const V3 owh = dot(lwh, lwo) > 0.0f ? lwh : -lwh;
    V3 lwi;
    if(!disney_refract(lwo, owh, eta, lwi))
        return v3s(0.0f);
    if(lwi.z * lwo.z > 0.0f || ((lwi.z > 0.0f) != (dot(lwh, lwi) > 0.0f)))
        return v3s(0.0f);
    return normalize(lwi);
}

BTRC_HD V3 disney_sample_inner_refl(const Disney &d, V3 lwo, float s0, float s1)
{
    const V3 lwh = mf_sample_anisotropic_gtr2(d.trans_ax, d.trans_ay, s0, s1);
    if(lwh.z <= 0.0f)
        return v3s(0.0f);
    V3 lwi = 2.0f * dot(lwo, lwh) * lwh - lwo;
    if(lwi.z > 0.0f)
        return v3s(0.0f);
    return normalize(lwi);
}

// disney.cpp:339-367
BTRC_HD void disney_pdf_specular_clearcoat(const Disney &d, V3 lwi, V3 lwo, float *pdf_spec, float *pdf_cc)
{
    const V3 lwh = normalize(lwi + lwo);
    float sin_phi_h, cos_phi_h;
    sincos_phi(lwh, &sin_phi_h, &cos_phi_h);
    const float cos_theta_h = lwh.z;
    const float sin_theta_h = cos2sin(cos_theta_h);
    const float cos_theta_d = dot(lwi, lwh);

    float sin_phi_o_unused, cos_phi_o;
    sincos_phi(lwo, &sin_phi_o_unused, &cos_phi_o);
    const float sin_phi_o = cos2sin(cos_phi_o);
    const float tan_theta_o = tan_theta(lwo);

    const float specular_D = mf_anisotropic_gtr2(sin_phi_h, cos_phi_h, sin_theta_h, cos_theta_h, d.ax, d.ay);
    *pdf_spec = mf_smith_anisotropic_gtr2(cos_phi_o, sin_phi_o, d.ax, d.ay, tan_theta_o)
              * specular_D / (4.0f * lwo.z);
    const float clearcoat_D = mf_gtr1(sin_theta_h, cos_theta_h, d.clearcoat_roughness);
    *pdf_cc = cos_theta_h * clearcoat_D / (4.0f * cos_theta_d);
}

// disney.cpp:369-402
BTRC_HD float disney_pdf_transmission(const Disney &d, V3 lwi, V3 lwo)
{
    const float eta = lwo.z > 0.0f ? d.ior : 1.0f / d.ior;
    V3 lwh = normalize(lwo + eta * lwi);
    if(lwh.z < 0.0f)
        lwh = -lwh;
    if(((lwo.z > 0.0f) != (dot(lwh, lwo) > 0.0f)) || ((lwi.z > 0.0f) != (dot(lwh, lwi) > 0.0f)))
        return 0.0f;
    const float sdem = dot(lwo, lwh) + eta * dot(lwi, lwh);
    const float dwh_to_dwi = eta * eta * dot(lwi, lwh) / (sdem * sdem);
    float sin_phi_h, cos_phi_h;
    sincos_phi(lwh, &sin_phi_h, &cos_phi_h);
    const float cos_theta_h = lwh.z;
    const float sin_theta_h = cos2sin(cos_theta_h);
    const float D = mf_anisotropic_gtr2(sin_phi_h, cos_phi_h, sin_theta_h, cos_theta_h, d.trans_ax, d.trans_ay);
    return fabs_(dot(lwi, lwh) * D * dwh_to_dwi);
}

// disney.cpp:404-420
BTRC_HD float disney_pdf_inner_refl(const Disney &d, V3 lwi, V3 lwo)
{
    const V3 lwh = -normalize(lwi + lwo);
    float sin_phi_h, cos_phi_h;
    sincos_phi(lwh, &sin_phi_h, &cos_phi_h);
    const float cos_theta_h = lwh.z;
    const float sin_theta_h = cos2sin(cos_theta_h);
    const float cos_theta_d = dot(lwi, lwh);
    const float D = mf_anisotropic_gtr2(sin_phi_h, cos_phi_h, sin_theta_h, cos_theta_h, d.trans_ax, d.trans_ay);
    return fabs_(cos_theta_h * D / (4.0f * cos_theta_d));
}

// DisneyBSDF::eval, disney.cpp:477-558
BTRC_HD V3 disney_eval(const Disney &d, V3 lwi, V3 lwo)
{
    if(fabs_(lwi.z) < 1e-4f || fabs_(lwo.z) < 1e-4f)
        return v3s(0.0f);
    if(lwi.z * lwo.z < 0.0f)
    {
        if(d.transmission == 0.0f)
            return v3s(0.0f);
        return disney_eval_trans(d, lwi, lwo);
    }
    if(lwi.z < 0.0f && lwo.z < 0.0f)
    {
        if(d.transmission == 0.0f)
            return v3s(0.0f);
        return disney_eval_inner_refl(d, lwi, lwo);
    }
    if(lwi.z <= 0.0f || lwo.z <= 0.0f)
        return v3s(0.0f);

    const float cos_theta_i = lwi.z;
    const float cos_theta_o = lwo.z;
    const V3 lwh = normalize(lwi + lwo);
    const float cos_theta_d = dot(lwi, lwh);

    V3 diffuse = v3s(0.0f), sheen = v3s(0.0f);
    if(d.metallic < 1.0f)
    {
        diffuse = disney_eval_diffuse(d, cos_theta_i, cos_theta_o, cos_theta_d);
        if(d.sheen > 0.0f)
            sheen = disney_eval_sheen(d, cos_theta_d);
    }
    const V3 specular = disney_eval_specular(d, lwi, lwo);
    V3 clearcoat = v3s(0.0f);
    if(d.clearcoat > 0.0f)
    {
        const float tan_theta_i = tan_theta(lwi);
        const float tan_theta_o = tan_theta(lwo);
        const float cos_theta_h = lwh.z;
        const float sin_theta_h = cos2sin(cos_theta_h);
        clearcoat = disney_eval_clearcoat(
            d, cos_theta_i, cos_theta_o, tan_theta_i, tan_theta_o, sin_theta_h, cos_theta_h, cos_theta_d);
    }
    return (1.0f - d.metallic) * (1.0f - d.transmission) * (diffuse + sheen) + specular + clearcoat;
}

// DisneyBSDF::pdf, disney.cpp:682-730
BTRC_HD float disney_pdf(const Disney &d, V3 lwi, V3 lwo)
{
    if(fabs_(lwi.z) < 1e-4f || fabs_(lwo.z) < 1e-4f)
        return 0.0f;
    if(lwo.z < 0.0f)
    {
        if(d.transmission == 0.0f)
            return 0.0f;
        const float macro_F = clampf(dielectric_fresnel(d.ior, 1.0f, lwo.z), 0.1f, 0.9f);
        if(lwi.z > 0.0f)
            return (1.0f - macro_F) * disney_pdf_transmission(d, lwi, lwo);
        return macro_F * disney_pdf_inner_refl(d, lwi, lwo);
    }
    if(lwi.z < 0.0f)
        return d.transmission_weight * disney_pdf_transmission(d, lwi, lwo);
    const float diffuse = pdf_sample_hemisphere_zweighted(lwi);
    float specular, clearcoat;
    disney_pdf_specular_clearcoat(d, lwi, lwo, &specular, &clearcoat);
    return d.diffuse_weight * diffuse + d.specular_weight * specular + d.clearcoat_weight * clearcoat;
}

// DisneyBSDF::sample, disney.cpp:560-658
BTRC_HD void disney_sample(const Disney &d, V3 lwo, float s0, float s1, float s2, BSDFSample &out)
{
    bsdf_sample_clear(out);
    if(fabs_(lwo.z) < 1e-4f)
        return;

    V3 lwi;
    if(lwo.z < 0.0f)
    {
        if(d.transmission == 0.0f)
            return;
        const float macro_F = clampf(dielectric_fresnel(d.ior, 1.0f, lwo.z), 0.1f, 0.9f);
        if(s0 >= macro_F)
            lwi = disney_sample_transmission(d, lwo, s1, s2);
        else
            lwi = disney_sample_inner_refl(d, lwo, s1, s2);
    }
    else
    {
        float sel = s0;
        if(sel < d.diffuse_weight)
            lwi = sample_hemisphere_zweighted(s1, s2);
        else
        {
            sel = sel - d.diffuse_weight;
            if(sel < d.transmission_weight)
                lwi = disney_sample_transmission(d, lwo, s1, s2);
            else
            {
                sel = sel - d.transmission_weight;
                if(sel < d.specular_weight)
                    lwi = disney_sample_specular(d, lwo, s1, s2);
                else
                    lwi = disney_sample_clearcoat(d, lwo, s1, s2);
            }
        }
    }
    if(lwi.x == 0.0f && lwi.y == 0.0f && lwi.z == 0.0f)
        return;
    out.dir = lwi;
    out.bsdf = disney_eval(d, lwi, lwo);
    out.pdf = disney_pdf(d, lwi, lwo);
    // handle_bad_sample, disney.cpp:566-573
    const V3 v = out.bsdf * fabs_(out.dir.z) / out.pdf;
    if(fmax_(fmax_(v.x, v.y), v.z) > 10.0f)
        bsdf_sample_clear(out);
}

// ---- shading frame and aggregate, shader_frame.cpp / aggregate.cpp ----

struct ShaderFrame
{
    Frame geometry;
    Frame shading;
};

BTRC_HD bool frame_is_black_fringes(const ShaderFrame &f, V3 w)
{
    return in_positive_z_hemisphere(f.geometry, w) != in_positive_z_hemisphere(f.shading, w);
}

BTRC_HD ShaderFrame frame_flip_for_black_fringes(const ShaderFrame &f, V3 wo)
{
    ShaderFrame r;
    r.geometry = f.geometry;
    r.shading = frame_is_black_fringes(f, wo) ? f.geometry : f.shading;
    return r;
}

BTRC_HD float frame_correct_shading_energy(const ShaderFrame &f, V3 wi)
{
    const float down = fabs_(dot(f.geometry.z, wi));
    if(down > 1e-4f)
    {
        const float up = fabs_(dot(f.shading.z, wi));
        return up / down;
    }
    return 1.0f;
}

// aggregate.cpp:221-229
BTRC_HD float shadow_terminator_term(const ShaderFrame &f, bool enabled, V3 wi)
{
    if(!enabled)
        return 1.0f;
    const float wgwi = dot(f.geometry.z, wi);
    const float wswi = dot(f.shading.z, wi);
    const float wgws = dot(f.geometry.z, f.shading.z);
    const float g = saturate(wgwi / (wswi * wgws));
    return -g * g * g + g * g + g;
}

// A material instantiated at a surface point (Material::create_shader)
struct Shader
{
    ShaderFrame frame;
    int32_t type;
    bool shadow_terminator;
    V3 albedo;  // Lambert albedo; Disney base colour; the g-buffer albedo
    Disney disney;
};

BTRC_HD Shader make_shader(const MaterialRec &m, const Frame &geometry_frame, V3 interp_z)
{
    Shader s;
    s.frame.geometry = geometry_frame;
    s.frame.shading = rotate_to_new_z(geometry_frame, interp_z);
    if(m.normal_tex >= 0 && m.type != MATERIAL_INVISIBLE)
    {
        // NormalMap::adjust_frame, material/utils/normal_map.cpp:13-21
        const V3 local = v3(2.0f * m.normal_rgb[0] - 1.0f, 2.0f * m.normal_rgb[1] - 1.0f, 2.0f * m.normal_rgb[2] - 1.0f);
        s.frame.shading = rotate_to_new_z(s.frame.shading, normalize(local_to_global(s.frame.shading, local)));
    }
    s.type = m.type;
    s.shadow_terminator = m.shadow_terminator != 0;
    s.albedo = m.type == MATERIAL_INVISIBLE ? v3s(1.0f) : v3(m.color[0], m.color[1], m.color[2]);
    if(m.type == MATERIAL_DISNEY)
        s.disney = disney_make(m);
    return s;
}

BTRC_HD V3 component_eval(const Shader &s, V3 lwi, V3 lwo)
{
    if(s.type == MATERIAL_DISNEY)
        return disney_eval(s.disney, lwi, lwo);
    return lambert_eval(s.albedo, lwi, lwo);
}

BTRC_HD float component_pdf(const Shader &s, V3 lwi, V3 lwo)
{
    if(s.type == MATERIAL_DISNEY)
        return disney_pdf(s.disney, lwi, lwo);
    return lambert_pdf(lwi, lwo);
}

// g-buffer normal: the shading normal; an invisible surface reports its
// geometric normal (invisible.cpp:46-55)
BTRC_HD V3 shader_normal(const Shader &s)
{
    return s.type == MATERIAL_INVISIBLE ? s.frame.geometry.z : s.frame.shading.z;
}

// BSDFAggregate::sample, aggregate.cpp:32-103; InvisibleShader passes the
// ray straight through as a delta lobe (invisible.cpp:13-26)
BTRC_HD void shader_sample(const Shader &s, V3 wo, float s0, float s1, float s2, BSDFSample &out)
{
    if(s.type == MATERIAL_INVISIBLE)
    {
        out.bsdf = v3s(1.0f) / fabs_(dot(s.frame.geometry.z, normalize(wo)));
        out.dir = -wo;
        out.pdf = 1.0f;
        out.is_delta = true;
        return;
    }
    const ShaderFrame frame = frame_flip_for_black_fringes(s.frame, wo);
    const V3 lwo = normalize(global_to_local(frame.shading, wo));
    BSDFSample comp;
    if(s.type == MATERIAL_DISNEY)
        disney_sample(s.disney, lwo, s0, s1, s2, comp);
    else
        lambert_sample(s.albedo, lwo, s0, s1, comp);
    bsdf_sample_clear(out);
    if(comp.pdf <= 0.0f)
        return;
    const V3 dir = local_to_global(frame.shading, comp.dir);
    if(frame_is_black_fringes(s.frame, dir))
        return;
    out.dir = dir;
    out.pdf = comp.pdf;
    const float corr = frame_correct_shading_energy(frame, dir);
    const float st = shadow_terminator_term(s.frame, s.shadow_terminator, dir);
    out.bsdf = comp.bsdf * corr * st;
    out.is_delta = false;
}

// BSDFAggregate::eval, aggregate.cpp:189-202
BTRC_HD V3 shader_eval(const Shader &s, V3 wi, V3 wo)
{
    if(s.type == MATERIAL_INVISIBLE)
        return v3s(0.0f);
    const ShaderFrame frame = frame_flip_for_black_fringes(s.frame, wo);
    const V3 lwi = normalize(global_to_local(frame.shading, wi));
    const V3 lwo = normalize(global_to_local(frame.shading, wo));
    const V3 r = component_eval(s, lwi, lwo);
    const float corr = frame_correct_shading_energy(frame, wi);
    const float st = shadow_terminator_term(s.frame, s.shadow_terminator, wi);
    return r * corr * st;
}

// BSDFAggregate::pdf, aggregate.cpp:204-214
BTRC_HD float shader_pdf(const Shader &s, V3 wi, V3 wo)
{
    if(s.type == MATERIAL_INVISIBLE)
        return 0.0f;
    const ShaderFrame frame = frame_flip_for_black_fringes(s.frame, wo);
    const V3 lwi = normalize(global_to_local(frame.shading, wi));
    const V3 lwo = normalize(global_to_local(frame.shading, wo));
    return component_pdf(s, lwi, lwo);
}

} // namespace btrc
