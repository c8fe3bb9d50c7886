/*
 * tr_shading.h - BSDFs, light sampling and geometry sampling of the kernels.
 *
 * Each function restates the reference function named beside it, in the
 * reference's FP32 operation order.  Deviations from the reference, all
 * deliberate and matched by the CPU oracle:
 *  - random numbers come from Philox (tr_math.h) instead of FastFloatRand;
 *  - sin/cos/log/exp come from tr_fmath.h instead of FastSinCos / libm, and
 *    angles are taken in turns (sin(2*pi*u) is tr_sincos_turns(u));
 *  - pow(x, 5) in FresnelSchlick is x^2*x^2*x;
 *  - a mesh light's triangle is picked as floor(u32 * count / 2^32) instead
 *    of Hash32(seed) % count (functions.hh:243-247);
 *  - the microfacet specular term returns 0 instead of 0/0 = NaN when a
 *    direction is exactly grazing (cos = 0).
 */
#ifndef TR_SHADING_H
#define TR_SHADING_H

#include "tr_math.h"
#include "tr_scene.h"
#include "tr_texture.h"
#include "tr_traverse.h"

namespace trk {

/* SampleData (ray-tracing-common.hh:184-198), the parts the models read. */
struct Sample {
    const tr_material* mat;
    V3 wo, wi, tangent, binormal, normal;
    float dir_density;
    float pdf;
    float tcx, tcy;  /* TexCoord */
};

struct Strata {
    float x, y, tx, ty;
};

TR_HD float geom_factor_sq(float power) { return 2.0f / (power + 2.0f); } /* microfacet-common.hh:109-112 */
TR_HD float fresnel_schlick(float refl, float c) { return refl + (1.0f - refl) * pow5(1.0f - c); } /* functions.hh:262 */
TR_HD float cos_transmittance(float eta, float ci) { return sqrtf(1.0f - (1.0f - ci * ci) / (eta * eta)); } /* :268 */

/* FresnelConductor, functions.hh:275-294 */
TR_HD float fresnel_conductor(float eta, float k, float cos_inc)
{
    float cos_inc2 = cos_inc * cos_inc;
    float sin_inc2 = 1.0f - cos_inc2;
    float eta2 = eta * eta;
    float etak2 = k * k;
    float t0 = eta2 - etak2 - sin_inc2;
    float a2plusb2 = sqrtf(t0 * t0 + 4.0f * eta2 * etak2);
    float t1 = a2plusb2 + cos_inc2;
    float a = sqrtf(0.5f * (a2plusb2 + t0));
    float t2 = 2.0f * a * cos_inc;
    float r_perp = (t1 - t2) / (t1 + t2);
    float t3 = cos_inc2 * a2plusb2 + sin_inc2 * sin_inc2;
    float t4 = t2 * sin_inc2;
    float r_parallel = r_perp * (t3 - t4) / (t3 + t4);
    return 0.5f * (r_perp + r_parallel);
}

/* FresnelDielectric, functions.hh:296-301 */
TR_HD float fresnel_dielectric(float eta, float ci, float ct)
{
    float r_parallel = (eta * ci - ct) / (eta * ci + ct);
    float r_perp = (ci - eta * ct) / (ci + eta * ct);
    return 0.5f * (r_parallel * r_parallel + r_perp * r_perp);
}

/* GGXMicrofacetDistribution, ggx-models-impl.hh:126-139 */
TR_HD float ggx_d(float g2, float cos_nh)
{
    if (cos_nh <= 0.0f) return 0.0f;
    float c2 = cos_nh * cos_nh;
    float c4 = c2 * c2;
    float tan2 = (1.0f - c2) / c2;
    float den = g2 + tan2;
    return g2 / (kPi * c4 * den * den);
}

/* G1_GGX, ggx-models-impl.hh:141-151 */
TR_HD float ggx_g1(float g2, float cos_vm, float cos_vn)
{
    float ratio = cos_vm / cos_vn;
    if (ratio <= 0.0f) return 0.0f;
    float c2 = cos_vn * cos_vn;
    float tan2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + g2 * tan2));
}

/* BeckmannMicrofacetDistribution, beckmann-models-impl.hh:57-70 */
TR_HD float beckmann_d(float g2, float cos_nh)
{
    if (cos_nh <= 0.0f) return 0.0f;
    float c2 = cos_nh * cos_nh;
    float c4 = c2 * c2;
    float tan2 = (1.0f - c2) / c2;
    return tr_expf(-tan2 / g2) / (kPi * g2 * c4);
}

/* G1_Beckmann, rational approximation, beckmann-models-impl.hh:34-55 */
TR_HD float beckmann_g1(float g2, float cos_vm, float cos_vn)
{
    float ratio = cos_vm / cos_vn;
    if (ratio <= 0.0f) return 0.0f;
    float c2 = cos_vn * cos_vn;
    float tan2 = (1.0f - c2) / c2;
    float a2 = 1.0f / (g2 * tan2);
    float a = sqrtf(a2);
    return a < 1.6f ? (3.535f * a + 2.181f * a2) / (1.0f + 2.276f * a + 2.577f * a2) : 1.0f;
}

TR_HD bool is_beckmann(uint32_t m) { return m >= TR_MODEL_BECKMANN_MICROFACET && m <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC; }
TR_HD bool is_ggx(uint32_t m) { return m >= TR_MODEL_GGX_MICROFACET && m <= TR_MODEL_GGX_MICROFACET_DIELECTRIC; }
TR_HD bool is_microfacet(uint32_t m) { return is_ggx(m) || is_beckmann(m); }

/* GGXMicrofacet{,Conductor,Dielectric}BRDF (ggx-models-impl.hh:156-209) and
 * the Beckmann twins (beckmann-models-impl.hh:121-172): BRDF * cos. */
TR_HD V3 microfacet_brdf(const Sample& s)
{
    const tr_material* m = s.mat;
    uint32_t model = m->model;
    bool beck = is_beckmann(model);
    /* ComputeMicrofacetAngles, microfacet-common.hh:74-84 */
    V3 h = normalize(add(s.wi, s.wo));
    float cos_in = dot(s.wi, s.normal);
    float cos_on = dot(s.wo, s.normal);
    float cos_im = dot(s.wi, h);
    float cos_mn = dot(s.normal, h);
    uint32_t variant = model - (beck ? (uint32_t)TR_MODEL_BECKMANN_MICROFACET : (uint32_t)TR_MODEL_GGX_MICROFACET);
    float fresnel;
    if (variant == 0) {
        fresnel = fresnel_schlick(m->fresnel[0], cos_im);
    } else if (variant == 1) {
        fresnel = fresnel_conductor(m->fresnel[0], m->fresnel[1], cos_im);
    } else {
        float ct = cos_transmittance(m->fresnel[0], cos_im);
        fresnel = fresnel_dielectric(m->fresnel[0], cos_im, ct);
    }
    float g2 = geom_factor_sq(m->specular_power[0]);
    float shadow, d;
    if (beck) {
        shadow = beckmann_g1(g2, cos_im, cos_in) * beckmann_g1(g2, cos_im, cos_on);
        d = beckmann_d(g2, cos_mn);
    } else {
        shadow = ggx_g1(g2, cos_im, cos_in) * ggx_g1(g2, cos_im, cos_on);
        d = ggx_d(g2, cos_mn);
    }
    /* At exactly grazing incidence (cos = 0) the reference's
     * shadow*D/(4|cos_in||cos_on|) is 0/0; G1 is 0 there and the limit 0 is
     * used instead of NaN. */
    float den = 4.0f * fabsf(cos_in) * fabsf(cos_on);
    float spec = den > 0.0f ? shadow * d / den : 0.0f;
    float brdf = fresnel * spec;
    return muls(add(muls(ld3(m->diffuse), 1.0f / kPi), muls(ld3(m->specular), brdf)), cos_in);
}

/* GGXMicrofacetPDF / BeckmannMicrofacetPDF (ggx-models-impl.hh:211-237,
 * beckmann-models-impl.hh:240-266) */
TR_HD float microfacet_pdf(const Sample& s)
{
    const tr_material* m = s.mat;
    float cos_ln = dot(s.wi, s.normal);
    if (cos_ln < 0.0f) return 0.0f;
    V3 h = normalize(add(s.wi, s.wo));
    float cos_hn = dot(h, s.normal);
    float g2 = geom_factor_sq(m->specular_power[0]);
    float norm_pdf = (is_beckmann(m->model) ? beckmann_d(g2, cos_hn) : ggx_d(g2, cos_hn)) * cos_hn;
    if (cos_hn <= 0.0f || norm_pdf <= 0.0f) return 0.0f;
    float jacobian = 1.0f / (4.0f * dot(h, s.wo));
    float pdf0 = norm_pdf * jacobian;
    float mv = max_value(add(ld3(m->diffuse), ld3(m->specular)));
    float ratio = mv != 0.0f ? max_value(ld3(m->specular)) / mv : 0.0f;
    float pdf1 = cos_ln * (1.0f / kPi);
    return (pdf0 - pdf1) * ratio + pdf1;
}

/* GGXMicrofacetSampleIncidentLight / BeckmannMicrofacetSampleIncidentLight
 * (ggx-models-impl.hh:239-282, beckmann-models-impl.hh:268-311) with
 * CosineSampleHemisphere (sampling3.hh:69-76). */
TR_HD void microfacet_sample(const Strata& st, Sample* s, float xi0, float xi1, float xi2)
{
    const tr_material* m = s->mat;
    float x_ray = (st.x + xi0) / st.tx;
    float y_ray = (st.y + xi1) / st.ty;
    float mv = max_value(add(ld3(m->diffuse), ld3(m->specular)));
    float ratio = mv != 0.0f ? max_value(ld3(m->specular)) / mv : 0.0f;
    M3 basis = make_basis(s->normal);
    float sin_phi, cos_phi;
    tr_sincos_turns(y_ray, &sin_phi, &cos_phi);
    if (xi2 <= ratio) {
        float g2 = geom_factor_sq(m->specular_power[0]);
        float g = sqrtf(g2);
        float r = is_beckmann(m->model) ? -tr_logf(x_ray) : x_ray / (1.0f - x_ray);
        float sqrt_ratio = sqrtf(r);
        float rcp_len = 1.0f / sqrtf(r * g2 + 1.0f);
        float common = sqrt_ratio * g * rcp_len;
        V3 mn = rcp_len > 1e-6f ? mk(-common * cos_phi, -common * sin_phi, rcp_len) : mk(0.0f, 0.0f, 1.0f);
        mn = transform(basis, mn);
        s->wi = reflect(s->wo, mn);
    } else {
        float r = sqrtf(x_ray);
        V3 d = mk(cos_phi * r, sin_phi * r, sqrtf(fmax_(0.0f, 1.0f - x_ray)));
        s->wi = normalize(transform(basis, d));
    }
    s->pdf = microfacet_pdf(*s);
}

/* ---------------- anisotropic GGX / Beckmann ---------------- */

TR_HD bool is_aniso(uint32_t m)
{
    return (m >= TR_MODEL_GGX_MICROFACET_ANISOTROPIC && m <= TR_MODEL_GGX_MICROFACET_DIELECTRIC_ANISOTROPIC) ||
           (m >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC && m <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC_ANISOTROPIC);
}

/* NormalizeSelf of a Vector2 (math/vector2.hh:60-74): divides by the length,
 * leaves (0, 0) alone. */
TR_HD void normalize2(float* x, float* y)
{
    float l = sqrtf(*x * *x + *y * *y);
    if (l != 0.0f) {
        *x /= l;
        *y /= l;
    }
}

/* GGXMicrofacetDistributionAnisotropic (ggx-models-impl.hh:111-127) and
 * BeckmannMicrofacetDistributionAnisotropic (beckmann-models-impl.hh:69-85).
 * (ct, st): the half vector's direction in the tangent plane.  The reference
 * has no cos <= 0 guard; every caller then multiplies by a zero shadowing
 * term or rejects the non-positive PDF, so the guard changes only the NaN of
 * cos = 0 into 0. */
TR_HD float aniso_d(bool beck, float ax, float ay, float cos_nh, float ct, float st)
{
    if (cos_nh <= 0.0f) return 0.0f;
    float c2 = cos_nh * cos_nh;
    float c4 = c2 * c2;
    float tan2 = (1.0f - c2) / c2;
    float r1 = ct / ax, r2 = st / ay;
    float aniso = r1 * r1 + r2 * r2;
    if (beck) return tr_expf(-tan2 * aniso) / (kPi * ax * ay * c4);
    float den = 1.0f + tan2 * aniso;
    return 1.0f / (kPi * c4 * ax * ay * den * den);
}

TR_HD void aniso_alphas(const tr_material* m, float* ax, float* ay)
{
    *ax = sqrtf(geom_factor_sq(m->specular_power[0]));
    *ay = sqrtf(geom_factor_sq(m->specular_power[1]));
}

/* GGXMicrofacetAnisotropic{,Conductor,Dielectric}BRDF with
 * GGXSpecularAnisotropicBRDF (ggx-models-impl.hh:283-341) and the Beckmann
 * twins (beckmann-models-impl.hh:174-238); ComputeMicrofacetAnglesAnisotropic
 * and ComputeRoughnessProjectedOnDirectionSq, microfacet-common.hh:86-107.
 * Both G1 terms use the roughness projected on the outgoing direction. */
TR_HD V3 microfacet_aniso_brdf(const Sample& s)
{
    const tr_material* m = s.mat;
    uint32_t model = m->model;
    bool beck = model >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC;
    V3 h = normalize(add(s.wi, s.wo));
    float cos_in = dot(s.wi, s.normal);
    float cos_on = dot(s.wo, s.normal);
    float cos_im = dot(s.wi, h);
    float cos_mn = dot(s.normal, h);
    float ht = dot(s.tangent, h), hb = dot(s.binormal, h);
    normalize2(&ht, &hb);
    uint32_t variant =
        model - (beck ? (uint32_t)TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC : (uint32_t)TR_MODEL_GGX_MICROFACET_ANISOTROPIC);
    float fresnel;
    if (variant == 0) {
        fresnel = fresnel_schlick(m->fresnel[0], cos_im);
    } else if (variant == 1) {
        fresnel = fresnel_conductor(m->fresnel[0], m->fresnel[1], cos_im);
    } else {
        float ct = cos_transmittance(m->fresnel[0], cos_im);
        fresnel = fresnel_dielectric(m->fresnel[0], cos_im, ct);
    }
    float ot = dot(s.wo, s.tangent), ob = dot(s.wo, s.binormal);
    normalize2(&ot, &ob);
    float ax, ay;
    aniso_alphas(m, &ax, &ay);
    float tt = ax * ot, bt = ay * ob;
    float g2o = fabsf(tt * tt + bt * bt);
    float shadow = beck ? beckmann_g1(g2o, cos_im, cos_in) * beckmann_g1(g2o, cos_im, cos_on)
                        : ggx_g1(g2o, cos_im, cos_in) * ggx_g1(g2o, cos_im, cos_on);
    float d = aniso_d(beck, ax, ay, cos_mn, ht, hb);
    float den = 4.0f * fabsf(cos_in) * fabsf(cos_on);
    float spec = den > 0.0f ? shadow * d / den : 0.0f;
    float brdf = fresnel * spec;
    return muls(add(muls(ld3(m->diffuse), 1.0f / kPi), muls(ld3(m->specular), brdf)), cos_in);
}

/* GGXMicrofacetAnisotropicPDF / BeckmannMicrofacetAnisotropicPDF
 * (ggx-models-impl.hh:343-375, beckmann-models-impl.hh:313-345) */
TR_HD float microfacet_aniso_pdf(const Sample& s)
{
    const tr_material* m = s.mat;
    float cos_ln = dot(s.wi, s.normal);
    if (cos_ln < 0.0f) return 0.0f;
    V3 h = normalize(add(s.wi, s.wo));
    float cos_hn = dot(h, s.normal);
    float ht = dot(s.tangent, h), hb = dot(s.binormal, h);
    normalize2(&ht, &hb);
    float ax, ay;
    aniso_alphas(m, &ax, &ay);
    float norm_pdf = aniso_d(m->model >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC, ax, ay, cos_hn, ht, hb) * cos_hn;
    if (norm_pdf <= 0.0f) return 0.0f;
    float jacobian = 1.0f / (4.0f * dot(h, s.wo));
    float pdf0 = norm_pdf * jacobian;
    float mv = max_value(add(ld3(m->diffuse), ld3(m->specular)));
    float ratio = mv != 0.0f ? max_value(ld3(m->specular)) / mv : 0.0f;
    float pdf1 = cos_ln * (1.0f / kPi);
    return (pdf0 - pdf1) * ratio + pdf1;
}

/* GGXMicrofacetAnisotropicSampleIncidentLight with
 * GGXSampleMicrofacetAnisotropicDistribution (ggx-models-impl.hh:377-432)
 * and the Beckmann twins (beckmann-models-impl.hh:347-398).  The specular
 * branch draws a second stratified pair (xi3, xi4); the micro-normal is
 * built in makeBasis(Normal), not the tangent frame, as in the reference. */
TR_HD void microfacet_aniso_sample(const Strata& st, Sample* s, float xi0, float xi1, float xi2, float xi3, float xi4)
{
    const tr_material* m = s->mat;
    float x_ray = (st.x + xi0) / st.tx;
    float y_ray = (st.y + xi1) / st.ty;
    float mv = max_value(add(ld3(m->diffuse), ld3(m->specular)));
    float ratio = mv != 0.0f ? max_value(ld3(m->specular)) / mv : 0.0f;
    M3 basis = make_basis(s->normal);
    if (xi2 <= ratio) {
        float ax, ay;
        aniso_alphas(m, &ax, &ay);
        float xr = (st.x + xi3) / st.tx;
        float yr = (st.y + xi4) / st.ty;
        float sin_phi, cos_phi;
        tr_sincos_turns(yr, &sin_phi, &cos_phi);
        float sqrt_ratio =
            m->model >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC ? sqrtf(-tr_logf(xr)) : sqrtf(xr / (1.0f - xr));
        float x = ax * sqrt_ratio * cos_phi;
        float y = ay * sqrt_ratio * sin_phi;
        V3 mn = normalize(normalize(mk(-x, -y, 1.0f)));
        mn = normalize(transform(basis, mn));
        s->wi = reflect(s->wo, mn);
    } else {
        float sin_phi, cos_phi;
        tr_sincos_turns(y_ray, &sin_phi, &cos_phi);
        float r = sqrtf(x_ray);
        V3 d = mk(cos_phi * r, sin_phi * r, sqrtf(fmax_(0.0f, 1.0f - x_ray)));
        s->wi = normalize(transform(basis, d));
    }
    s->pdf = microfacet_aniso_pdf(*s);
}

/* powf(x, p) for the Blinn-Phong lobes, in the shared exp/log: x > 0 gives
 * exp(p log x); x = 0 gives 0 for p > 0 and 1 for p = 0; x < 0 gives
 * |x|^p with the sign of (-1)^p for an integer p and NaN otherwise. */
TR_HD float phong_pow(float x, float p)
{
    if (x > 0.0f) return tr_expf(p * tr_logf(x));
    if (p == 0.0f) return 1.0f;
    if (x == 0.0f) return p > 0.0f ? 0.0f : __builtin_inff();
    float fl = floorf(p);
    if (fl != p) return __builtin_nanf("");
    float m = tr_expf(p * tr_logf(-x));
    return fmodf(fl, 2.0f) != 0.0f ? -m : m;
}

/* Blinn-Phong's lobe choice: MaxValue(Specular) / MaxValue(albedo +
 * Specular), 0 for black. */
TR_HD float lobe_ratio(V3 albedo, V3 specular)
{
    float mv = max_value(add(albedo, specular));
    return mv != 0.0f ? max_value(specular) / mv : 0.0f;
}

/* Scenes without texture maps, cube map and pseudo-volume materials shade
 * through this view: every map lookup is the material's constant, at no
 * cost, and the pseudo-volume BRDF (hundreds of random walks per call) is
 * left out of those kernel instances. */
struct NoTex {};
template <class TV> struct FullModels {
    static constexpr bool value = true;
};
template <> struct FullModels<NoTex> {
    static constexpr bool value = false;
};

TR_HD V3 map_spectrum(const NoTex&, uint32_t, const float* constant, const Sample&) { return ld3(constant); }

/* A spectrum map (SampleSpectrum at the TexCoord) or the constant. */
TR_HD V3 map_spectrum(const TexView& tv, uint32_t map, const float* constant, const Sample& s)
{
    if (map == 0u) return ld3(constant);
    V4f c = tex_sample(tv, map, s.tcx, s.tcy, TEX_SPECTRUM);
    return mk(c.x, c.y, c.z);
}

/* BlinnPhongBRDF (illumination-models.hh:679-698): BRDF * cos; the albedo
 * from DiffuseMap when there is one. */
template <class TV> TR_HD V3 blinn_phong_brdf(const Sample& s, const TV& tv)
{
    const tr_material* m = s.mat;
    V3 albedo = map_spectrum(tv, m->diffuse_map, m->diffuse, s);
    float power = m->specular_power[0];
    float cos_ln = fmax_(0.0f, dot(s.normal, s.wi));
    V3 h = normalize(add(s.wi, s.wo));
    float lobe = phong_pow(fmax_(0.0f, dot(s.normal, h)), power);
    float cos_theta = fmin_(fmax_(dot(s.wi, h), 0.0f), 1.0f);
    float fresnel = fresnel_schlick(m->fresnel[0], cos_theta);
    V3 spec = muls(muls(muls(muls(ld3(m->specular), lobe), fresnel), power + 8.0f), 1.0f / 8.0f);
    return muls(muls(add(albedo, spec), cos_ln), 1.0f / kPi);
}

/* BlinnPhongPDF (illumination-models.hh:618-643) */
template <class TV> TR_HD float blinn_phong_pdf(const Sample& s, const TV& tv)
{
    const tr_material* m = s.mat;
    V3 h = normalize(add(s.wi, s.wo));
    V3 albedo = map_spectrum(tv, m->diffuse_map, m->diffuse, s);
    float cos_ln = dot(s.wi, s.normal);
    if (cos_ln < 0.0f) return 0.0f;
    float ratio = lobe_ratio(albedo, ld3(m->specular));
    float power = m->specular_power[0];
    float pdf0 = (power + 1.0f) * phong_pow(dot(h, s.normal), power) * (1.0f / (2.0f * kPi));
    float jac = 4.0f * dot(h, s.wo);
    pdf0 = jac > 0.0f && pdf0 > 0.0f ? pdf0 / jac : 0.0f;
    float pdf1 = cos_ln * (1.0f / kPi);
    return (pdf0 - pdf1) * ratio + pdf1;
}

/* BlinnPhongSampleIncidentLight (illumination-models.hh:645-677) with
 * PowerCosineSampleHemisphere (sampling3.hh:96-104) and
 * CosineSampleHemisphere (:69-76). */
template <class TV>
TR_HD void blinn_phong_sample(const Strata& st, Sample* s, float xi0, float xi1, float xi2, const TV& tv)
{
    const tr_material* m = s->mat;
    float x_ray = (st.x + xi0) / st.tx;
    float y_ray = (st.y + xi1) / st.ty;
    float ratio = lobe_ratio(map_spectrum(tv, m->diffuse_map, m->diffuse, *s), ld3(m->specular));
    M3 basis = make_basis(s->normal);
    float sin_phi, cos_phi;
    tr_sincos_turns(y_ray, &sin_phi, &cos_phi);
    if (xi2 <= ratio) {
        float cos_t = phong_pow(x_ray, 1.0f / (m->specular_power[0] + 1.0f));
        float sin_t = sqrtf(fmax_(0.0f, 1.0f - cos_t * cos_t));
        V3 mn = normalize(transform(basis, mk(cos_phi * sin_t, sin_phi * sin_t, cos_t)));
        s->wi = reflect(s->wo, mn);
    } else {
        float r = sqrtf(x_ray);
        V3 d = mk(cos_phi * r, sin_phi * r, sqrtf(fmax_(0.0f, 1.0f - x_ray)));
        s->wi = normalize(transform(basis, d));
    }
    s->pdf = blinn_phong_pdf(*s, tv);
}

/* UniformHemispherePDF, illumination-models.hh:723-727 */
TR_HD float uniform_hemisphere_pdf(const Sample& s)
{
    return dot(s.normal, s.wi) >= 0.0f ? 1.0f / (2.0f * kPi) : 0.0f;
}

/* UniformSampleIncidentLight, illumination-models.hh:729-738, with
 * UniformSampleHemisphere (sampling3.hh:55-62) */
TR_HD void uniform_hemisphere_sample(const Strata& st, Sample* s, float xi0, float xi1)
{
    M3 basis = make_basis(s->normal);
    float ra = (st.x + xi0) / st.tx;
    float rb = (st.y + xi1) / st.ty;
    float r = sqrtf(1.0f - ra * ra);
    float sin_phi, cos_phi;
    tr_sincos_turns(rb, &sin_phi, &cos_phi);
    V3 d = mk(cos_phi * r, sin_phi * r, ra);
    s->pdf = 1.0f / (2.0f * kPi);
    s->wi = normalize(transform(basis, d));
}

/* SGGXProjectedArea, ggx-models-impl.hh:99-102 */
TR_HD float sggx_projected_area(V3 stddev, const M3& basis, V3 dir)
{
    return length(mulv(transform_rotation_inverse(basis, dir), stddev));
}

/* SGGXMicroFlakeNDF(stddev, basis, normal), ggx-models-impl.hh:48-57 */
TR_HD float sggx_ndf(V3 stddev, const M3& basis, V3 n)
{
    V3 sq = mulv(stddev, stddev);
    M3 inv_cov = transform_covariance(basis, mk(1.0f / sq.x, 1.0f / sq.y, 1.0f / sq.z));
    float det = determinant(inv_cov);
    if (det < 1e-6f) return 0.0f;
    float den = dot(n, transform(inv_cov, n));
    return 0.5f * sqrtf(det) / (kPi * den * den);
}

/* The SGGX surface's StandardDeviation (x, y; z = 1) and SGGXBasis
 * quaternion, from StandardDeviationMap (SampleRG) and BasisMap
 * (SampleRGBA) when present (ggx-models-impl.hh:495-517, 593-608). */
TR_HD V3 sggx_surface_stddev(const Sample& s, const TexView& tv)
{
    const tr_material* m = s.mat;
    if (m->stddev_map == 0u) return mk(m->specular_power[0], m->specular_power[1], 1.0f);
    V4f c = tex_sample(tv, m->stddev_map, s.tcx, s.tcy, TEX_RG);
    return mk(c.x, c.y, 1.0f);
}
TR_HD V3 sggx_surface_stddev(const Sample& s, const NoTex&)
{
    return mk(s.mat->specular_power[0], s.mat->specular_power[1], 1.0f);
}
TR_HD V4f sggx_surface_quat(const Sample& s, const NoTex&)
{
    V4f q = {s.mat->sggx_basis[0], s.mat->sggx_basis[1], s.mat->sggx_basis[2], s.mat->sggx_basis[3]};
    return q;
}
TR_HD V4f sggx_surface_quat(const Sample& s, const TexView& tv)
{
    const tr_material* m = s.mat;
    if (m->basis_map == 0u) {
        V4f q = {m->sggx_basis[0], m->sggx_basis[1], m->sggx_basis[2], m->sggx_basis[3]};
        return q;
    }
    return tex_sample(tv, m->basis_map, s.tcx, s.tcy, TEX_RGBA);
}

/* SGGXMicroFlakeSurfaceDensity via SGGXSurfaceCache, ggx-models-impl.hh:589-619
 * (the quaternion is used as it is, not normalised) */
template <class TV> TR_HD float sggx_surface_density(const Sample& s, const TV& tv)
{
    M3 orient = m3(s.tangent, s.binormal, s.normal);
    V4f q = sggx_surface_quat(s, tv);
    M3 basis = mul(orient, quat_to_m3(q.x, q.y, q.z, q.w));
    return sggx_projected_area(sggx_surface_stddev(s, tv), basis, s.wi);
}

/* SGGXMicroFlakeSurfaceBRDF, ggx-models-impl.hh:476-551 */
template <class TV> TR_HD V3 sggx_surface_brdf(const Sample& s, const TV& tv)
{
    const tr_material* m = s.mat;
    float din = dot(s.normal, s.wi);
    float don = dot(s.normal, s.wo);
    if (din < 0.0f || don < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    V3 result = muls(muls(map_spectrum(tv, m->diffuse_map, m->diffuse, s), din), 1.0f / kPi);
    V3 stddev = sggx_surface_stddev(s, tv);
    if (stddev.x >= 1e-6f && stddev.y >= 1e-6f && stddev.z >= 1e-6f && s.dir_density >= 1e-6f) {
        V3 mn = normalize(add(s.wi, s.wo));
        M3 orient = m3(s.tangent, s.binormal, s.normal);
        V4f q = sggx_surface_quat(s, tv);
        float qx = q.x, qy = q.y, qz = q.z, qw = q.w;
        float ql = sqrtf(qx * qx + qy * qy + qz * qz + qw * qw);
        if (ql == 0.0f) { qx = qy = qz = qw = 0.0f; }
        else { qx = qx / ql; qy = qy / ql; qz = qz / ql; qw = qw / ql; }
        M3 basis = mul(orient, quat_to_m3(qx, qy, qz, qw));
        float cos_micro = dot(mn, s.wi);
        float fresnel;
        float eta = m->fresnel[0];
        if (eta == 1.0f) {
            fresnel = 1.0f;
        } else {
            float ct = cos_transmittance(eta, cos_micro);
            fresnel = fresnel_dielectric(eta, cos_micro, ct);
            fresnel = fresnel < 0.0f ? 0.0f : (fresnel > 1.0f ? 1.0f : fresnel);
        }
        float in_density = sggx_projected_area(stddev, basis, s.wi);
        float contrib = 1.0f / (s.dir_density + in_density * don / din);
        V3 spec = map_spectrum(tv, m->specular_map, m->specular, s);
        result = add(result, muls(muls(muls(muls(smul(contrib, spec), sggx_ndf(stddev, basis, mn)), fresnel), 2.0f), 0.25f));
    }
    return result;
}

/* ---------------- SGGX micro-flake volume phase function ---------------- */

/* Matrix3::transpose and choleskyDecomposition (matrix3.hh:92-97, 173-191). */
TR_HD M3 transpose(const M3& m)
{
    return m3(mk(m.c0.x, m.c1.x, m.c2.x), mk(m.c0.y, m.c1.y, m.c2.y), mk(m.c0.z, m.c1.z, m.c2.z));
}
TR_HD M3 cholesky(const M3& m)
{
    float l00 = sqrtf(m.c0.x);
    float rcp_l00 = 1.0f / l00;
    float l10 = rcp_l00 * m.c0.y;
    float l20 = rcp_l00 * m.c0.z;
    float l11 = sqrtf(m.c1.y - l10 * l10);
    float rcp_l11 = 1.0f / l11;
    float l21 = rcp_l11 * (m.c1.z - l20 * l10);
    float l22 = sqrtf(m.c2.z - l21 * l21 - l20 * l20);
    return m3(mk(l00, l10, l20), mk(0.0f, l11, l21), mk(0.0f, 0.0f, l22));
}

TR_HD M3 frame_of(const Sample& s) { return m3(s.tangent, s.binormal, s.normal); }

/* SGGXMicroFlakeBRDF (ggx-models-impl.hh:60-68, 466-474): the phase
 * function 2 D(h) / (4 sigma(wo)), h = normalize(wi + wo). */
TR_HD float sggx_flake_phase(const Sample& s)
{
    V3 mn = normalize(add(s.wi, s.wo));
    return sggx_ndf(ld3(s.mat->sggx_stddev), frame_of(s), mn) * 2.0f / (4.0f * s.dir_density);
}

/* SGGXMicroFlakePDF (ggx-models-impl.hh:447-464): the same expression,
 * evaluated as 2 * D * (1 / (4 sigma)). */
TR_HD float sggx_flake_pdf_h(float dir_density, V3 stddev, const M3& basis, V3 mn)
{
    float ndf = sggx_ndf(stddev, basis, mn);
    float proj_area = 1.0f / (4.0f * dir_density);
    return 2.0f * ndf * proj_area;
}
TR_HD float sggx_flake_pdf(const Sample& s)
{
    return sggx_flake_pdf_h(s.dir_density, ld3(s.mat->sggx_stddev), frame_of(s), normalize(add(s.wi, s.wo)));
}

/* The visible flake normal of SGGXMicroFlakeSampleIncidentLight
 * (ggx-models-impl.hh:554-570): a uniform sphere sample (UniformSampleSphere,
 * sampling3.hh:106-116) mapped through the Cholesky factor of the covariance
 * S expressed in the frame of wo. */
TR_HD V3 sggx_visible_normal(V3 wo, const M3& S, float ra, float rb)
{
    float sin_phi, cos_phi;
    tr_sincos_turns(ra, &sin_phi, &cos_phi);
    float cos_theta = 2.0f * rb - 1.0f;
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    V3 u = mk(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta);
    M3 ob = make_basis(wo);
    M3 So = mul(mul(ob, S), transpose(ob));
    V3 vn = normalize(transform_rotation_inverse(cholesky(So), u));
    return transform(ob, vn);
}

/* SGGXMicroFlakeSampleIncidentLight (ggx-models-impl.hh:554-580): wo
 * reflected about the visible flake normal. */
TR_HD void sggx_flake_sample(Sample* s, float ra, float rb)
{
    V3 sd = ld3(s->mat->sggx_stddev);
    M3 basis = frame_of(*s);
    M3 S = transform_covariance(basis, mulv(sd, sd));
    s->wi = reflect(s->wo, sggx_visible_normal(s->wo, S, ra, rb));
    s->pdf = sggx_flake_pdf_h(s->dir_density, sd, basis, normalize(add(s->wi, s->wo)));
}

/* ---------------- SGGX pseudo-volume surface (model 25) ---------------- */

/* FloatIntelFastRand over IntelFastRand (math/functions.hh:214-218,
 * 249-252): the reference's 15-bit LCG.  The pseudo-volume BRDF seeds it
 * from the TexCoord, so its sequence is a function of the hit point and is
 * reproduced exactly. */
struct Lcg15 {
    uint32_t seed;
    TR_HD float next()
    {
        seed = 214013u * seed + 2531011u;
        return (float)((seed >> 16) & 0x7FFFu) / 32767.0f;
    }
};

/* (uint32_t)x of the seed formula: negative and NaN give 0, values past
 * 2^32 - 1 saturate (the C++ conversion is undefined there). */
TR_HD uint32_t tc_to_u32(float x) { return x > 0.0f ? (x < 4294967296.0f ? (uint32_t)x : 0xFFFFFFFFu) : 0u; }

/* sggx_surface_orientation * ToMatrix3(quat) with the quaternion as given
 * (ggx-models-impl.hh:646-659, 780-792, 817-828). */
template <class TV> TR_HD M3 sggx_pseudo_basis(const Sample& s, const TV& tv)
{
    V4f q = sggx_surface_quat(s, tv);
    return mul(frame_of(s), quat_to_m3(q.x, q.y, q.z, q.w));
}

/* SGGXMicroFlakePseudoVolumeBRDF, ggx-models-impl.hh:621-758: single
 * scattering of the surface's flakes plus SampleCount random walks of up to
 * Depth flake scatterings below the surface, each scattering attenuated by
 * the extinction back along the incident direction.  Specular and Diffuse
 * are the material's constants here (the reference reads no maps for
 * them); StandardDeviation and the basis come from their maps. */
template <class TV> TR_HD V3 sggx_pseudo_volume_brdf(const Sample& s, const TV& tv)
{
    const tr_material* m = s.mat;
    float din = dot(s.normal, s.wi);
    float don = dot(s.normal, s.wo);
    if (don < 0.0f || din < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    V3 sd = sggx_surface_stddev(s, tv);
    if (sd.x < 1e-6f || sd.y < 1e-6f) return mk(0.0f, 0.0f, 0.0f);
    M3 basis = sggx_pseudo_basis(s, tv);
    Lcg15 rng{(tc_to_u32(s.tcy * 65535.0f) << 16) + tc_to_u32(s.tcx * 65535.0f)};
    rng.next();
    const uint32_t depth = m->sggx_depth, count = m->sggx_sample_count;
    const float start = s.dir_density;
    const float in_density = sggx_projected_area(sd, basis, s.wi);
    const V3 spec = ld3(m->specular);
    V3 mn = normalize(add(s.wo, s.wi));
    float contrib = (float)count / (start + in_density * don / din);
    V3 radiance = muls(muls(muls(smul(contrib, spec), sggx_ndf(sd, basis, mn)), 2.0f), 0.25f);
    const M3 S = transform_covariance(basis, mulv(sd, sd));
    for (uint32_t k = 0; k < count; ++k) {
        V3 thr = spec;
        V3 out = s.wo;
        float pen = (-tr_logf(rng.next()) / start) * don;
        for (uint32_t it = 0; thr.x > 1e-3f && thr.y > 1e-3f && thr.z > 1e-3f && it < depth; ++it) {
            float ra = rng.next();
            float rb = rng.next();
            V3 inc = reflect(out, sggx_visible_normal(out, S, ra, rb));
            float dd = sggx_projected_area(sd, basis, inc);
            out = neg(inc);
            pen += (-tr_logf(rng.next()) / dd) * dot(s.normal, out);
            if (pen < 0.0f) break;
            float ext = tr_expf(-in_density * (pen / din));
            if (ext < 1e-9f) break;
            V3 mn2 = normalize(add(out, s.wi));
            thr = mulv(thr, smul(1.0f, spec));
            float phase = sggx_ndf(sd, basis, mn2) * 2.0f / (4.0f * dd);
            radiance = add(radiance, smul(phase, smul(ext, thr)));
        }
    }
    return add(divs(radiance, (float)count), muls(muls(ld3(m->diffuse), din), 1.0f / kPi));
}

/* SGGXMicroFlakeSurfacePDF, ggx-models-impl.hh:760-794 */
template <class TV> TR_HD float sggx_pseudo_volume_pdf(const Sample& s, const TV& tv)
{
    if (dot(s.wi, s.normal) < 0.0f || dot(s.wo, s.normal) < 0.0f) return 0.0f;
    V3 mn = normalize(add(s.wi, s.wo));
    return 2.0f * sggx_flake_pdf_h(s.dir_density, sggx_surface_stddev(s, tv), sggx_pseudo_basis(s, tv), mn);
}

/* SGGXMicroFlakeSurfaceSampleIncidentLight, ggx-models-impl.hh:796-842,
 * with its flip of every flake normal not exactly orthogonal to the normal
 * (`if(Dot(micro_norm, Normal))`, :836; the reflection does not depend on
 * the sign). */
template <class TV> TR_HD void sggx_pseudo_volume_sample(Sample* s, float ra, float rb, const TV& tv)
{
    V3 sd = sggx_surface_stddev(*s, tv);
    M3 basis = sggx_pseudo_basis(*s, tv);
    M3 S = transform_covariance(basis, mulv(sd, sd));
    V3 mn = sggx_visible_normal(s->wo, S, ra, rb);
    if (dot(mn, s->normal) != 0.0f) mn = neg(mn);
    s->wi = reflect(s->wo, mn);
    s->pdf = dot(s->wi, s->normal) > 0.0f ? 2.0f * sggx_flake_pdf_h(s->dir_density, sd, basis, mn) : 0.0f;
}

TR_HD bool is_emissive(uint32_t m) { return m == TR_MODEL_EMISSIVE; }
TR_HD bool is_volume_model(uint32_t m) { return m == TR_MODEL_SGGX_MICRO_FLAKE; }

/* TransmittanceLookup / PDFLookup / SampleIncidentLightLookup /
 * MaterialCacheLookup (illumination-models.hh:233-492) for the supported
 * models: Blinn-Phong, GGX and Beckmann (plain, conductor, dielectric;
 * isotropic and anisotropic), SGGX surface and pseudo-volume, SGGX
 * micro-flake volume. */
template <class TV> TR_HD V3 bsdf_eval(const Sample& s, const TV& tv)
{
    uint32_t m = s.mat->model;
    if (is_microfacet(m)) return microfacet_brdf(s);
    if (is_aniso(m)) return microfacet_aniso_brdf(s);
    if (m == TR_MODEL_BLINN_PHONG) return blinn_phong_brdf(s, tv);
    if (m == TR_MODEL_SGGX_SURFACE) return sggx_surface_brdf(s, tv);
    if constexpr (FullModels<TV>::value) {
        if (m == TR_MODEL_SGGX_PSEUDO_VOLUME) return sggx_pseudo_volume_brdf(s, tv);
    }
    if (m == TR_MODEL_SGGX_MICRO_FLAKE) {
        float p = sggx_flake_phase(s);
        return mk(p, p, p);
    }
    return mk(0.0f, 0.0f, 0.0f);
}
template <class TV> TR_HD float bsdf_pdf(const Sample& s, const TV& tv)
{
    uint32_t m = s.mat->model;
    if (is_microfacet(m)) return microfacet_pdf(s);
    if (is_aniso(m)) return microfacet_aniso_pdf(s);
    if (m == TR_MODEL_BLINN_PHONG) return blinn_phong_pdf(s, tv);
    if (m == TR_MODEL_SGGX_SURFACE) return uniform_hemisphere_pdf(s);
    if constexpr (FullModels<TV>::value) {
        if (m == TR_MODEL_SGGX_PSEUDO_VOLUME) return sggx_pseudo_volume_pdf(s, tv);
    }
    if (m == TR_MODEL_SGGX_MICRO_FLAKE) return sggx_flake_pdf(s);
    return 0.0f;
}
/* xi0..xi2: the vertex's BSDF draw; the anisotropic microfacet models take
 * their second stratified pair from the same stream's sub-counter 1. */
template <class TV>
TR_HD void bsdf_sample(const Strata& st, Sample* s, float xi0, float xi1, float xi2, const TV& tv, const Rng& rng,
                       uint32_t vertex)
{
    uint32_t m = s->mat->model;
    if (is_microfacet(m)) {
        microfacet_sample(st, s, xi0, xi1, xi2);
    } else if (is_aniso(m)) {
        U4 x2 = rng.draw(vertex, STREAM_BSDF, 1u);
        microfacet_aniso_sample(st, s, xi0, xi1, xi2, u01(x2.a), u01(x2.b));
    }
    else if (m == TR_MODEL_BLINN_PHONG) blinn_phong_sample(st, s, xi0, xi1, xi2, tv);
    else if (m == TR_MODEL_SGGX_SURFACE) uniform_hemisphere_sample(st, s, xi0, xi1);
    else if (FullModels<TV>::value && m == TR_MODEL_SGGX_PSEUDO_VOLUME) sggx_pseudo_volume_sample(s, xi0, xi1, tv);
    else if (m == TR_MODEL_SGGX_MICRO_FLAKE) sggx_flake_sample(s, xi0, xi1);
    else s->pdf = 0.0f;
}
template <class TV> TR_HD void material_cache(Sample* s, const TV& tv)
{
    if (s->mat->model == TR_MODEL_SGGX_SURFACE || s->mat->model == TR_MODEL_SGGX_PSEUDO_VOLUME)
        s->dir_density = sggx_surface_density(*s, tv);
}

/* ---------------- geometry sampling ---------------- */

/* IntersectSphere's normal (intersect.hh:181-182) */
TR_HD V3 sphere_normal(V3 org, V3 dir, const float* c, float t)
{
    V3 dc = sub(ld3(c), org);
    return normalize(sub(smul(t, dir), dc));
}

/* The sample() of the user shapes: the geometric normal is the one the
 * shape's intersection wrote (recomputed from the same ray, so the same
 * bits).
 *  - RectGeometry::sample (ray-tracing.cc:1246-1272): TexCoord = start +
 *    (u, v) * multiplier; the frame of ToMatrix3(Orientation), or with a
 *    tangent map Orientation * makeBasisTangent(normalized sampleRGB).
 *  - DiskGeometry::sample (:1274-1283): Binormal towards the centre.
 *  - CylinderGeometry::sample (:1142-1164): on a cap the tangent points
 *    away from the axis, on the side it is +z; Binormal = Normal x Tangent.
 *    Oblique cylinders take the same frame around their own normal (their
 *    reference sample() leaves it zero, :1210-1237).
 *  - EllipsoidGeometry::sample (:1182-1208): Binormal = ToNormal(q) x N,
 *    or makeBasis(N) where the two are parallel. */
template <bool TC>
TR_HD void shape_sample(const SceneView& sc, const ShapeRec& sh, float t, float u, float v, V3 org, V3 dir, Sample* sd)
{
    float st, su, sv;
    V3 n = mk(0.0f, 0.0f, 1.0f);
    shape_intersect(sh, org, dir, &st, &su, &sv, &n);
    sd->tcx = 0.0f;
    sd->tcy = 0.0f;
    const float* f = sh.f;
    if (sh.kind == SHAPE_RECT) {
        sd->tcx = f[14] + u * f[16];
        sd->tcy = f[15] + v * f[17];
        M3 o = m3(mk(f[3], f[4], f[5]), mk(f[6], f[7], f[8]), mk(f[9], f[10], f[11]));
        if (TC && sh.tangent_map) {
            V4f c = tex_sample(sc.tex, sh.tangent_map, sd->tcx, sd->tcy, TEX_RGB);
            o = mul(o, make_basis_tangent(normalize(mk(c.x, c.y, c.z))));
        }
        sd->tangent = o.c0;
        sd->binormal = o.c1;
        sd->normal = o.c2;
    } else if (sh.kind == SHAPE_DISK) {
        sd->binormal = normalize(sub(mk(f[0], f[1], f[2]), add(smul(t, dir), org)));
        sd->tangent = cross(sd->binormal, n);
        sd->normal = n;
    } else if (sh.kind == SHAPE_ELLIPSOID) {
        const V3 qn = mk(f[14], f[15], f[16]);
        sd->normal = n;
        if (fabsf(dot(qn, n)) < 1e-6f) {
            const M3 b = make_basis(n);
            sd->binormal = b.c1;
            sd->tangent = b.c0;
        } else {
            sd->binormal = normalize(cross(qn, n));
            sd->tangent = cross(sd->binormal, sd->normal);
        }
    } else {
        sd->normal = n;
        if (fabsf(n.z) == 1.0f) {
            float px = org.x + t * dir.x, py = org.y + t * dir.y;
            float tx = px - f[0], ty = py - f[1];
            const float l = sqrtf(tx * tx + ty * ty);
            if (l != 0.0f) { tx = tx / l; ty = ty / l; }
            sd->tangent = mk(tx, ty, 0.0f);
        } else {
            sd->tangent = mk(0.0f, 0.0f, 1.0f);
        }
        sd->binormal = cross(n, sd->tangent);
    }
}

/* RTMesh::sample + DefaultMeshSample (ray-tracing.cc:375-461),
 * SphereGeometry::sample (:1166-1180), SphereAreaLightGeometry::sample
 * (:1775-1783).  Tangent frames of vertex formats without tangents are
 * zero (the reference leaves the previous values in place). */
template <bool TC = true>
TR_HD void sample_geometry(const SceneView& sc, uint32_t geom, uint32_t prim, float t, float u, float v, V3 org, V3 dir,
                           Sample* sd)
{
    const GeomRec& g = sc.geoms[geom];
    sd->mat = &sc.mats[g.material];
    if (g.type == GEOM_MESH) {
        const int32_t* ib = sc.ipool + g.index_base + (size_t)prim * 3;
        const uint8_t* base = sc.vpool + g.vbase;
        const uint8_t* p0 = base + (size_t)ib[0] * g.stride;
        const uint8_t* p1 = base + (size_t)ib[1] * g.stride;
        const uint8_t* p2 = base + (size_t)ib[2] * g.stride;
        uint32_t no = g.stride - 12;
        float w = (1.0f - u) - v;
        V3 n0 = ld3(reinterpret_cast<const float*>(p0 + no));
        V3 n1 = ld3(reinterpret_cast<const float*>(p1 + no));
        V3 n2 = ld3(reinterpret_cast<const float*>(p2 + no));
        V3 norm = normalize(add(add(muls(n0, w), muls(n1, u)), muls(n2, v)));
        sd->tangent = mk(0.0f, 0.0f, 0.0f);
        sd->binormal = mk(0.0f, 0.0f, 0.0f);
        sd->tcx = 0.0f;
        sd->tcy = 0.0f;
        if (TC && g.stride >= 32) {
            /* TexCoord = tc0 * w + tc1 * u + tc2 * v (PTNFormat and wider) */
            const float* t0 = reinterpret_cast<const float*>(p0 + 12);
            const float* t1 = reinterpret_cast<const float*>(p1 + 12);
            const float* t2 = reinterpret_cast<const float*>(p2 + 12);
            sd->tcx = (t0[0] * w + t1[0] * u) + t2[0] * v;
            sd->tcy = (t0[1] * w + t1[1] * u) + t2[1] * v;
        }
        if (g.stride >= 44) {
            V3 t0 = ld3(reinterpret_cast<const float*>(p0 + 20));
            V3 t1 = ld3(reinterpret_cast<const float*>(p1 + 20));
            V3 t2 = ld3(reinterpret_cast<const float*>(p2 + 20));
            V3 tan = add(add(muls(t0, w), muls(t1, u)), muls(t2, v));
            M3 b = make_basis_orthogonalize(tan, norm);
            sd->tangent = b.c0;
            sd->binormal = b.c1;
        }
        sd->normal = norm;
    } else if (g.type == GEOM_SHAPE) {
        shape_sample<TC>(sc, sc.shapes[g.index_base], t, u, v, org, dir, sd);
    } else if (g.type == GEOM_SPHERE) {
        V3 n = sphere_normal(org, dir, g.sphere, t);
        sd->normal = n;
        sd->tcx = 0.0f;
        sd->tcy = 0.0f;
        sd->binormal = normalize(cross(mk(0.0f, 0.0f, 1.0f), n));
        sd->tangent = cross(sd->binormal, sd->normal);
    } else {
        sd->normal = normalize(sub(add(org, muls(dir, t)), ld3(g.sphere)));
        sd->tangent = mk(0.0f, 0.0f, 0.0f);
        sd->binormal = mk(0.0f, 0.0f, 0.0f);
        sd->tcx = 0.0f;
        sd->tcy = 0.0f;
    }
}

/* ---------------- lights ---------------- */

struct LightSample {
    uint32_t geom;
    V3 dir;
    float dist;
    V3 radiance;
    float pdf;
};

/* TriangleArea of projected vertices, sampling3.hh:206-212 */
TR_HD float projected_triangle_area(V3 pos, V3 p0, V3 p1, V3 p2)
{
    V3 a = normalize(sub(p0, pos)), b = normalize(sub(p1, pos)), c = normalize(sub(p2, pos));
    return 0.5f * length(cross(sub(b, a), sub(c, a)));
}

TR_HD V3 light_vertex(const SceneView& sc, const GeomRec& g, uint32_t tri, int k)
{
    int32_t idx = sc.ipool[g.index_base + (size_t)tri * 3 + k];
    return ld3(reinterpret_cast<const float*>(sc.vpool + g.vbase + (size_t)idx * g.stride));
}

/* Unbiased estimator: solid-angle density of area-uniform sampling of a
 * mesh-light triangle with n = e0 x e1 (twice its area), seen along dir at
 * distance dist: dist^2 / (cos * area) / triangles; 0 from behind, by the
 * reference's facing test (sampling3.hh:220-226). */
TR_HD float mesh_light_pdf_u(V3 n, V3 dir, float dist, uint32_t tri_count)
{
    float c = dot(n, dir);
    if (!(c < -1e-6f)) return 0.0f;
    return (dist * dist) / (-0.5f * c) / (float)tri_count;
}

/* ComputeSphereAreaPDF (sampling3.hh:229-262) */
TR_HD float sphere_light_pdf(V3 center, float radius, V3 pos, V3 dir)
{
    V3 dc = sub(center, pos);
    float dist = sqrtf(dot(dc, dc));
    if (dist <= radius) return 1.0f / (4.0f * kPi);
    V3 to_c = divs(dc, dist);
    float sin_rc = radius / dist;
    float cos_rc = sqrtf(1.0f - sin_rc * sin_rc);
    float cos_theta = dot(dir, to_c);
    if (cos_theta < cos_rc) return 0.0f;
    return 1.0f / (2.0f * kPi * (1.0f - cos_rc));
}

/* Lights sampled SamplesLocalAreaLight times: IsAreaLight (sphere lights,
 * ray-tracing-common.hh:546-551) in the reference estimator; sphere and
 * mesh lights in the unbiased one. */
TR_HD bool multi_sample_light(const LightRec& L, bool unbiased)
{
    return L.area || (unbiased && L.src.type == TR_LIGHT_MESH_AREA);
}

/* SampleLightSourceFunctionLookup (ray-tracing-common.hh:567-647) with
 * SampleSphereArea / UniformSampleSphereArea / UniformSampleTriangleArea
 * (sampling3.hh:142-227).  x0: triangle pick, x1/x2: the two coordinates.
 * The unbiased estimator reports the mesh-light triangle's true
 * solid-angle density. */
TR_HD LightSample sample_light(const SceneView& sc, const LightRec& L, V3 pos, uint32_t x0, uint32_t x1, uint32_t x2,
                               bool unbiased = false, const float* dir_override = nullptr)
{
    LightSample r;
    const tr_light& l = L.src;
    float ra = u01(x1), rb = u01(x2);
    if (l.type == TR_LIGHT_DIRECTIONAL) {
        r.dist = INFINITY;
        r.dir = ld3(dir_override ? dir_override : l.direction);
        r.radiance = ld3(l.radiance);
        r.pdf = 1.0f;
        r.geom = TR_INVALID_GEOMETRY;
    } else if (l.type == TR_LIGHT_POINT) {
        V3 dv = sub(ld3(l.position), pos);
        float dist = length(dv);
        r.dist = INFINITY;
        r.dir = divs(dv, dist);
        r.radiance = divs(ld3(l.radiance), kTau * dist * dist);
        r.pdf = 1.0f;
        r.geom = TR_INVALID_GEOMETRY;
    } else if (l.type == TR_LIGHT_SPHERE_AREA) {
        V3 c = ld3(l.sphere_center);
        float radius = l.sphere_radius;
        V3 dc = sub(c, pos);
        float dist2 = dot(dc, dc);
        float dist = sqrtf(dist2);
        r.radiance = ld3(l.radiance);
        r.geom = L.geom;
        if (dist <= radius) {
            float sin_phi, cos_phi;
            tr_sincos_turns(rb, &sin_phi, &cos_phi);
            float r2 = 2.0f * radius;
            float q = sqrtf(ra * (1.0f - ra));
            V3 res = add(sub(c, pos), mk(r2 * cos_phi * q, r2 * sin_phi * q, radius * (1.0f - 2.0f * ra)));
            float t = length(res);
            r.dir = divs(res, t);
            r.dist = t;
            r.pdf = 1.0f / (4.0f * kPi);
            return r;
        }
        V3 to_c = divs(dc, dist);
        float sin_rc = radius / dist;
        float cos_rc = sqrtf(1.0f - sin_rc * sin_rc);
        float sin_phi, cos_phi;
        tr_sincos_turns(ra, &sin_phi, &cos_phi);
        float cos_theta = 1.0f - rb + rb * cos_rc;
        float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        M3 basis = make_basis(to_c);
        V3 d = normalize(transform(basis, mk(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta)));
        float b = dist / cos_theta;
        float cc = dist2 - radius * radius;
        float D = b * b - cc;
        float t = b;
        if (D >= 0.0f) t -= sqrtf(D);
        r.dir = d;
        r.dist = t;
        r.pdf = 1.0f / (2.0f * kPi * (1.0f - cos_rc));
    } else {
        const GeomRec& g = sc.geoms[L.geom];
        uint32_t tri = mulhi32(x0, g.tri_count);
        V3 p0 = light_vertex(sc, g, tri, 0), p1 = light_vertex(sc, g, tri, 1), p2 = light_vertex(sc, g, tri, 2);
        float sqrt_r = sqrtf(ra);
        V3 e0 = sub(p1, p0), e1 = sub(p2, p0);
        V3 point = add(add(p0, smul(sqrt_r * rb, e0)), smul(1.0f - sqrt_r, e1));
        V3 du = sub(point, pos);
        float len = length(du);
        r.dir = divs(du, len);
        r.dist = len;
        if (unbiased) {
            r.pdf = mesh_light_pdf_u(cross(e0, e1), r.dir, r.dist, g.tri_count);
        } else {
            float pdf = dot(cross(e0, e1), r.dir) < -1e-6f ? 1.0f / projected_triangle_area(pos, p0, p1, p2) : 0.0f;
            r.pdf = pdf / (float)g.tri_count;
        }
        r.radiance = ld3(sc.mats[g.material].diffuse);
        r.geom = L.geom;
    }
    return r;
}

/* LightSourceComputePDFLookup (ray-tracing-common.hh:582-687) */
TR_HD float light_pdf(const SceneView& sc, const LightRec& L, V3 pos, V3 dir)
{
    const tr_light& l = L.src;
    if (l.type == TR_LIGHT_DIRECTIONAL || l.type == TR_LIGHT_POINT) return 1.0f;
    if (l.type == TR_LIGHT_SPHERE_AREA) return sphere_light_pdf(ld3(l.sphere_center), l.sphere_radius, pos, dir);
    const GeomRec& g = sc.geoms[L.geom];
    float total = 0.0f;
    for (uint32_t tri = 0; tri < g.tri_count; ++tri) {
        V3 p0 = light_vertex(sc, g, tri, 0), p1 = light_vertex(sc, g, tri, 1), p2 = light_vertex(sc, g, tri, 2);
        V3 un = cross(sub(p2, p1), sub(p0, p1));
        if (dot(dir, un) >= 0.0f) continue;
        total = total + projected_triangle_area(pos, p0, p1, p2);
    }
    return total != 0.0f ? 1.0f / total : 0.0f;
}

/* Unbiased estimator: density with which the light of an emitter hit
 * would have produced the hit point from `pos` (ray origin) along dir at
 * distance t; 0 for emitters that are not lights. */
TR_HD float hit_light_pdf(const SceneView& sc, uint32_t geom, uint32_t prim, V3 pos, V3 dir, float t)
{
    const GeomRec& g = sc.geoms[geom];
    if (g.type == GEOM_SPHERE_LIGHT) return sphere_light_pdf(mk(g.sphere[0], g.sphere[1], g.sphere[2]), g.sphere[3], pos, dir);
    if (g.type != GEOM_MESH) return 0.0f;
    V3 p0 = light_vertex(sc, g, prim, 0), p1 = light_vertex(sc, g, prim, 1), p2 = light_vertex(sc, g, prim, 2);
    return mesh_light_pdf_u(cross(sub(p1, p0), sub(p2, p0)), dir, t, g.tri_count);
}

/* EvaluateTotalBalanceHeuristic (ray-tracing.cc:1791-1806) */
template <class TV>
TR_HD float total_balance(const SceneView& sc, float light_samples, float surface_samples, const Sample& s, V3 pos,
                          const TV& tv, bool lights = true)
{
    float total = 0.0f;
    const uint32_t n_lights = lights ? sc.n_lights : 0u;
    for (uint32_t i = 0; i < n_lights; ++i) {
        const LightRec& L = sc.lights[i];
        float n = L.area ? light_samples : 1.0f;
        total = total + light_pdf(sc, L, pos, s.wi) * n;
    }
    total = total + bsdf_pdf(s, tv) * surface_samples;
    return total;
}

}  // namespace trk

#endif
