/*
 * orc_bsdf.h - BSDF evaluation, PDFs and sampling of the oracle.
 *
 * ORACLE: test infrastructure only.  Each function restates the reference
 * function named in its comment, in the reference's operation order.
 * Deviations, applied identically in the GPU kernels:
 *  - random numbers come from Philox (three per surface sample: the two
 *    stratified coordinates and the lobe selector, in the reference's draw
 *    order; the anisotropic microfacet models take their specular pair from
 *    a second draw) instead of FastFloatRand;
 *  - sin/cos of 2*pi*u use tr_sincos_turns instead of FastSinCos
 *    (functions.hh:322-353), logf uses tr_logf, expf uses tr_expf,
 *    pow(x, 5) is x^2*x^2*x.
 */
#ifndef ORC_BSDF_H
#define ORC_BSDF_H

#include "orc_math.h"
#include "orc_texture.h"
#include "../include/tempest_rt.h"

namespace orc {

/* SampleData (ray-tracing-common.hh:184-198). */
template <class T> struct Sample {
    const tr_material* mat;
    V3<T> wo;       /* OutgoingLight */
    V3<T> wi;       /* IncidentLight */
    V3<T> tangent;
    V3<T> binormal;
    V3<T> normal;
    V3<T> position;
    T total_density;
    T dir_density;  /* DirectionalDensity */
    T pdf;
    T tcx = 0, tcy = 0;             /* TexCoord */
    const OrcTexSet* tex = nullptr; /* the scene's textures (none in unit calls) */
};

/* A spectrum map at the TexCoord (SampleSpectrum) or the constant. */
template <class T> inline V3<T> map_spectrum(const Sample<T>& s, uint32_t map, const float* constant)
{
    if (map == 0 || !s.tex) return v3((T)constant[0], (T)constant[1], (T)constant[2]);
    Rgba c = orc_tex_sample(*s.tex, map, (float)s.tcx, (float)s.tcy, ORC_TEX_SPECTRUM);
    return v3((T)c.v[0], (T)c.v[1], (T)c.v[2]);
}

template <class T> struct Strata {
    T x, y, tx, ty;
};

template <class T> inline V3<T> spec3(const float* f) { return v3((T)f[0], (T)f[1], (T)f[2]); }

/* SpecularPowerToGeometryFactorSq, microfacet-common.hh:109-112 */
template <class T> inline T geom_factor_sq(T power) { return (T)2 / (power + (T)2); }

/* FresnelSchlick, functions.hh:262-265 */
template <class T> inline T fresnel_schlick(T refl, T cos_theta) { return refl + ((T)1 - refl) * pow5((T)1 - cos_theta); }

/* CosTransmittance, functions.hh:268-271 */
template <class T> inline T cos_transmittance(T eta, T cos_inc)
{
    return xsqrt((T)1 - ((T)1 - cos_inc * cos_inc) / (eta * eta));
}

/* FresnelConductor, functions.hh:275-294 */
template <class T> inline T fresnel_conductor(T eta, T k, T cos_inc)
{
    T cos_inc2 = cos_inc * cos_inc;
    T sin_inc2 = (T)1 - cos_inc2;
    T eta2 = eta * eta;
    T etak2 = k * k;
    T t0 = eta2 - etak2 - sin_inc2;
    T a2plusb2 = xsqrt(t0 * t0 + (T)4 * eta2 * etak2);
    T t1 = a2plusb2 + cos_inc2;
    T a = xsqrt((T)0.5 * (a2plusb2 + t0));
    T t2 = (T)2 * a * cos_inc;
    T r_perp = (t1 - t2) / (t1 + t2);
    T t3 = cos_inc2 * a2plusb2 + sin_inc2 * sin_inc2;
    T t4 = t2 * sin_inc2;
    T r_parallel = r_perp * (t3 - t4) / (t3 + t4);
    return (T)0.5 * (r_perp + r_parallel);
}

/* FresnelDielectric, functions.hh:296-301 */
template <class T> inline T fresnel_dielectric(T eta, T cos_inc, T cos_trans)
{
    T r_parallel = (eta * cos_inc - cos_trans) / (eta * cos_inc + cos_trans);
    T r_perp = (cos_inc - eta * cos_trans) / (cos_inc + eta * cos_trans);
    return (T)0.5 * (r_parallel * r_parallel + r_perp * r_perp);
}

template <class T> struct Angles {
    T cos_in, cos_on, cos_im, cos_mn;
};

/* ComputeMicrofacetAngles, microfacet-common.hh:74-84 */
template <class T> inline Angles<T> microfacet_angles(const Sample<T>& s)
{
    V3<T> h = normalize(s.wi + s.wo);
    Angles<T> a;
    a.cos_in = dot(s.wi, s.normal);
    a.cos_on = dot(s.wo, s.normal);
    a.cos_im = dot(s.wi, h);
    a.cos_mn = dot(s.normal, h);
    return a;
}

/* GGXMicrofacetDistribution, ggx-models-impl.hh:126-139 */
template <class T> inline T ggx_d(T g2, T cos_nh)
{
    if (cos_nh <= (T)0) return (T)0;
    T c2 = cos_nh * cos_nh;
    T c4 = c2 * c2;
    T tan2 = ((T)1 - c2) / c2;
    T den = g2 + tan2;
    return g2 / (C<T>::pi * c4 * den * den);
}

/* G1_GGX, ggx-models-impl.hh:141-151 */
template <class T> inline T ggx_g1(T g2, T cos_vm, T cos_vn)
{
    T ratio = cos_vm / cos_vn;
    if (ratio <= (T)0) return (T)0;
    T c2 = cos_vn * cos_vn;
    T tan2 = ((T)1 - c2) / c2;
    return (T)2 / ((T)1 + xsqrt((T)1 + g2 * tan2));
}

/* BeckmannMicrofacetDistribution, beckmann-models-impl.hh:57-70 */
template <class T> inline T beckmann_d(T g2, T cos_nh)
{
    if (cos_nh <= (T)0) return (T)0;
    T c2 = cos_nh * cos_nh;
    T c4 = c2 * c2;
    T tan2 = ((T)1 - c2) / c2;
    return xexp(-tan2 / g2) / (C<T>::pi * g2 * c4);
}

/* G1_Beckmann (rational approximation), beckmann-models-impl.hh:34-55 */
template <class T> inline T beckmann_g1(T g2, T cos_vm, T cos_vn)
{
    T ratio = cos_vm / cos_vn;
    if (ratio <= (T)0) return (T)0;
    T c2 = cos_vn * cos_vn;
    T tan2 = ((T)1 - c2) / c2;
    T a2 = (T)1 / (g2 * tan2);
    T a = xsqrt(a2);
    return a < (T)1.6f ? ((T)3.535f * a + (T)2.181f * a2) / ((T)1 + (T)2.276f * a + (T)2.577f * a2) : (T)1;
}

inline bool is_beckmann(uint32_t m) { return m >= TR_MODEL_BECKMANN_MICROFACET && m <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC; }
inline bool is_ggx(uint32_t m) { return m >= TR_MODEL_GGX_MICROFACET && m <= TR_MODEL_GGX_MICROFACET_DIELECTRIC; }
inline bool is_microfacet(uint32_t m) { return is_ggx(m) || is_beckmann(m); }

/* GGXSpecularBRDF / BeckmannSpecularBRDF (ggx-models-impl.hh:156-166,
 * beckmann-models-impl.hh:121-131) */
template <class T> inline T specular_brdf(const Sample<T>& s, const Angles<T>& a, bool beckmann)
{
    T g2 = geom_factor_sq((T)s.mat->specular_power[0]);
    T shadow, d;
    if (beckmann) {
        shadow = beckmann_g1(g2, a.cos_im, a.cos_in) * beckmann_g1(g2, a.cos_im, a.cos_on);
        d = beckmann_d(g2, a.cos_mn);
    } else {
        shadow = ggx_g1(g2, a.cos_im, a.cos_in) * ggx_g1(g2, a.cos_im, a.cos_on);
        d = ggx_d(g2, a.cos_mn);
    }
    /* Deviation: at exactly grazing incidence (cos = 0) the reference's
     * shadow*D/(4|cos_in||cos_on|) is 0/0 = NaN; G1 is 0 there, so the limit
     * value 0 is returned instead (the GPU kernels do the same). */
    T den = (T)4 * xabs(a.cos_in) * xabs(a.cos_on);
    return den > (T)0 ? shadow * d / den : (T)0;
}

/* GGXMicrofacet{,Conductor,Dielectric}BRDF (ggx-models-impl.hh:168-209) and
 * the Beckmann twins (beckmann-models-impl.hh:133-172): BRDF times cos. */
template <class T> inline V3<T> microfacet_brdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    uint32_t model = m->model;
    Angles<T> a = microfacet_angles(s);
    T fresnel;
    uint32_t variant = model - (is_beckmann(model) ? TR_MODEL_BECKMANN_MICROFACET : TR_MODEL_GGX_MICROFACET);
    if (variant == 0) {
        fresnel = fresnel_schlick((T)m->fresnel[0], a.cos_im);
    } else if (variant == 1) {
        fresnel = fresnel_conductor((T)m->fresnel[0], (T)m->fresnel[1], a.cos_im);
    } else {
        T cos_trans = cos_transmittance((T)m->fresnel[0], a.cos_im);
        fresnel = fresnel_dielectric((T)m->fresnel[0], a.cos_im, cos_trans);
    }
    T brdf = fresnel * specular_brdf(s, a, is_beckmann(model));
    return (spec3<T>(m->diffuse) * ((T)1 / C<T>::pi) + spec3<T>(m->specular) * brdf) * a.cos_in;
}

/* GGXMicrofacetPDF / BeckmannMicrofacetPDF (ggx-models-impl.hh:211-237,
 * beckmann-models-impl.hh:240-266) */
template <class T> inline T microfacet_pdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    T cos_ln = dot(s.wi, s.normal);
    if (cos_ln < (T)0) return (T)0;
    V3<T> h = normalize(s.wi + s.wo);
    T cos_hn = dot(h, s.normal);
    T g2 = geom_factor_sq((T)m->specular_power[0]);
    T norm_pdf = (is_beckmann(m->model) ? beckmann_d(g2, cos_hn) : ggx_d(g2, cos_hn)) * cos_hn;
    if (cos_hn <= (T)0 || norm_pdf <= (T)0) return (T)0;
    T jacobian = (T)1 / ((T)4 * dot(h, s.wo));
    T pdf0 = norm_pdf * jacobian;
    T mv = max_value(spec3<T>(m->diffuse) + spec3<T>(m->specular));
    T ratio = mv != (T)0 ? max_value(spec3<T>(m->specular)) / mv : (T)0;
    T pdf1 = cos_ln * ((T)1 / C<T>::pi);
    return (pdf0 - pdf1) * ratio + pdf1;
}

/* GGXMicrofacetSampleIncidentLight / BeckmannMicrofacetSampleIncidentLight
 * (ggx-models-impl.hh:239-282, beckmann-models-impl.hh:268-311).
 * xi[0], xi[1]: the stratified coordinates; xi[2]: the lobe selector. */
template <class T> inline void microfacet_sample(const Strata<T>& st, Sample<T>* s, const T xi[3])
{
    const tr_material* m = s->mat;
    T x_ray = (st.x + xi[0]) / st.tx;
    T y_ray = (st.y + xi[1]) / st.ty;
    T mv = max_value(spec3<T>(m->diffuse) + spec3<T>(m->specular));
    T ratio = mv != (T)0 ? max_value(spec3<T>(m->specular)) / mv : (T)0;
    M3<T> basis = make_basis(s->normal);
    if (xi[2] <= ratio) {
        T g2 = geom_factor_sq((T)m->specular_power[0]);
        T g = xsqrt(g2);
        T sin_phi, cos_phi;
        sincos_turns(y_ray, &sin_phi, &cos_phi);
        T r = is_beckmann(m->model) ? -xlog(x_ray) : x_ray / ((T)1 - x_ray);
        T sqrt_ratio = xsqrt(r);
        T rcp_len = (T)1 / xsqrt(r * g2 + (T)1);
        T common = sqrt_ratio * g * rcp_len;
        V3<T> mn = rcp_len > (T)1e-6f ? v3(-common * cos_phi, -common * sin_phi, rcp_len) : v3((T)0, (T)0, (T)1);
        mn = transform(basis, mn);
        s->wi = reflect(s->wo, mn);
    } else {
        /* CosineSampleHemisphere, sampling3.hh:69-76 */
        T sin_phi, cos_phi;
        sincos_turns(y_ray, &sin_phi, &cos_phi);
        T r = xsqrt(x_ray);
        V3<T> d = v3(cos_phi * r, sin_phi * r, xsqrt(xmax((T)0, (T)1 - x_ray)));
        s->wi = normalize(transform(basis, d));
    }
    s->pdf = microfacet_pdf(*s);
}

/* ---- anisotropic GGX and Beckmann (models 15-17 and 31-33) ---- */

inline bool is_aniso(uint32_t m)
{
    return (m >= TR_MODEL_GGX_MICROFACET_ANISOTROPIC && m <= TR_MODEL_GGX_MICROFACET_DIELECTRIC_ANISOTROPIC) ||
           (m >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC && m <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC_ANISOTROPIC);
}
inline bool aniso_is_beckmann(uint32_t m) { return m >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC; }

/* A Vector2 and its NormalizeSelf (math/vector2.hh:60-74). */
template <class T> struct Dir2 {
    T x, y;
};
template <class T> inline Dir2<T> unit2(T x, T y)
{
    T l = xsqrt(x * x + y * y);
    if (l != (T)0) {
        x = x / l;
        y = y / l;
    }
    return Dir2<T>{x, y};
}

/* (alpha_x, alpha_y): the square roots of SpecularPowerToGeometryFactorSq of
 * SpecularPower.x and .y */
template <class T> inline Dir2<T> aniso_alpha(const tr_material* m)
{
    return Dir2<T>{xsqrt(geom_factor_sq((T)m->specular_power[0])), xsqrt(geom_factor_sq((T)m->specular_power[1]))};
}

/* GGXMicrofacetDistributionAnisotropic, ggx-models-impl.hh:111-127, and
 * BeckmannMicrofacetDistributionAnisotropic, beckmann-models-impl.hh:69-85;
 * `th` is the half vector's normalized tangent-plane direction.
 * Deviation: 0 for cos_nh <= 0 (the reference has no guard; its callers
 * multiply such a value by a zero shadowing term or drop the non-positive
 * PDF, so only the NaN at cos_nh = 0 changes). */
template <class T> inline T aniso_distribution(bool beckmann, Dir2<T> alpha, T cos_nh, Dir2<T> th)
{
    if (cos_nh <= (T)0) return (T)0;
    T c2 = cos_nh * cos_nh;
    T c4 = c2 * c2;
    T tan2 = ((T)1 - c2) / c2;
    T r1 = th.x / alpha.x;
    T r2 = th.y / alpha.y;
    T aniso = r1 * r1 + r2 * r2;
    if (beckmann) return xexp(-tan2 * aniso) / (C<T>::pi * alpha.x * alpha.y * c4);
    T den = (T)1 + tan2 * aniso;
    return (T)1 / (C<T>::pi * c4 * alpha.x * alpha.y * den * den);
}

/* {GGX,Beckmann}MicrofacetAnisotropic{,Conductor,Dielectric}BRDF
 * (ggx-models-impl.hh:283-341, beckmann-models-impl.hh:174-238) with
 * ComputeMicrofacetAnglesAnisotropic and
 * ComputeRoughnessProjectedOnDirectionSq (microfacet-common.hh:86-107):
 * BRDF times cos. */
template <class T> inline V3<T> microfacet_aniso_brdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    const bool beck = aniso_is_beckmann(m->model);
    Angles<T> a = microfacet_angles(s);
    V3<T> h = normalize(s.wi + s.wo);
    Dir2<T> th = unit2(dot(s.tangent, h), dot(s.binormal, h));
    T fresnel;
    uint32_t variant =
        m->model - (beck ? TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC : TR_MODEL_GGX_MICROFACET_ANISOTROPIC);
    if (variant == 0) {
        fresnel = fresnel_schlick((T)m->fresnel[0], a.cos_im);
    } else if (variant == 1) {
        fresnel = fresnel_conductor((T)m->fresnel[0], (T)m->fresnel[1], a.cos_im);
    } else {
        T cos_trans = cos_transmittance((T)m->fresnel[0], a.cos_im);
        fresnel = fresnel_dielectric((T)m->fresnel[0], a.cos_im, cos_trans);
    }
    Dir2<T> to = unit2(dot(s.wo, s.tangent), dot(s.wo, s.binormal));
    Dir2<T> alpha = aniso_alpha<T>(m);
    T px = alpha.x * to.x, py = alpha.y * to.y;
    T g2_out = xabs(px * px + py * py);
    T shadow = beck ? beckmann_g1(g2_out, a.cos_im, a.cos_in) * beckmann_g1(g2_out, a.cos_im, a.cos_on)
                    : ggx_g1(g2_out, a.cos_im, a.cos_in) * ggx_g1(g2_out, a.cos_im, a.cos_on);
    T d = aniso_distribution(beck, alpha, a.cos_mn, th);
    /* the grazing-incidence limit 0, as for the isotropic models */
    T den = (T)4 * xabs(a.cos_in) * xabs(a.cos_on);
    T spec = den > (T)0 ? shadow * d / den : (T)0;
    T brdf = fresnel * spec;
    return (spec3<T>(m->diffuse) * ((T)1 / C<T>::pi) + spec3<T>(m->specular) * brdf) * a.cos_in;
}

/* {GGX,Beckmann}MicrofacetAnisotropicPDF (ggx-models-impl.hh:343-375,
 * beckmann-models-impl.hh:313-345) */
template <class T> inline T microfacet_aniso_pdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    T cos_ln = dot(s.wi, s.normal);
    if (cos_ln < (T)0) return (T)0;
    V3<T> h = normalize(s.wi + s.wo);
    T cos_hn = dot(h, s.normal);
    Dir2<T> th = unit2(dot(s.tangent, h), dot(s.binormal, h));
    T norm_pdf = aniso_distribution(aniso_is_beckmann(m->model), aniso_alpha<T>(m), cos_hn, th) * cos_hn;
    if (norm_pdf <= (T)0) return (T)0;
    T jacobian = (T)1 / ((T)4 * dot(h, s.wo));
    T pdf0 = norm_pdf * jacobian;
    T mv = max_value(spec3<T>(m->diffuse) + spec3<T>(m->specular));
    T ratio = mv != (T)0 ? max_value(spec3<T>(m->specular)) / mv : (T)0;
    T pdf1 = cos_ln * ((T)1 / C<T>::pi);
    return (pdf0 - pdf1) * ratio + pdf1;
}

/* {GGX,Beckmann}MicrofacetAnisotropicSampleIncidentLight with
 * {GGX,Beckmann}SampleMicrofacetAnisotropicDistribution
 * (ggx-models-impl.hh:377-432, beckmann-models-impl.hh:347-398).
 * xi[0], xi[1]: the stratified coordinates of the cosine lobe; xi[2]: the
 * lobe selector; xi[3], xi[4]: the specular lobe's own stratified pair.
 * The micro-normal lives in makeBasis(Normal), not in the tangent frame. */
template <class T> inline void microfacet_aniso_sample(const Strata<T>& st, Sample<T>* s, const T xi[5])
{
    const tr_material* m = s->mat;
    T x_ray = (st.x + xi[0]) / st.tx;
    T y_ray = (st.y + xi[1]) / st.ty;
    T mv = max_value(spec3<T>(m->diffuse) + spec3<T>(m->specular));
    T ratio = mv != (T)0 ? max_value(spec3<T>(m->specular)) / mv : (T)0;
    M3<T> basis = make_basis(s->normal);
    if (xi[2] <= ratio) {
        Dir2<T> alpha = aniso_alpha<T>(m);
        T r0 = (st.x + xi[3]) / st.tx;
        T r1 = (st.y + xi[4]) / st.ty;
        T sin_phi, cos_phi;
        sincos_turns(r1, &sin_phi, &cos_phi);
        T sqrt_ratio = aniso_is_beckmann(m->model) ? xsqrt(-xlog(r0)) : xsqrt(r0 / ((T)1 - r0));
        T x = alpha.x * sqrt_ratio * cos_phi;
        T y = alpha.y * sqrt_ratio * sin_phi;
        V3<T> mn = normalize(v3(-x, -y, (T)1));
        mn = normalize(mn);  /* normalized once more by the caller */
        mn = normalize(transform(basis, mn));
        s->wi = reflect(s->wo, mn);
    } else {
        T sin_phi, cos_phi;
        sincos_turns(y_ray, &sin_phi, &cos_phi);
        T r = xsqrt(x_ray);
        V3<T> d = v3(cos_phi * r, sin_phi * r, xsqrt(xmax((T)0, (T)1 - x_ray)));
        s->wi = normalize(transform(basis, d));
    }
    s->pdf = microfacet_aniso_pdf(*s);
}

/* powf as the Blinn-Phong lobes use it.  FP32: exp(p log x) in the shared
 * tr_expf/tr_logf for x > 0; x = 0 gives 0 (p > 0) or 1 (p = 0); a negative
 * base gives |x|^p signed by (-1)^p for integer p, NaN otherwise (C99
 * pow).  FP64: libm pow. */
inline float xpow(float x, float p)
{
    if (x > 0.0f) return tr_expf(p * tr_logf(x));
    if (p == 0.0f) return 1.0f;
    if (x == 0.0f) return p > 0.0f ? 0.0f : HUGE_VALF;
    if (floorf(p) != p) return NAN;
    float m = tr_expf(p * tr_logf(-x));
    return fmodf(p, 2.0f) != 0.0f ? -m : m;
}
inline double xpow(double x, double p) { return pow(x, p); }

template <class T> inline T lobe_ratio(V3<T> albedo, V3<T> specular)
{
    T mv = max_value(albedo + specular);
    return mv != (T)0 ? max_value(specular) / mv : (T)0;
}

/* BlinnPhongBRDF (illumination-models.hh:679-698): BRDF times cos, the
 * albedo from DiffuseMap when there is one. */
template <class T> inline V3<T> blinn_phong_brdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    V3<T> albedo = map_spectrum(s, m->diffuse_map, m->diffuse);
    T power = (T)m->specular_power[0];
    T dot_nl = xmax((T)0, dot(s.normal, s.wi));
    V3<T> h = normalize(s.wi + s.wo);
    V3<T> spec = spec3<T>(m->specular) * xpow(xmax((T)0, dot(s.normal, h)), power);
    T cos_theta = xmin(xmax(dot(s.wi, h), (T)0), (T)1);
    T fresnel = fresnel_schlick((T)m->fresnel[0], cos_theta);
    return (albedo + spec * fresnel * (power + (T)8) * ((T)1 / (T)8)) * dot_nl * ((T)1 / C<T>::pi);
}

/* BlinnPhongPDF (illumination-models.hh:618-643) */
template <class T> inline T blinn_phong_pdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    V3<T> h = normalize(s.wi + s.wo);
    T dot_ln = dot(s.wi, s.normal);
    if (dot_ln < (T)0) return (T)0;
    T ratio = lobe_ratio<T>(map_spectrum(s, m->diffuse_map, m->diffuse), spec3<T>(m->specular));
    T power = (T)m->specular_power[0];
    T pdf0 = (power + (T)1) * xpow(dot(h, s.normal), power) * ((T)1 / ((T)2 * C<T>::pi));
    T transform4 = (T)4 * dot(h, s.wo);
    pdf0 = transform4 > (T)0 && pdf0 > (T)0 ? pdf0 / transform4 : (T)0;
    T pdf1 = dot_ln * ((T)1 / C<T>::pi);
    return (pdf0 - pdf1) * ratio + pdf1;
}

/* BlinnPhongSampleIncidentLight (illumination-models.hh:645-677):
 * PowerCosineSampleHemisphere (sampling3.hh:96-104) for the reflected
 * half vector when xi[2] <= ratio, else CosineSampleHemisphere
 * (sampling3.hh:69-76). */
template <class T> inline void blinn_phong_sample(const Strata<T>& st, Sample<T>* s, const T xi[3])
{
    const tr_material* m = s->mat;
    T x_ray = (st.x + xi[0]) / st.tx;
    T y_ray = (st.y + xi[1]) / st.ty;
    T ratio = lobe_ratio<T>(map_spectrum(*s, m->diffuse_map, m->diffuse), spec3<T>(m->specular));
    M3<T> basis = make_basis(s->normal);
    T sin_phi, cos_phi;
    sincos_turns(y_ray, &sin_phi, &cos_phi);
    if (xi[2] <= ratio) {
        T cos_t = xpow(x_ray, (T)1 / ((T)m->specular_power[0] + (T)1));
        T sin_t = xsqrt(xmax((T)0, (T)1 - cos_t * cos_t));
        V3<T> mn = normalize(transform(basis, v3(cos_phi * sin_t, sin_phi * sin_t, cos_t)));
        s->wi = reflect(s->wo, mn);
    } else {
        T r = xsqrt(x_ray);
        V3<T> d = v3(cos_phi * r, sin_phi * r, xsqrt(xmax((T)0, (T)1 - x_ray)));
        s->wi = normalize(transform(basis, d));
    }
    s->pdf = blinn_phong_pdf(*s);
}

/* UniformHemispherePDF, illumination-models.hh:723-727 */
template <class T> inline T uniform_hemisphere_pdf(const Sample<T>& s)
{
    return dot(s.normal, s.wi) >= (T)0 ? (T)1 / ((T)2 * C<T>::pi) : (T)0;
}

/* UniformSampleIncidentLight, illumination-models.hh:729-738, with
 * UniformSampleHemisphere, sampling3.hh:55-62 */
template <class T> inline void uniform_hemisphere_sample(const Strata<T>& st, Sample<T>* s, const T xi[3])
{
    M3<T> basis = make_basis(s->normal);
    T ra = (st.x + xi[0]) / st.tx;
    T rb = (st.y + xi[1]) / st.ty;
    T r = xsqrt((T)1 - ra * ra);
    T sin_phi, cos_phi;
    sincos_turns(rb, &sin_phi, &cos_phi);
    V3<T> d = v3(cos_phi * r, sin_phi * r, ra);
    s->pdf = (T)1 / ((T)2 * C<T>::pi);
    s->wi = normalize(transform(basis, d));
}

/* SGGXProjectedArea, ggx-models-impl.hh:99-102 */
template <class T> inline T sggx_projected_area(V3<T> stddev, const M3<T>& basis, V3<T> dir)
{
    return length(transform_rotation_inverse(basis, dir) * stddev);
}

/* SGGXMicroFlakeNDF(stddev, basis, normal), ggx-models-impl.hh:48-57 */
template <class T> inline T sggx_ndf(V3<T> stddev, const M3<T>& basis, V3<T> n)
{
    V3<T> sq = stddev * stddev;
    M3<T> inv_cov = transform_covariance(basis, v3((T)1 / sq.x, (T)1 / sq.y, (T)1 / sq.z));
    T det = determinant(inv_cov);
    if (det < (T)1e-6f) return (T)0;
    T den = dot(n, transform(inv_cov, n));
    return (T)0.5 * xsqrt(det) / (C<T>::pi * den * den);
}

/* StandardDeviation (x, y, 1) from StandardDeviationMap (SampleRG) or
 * the material; SGGXBasis from BasisMap (SampleRGBA) or the material
 * (ggx-models-impl.hh:495-517, 593-608). */
template <class T> inline V3<T> sggx_surface_stddev(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    if (m->stddev_map == 0 || !s.tex) return v3((T)m->specular_power[0], (T)m->specular_power[1], (T)1);
    Rgba c = orc_tex_sample(*s.tex, m->stddev_map, (float)s.tcx, (float)s.tcy, ORC_TEX_RG);
    return v3((T)c.v[0], (T)c.v[1], (T)1);
}
template <class T> inline void sggx_surface_quat(const Sample<T>& s, T q[4])
{
    const tr_material* m = s.mat;
    if (m->basis_map == 0 || !s.tex) {
        for (int i = 0; i < 4; ++i) q[i] = (T)m->sggx_basis[i];
        return;
    }
    Rgba c = orc_tex_sample(*s.tex, m->basis_map, (float)s.tcx, (float)s.tcy, ORC_TEX_RGBA);
    for (int i = 0; i < 4; ++i) q[i] = (T)c.v[i];
}

/* SGGXMicroFlakeSurfaceDensity via SGGXSurfaceCache, ggx-models-impl.hh:589-619
 * (the quaternion is not normalised there) */
template <class T> inline T sggx_surface_density(const Sample<T>& s)
{
    T q[4];
    sggx_surface_quat(s, q);
    M3<T> orient = m3(s.tangent, s.binormal, s.normal);
    M3<T> basis = mul(orient, quat_to_m3(q));
    return sggx_projected_area(sggx_surface_stddev(s), basis, s.wi);
}

/* SGGXMicroFlakeSurfaceBRDF, ggx-models-impl.hh:476-551. */
template <class T> inline V3<T> sggx_surface_brdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    T din = dot(s.normal, s.wi);
    T don = dot(s.normal, s.wo);
    if (din < (T)0 || don < (T)0) return v3((T)0, (T)0, (T)0);
    V3<T> result = map_spectrum(s, m->diffuse_map, m->diffuse) * din * ((T)1 / C<T>::pi);
    V3<T> stddev = sggx_surface_stddev(s);
    if (stddev.x >= (T)1e-6f && stddev.y >= (T)1e-6f && stddev.z >= (T)1e-6f && s.dir_density >= (T)1e-6f) {
        V3<T> mn = normalize(s.wi + s.wo);
        M3<T> orient = m3(s.tangent, s.binormal, s.normal);
        T q[4];
        sggx_surface_quat(s, q);
        T qn[4];
        quat_normalize(q, qn);
        M3<T> basis = mul(orient, quat_to_m3(qn));
        T cos_micro = dot(mn, s.wi);
        T fresnel;
        T eta = (T)m->fresnel[0];
        if (eta == (T)1) {
            fresnel = (T)1;
        } else {
            T cos_trans = cos_transmittance(eta, cos_micro);
            fresnel = fresnel_dielectric(eta, cos_micro, cos_trans);
            fresnel = fresnel < (T)0 ? (T)0 : (fresnel > (T)1 ? (T)1 : fresnel);
        }
        T out_density = s.dir_density;
        T in_density = sggx_projected_area(stddev, basis, s.wi);
        T contrib = (T)1 / (out_density + in_density * don / din);
        result += contrib * map_spectrum(s, m->specular_map, m->specular) * sggx_ndf(stddev, basis, mn) * fresnel *
                  (T)2 * (T)0.25;
    }
    return result;
}

/* ---------------- SGGX micro-flake phase function ---------------- */

template <class T> inline M3<T> transpose(const M3<T>& m)
{
    return m3(v3(m.c0.x, m.c1.x, m.c2.x), v3(m.c0.y, m.c1.y, m.c2.y), v3(m.c0.z, m.c1.z, m.c2.z));
}

/* Matrix3::choleskyDecomposition (matrix3.hh:173-191) */
template <class T> inline M3<T> cholesky(const M3<T>& m)
{
    T l00 = xsqrt(m.c0.x);
    T rcp_l00 = (T)1 / l00;
    T l10 = rcp_l00 * m.c0.y;
    T l20 = rcp_l00 * m.c0.z;
    T l11 = xsqrt(m.c1.y - l10 * l10);
    T rcp_l11 = (T)1 / l11;
    T l21 = rcp_l11 * (m.c1.z - l20 * l10);
    T l22 = xsqrt(m.c2.z - l21 * l21 - l20 * l20);
    return m3(v3(l00, l10, l20), v3((T)0, l11, l21), v3((T)0, (T)0, l22));
}

template <class T> inline M3<T> sample_frame(const Sample<T>& s) { return m3(s.tangent, s.binormal, s.normal); }

/* SGGXMicroFlakeBRDF (ggx-models-impl.hh:60-68, 466-474) */
template <class T> inline T sggx_flake_brdf(const Sample<T>& s)
{
    V3<T> mn = normalize(s.wi + s.wo);
    return sggx_ndf(spec3<T>(s.mat->sggx_stddev), sample_frame(s), mn) * (T)2 / ((T)4 * s.dir_density);
}

/* SGGXMicroFlakePDF (ggx-models-impl.hh:447-464) */
template <class T> inline T sggx_flake_pdf(T dir_density, V3<T> stddev, const M3<T>& basis, V3<T> mn)
{
    T ndf = sggx_ndf(stddev, basis, mn);
    T proj_area = (T)1 / ((T)4 * dir_density);
    return (T)2 * ndf * proj_area;
}

/* SGGXMicroFlakeSampleIncidentLight (ggx-models-impl.hh:554-580) with
 * UniformSampleSphere (sampling3.hh:106-116). */
template <class T> inline void sggx_flake_sample(Sample<T>* s, T ra, T rb)
{
    T sin_phi, cos_phi;
    sincos_turns(ra, &sin_phi, &cos_phi);
    T cos_theta = (T)2 * rb - (T)1;
    T sin_theta = xsqrt((T)1 - cos_theta * cos_theta);
    V3<T> us = v3(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta);
    V3<T> sd = spec3<T>(s->mat->sggx_stddev);
    M3<T> basis = sample_frame(*s);
    M3<T> S = transform_covariance(basis, sd * sd);
    M3<T> ob = make_basis(s->wo);
    S = mul(mul(ob, S), transpose(ob));
    V3<T> vn = normalize(transform_rotation_inverse(cholesky(S), us));
    vn = transform(ob, vn);
    s->wi = reflect(s->wo, vn);
    s->pdf = sggx_flake_pdf(s->dir_density, sd, basis, normalize(s->wi + s->wo));
}

/* ---- SGGX pseudo-volume surface (model 25) ---- */

/* The flake normal of SGGXMicroFlakeSampleIncidentLight(out_light, ...)
 * (ggx-models-impl.hh:554-570) for the covariance S = basis
 * transformCovariance(stddev^2): Cholesky factor of S in the frame of wo. */
template <class T> inline V3<T> sggx_flake_normal(V3<T> wo, const M3<T>& S, T ra, T rb)
{
    T sin_phi, cos_phi;
    sincos_turns(ra, &sin_phi, &cos_phi);
    T cos_theta = (T)2 * rb - (T)1;
    T sin_theta = xsqrt((T)1 - cos_theta * cos_theta);
    V3<T> us = v3(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta);
    M3<T> ob = make_basis(wo);
    M3<T> Sw = mul(mul(ob, S), transpose(ob));
    return transform(ob, normalize(transform_rotation_inverse(cholesky(Sw), us)));
}

/* IntelFastRand / FloatIntelFastRand, math/functions.hh:214-218, 249-252 */
struct OrcLcg {
    uint32_t seed;
    float next()
    {
        seed = 214013u * seed + 2531011u;
        return (float)((seed >> 16) & 0x7FFFu) / 32767.0f;
    }
};

/* (uint32_t)(tc * USHRT_MAX) of the seed: 0 for negative and NaN, saturated
 * past 2^32 - 1 (undefined in C++). */
inline uint32_t orc_tc_seed_part(float x) { return x > 0.0f ? (x < 4294967296.0f ? (uint32_t)x : 0xFFFFFFFFu) : 0u; }

/* surface frame * ToMatrix3(quaternion as given) */
template <class T> inline M3<T> sggx_pseudo_basis(const Sample<T>& s)
{
    T q[4];
    sggx_surface_quat(s, q);
    return mul(sample_frame(s), quat_to_m3(q));
}

/* SGGXMicroFlakePseudoVolumeBRDF, ggx-models-impl.hh:621-758.  The
 * internal random walks use the reference's LCG seeded from the TexCoord;
 * the two uniforms of each flake sample are drawn left to right. */
template <class T> inline V3<T> sggx_pseudo_volume_brdf(const Sample<T>& s)
{
    const tr_material* m = s.mat;
    T din = dot(s.normal, s.wi);
    T don = dot(s.normal, s.wo);
    if (don < (T)0 || din < (T)0) return v3((T)0, (T)0, (T)0);
    V3<T> sd = sggx_surface_stddev(s);
    if (sd.x < (T)1e-6f || sd.y < (T)1e-6f) return v3((T)0, (T)0, (T)0);
    M3<T> basis = sggx_pseudo_basis(s);
    OrcLcg lcg{(orc_tc_seed_part((float)s.tcy * 65535.0f) << 16) + orc_tc_seed_part((float)s.tcx * 65535.0f)};
    lcg.next();
    const uint32_t depth = m->sggx_depth, count = m->sggx_sample_count;
    const T start = s.dir_density;
    const T in_density = sggx_projected_area(sd, basis, s.wi);
    const V3<T> spec = spec3<T>(m->specular);
    V3<T> mn = normalize(s.wo + s.wi);
    T contrib = (T)(float)count / (start + in_density * don / din);
    V3<T> radiance = spec * contrib * sggx_ndf(sd, basis, mn) * (T)2 * (T)0.25;
    const M3<T> S = transform_covariance(basis, sd * sd);
    for (uint32_t k = 0; k < count; ++k) {
        V3<T> thr = spec;
        V3<T> out = s.wo;
        T pen = (-xlog((T)lcg.next()) / start) * don;
        for (uint32_t it = 0; it < depth; ++it) {
            if (!(thr.x > (T)1e-3f && thr.y > (T)1e-3f && thr.z > (T)1e-3f)) break;
            T ra = (T)lcg.next();
            T rb = (T)lcg.next();
            V3<T> inc = reflect(out, sggx_flake_normal(out, S, ra, rb));
            T dd = sggx_projected_area(sd, basis, inc);
            out = inc * (T)-1;
            pen = pen + (-xlog((T)lcg.next()) / dd) * dot(s.normal, out);
            if (pen < (T)0) break;
            T ext = xexp(-in_density * (pen / din));
            if (ext < (T)1e-9f) break;
            V3<T> mn2 = normalize(out + s.wi);
            thr = v3(thr.x * spec.x, thr.y * spec.y, thr.z * spec.z);
            T phase = sggx_ndf(sd, basis, mn2) * (T)2 / ((T)4 * dd);
            radiance = radiance + thr * ext * phase;
        }
    }
    T rc = (T)1 / (T)(float)count;
    return radiance * rc + spec3<T>(m->diffuse) * din * ((T)1 / C<T>::pi);
}

/* SGGXMicroFlakeSurfacePDF, ggx-models-impl.hh:760-794 */
template <class T> inline T sggx_pseudo_volume_pdf(const Sample<T>& s)
{
    if (dot(s.wi, s.normal) < (T)0 || dot(s.wo, s.normal) < (T)0) return (T)0;
    V3<T> mn = normalize(s.wi + s.wo);
    return (T)2 * sggx_flake_pdf(s.dir_density, sggx_surface_stddev(s), sggx_pseudo_basis(s), mn);
}

/* SGGXMicroFlakeSurfaceSampleIncidentLight, ggx-models-impl.hh:796-842,
 * including its flip of any flake normal with a non-zero cosine to the
 * normal (:836), which leaves the reflected direction unchanged. */
template <class T> inline void sggx_pseudo_volume_sample(Sample<T>* s, T ra, T rb)
{
    V3<T> sd = sggx_surface_stddev(*s);
    M3<T> basis = sggx_pseudo_basis(*s);
    V3<T> mn = sggx_flake_normal(s->wo, transform_covariance(basis, sd * sd), ra, rb);
    if (dot(mn, s->normal) != (T)0) mn = mn * (T)-1;
    s->wi = reflect(s->wo, mn);
    s->pdf = dot(s->wi, s->normal) > (T)0 ? (T)2 * sggx_flake_pdf(s->dir_density, sd, basis, mn) : (T)0;
}

inline bool is_emissive(uint32_t m) { return m == TR_MODEL_EMISSIVE; }

/* TransmittanceLookup (illumination-models.hh:279-324) for the models the
 * tracer supports (SGGXMicroFlake only inside grid volumes). */
template <class T> inline V3<T> bsdf_eval(const Sample<T>& s)
{
    uint32_t m = s.mat->model;
    if (is_microfacet(m)) return microfacet_brdf(s);
    if (is_aniso(m)) return microfacet_aniso_brdf(s);
    if (m == TR_MODEL_BLINN_PHONG) return blinn_phong_brdf(s);
    if (m == TR_MODEL_SGGX_SURFACE) return sggx_surface_brdf(s);
    if (m == TR_MODEL_SGGX_PSEUDO_VOLUME) return sggx_pseudo_volume_brdf(s);
    if (m == TR_MODEL_SGGX_MICRO_FLAKE) {
        T p = sggx_flake_brdf(s);
        return v3(p, p, p);
    }
    return v3((T)0, (T)0, (T)0);
}

/* PDFLookup (illumination-models.hh:233-277) */
template <class T> inline T bsdf_pdf(const Sample<T>& s)
{
    uint32_t m = s.mat->model;
    if (is_microfacet(m)) return microfacet_pdf(s);
    if (is_aniso(m)) return microfacet_aniso_pdf(s);
    if (m == TR_MODEL_BLINN_PHONG) return blinn_phong_pdf(s);
    if (m == TR_MODEL_SGGX_SURFACE) return uniform_hemisphere_pdf(s);
    if (m == TR_MODEL_SGGX_PSEUDO_VOLUME) return sggx_pseudo_volume_pdf(s);
    if (m == TR_MODEL_SGGX_MICRO_FLAKE)
        return sggx_flake_pdf(s.dir_density, spec3<T>(s.mat->sggx_stddev), sample_frame(s), normalize(s.wi + s.wo));
    return (T)0;
}

/* SampleIncidentLightLookup (illumination-models.hh:325-370); xi[3] and
 * xi[4] are read by the anisotropic microfacet models only. */
template <class T> inline void bsdf_sample(const Strata<T>& st, Sample<T>* s, const T xi[5])
{
    uint32_t m = s->mat->model;
    if (is_microfacet(m)) microfacet_sample(st, s, xi);
    else if (is_aniso(m)) microfacet_aniso_sample(st, s, xi);
    else if (m == TR_MODEL_BLINN_PHONG) blinn_phong_sample(st, s, xi);
    else if (m == TR_MODEL_SGGX_SURFACE) uniform_hemisphere_sample(st, s, xi);
    else if (m == TR_MODEL_SGGX_PSEUDO_VOLUME) sggx_pseudo_volume_sample(s, xi[0], xi[1]);
    else if (m == TR_MODEL_SGGX_MICRO_FLAKE) sggx_flake_sample(s, xi[0], xi[1]);
    else s->pdf = (T)0;
}

/* MaterialCacheLookup (illumination-models.hh:455-492) */
template <class T> inline void material_cache(Sample<T>* s)
{
    if (s->mat->model == TR_MODEL_SGGX_SURFACE || s->mat->model == TR_MODEL_SGGX_PSEUDO_VOLUME)
        s->dir_density = sggx_surface_density(*s);
}

}  // namespace orc

#endif
