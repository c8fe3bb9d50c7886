/*
 * tr_fmath.h - deterministic FP32 elementary functions shared by host and
 * device code.
 *
 * The path tracer takes discrete decisions (lobe selection, hit/miss,
 * Russian roulette) from FP32 values.  For the GPU kernels and the CPU
 * oracle to take the same decisions on the same random numbers, every
 * transcendental they use must round identically on gfx950 and on x86-64.
 * Library sinf/logf/expf differ between ROCm's device library and glibc, so
 * both sides use the functions below, which are built only from IEEE-exact
 * operations (add, mul, fma, division, sqrt, rint, integer bit work).  Code
 * that includes this header must be compiled with -ffp-contract=off so that
 * the compiler adds no fused operations of its own.
 *
 * Methods:
 *  - sin/cos of 2*pi*t: exact reduction of t to r in [-1/8, 1/8] turns and a
 *    quadrant (Cody-Waite style reduction, exact here because the period is
 *    a power of two in turns), then odd/even polynomials in r.
 *  - log: x = 2^e * m with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *    log(m) = 2 atanh(s) = 2s + s*z*R(z), z = s^2 (the fdlibm structure),
 *    e*ln2 added in two parts (Cody-Waite split of ln2).
 *  - exp: x = k ln2 + r with Cody-Waite reduction, exp(r) = 1 + r + r^2 E(r),
 *    result scaled by 2^k through the exponent bits.
 *
 * Polynomial coefficients come from fit_fmath.py (weighted least squares on
 * Chebyshev nodes, rounded to float32).  The test suite bounds the error of
 * each function against float64 libm over its domain.
 *
 * The reference renderer uses a tan-half-angle sincos on the CPU and
 * __sincosf on CUDA (include/tempest/math/functions.hh:322-353), and
 * std::log/std::exp elsewhere; these functions replace all of them.
 */
#ifndef TR_FMATH_H
#define TR_FMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define TR_FN __host__ __device__ static inline
#else
#define TR_FN static inline
#endif

TR_FN uint32_t tr_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
TR_FN float tr_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* 2*pi rounded to float32 and the remainder, for the leading sin term. */
#define TR_TWO_PI_HI 6.28318548e+00f
#define TR_TWO_PI_LO -1.74845553e-07f

/* sin(2*pi*t) and cos(2*pi*t).  Exact argument reduction for any finite t. */
TR_FN void tr_sincos_turns(float t, float* s_out, float* c_out)
{
    float tf = t - floorf(t);           /* exact: t in [0, 1) afterwards   */
    float q = rintf(tf * 4.0f);         /* tf*4 is exact (power of two)    */
    float r = fmaf(-q, 0.25f, tf);      /* exact: |r| <= 1/8               */
    int quadrant = ((int)q) & 3;
    float y = r * r;
    /* sin(2 pi r) = r * (2pi + y * P(y)) */
    float ps = fmaf(y, 4.148257446e+01f, -7.669820404e+01f);
    ps = fmaf(y, ps, 8.160520935e+01f);
    ps = fmaf(y, ps, -4.134170151e+01f);
    float sr = fmaf(r, TR_TWO_PI_HI, r * fmaf(y, ps, TR_TWO_PI_LO));
    /* cos(2 pi r) = 1 + y * Q(y) */
    float pc = fmaf(y, -2.606242943e+01f, 6.023838043e+01f);
    pc = fmaf(y, pc, -8.545677185e+01f);
    pc = fmaf(y, pc, 6.493939209e+01f);
    pc = fmaf(y, pc, -1.973920822e+01f);
    float cr = fmaf(y, pc, 1.0f);
    float s, c;
    switch (quadrant) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
    *s_out = s;
    *c_out = c;
}

#define TR_LN2_HI 6.93145752e-01f  /* 0x3f317200: exact in few bits */
#define TR_LN2_LO 1.42860677e-06f

/* Natural logarithm.  log(0) = -inf, log(negative or NaN) = NaN, log(inf) = inf. */
TR_FN float tr_logf(float x)
{
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x *= 8388608.0f; e = -23; } /* denormal: scale by 2^23 */
    uint32_t bits = tr_f2u(x);
    e += (int)((bits >> 23) & 0xffu) - 127;
    float m = tr_u2f((bits & 0x7fffffu) | 0x3f800000u); /* m in [1, 2) */
    if (m > 1.41421354f) { m *= 0.5f; e += 1; }
    float f = m - 1.0f;                /* exact (Sterbenz) */
    float s = f / (2.0f + f);
    float z = s * s;
    float R = fmaf(z, 2.356793135e-01f, 2.853813767e-01f);
    R = fmaf(z, R, 4.000031948e-01f);
    R = fmaf(z, R, 6.666666865e-01f);
    R = z * R;
    /* log(1+f) = f - hfsq + s*(hfsq + R), hfsq = f*f/2 */
    float hfsq = 0.5f * f * f;
    float ef = (float)e;
    float lo = fmaf(s, hfsq + R, ef * TR_LN2_LO);
    return ef * TR_LN2_HI + ((f - hfsq) + lo);
}

#define TR_LOG2E 1.44269502e+00f

/* Exponential.  Overflow gives +inf, deep underflow gives 0. */
TR_FN float tr_expf(float x)
{
    if (x != x) return x;
    if (x > 88.7228317f) return INFINITY;
    if (x < -103.972084f) return 0.0f;
    float k = rintf(x * TR_LOG2E);
    float r = fmaf(-k, TR_LN2_HI, x);
    r = fmaf(-k, TR_LN2_LO, r);
    float p = fmaf(r, 1.982649555e-04f, 1.394602121e-03f);
    p = fmaf(r, p, 8.333429694e-03f);
    p = fmaf(r, p, 4.166626930e-02f);
    p = fmaf(r, p, 1.666666567e-01f);
    p = fmaf(r, p, 5.000000000e-01f);
    float er = fmaf(r * r, p, r) + 1.0f;
    int ki = (int)k;
    if (ki >= -126) {
        if (ki > 127) return er * 2.0f * tr_u2f((uint32_t)(ki - 1 + 127) << 23);
        return er * tr_u2f((uint32_t)(ki + 127) << 23);
    }
    /* subnormal result: scale in two steps */
    return (er * tr_u2f((uint32_t)(ki + 64 + 127) << 23)) * 5.42101086e-20f; /* 2^-64 */
}

#endif /* TR_FMATH_H */
