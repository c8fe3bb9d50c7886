/*
 * tr_math.h - FP32 vector math, bases and the Philox generator used by the
 * kernels (and by the host-side builder and debug exports).
 *
 * Operation order follows the reference math headers so that results round
 * as the reference's do: Dot/Cross/Length (include/tempest/math/
 * vector3.hh:72-104), normalisation and vector/scalar as a multiplication by
 * the reciprocal (vector3.hh:78-88, 185-218), Matrix3 makeBasis /
 * makeBasisOrthogonalize / transform / transformRotationInverse /
 * transformCovariance / determinant / product (matrix3.hh:99-227, 355-360),
 * quaternion -> matrix (quaternion.hh:122-142).  Compile with
 * -ffp-contract=off: the compiler must not fuse these operations.
 */
#ifndef TR_MATH_H
#define TR_MATH_H

#include <stdint.h>
#include "tr_fmath.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define TR_HD __host__ __device__ inline
#else
#define TR_HD inline
#endif

namespace trk {

/* MathPi of the reference is truncated (functions.hh:55); kept. */
constexpr float kPi = 3.141592f;
constexpr float kTau = 6.283185f;

struct V3 {
    float x, y, z;
};
TR_HD V3 mk(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
TR_HD V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
TR_HD V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
TR_HD V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
TR_HD V3 mulv(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
TR_HD V3 muls(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
TR_HD V3 smul(float s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
TR_HD V3 divs(V3 a, float s) { float r = 1.0f / s; return mk(a.x * r, a.y * r, a.z * r); }
TR_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TR_HD V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
TR_HD float length(V3 a) { return sqrtf(dot(a, a)); }
TR_HD V3 normalize(V3 a)
{
    float l = length(a);
    return l != 0.0f ? divs(a, l) : a;
}
TR_HD float fmax_(float a, float b) { return a > b ? a : b; }
TR_HD float fmin_(float a, float b) { return a < b ? a : b; }
TR_HD float max_value(V3 a) { return fmax_(fmax_(a.x, a.y), a.z); }
TR_HD V3 reflect(V3 inc, V3 n) { return sub(smul(2.0f * dot(n, inc), n), inc); }
TR_HD V3 ld3(const float* f) { return mk(f[0], f[1], f[2]); }
TR_HD float pow5(float x) { float x2 = x * x; return x2 * x2 * x; }

struct M3 {
    V3 c0, c1, c2;
};
TR_HD M3 m3(V3 a, V3 b, V3 c) { M3 m; m.c0 = a; m.c1 = b; m.c2 = c; return m; }
TR_HD M3 make_basis(V3 n)
{
    V3 d0 = cross(mk(1.0f, 0.0f, 0.0f), n), d1 = cross(mk(0.0f, 1.0f, 0.0f), n);
    float d0_len = dot(d0, d0), d1_len = dot(d1, d1);
    M3 m;
    m.c0 = d0_len > d1_len ? divs(d0, sqrtf(d0_len)) : divs(d1, sqrtf(d1_len));
    m.c1 = cross(n, m.c0);
    m.c2 = n;
    return m;
}
/* Matrix3::makeBasisTangent (matrix3.hh:110-120). */
TR_HD M3 make_basis_tangent(V3 tan)
{
    V3 d0 = cross(mk(0.0f, 1.0f, 0.0f), tan), d1 = cross(mk(0.0f, 0.0f, 1.0f), tan);
    float d0_len = dot(d0, d0), d1_len = dot(d1, d1);
    V3 b = d0_len > d1_len ? divs(d0, sqrtf(d0_len)) : divs(d1, sqrtf(d1_len));
    return m3(tan, b, cross(tan, b));
}
TR_HD M3 make_basis_orthogonalize(V3 tan, V3 n)
{
    V3 binorm = normalize(cross(n, tan));
    V3 ortho_tan = cross(binorm, n);
    return m3(ortho_tan, binorm, n);
}
TR_HD V3 transform(const M3& m, V3 v) { return add(add(smul(v.x, m.c0), smul(v.y, m.c1)), smul(v.z, m.c2)); }
TR_HD V3 transform_rotation_inverse(const M3& m, V3 v) { return mk(dot(v, m.c0), dot(v, m.c1), dot(v, m.c2)); }
TR_HD M3 transform_covariance(const M3& m, V3 var)
{
    V3 t = add(add(muls(muls(m.c0, var.x), m.c0.x), muls(muls(m.c1, var.y), m.c1.x)), muls(muls(m.c2, var.z), m.c2.x));
    V3 b = add(add(muls(muls(m.c0, var.x), m.c0.y), muls(muls(m.c1, var.y), m.c1.y)), muls(muls(m.c2, var.z), m.c2.y));
    V3 n = add(add(muls(muls(m.c0, var.x), m.c0.z), muls(muls(m.c1, var.y), m.c1.z)), muls(muls(m.c2, var.z), m.c2.z));
    return m3(t, b, n);
}
/* determinant with Elements[col][row], matrix3.hh:193-201 */
TR_HD float determinant(const M3& m)
{
    return m.c0.x * (m.c1.y * m.c2.z - m.c1.z * m.c2.y) - m.c1.x * (m.c0.y * m.c2.z - m.c0.z * m.c2.y) +
           m.c2.x * (m.c0.y * m.c1.z - m.c0.z * m.c1.y);
}
TR_HD M3 mul(const M3& l, const M3& r) { return m3(transform(l, r.c0), transform(l, r.c1), transform(l, r.c2)); }
TR_HD M3 quat_to_m3(float x, float y, float z, float w)
{
    float xx = x * x, xy = x * y, xz = x * z, xw = x * w, yy = y * y, yz = y * z, yw = y * w, zz = z * z, zw = z * w;
    return m3(mk(1.0f - 2.0f * (yy + zz), 2.0f * (xy + zw), 2.0f * (xz - yw)),
              mk(2.0f * (xy - zw), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + xw)),
              mk(2.0f * (xz + yw), 2.0f * (yz - xw), 1.0f - 2.0f * (xx + yy)));
}

/* Philox-4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53 and
 * 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85. */
struct U4 {
    uint32_t a, b, c, d;
};
TR_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32); }
TR_HD U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    U4 o; o.a = c0; o.b = c1; o.c = c2; o.d = c3;
    return o;
}
/* Uniform float in (0, 1): top 23 bits as an odd multiple of 2^-24. */
TR_HD float u01(uint32_t x) { return (float)((x >> 8) | 1u) * 5.9604644775390625e-08f; }

/* Counter layout: (pixel, camera sample, split<<16 | vertex<<8 | stream, sub). */
enum { STREAM_CAMERA = 0, STREAM_BSDF = 1, STREAM_LIGHT = 2, STREAM_RR = 3 };
struct Rng {
    uint32_t k0, k1, pixel, sample, split;
    TR_HD U4 draw(uint32_t vertex, uint32_t stream, uint32_t sub) const
    {
        return philox(pixel, sample, (split << 16) | (vertex << 8) | stream, sub, k0, k1);
    }
};

}  // namespace trk

#endif
