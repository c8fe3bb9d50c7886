/*
 * orc_math.h - vector, basis and elementary-function helpers of the oracle.
 *
 * ORACLE: test infrastructure only.  Operation order follows the reference
 * math headers so that the FP32 instantiation rounds like the reference:
 *  - Dot/Cross/Length: include/tempest/math/vector3.hh:72-104
 *  - Normalize (multiply by the reciprocal length): vector3.hh:78-88, 215-218
 *  - vector / scalar as multiplication by the reciprocal: vector3.hh:185-199
 *  - Matrix3 makeBasis / makeBasisOrthogonalize / transform /
 *    transformRotationInverse / transformCovariance / determinant:
 *    include/tempest/math/matrix3.hh:99-227
 *  - Matrix3 * Matrix3: matrix3.hh:355-360
 *  - quaternion -> Matrix3 and quaternion normalisation:
 *    include/tempest/math/quaternion.hh:122-142, 253-270
 * Elementary functions: FP32 uses tr_fmath.h (shared with the GPU so both
 * take identical decisions); FP64 uses libm.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>
#include <stdint.h>
#include "../tempestrenderer_amd/csrc/tr_fmath.h"

namespace orc {

/* MathPi of the reference is truncated (functions.hh:55); kept. */
template <class T> struct C {
    static constexpr T pi = (T)3.141592f;
    static constexpr T tau = (T)6.283185f;
};

inline void sincos_turns(float t, float* s, float* c) { tr_sincos_turns(t, s, c); }
inline void sincos_turns(double t, double* s, double* c)
{
    double a = 2.0 * M_PI * t;
    *s = sin(a);
    *c = cos(a);
}
inline float xlog(float x) { return tr_logf(x); }
inline double xlog(double x) { return log(x); }
inline float xexp(float x) { return tr_expf(x); }
inline double xexp(double x) { return exp(x); }
inline float xsqrt(float x) { return sqrtf(x); }
inline double xsqrt(double x) { return sqrt(x); }
inline float xabs(float x) { return fabsf(x); }
inline double xabs(double x) { return fabs(x); }
template <class T> inline T xmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T xmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T pow5(T x) { T x2 = x * x; return x2 * x2 * x; }

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator-(V3<T> a) { return v3(-a.x, -a.y, -a.z); }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return v3(s * a.x, s * a.y, s * a.z); }
/* vector / scalar multiplies by the reciprocal (vector3.hh:185-199) */
template <class T> inline V3<T> operator/(V3<T> a, T s) { T r = (T)1 / s; return v3(a.x * r, a.y * r, a.z * r); }
template <class T> inline V3<T>& operator+=(V3<T>& a, V3<T> b) { a = a + b; return a; }
template <class T> inline V3<T>& operator*=(V3<T>& a, V3<T> b) { a = a * b; return a; }
template <class T> inline V3<T>& operator*=(V3<T>& a, T s) { a = a * s; return a; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T length(V3<T> a) { return xsqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a)
{
    T l = length(a);
    return l != (T)0 ? a / l : a;
}
template <class T> inline T max_value(V3<T> a) { return xmax(xmax(a.x, a.y), a.z); }
/* Reflect (vector3.hh:420) */
template <class T> inline V3<T> reflect(V3<T> inc, V3<T> n) { return ((T)2 * dot(n, inc)) * n - inc; }

template <class T> struct M3 {
    V3<T> c0, c1, c2;
    T e(int col, int row) const
    {
        const V3<T>& c = col == 0 ? c0 : (col == 1 ? c1 : c2);
        return row == 0 ? c.x : (row == 1 ? c.y : c.z);
    }
};
template <class T> inline M3<T> m3(V3<T> a, V3<T> b, V3<T> c) { M3<T> m; m.c0 = a; m.c1 = b; m.c2 = c; return m; }

template <class T> inline M3<T> make_basis(V3<T> n)
{
    V3<T> d0 = cross(v3((T)1, (T)0, (T)0), n), d1 = cross(v3((T)0, (T)1, (T)0), n);
    T d0_len = dot(d0, d0), d1_len = dot(d1, d1);
    M3<T> m;
    m.c0 = d0_len > d1_len ? d0 / xsqrt(d0_len) : d1 / xsqrt(d1_len);
    m.c1 = cross(n, m.c0);
    m.c2 = n;
    return m;
}
template <class T> inline M3<T> make_basis_orthogonalize(V3<T> tan, V3<T> n)
{
    V3<T> binorm = normalize(cross(n, tan));
    V3<T> ortho_tan = cross(binorm, n);
    return m3(ortho_tan, binorm, n);
}
template <class T> inline V3<T> transform(const M3<T>& m, V3<T> v) { return v.x * m.c0 + v.y * m.c1 + v.z * m.c2; }
template <class T> inline V3<T> transform_rotation_inverse(const M3<T>& m, V3<T> v)
{
    return v3(dot(v, m.c0), dot(v, m.c1), dot(v, m.c2));
}
template <class T> inline M3<T> transform_covariance(const M3<T>& m, V3<T> var)
{
    V3<T> t = m.c0 * var.x * m.c0.x + m.c1 * var.y * m.c1.x + m.c2 * var.z * m.c2.x;
    V3<T> b = m.c0 * var.x * m.c0.y + m.c1 * var.y * m.c1.y + m.c2 * var.z * m.c2.y;
    V3<T> n = m.c0 * var.x * m.c0.z + m.c1 * var.y * m.c1.z + m.c2 * var.z * m.c2.z;
    return m3(t, b, n);
}
template <class T> inline T determinant(const M3<T>& m)
{
    return m.e(0, 0) * (m.e(1, 1) * m.e(2, 2) - m.e(1, 2) * m.e(2, 1)) -
           m.e(1, 0) * (m.e(0, 1) * m.e(2, 2) - m.e(0, 2) * m.e(2, 1)) +
           m.e(2, 0) * (m.e(0, 1) * m.e(1, 2) - m.e(0, 2) * m.e(1, 1));
}
template <class T> inline M3<T> mul(const M3<T>& l, const M3<T>& r)
{
    return m3(transform(l, r.c0), transform(l, r.c1), transform(l, r.c2));
}
/* quaternion (x, y, z, w) -> rotation matrix, quaternion.hh:122-142 */
template <class T> inline M3<T> quat_to_m3(const T q[4])
{
    T x = q[0], y = q[1], z = q[2], w = q[3];
    T xx = x * x, xy = x * y, xz = x * z, xw = x * w, yy = y * y, yz = y * z, yw = y * w, zz = z * z,
      zw = z * w;
    return m3(v3((T)1 - (T)2 * (yy + zz), (T)2 * (xy + zw), (T)2 * (xz - yw)),
              v3((T)2 * (xy - zw), (T)1 - (T)2 * (xx + zz), (T)2 * (yz + xw)),
              v3((T)2 * (xz + yw), (T)2 * (yz - xw), (T)1 - (T)2 * (xx + yy)));
}
/* quaternion normalisation by division, quaternion.hh:253-270 */
template <class T> inline void quat_normalize(const T q[4], T out[4])
{
    T l = xsqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (l == (T)0) { out[0] = out[1] = out[2] = out[3] = (T)0; return; }
    for (int i = 0; i < 4; ++i) out[i] = q[i] / l;
}

}  // namespace orc

#endif
