/*
 * orc_geom.h - ray/primitive intersection of the oracle.
 *
 * ORACLE: test infrastructure only.
 *  - Triangles: the watertight test of S. Woop, C. Benthin, I. Wald,
 *    "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013, with its
 *    float64 fallback when an edge function is exactly zero.  The
 *    reference's CPU path lets Embree decide hits and culls back faces
 *    with BackfaceCulling (ray-tracing.cc:1303-1307: a hit is kept iff
 *    N_ccw . d < 0, N_ccw = (v1-v0) x (v2-v0)); two-sided meshes use
 *    TwoSidedBias (ray-tracing.cc:1309-1313: hits closer than 1e-3 are
 *    dropped).  Both rules are kept here.  Barycentrics follow Embree:
 *    u weights v1, v weights v2.
 *  - Spheres: IntersectSphere (include/tempest/math/intersect.hh:163-190),
 *    plus the range check t in (tnear, tfar) that the CPU callback omits
 *    (ray-tracing.cc:865-875) and the CUDA path restores
 *    (ray-tracing.cu:834-843).
 *  - Closest hit: the smallest (t, geom, prim) in lexicographic order among
 *    hits with tnear < t < tfar, so that any correct traversal returns the
 *    same ids.
 */
#ifndef ORC_GEOM_H
#define ORC_GEOM_H

#include "orc_math.h"

namespace orc {

struct WatertightRay {
    int kx, ky, kz;
    float sx, sy, sz;
};

inline WatertightRay watertight_setup(const float dir[3])
{
    WatertightRay r;
    float ax = fabsf(dir[0]), ay = fabsf(dir[1]), az = fabsf(dir[2]);
    r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
    r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
    if (dir[r.kz] < 0.0f) { int t = r.kx; r.kx = r.ky; r.ky = t; }
    r.sx = dir[r.kx] / dir[r.kz];
    r.sy = dir[r.ky] / dir[r.kz];
    r.sz = 1.0f / dir[r.kz];
    return r;
}

/* Returns 1 and writes t, u, v on a hit in (tnear, tfar). */
inline int watertight_triangle(const WatertightRay& wr, const float org[3], const float v0[3],
                               const float v1[3], const float v2[3], int two_sided, float tnear,
                               float tfar, float* t_out, float* u_out, float* v_out)
{
    float A[3] = {v0[0] - org[0], v0[1] - org[1], v0[2] - org[2]};
    float B[3] = {v1[0] - org[0], v1[1] - org[1], v1[2] - org[2]};
    float Cc[3] = {v2[0] - org[0], v2[1] - org[1], v2[2] - org[2]};
    float Ax = A[wr.kx] - wr.sx * A[wr.kz];
    float Ay = A[wr.ky] - wr.sy * A[wr.kz];
    float Bx = B[wr.kx] - wr.sx * B[wr.kz];
    float By = B[wr.ky] - wr.sy * B[wr.kz];
    float Cx = Cc[wr.kx] - wr.sx * Cc[wr.kz];
    float Cy = Cc[wr.ky] - wr.sy * Cc[wr.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if (two_sided) {
        if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    } else {
        if (U < 0.0f || V < 0.0f || W < 0.0f) return 0;
    }
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float Az = wr.sz * A[wr.kz];
    float Bz = wr.sz * B[wr.kz];
    float Cz = wr.sz * Cc[wr.kz];
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t > tnear) || !(t < tfar)) return 0;
    if (two_sided && t < 1e-3f) return 0;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return 1;
}

/* IntersectSphere (intersect.hh:163-190) with the (tnear, tfar) range. */
inline int sphere_hit(const float org[3], const float dir[3], const float c[3], float radius,
                      float tnear, float tfar, float* t_out)
{
    float dc[3] = {c[0] - org[0], c[1] - org[1], c[2] - org[2]};
    float b = dc[0] * dir[0] + dc[1] * dir[1] + dc[2] * dir[2];
    float cc = (dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) - radius * radius;
    float D = b * b - cc;
    if (D < 0.0f || b < 0.0f) return 0;
    float sqrt_D = sqrtf(D);
    float t = b + (2.0f * (float)(b < sqrt_D) - 1.0f) * sqrt_D;
    float n[3] = {t * dir[0] - dc[0], t * dir[1] - dc[1], t * dir[2] - dc[2]};
    float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0.0f) {
        float r = 1.0f / l;
        n[0] *= r; n[1] *= r; n[2] *= r;
    }
    if (n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2] > 0.0f) return 0;
    if (!(t > tnear) || !(t < tfar)) return 0;
    *t_out = t;
    return 1;
}

/* The surface normal of a sphere hit, as IntersectSphere computes it. */
inline void sphere_normal(const float org[3], const float dir[3], const float c[3], float t, float n[3])
{
    float dc[3] = {c[0] - org[0], c[1] - org[1], c[2] - org[2]};
    n[0] = t * dir[0] - dc[0];
    n[1] = t * dir[1] - dc[1];
    n[2] = t * dir[2] - dc[2];
    float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0.0f) {
        float r = 1.0f / l;
        n[0] *= r; n[1] *= r; n[2] *= r;
    }
}

/* ---------------- user shapes ----------------
 * Restated from the reference's intersections in FP32, operation by
 * operation; each returns t, (u, v) and the geometric normal, and the
 * caller applies the (tnear, tfar) range the reference's callbacks omit.
 * Parameters per kind (precomputed in FP32 when the shape is added):
 *   rect (and blocker): f[0-2] centre, f[3-11] the columns of
 *     ToMatrix3(Orientation), f[12-13] half sizes, f[14-17] texture start
 *     and multiplier; disk: f[0-2] centre, f[3-5] normal, f[6-7] inner and
 *     outer radius; cylinder: f[0-2] centre, f[3] radius, f[4] half height,
 *     f[5-6] bending (0 for the upright one); ellipsoid: f[0-2] centre,
 *     f[3-5] scale, f[6-9] orientation, f[10-13] its conjugate, f[14-16]
 *     ToNormal(orientation). */
enum { ORC_RECT = 1, ORC_DISK = 2, ORC_CYLINDER = 3, ORC_OBLIQUE = 4, ORC_ELLIPSOID = 5 };

struct OrcShape {
    uint32_t kind;
    uint32_t tangent_map;
    float f[20];
};

inline float dot3f(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3f(const float a[3], const float b[3], float o[3])
{
    float x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}

/* IntersectRect3 (include/tempest/math/intersect.hh:105-131) */
inline int orc_rect(const float* f, const float o[3], const float d[3], float* t, float* u, float* v, float n[3])
{
    const float* N = f + 9;
    float denom = dot3f(d, N);
    if (denom > -1e-6f) return 0;
    float dist[3] = {f[0] - o[0], f[1] - o[1], f[2] - o[2]};
    float it = dot3f(dist, N) / denom;
    if (it < 0.0f) return 0;
    float pp[3];
    for (int a = 0; a < 3; ++a) pp[a] = (it * d[a] + o[a]) - f[a];
    float x = dot3f(pp, f + 3), y = dot3f(pp, f + 6);
    if (fabsf(x) >= f[12] || fabsf(y) >= f[13]) return 0;
    *t = it;
    *u = (x / f[12]) * 0.5f + 0.5f;
    *v = (y / f[13]) * 0.5f + 0.5f;
    n[0] = N[0]; n[1] = N[1]; n[2] = N[2];
    return 1;
}

/* IntersectDisk3 (intersect.hh:135-161) */
inline int orc_disk(const float* f, const float o[3], const float d[3], float* t, float n[3])
{
    const float* N = f + 3;
    float denom = dot3f(d, N);
    if (denom > -1e-6f) return 0;
    float dist[3] = {f[0] - o[0], f[1] - o[1], f[2] - o[2]};
    float it = dot3f(dist, N) / denom;
    if (it < 0.0f) return 0;
    float pp[3];
    for (int a = 0; a < 3; ++a) pp[a] = f[a] - (it * d[a] + o[a]);
    float len = sqrtf(dot3f(pp, pp));
    if (f[6] > len || len > f[7]) return 0;
    *t = it;
    n[0] = N[0]; n[1] = N[1]; n[2] = N[2];
    return 1;
}

/* CylinderShapeIntersectFunction and ObliqueCylinderIntersectFunction
 * (src/graphics/ray-tracing/ray-tracing.cc:892-960, 992-1046). */
inline int orc_cylinder(const float* f, float bx, float by, const float o[3], const float d[3], float* t_out,
                        float n[3])
{
    float cx = f[0] - o[0], cy = f[1] - o[1];
    cx += bx * (o[2] - f[2]);
    cy += by * (o[2] - f[2]);
    float ax = d[0] - bx * d[2], ay = d[1] - by * d[2];
    float b = cx * ax + cy * ay;
    float r = f[3], hh = f[4];
    float c = (cx * cx + cy * cy) - r * r;
    float a = ax * ax + ay * ay;
    float D = b * b - c * a;
    if (D < 0.0f || b < 0.0f) return 0;
    float sqrt_D = sqrtf(D);
    float t = (b + (2.0f * (float)(b < sqrt_D) - 1.0f) * sqrt_D) / a;
    float dz = f[2] - o[2];
    if (b > sqrt_D && fabsf(t * d[2] - dz) > hh) t = (b + sqrt_D) / a;
    float nv[2] = {t * ax - cx, t * ay - cy};
    float l = sqrtf(nv[0] * nv[0] + nv[1] * nv[1]);
    if (l != 0.0f) { nv[0] /= l; nv[1] /= l; }
    float nn[3] = {nv[0], nv[1], 0.0f};
    float r2 = r * r;
    float t_top = (dz + hh) / d[2];
    float tv[2] = {t_top * ax - cx, t_top * ay - cy};
    if (t_top >= 0.0f && t_top < t && tv[0] * tv[0] + tv[1] * tv[1] < r2) {
        t = t_top;
        nn[0] = 0.0f; nn[1] = 0.0f; nn[2] = 1.0f;
    }
    float t_bottom = (dz - hh) / d[2];
    float bv[2] = {t_bottom * ax - cx, t_bottom * ay - cy};
    if (t_bottom >= 0.0f && t_bottom < t && bv[0] * bv[0] + bv[1] * bv[1] < r2) {
        t = t_bottom;
        nn[0] = 0.0f; nn[1] = 0.0f; nn[2] = -1.0f;
    }
    if (fabsf(t * d[2] - dz) > hh) return 0;
    if (dot3f(nn, d) > 0.0f) return 0;
    *t_out = t;
    n[0] = nn[0]; n[1] = nn[1]; n[2] = nn[2];
    return 1;
}

/* Transform(Quaternion, Vector3) (quaternion.hh:347-350) */
inline void orc_quat_transform(const float* q, const float v[3], float out[3])
{
    float c1[3];
    cross3f(q, v, c1);
    float in[3] = {c1[0] + q[3] * v[0], c1[1] + q[3] * v[1], c1[2] + q[3] * v[2]};
    float c2[3];
    cross3f(q, in, c2);
    for (int a = 0; a < 3; ++a) out[a] = v[a] + 2.0f * c2[a];
}

/* IntersectEllipsoid (intersect.hh:192-221) */
inline int orc_ellipsoid(const float* f, const float o[3], const float d[3], float* t_out, float n[3])
{
    const float* sc = f + 3;
    float qd[3], qc[3];
    orc_quat_transform(f + 10, d, qd);
    float dc[3] = {f[0] - o[0], f[1] - o[1], f[2] - o[2]};
    orc_quat_transform(f + 10, dc, qc);
    float de[3], ce[3];
    for (int k = 0; k < 3; ++k) { de[k] = qd[k] / sc[k]; ce[k] = qc[k] / sc[k]; }
    float a = dot3f(de, de), b = dot3f(ce, de), c = dot3f(ce, ce) - 1.0f;
    float D = b * b - c * a;
    if (D < 0.0f || b < 0.0f) return 0;
    float sqrt_D = sqrtf(D);
    float t = (b + (2.0f * (float)(b < sqrt_D) - 1.0f) * sqrt_D) / a;
    float m[3];
    for (int k = 0; k < 3; ++k) m[k] = (t * de[k] - ce[k]) * sc[k];
    float nn[3];
    orc_quat_transform(f + 6, m, nn);
    float l = sqrtf(dot3f(nn, nn));
    if (l != 0.0f) {
        float rl = 1.0f / l;
        nn[0] *= rl; nn[1] *= rl; nn[2] *= rl;
    }
    if (dot3f(nn, d) > 0.0f) return 0;
    *t_out = t;
    n[0] = nn[0]; n[1] = nn[1]; n[2] = nn[2];
    return 1;
}

inline int orc_shape_hit(const OrcShape& sh, const float o[3], const float d[3], float* t, float* u, float* v,
                         float n[3])
{
    *u = 0.0f;
    *v = 0.0f;
    switch (sh.kind) {
    case ORC_RECT: return orc_rect(sh.f, o, d, t, u, v, n);
    case ORC_DISK: return orc_disk(sh.f, o, d, t, n);
    case ORC_CYLINDER: return orc_cylinder(sh.f, 0.0f, 0.0f, o, d, t, n);
    case ORC_OBLIQUE: return orc_cylinder(sh.f, sh.f[5], sh.f[6], o, d, t, n);
    case ORC_ELLIPSOID: return orc_ellipsoid(sh.f, o, d, t, n);
    default: return 0;
    }
}

inline bool hit_better(float t, uint32_t g, uint32_t p, float bt, uint32_t bg, uint32_t bp)
{
    if (t != bt) return t < bt;
    if (g != bg) return g < bg;
    return p < bp;
}

}  // namespace orc

#endif
