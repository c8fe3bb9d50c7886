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

inline bool hit_better(float t, uint32_t g, uint32_t p, float bt, uint32_t bg, uint32_t bp)
{
    if (t != bt) return t < bt;
    if (g != bg) return g < bg;
    return p < bp;
}

}  // namespace orc

#endif
