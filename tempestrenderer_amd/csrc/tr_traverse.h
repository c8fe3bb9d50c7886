/*
 * tr_traverse.h - closest-hit traversal of the 8-wide BVH with watertight
 * triangle and sphere tests.
 *
 * Closest hit = the smallest (t, geom, prim), lexicographically, over all
 * primitive hits with tnear < t < tfar.  The boxes are conservative (the
 * builder rounds outward, the slab test scales its far bound by 1 + 2^-20),
 * so any traversal order returns the same ids as a brute-force loop.
 *
 * Triangle test: S. Woop, C. Benthin, I. Wald, "Watertight Ray/Triangle
 * Intersection", JCGT 2(1), 2013, with the FP64 recomputation of zero edge
 * functions.  Culling follows the reference's Embree filters:
 * BackfaceCulling keeps a hit iff N_ccw . d < 0 (ray-tracing.cc:1303-1307);
 * TwoSidedBias drops hits closer than 1e-3 (ray-tracing.cc:1309-1313).
 * Sphere test: IntersectSphere (include/tempest/math/intersect.hh:163-190)
 * plus the (tnear, tfar) range check.  User shapes (rect, disk, cylinder,
 * oblique cylinder, ellipsoid) likewise.
 */
#ifndef TR_TRAVERSE_H
#define TR_TRAVERSE_H

#include "tr_math.h"
#include "tr_scene.h"

namespace trk {

struct Hit {
    float t, u, v;
    uint32_t geom, prim;
};

/* A ray prepared for traversal.  The watertight test's axis permutation
 * (kx, ky, kz) and the direction octant share one word: kx in bits 0-1,
 * ky in 2-3, kz in 4-5, octant in 8-10, so that the traversal loop holds
 * fewer registers. */
struct RayPre {
    float org[3], dir[3];
    float inv[3];
    float sx, sy, sz;
    uint32_t packed;
    TR_HD int kx() const { return (int)(packed & 3u); }
    TR_HD int ky() const { return (int)((packed >> 2) & 3u); }
    TR_HD int kz() const { return (int)((packed >> 4) & 3u); }
    TR_HD uint32_t octant() const { return (packed >> 8) & 7u; }
};

TR_HD RayPre ray_setup(const float org[3], const float dir[3])
{
    RayPre r;
    for (int a = 0; a < 3; ++a) {
        r.org[a] = org[a];
        r.dir[a] = dir[a];
        float d = dir[a];
        if (fabsf(d) < 1e-20f) d = d < 0.0f ? -1e-20f : 1e-20f;
        r.inv[a] = 1.0f / d;
    }
    float ax = fabsf(dir[0]), ay = fabsf(dir[1]), az = fabsf(dir[2]);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (dir[kz] < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.sx = dir[kx] / dir[kz];
    r.sy = dir[ky] / dir[kz];
    r.sz = 1.0f / dir[kz];
    const uint32_t octant = (dir[0] < 0.0f ? 1u : 0u) | (dir[1] < 0.0f ? 2u : 0u) | (dir[2] < 0.0f ? 4u : 0u);
    r.packed = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4) | (octant << 8);
    return r;
}

TR_HD float pick(const float* a, int k) { return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]); }

TR_HD bool watertight_triangle(const RayPre& r, const float* v0, const float* v1, const float* v2, bool two_sided,
                               float tnear, float tfar, float* t_out, float* u_out, float* v_out)
{
    float A[3] = {v0[0] - r.org[0], v0[1] - r.org[1], v0[2] - r.org[2]};
    float B[3] = {v1[0] - r.org[0], v1[1] - r.org[1], v1[2] - r.org[2]};
    float Cc[3] = {v2[0] - r.org[0], v2[1] - r.org[1], v2[2] - r.org[2]};
    const int kx = r.kx(), ky = r.ky(), kz = r.kz();
    float Akz = pick(A, kz), Bkz = pick(B, kz), Ckz = pick(Cc, kz);
    float Ax = pick(A, kx) - r.sx * Akz;
    float Ay = pick(A, ky) - r.sy * Akz;
    float Bx = pick(B, kx) - r.sx * Bkz;
    float By = pick(B, ky) - r.sy * Bkz;
    float Cx = pick(Cc, kx) - r.sx * Ckz;
    float Cy = pick(Cc, ky) - r.sy * Ckz;
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
        if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    } else {
        if (U < 0.0f || V < 0.0f || W < 0.0f) return false;
    }
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz;
    float Bz = r.sz * Bkz;
    float Cz = r.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t > tnear) || !(t < tfar)) return false;
    if (two_sided && t < 1e-3f) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

TR_HD bool sphere_hit(const RayPre& r, const float* c, float radius, float tnear, float tfar, float* t_out)
{
    float dc[3] = {c[0] - r.org[0], c[1] - r.org[1], c[2] - r.org[2]};
    float b = dc[0] * r.dir[0] + dc[1] * r.dir[1] + dc[2] * r.dir[2];
    float cc = (dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) - radius * radius;
    float D = b * b - cc;
    if (D < 0.0f || b < 0.0f) return false;
    float sqrt_D = sqrtf(D);
    float t = b + (2.0f * (float)(b < sqrt_D) - 1.0f) * sqrt_D;
    float n[3] = {t * r.dir[0] - dc[0], t * r.dir[1] - dc[1], t * r.dir[2] - dc[2]};
    float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0.0f) {
        float rl = 1.0f / l;
        n[0] *= rl; n[1] *= rl; n[2] *= rl;
    }
    if (n[0] * r.dir[0] + n[1] * r.dir[1] + n[2] * r.dir[2] > 0.0f) return false;
    if (!(t > tnear) || !(t < tfar)) return false;
    *t_out = t;
    return true;
}

TR_HD bool hit_better(float t, uint32_t g, uint32_t p, const Hit& b)
{
    if (b.geom == TR_INVALID_GEOMETRY) return true;
    if (t != b.t) return t < b.t;
    if (g != b.geom) return g < b.geom;
    return p < b.prim;
}

/* A grid volume's box: the ray enters at max(slab entry, tnear) if the
 * slab interval ends beyond tnear and the entry lies before tfar.  The
 * entry is the hit distance and the slab exit is returned in *t_exit (the
 * hit's u), where delta tracking stops.  A ray that starts inside enters
 * at tnear, as the reference's volume intersection does at tnear
 * (ray-tracing.cc:551-571). */
TR_HD bool volume_box_hit(const RayPre& r, const float* lo, const float* hi, float tnear, float tfar, float* t_out,
                          float* t_exit)
{
    float t0 = -INFINITY, t1 = INFINITY;
    for (int a = 0; a < 3; ++a) {
        float ta = (lo[a] - r.org[a]) * r.inv[a];
        float tb = (hi[a] - r.org[a]) * r.inv[a];
        float mn = ta < tb ? ta : tb, mx = ta < tb ? tb : ta;
        t0 = mn > t0 ? mn : t0;
        t1 = mx < t1 ? mx : t1;
    }
    if (!(t0 <= t1) || !(t1 > tnear)) return false;
    float t = t0 > tnear ? t0 : tnear;
    if (!(t < tfar)) return false;
    *t_out = t;
    *t_exit = t1;
    return true;
}

/* ---------------- user shapes (ray-tracing.cc:858-1140) ----------------
 * Each test follows its reference function operation by operation, in FP32,
 * and returns the distance, (u, v) and the geometric normal without a range
 * check; the caller applies (tnear, tfar), which the reference's Embree
 * callbacks omit. */

/* IntersectRect3 (include/tempest/math/intersect.hh:105-131) over the
 * precomputed columns of ToMatrix3(Orientation). */
TR_HD bool rect_hit(const float* f, V3 o, V3 d, float* t, float* u, float* v, V3* n)
{
    const V3 c = mk(f[0], f[1], f[2]), T = mk(f[3], f[4], f[5]), B = mk(f[6], f[7], f[8]), N = mk(f[9], f[10], f[11]);
    const float denom = dot(d, N);
    if (denom > -1e-6f) return false;
    const float it = dot(sub(c, o), N) / denom;
    if (it < 0.0f) return false;
    const V3 pp = sub(add(smul(it, d), o), c);
    const float x = dot(pp, T), y = dot(pp, B);
    if (fabsf(x) >= f[12] || fabsf(y) >= f[13]) return false;
    *t = it;
    *u = (x / f[12]) * 0.5f + 0.5f;
    *v = (y / f[13]) * 0.5f + 0.5f;
    *n = N;
    return true;
}

/* IntersectDisk3 (intersect.hh:135-161). */
TR_HD bool disk_hit(const float* f, V3 o, V3 d, float* t, V3* n)
{
    const V3 c = mk(f[0], f[1], f[2]), N = mk(f[3], f[4], f[5]);
    const float denom = dot(d, N);
    if (denom > -1e-6f) return false;
    const float it = dot(sub(c, o), N) / denom;
    if (it < 0.0f) return false;
    const float dc = length(sub(c, add(smul(it, d), o)));
    if (f[6] > dc || dc > f[7]) return false;
    *t = it;
    *n = N;
    return true;
}

/* CylinderShapeIntersectFunction / ObliqueCylinderIntersectFunction
 * (ray-tracing.cc:892-1046): the side as a 2-D quadratic in the plane
 * perpendicular to z (the oblique form with bending (bx, by); bending 0
 * gives the upright cylinder bit for bit), then the two caps. */
TR_HD bool cylinder_hit(const float* f, float bx, float by, V3 o, V3 d, float* t_out, V3* n_out)
{
    float dcx = f[0] - o.x, dcy = f[1] - o.y;
    dcx = dcx + bx * (o.z - f[2]);
    dcy = dcy + by * (o.z - f[2]);
    const float ax = d.x - bx * d.z, ay = d.y - by * d.z;
    const float b = dcx * ax + dcy * ay;
    const float r = f[3], h = f[4];
    const float cq = (dcx * dcx + dcy * dcy) - r * r;
    const float a = ax * ax + ay * ay;
    const float D = b * b - cq * a;
    if (D < 0.0f || b < 0.0f) return false;
    const float dz = f[2] - o.z;
    const float sq = sqrtf(D);
    float t = (b + (2.0f * (float)(b < sq) - 1.0f) * sq) / a;
    if (b > sq && fabsf(t * d.z - dz) > h) t = (b + sq) / a;
    float nx = t * ax - dcx, ny = t * ay - dcy;
    const float l = sqrtf(nx * nx + ny * ny);
    if (l != 0.0f) { nx = nx / l; ny = ny / l; }
    V3 n = mk(nx, ny, 0.0f);
    const float r2 = r * r;
    const float t_top = (dz + h) / d.z;
    const float ux = t_top * ax - dcx, uy = t_top * ay - dcy;
    if (t_top >= 0.0f && t_top < t && ux * ux + uy * uy < r2) { t = t_top; n = mk(0.0f, 0.0f, 1.0f); }
    const float t_bot = (dz - h) / d.z;
    const float wx = t_bot * ax - dcx, wy = t_bot * ay - dcy;
    if (t_bot >= 0.0f && t_bot < t && wx * wx + wy * wy < r2) { t = t_bot; n = mk(0.0f, 0.0f, -1.0f); }
    if (fabsf(t * d.z - dz) > h) return false;
    if (dot(n, d) > 0.0f) return false;
    *t_out = t;
    *n_out = n;
    return true;
}

/* Transform(Quaternion, Vector3) (quaternion.hh:347-350): v + 2 q x (q x v + w v). */
TR_HD V3 quat_rotate(const float* q, V3 v)
{
    const V3 qv = mk(q[0], q[1], q[2]);
    return add(v, smul(2.0f, cross(qv, add(cross(qv, v), smul(q[3], v)))));
}

/* IntersectEllipsoid (intersect.hh:192-221). */
TR_HD bool ellipsoid_hit(const float* f, V3 o, V3 d, float* t_out, V3* n_out)
{
    const V3 sc = mk(f[3], f[4], f[5]);
    const V3 qd = quat_rotate(f + 10, d), qc = quat_rotate(f + 10, sub(mk(f[0], f[1], f[2]), o));
    const V3 de = mk(qd.x / sc.x, qd.y / sc.y, qd.z / sc.z);
    const V3 ce = mk(qc.x / sc.x, qc.y / sc.y, qc.z / sc.z);
    const float a = dot(de, de), b = dot(ce, de), c = dot(ce, ce) - 1.0f;
    const float D = b * b - c * a;
    if (D < 0.0f || b < 0.0f) return false;
    const float sq = sqrtf(D);
    const float t = (b + (2.0f * (float)(b < sq) - 1.0f) * sq) / a;
    const V3 n = normalize(quat_rotate(f + 6, mulv(sub(smul(t, de), ce), sc)));
    if (dot(n, d) > 0.0f) return false;
    *t_out = t;
    *n_out = n;
    return true;
}

/* A user shape's hit: distance, (u, v) (the rect's, else 0) and normal. */
TR_HD bool shape_intersect(const ShapeRec& sh, V3 o, V3 d, float* t, float* u, float* v, V3* n)
{
    *u = 0.0f;
    *v = 0.0f;
    switch (sh.kind) {
    case SHAPE_RECT: return rect_hit(sh.f, o, d, t, u, v, n);
    case SHAPE_DISK: return disk_hit(sh.f, o, d, t, n);
    case SHAPE_CYLINDER: return cylinder_hit(sh.f, 0.0f, 0.0f, o, d, t, n);
    case SHAPE_OBLIQUE_CYLINDER: return cylinder_hit(sh.f, sh.f[5], sh.f[6], o, d, t, n);
    case SHAPE_ELLIPSOID: return ellipsoid_hit(sh.f, o, d, t, n);
    default: return false;
    }
}

/* SHAPES false leaves the user-shape intersections out of the code, for
 * scenes that have none; VOLUMES false the volume boxes. */
template <bool SHAPES = true, bool VOLUMES = true>
TR_HD void test_prim(const PrimRec& p, const ShapeRec* shapes, const RayPre& r, float tnear, float tfar, Hit* best)
{
    float t, u = 0.0f, v = 0.0f;
    bool hit;
    if (p.flags & PRIM_SPHERE) hit = sphere_hit(r, p.v0, p.v1[0], tnear, tfar, &t);
    else if (VOLUMES && (p.flags & PRIM_VOLUME)) hit = volume_box_hit(r, p.v0, p.v1, tnear, tfar, &t, &u);
    else if (SHAPES && (p.flags & PRIM_SHAPE)) {
        V3 n;
        hit = shape_intersect(shapes[tr_f2u(p.v0[0])], mk(r.org[0], r.org[1], r.org[2]), mk(r.dir[0], r.dir[1], r.dir[2]),
                              &t, &u, &v, &n) &&
              t > tnear && t < tfar;
    } else hit = watertight_triangle(r, p.v0, p.v1, p.v2, (p.flags & PRIM_TWO_SIDED) != 0, tnear, tfar, &t, &u, &v);
    if (hit && hit_better(t, p.geom, p.prim, *best)) {
        best->t = t; best->u = u; best->v = v; best->geom = p.geom; best->prim = p.prim;
    }
}

TR_HD float exp2i(int e) { return tr_u2f((uint32_t)(e + 127) << 23); }

/* A node as 32 dwords, fetched with eight 16-byte loads.  Field offsets
 * (little-endian): 0-2 origin, 3 exp[0..2] | imask << 24, 4 child_base,
 * 5 prim_base, 6-7 meta[8], 8-19 qlo[3][8], 20-31 qhi[3][8] (binary16). */
struct NodeRegs {
    uint32_t d[32];
};

TR_HD NodeRegs load_node(const Node8* nodes, uint32_t idx)
{
    NodeRegs r;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4* p = reinterpret_cast<const uint4*>(nodes + idx);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint4 v = p[k];
        r.d[4 * k + 0] = v.x;
        r.d[4 * k + 1] = v.y;
        r.d[4 * k + 2] = v.z;
        r.d[4 * k + 3] = v.w;
    }
#else
    memcpy(r.d, nodes + idx, sizeof(r.d));
#endif
    return r;
}

TR_HD uint32_t node_imask(const NodeRegs& n) { return n.d[3] >> 24; }
/* meta[s] without indexing the register array by a run-time value (which
 * would move the whole node to scratch memory) */
TR_HD uint32_t node_meta(const NodeRegs& n, uint32_t s)
{
    uint64_t m = (uint64_t)n.d[6] | ((uint64_t)n.d[7] << 32);
    return (uint32_t)(m >> (8u * s)) & 0xffu;
}

/* Binary16 code in the low or high half of a node word, as FP32.  On the
 * GPU the conversion folds into the fused multiply-add that uses it
 * (v_fma_mix_f32 with an op_sel for the half). */
TR_HD float half_lo(uint32_t w)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
#else
    return half_value((uint16_t)(w & 0xffffu));
#endif
}
TR_HD float half_hi(uint32_t w)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
#else
    return half_value((uint16_t)(w >> 16));
#endif
}

/* Slab test of the 8 children of a node; returns the hit mask in slot
 * order.  tmax is the current far bound (closest hit so far or tfar).  A
 * child plane at o + h s (h the binary16 code, s a power of two) is crossed
 * at t = h (s inv) + (o - org) inv: s inv is exact, so one fused
 * multiply-add per plane replaces the decode, subtraction and product,
 * with no more roundings than those three.  On an axis where the ray runs
 * backwards the hi plane is crossed first, so the words of the two plane
 * sets are exchanged once per node and each child takes the maximum of
 * its three near planes and the minimum of its three far planes.  Empty
 * slots carry inverted boxes and fail the test. */
TR_HD uint32_t intersect_children(const NodeRegs& n, const RayPre& r, float tnear, float tmax)
{
    const float ox = tr_u2f(n.d[0]), oy = tr_u2f(n.d[1]), oz = tr_u2f(n.d[2]);
    const float ax = exp2i((int)(int8_t)(n.d[3] & 0xffu)) * r.inv[0];
    const float ay = exp2i((int)(int8_t)((n.d[3] >> 8) & 0xffu)) * r.inv[1];
    const float az = exp2i((int)(int8_t)((n.d[3] >> 16) & 0xffu)) * r.inv[2];
    const float bx = (ox - r.org[0]) * r.inv[0];
    const float by = (oy - r.org[1]) * r.inv[1];
    const float bz = (oz - r.org[2]) * r.inv[2];
    const bool nx = r.inv[0] < 0.0f, ny = r.inv[1] < 0.0f, nz = r.inv[2] < 0.0f;
    uint32_t near_w[12], far_w[12];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int w = 0; w < 4; ++w) {
        near_w[w] = nx ? n.d[20 + w] : n.d[8 + w];
        far_w[w] = nx ? n.d[8 + w] : n.d[20 + w];
        near_w[4 + w] = ny ? n.d[24 + w] : n.d[12 + w];
        far_w[4 + w] = ny ? n.d[12 + w] : n.d[24 + w];
        near_w[8 + w] = nz ? n.d[28 + w] : n.d[16 + w];
        far_w[8 + w] = nz ? n.d[16 + w] : n.d[28 + w];
    }
    uint32_t mask = 0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (uint32_t s = 0; s < 8; ++s) {
        const uint32_t w = s >> 1;
        const bool up = (s & 1u) != 0;
        const float tx0 = fmaf(up ? half_hi(near_w[w]) : half_lo(near_w[w]), ax, bx);
        const float ty0 = fmaf(up ? half_hi(near_w[4 + w]) : half_lo(near_w[4 + w]), ay, by);
        const float tz0 = fmaf(up ? half_hi(near_w[8 + w]) : half_lo(near_w[8 + w]), az, bz);
        const float tx1 = fmaf(up ? half_hi(far_w[w]) : half_lo(far_w[w]), ax, bx);
        const float ty1 = fmaf(up ? half_hi(far_w[4 + w]) : half_lo(far_w[4 + w]), ay, by);
        const float tz1 = fmaf(up ? half_hi(far_w[8 + w]) : half_lo(far_w[8 + w]), az, bz);
        const float t0 = fmaxf(fmaxf(fmaxf(tx0, ty0), tz0), tnear);
        float t1 = fminf(fminf(fminf(tx1, ty1), tz1), tmax);
        t1 = t1 * 1.00000095f; /* 1 + 2^-20: keeps boxes touching a hit at t */
        if (t0 <= t1) mask |= 1u << s;
    }
    return mask;
}

/* Permutes a slot-order mask into visiting order: bit i <-> slot i ^ octant
 * (XOR by each octant bit swaps bits, bit pairs or nibbles). */
TR_HD uint32_t to_visit_order(uint32_t m, uint32_t oct)
{
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m & 0xAAu) >> 1);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m & 0xCCu) >> 2);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m & 0xF0u) >> 4);
    return m;
}

TR_HD uint32_t popc8(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__popc(x);
#else
    return (uint32_t)__builtin_popcount(x);
#endif
}
TR_HD uint32_t ctz32(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)(__ffs(x) - 1);
#else
    return (uint32_t)__builtin_ctz(x);
#endif
}

/* A primitive fetched with three 16-byte loads. */
TR_HD PrimRec load_prim(const PrimRec* prims, uint32_t idx)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4* p = reinterpret_cast<const uint4*>(prims + idx);
    uint4 a = p[0], b = p[1], c = p[2];
    PrimRec r;
    r.v0[0] = __uint_as_float(a.x); r.v0[1] = __uint_as_float(a.y); r.v0[2] = __uint_as_float(a.z); r.prim = a.w;
    r.v1[0] = __uint_as_float(b.x); r.v1[1] = __uint_as_float(b.y); r.v1[2] = __uint_as_float(b.z); r.geom = b.w;
    r.v2[0] = __uint_as_float(c.x); r.v2[1] = __uint_as_float(c.y); r.v2[2] = __uint_as_float(c.z); r.flags = c.w;
    return r;
#else
    return prims[idx];
#endif
}

/* Stack of node groups.  An entry is (child_base, visit-order hit mask |
 * imask << 8); its depth is bounded by the tree depth plus one.  The
 * kernels keep the first entries in LDS and spill deeper ones to global
 * memory; this host/device template takes any stack with push/pop/empty. */
template <class Stack>
TR_HD Hit traverse_closest(const Node8* nodes, const PrimRec* prims, const ShapeRec* shapes, const RayPre& r, float tnear,
                           float tfar, Stack& stack, uint32_t* node_visits, uint32_t* prim_tests)
{
    Hit best;
    best.t = tfar; best.u = 0.0f; best.v = 0.0f; best.geom = TR_INVALID_GEOMETRY; best.prim = TR_INVALID_GEOMETRY;
    uint32_t oct = r.octant();
    /* virtual root group: node 0 sits in slot 0 of an imask=1 group */
    stack.push(0u, (1u << oct) | (1u << 8));
    while (!stack.empty()) {
        uint32_t base, masks;
        stack.pop(&base, &masks);
        uint32_t vis = masks & 0xffu;
        uint32_t imask = masks >> 8;
        uint32_t i = ctz32(vis);
        vis &= vis - 1u;
        if (vis) stack.push(base, vis | (imask << 8));
        uint32_t slot = i ^ oct;
        uint32_t node_index = base + popc8(imask & ((1u << slot) - 1u));
        NodeRegs n = load_node(nodes, node_index);
        if (node_visits) ++*node_visits;
        float tmax = best.geom == TR_INVALID_GEOMETRY ? tfar : best.t;
        uint32_t hit = intersect_children(n, r, tnear, tmax);
        uint32_t nimask = node_imask(n);
        uint32_t leaves = hit & ~nimask;
        uint32_t inner = hit & nimask;
        uint32_t lv = to_visit_order(leaves, oct);
        while (lv) {
            uint32_t li = ctz32(lv);
            lv &= lv - 1u;
            uint32_t meta = node_meta(n, li ^ oct);
            uint32_t first = n.d[5] + (meta & 31u);
            uint32_t count = meta >> 5;
            for (uint32_t k = 0; k < count; ++k) {
                if (prim_tests) ++*prim_tests;
                test_prim(load_prim(prims, first + k), shapes, r, tnear, tfar, &best);
            }
        }
        if (inner) stack.push(n.d[4], to_visit_order(inner, oct) | (nimask << 8));
    }
    if (best.geom == TR_INVALID_GEOMETRY) best.t = tfar;
    return best;
}

/* A plain array stack for host-side use. */
struct HostStack {
    uint32_t base[256];
    uint32_t mask[256];
    int sp = 0;
    TR_HD void push(uint32_t b, uint32_t m) { base[sp] = b; mask[sp] = m; ++sp; }
    TR_HD void pop(uint32_t* b, uint32_t* m) { --sp; *b = base[sp]; *m = mask[sp]; }
    TR_HD bool empty() const { return sp == 0; }
};

}  // namespace trk

#endif
