/*
 * orc_volume.h - grid volumes of the oracle: sample lookup, fibre frame,
 * delta tracking and Woodcock shadow transmittance.
 *
 * ORACLE: test infrastructure only.  Restates, for grid volumes (a box
 * primitive holding node-centred binary16 samples), the reference's
 *   RTVolumeSet::sampleScattering   src/graphics/ray-tracing/ray-tracing.cc:245-340
 *   TrilinearInterpolation          :89-108
 *   SampleVolumePoint               :110-182
 *   tracking loops of subdraw       :2108-2123, 2259-2270
 *   Woodcock loop of EvaluateLightSources :1876-1926
 *   Matrix3::makeBasisTangent       include/tempest/math/matrix3.hh:110-120
 * (the SGGX micro-flake phase function is in orc_bsdf.h)
 * The decisions the reference makes in FP32 are made here in FP32 in the
 * same operation order (delta tracking always runs in FP32 on the FP32 ray).
 */
#ifndef ORC_VOLUME_H
#define ORC_VOLUME_H

#include "orc_bsdf.h"
#include "orc_philox.h"

#include <stdint.h>
#include <vector>

namespace orc {

/* IEEE binary16 -> binary32. */
inline float half_bits_to_float(uint16_t h)
{
    int sign = (h >> 15) & 1, ex = (h >> 10) & 31, man = h & 1023;
    float v;
    if (ex == 0) v = ldexpf((float)man, -24);
    else if (ex == 31) v = man ? NAN : INFINITY;
    else v = ldexpf((float)(1024 + man), ex - 25);
    return sign ? -v : v;
}

struct GridVolume {
    float lo[3], hi[3], scale[3];
    uint32_t dims[3];
    float rcp_rho_max, rcp_majorant, rcp_smax;
    float stddev[3];
    /* 1 for the reference's brick volumes (u8 density in rho, directions
     * decoded from the quantized angles): sampleScattering's own decision
     * arithmetic, ray-tracing.cc:316-337 */
    uint32_t brick = 0;
    std::vector<float> rho;   /* per sample */
    std::vector<float> dir;   /* 3 per sample */
};

/* Position of a point in the grid (ray-tracing.cc:283-295): the cell
 * corner indices and the fractional offset. */
template <class T> struct GridCell {
    size_t idx[8];  /* corner k: x from bit 0, y from bit 1, z from bit 2 */
    T f[3];
};

template <class T> GridCell<T> grid_cell(const GridVolume& g, V3<T> pos)
{
    GridCell<T> c;
    T p[3] = {(pos.x - (T)g.lo[0]) * (T)g.scale[0], (pos.y - (T)g.lo[1]) * (T)g.scale[1],
              (pos.z - (T)g.lo[2]) * (T)g.scale[2]};
    size_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        T dm = (T)(g.dims[a] - 1);
        T v = xmin(xmax(p[a], (T)0), dm);
        T fl = floor(v);
        lo[a] = (size_t)fl;
        hi[a] = lo[a] + 1 < g.dims[a] ? lo[a] + 1 : lo[a];
        c.f[a] = v - fl;
    }
    for (int k = 0; k < 8; ++k) {
        size_t x = (k & 1) ? hi[0] : lo[0], y = (k & 2) ? hi[1] : lo[1], z = (k & 4) ? hi[2] : lo[2];
        c.idx[k] = (z * g.dims[1] + y) * g.dims[0] + x;
    }
    return c;
}

/* TrilinearInterpolation (ray-tracing.cc:89-108): lanes (x0y0, x1y0, x0y1,
 * x1y1) of the two z layers, weights (wx * wy) * wz, horizontal sums. */
template <class T> T grid_density(const GridVolume& g, const GridCell<T>& c)
{
    T fx = c.f[0], fy = c.f[1], fz = c.f[2];
    T ix = (T)1 - fx, iy = (T)1 - fy, iz = (T)1 - fz;
    T mix[4] = {ix * iy, fx * iy, ix * fy, fx * fy};
    T lane[4];
    for (int k = 0; k < 4; ++k)
        lane[k] = (T)g.rho[c.idx[k]] * (mix[k] * iz) + (T)g.rho[c.idx[k + 4]] * (mix[k] * fz);
    return (lane[0] + lane[1]) + (lane[2] + lane[3]);
}

/* SampleVolumePoint (ray-tracing.cc:110-182): structure tensor with the
 * reference's corner factors (v3factor[0] = f for the lower corner, [1] =
 * 1 - f for the upper), eight power iterations from (1,1,1)/sqrt(3).  A
 * zero tangent (the reference's basis would be NaN) becomes +z. */
template <class T> V3<T> grid_orientation(const GridVolume& g, const GridCell<T>& c)
{
    V3<T> fac[2] = {v3(c.f[0], c.f[1], c.f[2]), v3((T)1 - c.f[0], (T)1 - c.f[1], (T)1 - c.f[2])};
    T t00 = 0, t01 = 0, t02 = 0, t11 = 0, t12 = 0, t22 = 0;
    for (int k = 0; k < 8; ++k) {
        T factor = fac[(k & 4) >> 2].z * fac[(k & 2) >> 1].y * fac[k & 1].x;
        const float* d = &g.dir[c.idx[k] * 3];
        T d0 = (T)d[0], d1 = (T)d[1], d2 = (T)d[2];
        t00 += factor * d0 * d0;
        t01 += factor * d0 * d1;
        t02 += factor * d0 * d2;
        t11 += factor * d1 * d1;
        t12 += factor * d1 * d2;
        t22 += factor * d2 * d2;
    }
    V3<T> tan = v3((T)0, (T)0, (T)0);
    if (t00 != 0 || t01 != 0 || t02 != 0 || t11 != 0 || t12 != 0 || t22 != 0) {
        T r = (T)0.5773502691f;
        tan = v3(r, r, r);
        for (int i = 0; i < 8; ++i) {
            tan = v3(t00 * tan.x + t01 * tan.y + t02 * tan.z, t01 * tan.x + t11 * tan.y + t12 * tan.z,
                     t02 * tan.x + t12 * tan.y + t22 * tan.z);
            T l = length(tan);
            if (l != (T)0) {
                T rl = (T)1 / l;
                tan = v3(tan.x * rl, tan.y * rl, tan.z * rl);
            }
        }
    }
    if (tan.x == 0 && tan.y == 0 && tan.z == 0) tan = v3((T)0, (T)0, (T)1);
    return tan;
}

/* Matrix3::makeBasisTangent (matrix3.hh:110-120) */
template <class T> M3<T> make_basis_tangent(V3<T> tan)
{
    V3<T> d0 = cross(v3((T)0, (T)1, (T)0), tan), d1 = cross(v3((T)0, (T)0, (T)1), tan);
    T d0_len = dot(d0, d0), d1_len = dot(d1, d1);
    V3<T> b = d0_len > d1_len ? d0 / xsqrt(d0_len) : d1 / xsqrt(d1_len);
    return m3(tan, b, cross(tan, b));
}

template <class T> M3<T> grid_frame(const GridVolume& g, V3<T> pos)
{
    return make_basis_tangent(grid_orientation(g, grid_cell(g, pos)));
}

/* sampleScattering's decision for one tentative collision: 0 rejected,
 * 1 accepted, 2 zero density (returns false). */
inline int grid_accept(const GridVolume& g, V3<float> pos, V3<float> dir, float threshold)
{
    GridCell<float> c = grid_cell(g, pos);
    float rho = grid_density(g, c);
    if (rho == 0.0f) return 2;
    if (g.brick) {
        /* max_density = density / 255 against the threshold, then
         * TotalDensity = (DensityScale / 255) * density * dir_density times
         * RcpMaxExtinction (the tracking loops) */
        if (rho / 255.0f < threshold) return 0;
        M3<float> basis = make_basis_tangent(grid_orientation(g, c));
        float sigma = sggx_projected_area(v3(g.stddev[0], g.stddev[1], g.stddev[2]), basis, dir);
        float total = (100.0f / 255.0f) * rho * sigma;
        return total * g.rcp_majorant >= threshold ? 1 : 0;
    }
    float dens = rho * g.rcp_rho_max;
    if (dens < threshold) return 0;
    M3<float> basis = make_basis_tangent(grid_orientation(g, c));
    float sigma = sggx_projected_area(v3(g.stddev[0], g.stddev[1], g.stddev[2]), basis, dir);
    return dens * sigma * g.rcp_smax >= threshold ? 1 : 0;
}

enum { ORC_STREAM_VOLUME = 4, ORC_STREAM_SHADOW_VOLUME = 5 };
const uint32_t kOrcMaxVolumeSteps = 1u << 20;
const uint32_t kOrcMaxShadowSteps = 1u << 17;
const uint32_t kOrcMaxVolumeCrossings = 256u;

/* Two uniform numbers of tentative collision k: Philox draw k/2 of the
 * stream, words (0, 1) for even k and (2, 3) for odd k. */
struct StepDraws {
    uint32_t ctr[4];
    uint32_t key[2];
    void get(uint32_t sub_hi, uint32_t k, float* free_flight, float* accept) const
    {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], sub_hi | (k >> 1)};
        uint32_t out[4];
        orc_philox4x32_10(c, key, out);
        *free_flight = orc_u01(out[(k & 1) * 2]);
        *accept = orc_u01(out[(k & 1) * 2 + 1]);
    }
};

/* Delta tracking from the entry t0 to the exit t1 of the box along the
 * FP32 ray (ray-tracing.cc:2259-2270 in the ray parameter). */
inline bool delta_track(const GridVolume& g, V3<float> org, V3<float> dir, float t0, float t1, const StepDraws& dr,
                        uint32_t* step, float* t_out)
{
    float t = t0;
    for (;;) {
        uint32_t k = *step;
        float xa, xb;
        dr.get(0u, k, &xa, &xb);
        *step = k + 1;
        t = t - tr_logf(xa) * g.rcp_majorant;
        if (t >= t1) { *t_out = t; return false; }
        if (k + 1 >= kOrcMaxVolumeSteps) { *t_out = t; return true; }
        if (grid_accept(g, org + t * dir, dir, xb) == 1) { *t_out = t; return true; }
    }
}

/* Woodcock repetitions of one crossing (ray-tracing.cc:1889-1916); each
 * starts at max(step_through, entry). */
inline uint32_t woodcock_crossing(const GridVolume& g, V3<float> org, V3<float> dir, float t0, float t1,
                                  const StepDraws& dr, uint32_t sub_hi, uint32_t* step, float* step_through)
{
    uint32_t success = 0;
    for (int rep = 0; rep < 2; ++rep) {
        float sd = xmax(*step_through, t0);
        for (;;) {
            uint32_t k = *step;
            if (k >= kOrcMaxShadowSteps) break;
            float xa, xb;
            dr.get(sub_hi, k, &xa, &xb);
            *step = k + 1;
            sd = sd - tr_logf(xa) * g.rcp_majorant;
            if (sd >= t1) { *step_through = sd; ++success; break; }
            int a = grid_accept(g, org + sd * dir, dir, xb);
            if (a == 2) { *step_through = sd; ++success; break; }
            if (a == 1) break;
        }
    }
    return success;
}

}  // namespace orc

#endif
