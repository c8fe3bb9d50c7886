/*
 * tr_volume.h - grid volumes: sample lookup, fibre orientation, delta
 * tracking and the ratio-tracking shadow estimate.
 *
 * Restates for the build's grid volumes (one box primitive per volume, the
 * samples in HBM as binary16 x4 per grid node) the reference's
 *   RTVolumeSet::sampleScattering  (src/graphics/ray-tracing/ray-tracing.cc:245-340)
 *   SampleVolumePoint              (:110-182, structure tensor + power iteration)
 *   the tracking loops of subdraw  (:2108-2123, 2259-2270)
 *   the shadow-ray Woodcock loop   (:1876-1926)
 * Tracking runs in the ray parameter t of the ray that entered the volume:
 * the tentative collision is org + t * dir, and the walk leaves the volume
 * when t reaches the box exit.  A point of zero density continues the walk
 * (the reference re-traces from it, which with a box primitive returns the
 * same volume at distance 0).
 */
#ifndef TR_VOLUME_H
#define TR_VOLUME_H

#include "tr_math.h"
#include "tr_scene.h"

namespace trk {

/* binary16 -> binary32, exact. */
TR_HD float h2f(uint32_t h)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, (unsigned short)h);
#else
    uint32_t s = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        float v = (float)m * 5.9604644775390625e-08f;
        return s ? -v : v;
    }
    if (e == 31) return tr_u2f(s | 0x7f800000u | (m << 13));
    return tr_u2f(s | ((e + 112u) << 23) | (m << 13));
#endif
}

/* A volume's voxel words lie in tiles of 4 x 2 x 2 samples (16 words, one
 * 128-byte cache line; x fastest inside a tile), the tiles in x-, y-, then
 * z-order, from a voxel_base that is a multiple of 16 words.  A trilinear
 * lookup's eight samples then span 1.25 x 1.5 x 1.5 = 2.8 lines on
 * average, against 4.25 with the grid in plain x-fastest order, where the
 * y and z neighbours always lie in other lines. */
TR_HD uint64_t voxel_tile_index(uint32_t x, uint32_t y, uint32_t z, const uint32_t dims[3])
{
    const uint64_t ntx = (dims[0] + 3u) >> 2, nty = (dims[1] + 1u) >> 1;
    return ((((uint64_t)(z >> 1) * nty + (y >> 1)) * ntx + (x >> 2)) << 4) + ((z & 1u) << 3) + ((y & 1u) << 2) +
           (x & 3u);
}

/* Words of a volume's tiles (the grid padded to whole tiles). */
TR_HD uint64_t voxel_tile_words(const uint32_t dims[3])
{
    return (uint64_t)((dims[0] + 3u) >> 2) * ((dims[1] + 1u) >> 1) * ((dims[2] + 1u) >> 1) * 16u;
}

/* The eight grid samples around a point and its position in the cell. */
struct VoxelCell {
    uint64_t s[8];   /* corner k: x from bit 0, y from bit 1, z from bit 2 */
    float f[3];      /* fractional position in the cell */
};

TR_HD VoxelCell fetch_cell(const uint64_t* voxels, const VolumeRec& V, V3 pos)
{
    VoxelCell c;
    float p[3] = {(pos.x - V.lo[0]) * V.scale[0], (pos.y - V.lo[1]) * V.scale[1], (pos.z - V.lo[2]) * V.scale[2]};
    uint32_t i0[3], i1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float dm = (float)(V.dims[a] - 1u);
        float v = p[a] > 0.0f ? p[a] : 0.0f;
        v = v < dm ? v : dm;
        float fl = floorf(v);
        i0[a] = (uint32_t)fl;
        i1[a] = i0[a] + 1u < V.dims[a] ? i0[a] + 1u : i0[a];
        c.f[a] = v - fl;
    }
    const uint64_t* base = voxels + V.voxel_base;
    /* voxel_tile_index, split into its x, y and z terms */
    const uint32_t ntx = (V.dims[0] + 3u) >> 2, nty = (V.dims[1] + 1u) >> 1;
    const uint64_t ox[2] = {((uint64_t)(i0[0] >> 2) << 4) + (i0[0] & 3u), ((uint64_t)(i1[0] >> 2) << 4) + (i1[0] & 3u)};
    const uint64_t oy[2] = {((uint64_t)(i0[1] >> 1) * ntx << 4) + ((i0[1] & 1u) << 2),
                            ((uint64_t)(i1[1] >> 1) * ntx << 4) + ((i1[1] & 1u) << 2)};
    const uint64_t oz[2] = {((uint64_t)(i0[2] >> 1) * nty * ntx << 4) + ((i0[2] & 1u) << 3),
                            ((uint64_t)(i1[2] >> 1) * nty * ntx << 4) + ((i1[2] & 1u) << 3)};
#pragma unroll
    for (int k = 0; k < 8; ++k) c.s[k] = base[oz[(k >> 2) & 1] + oy[(k >> 1) & 1] + ox[k & 1]];
    return c;
}

/* Trilinear density in the reference's evaluation order
 * (TrilinearInterpolation, ray-tracing.cc:89-108): corner weights
 * (wx * wy) * wz, the two z layers summed per (x, y) corner, then the four
 * corner sums pairwise. */
TR_HD float cell_density(const VoxelCell& c)
{
    float fx = c.f[0], fy = c.f[1], fz = c.f[2];
    float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    float wxy[4] = {gx * gy, fx * gy, gx * fy, fx * fy};
    float lane[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float r0 = h2f((uint32_t)(c.s[k] >> 48)), r1 = h2f((uint32_t)(c.s[k + 4] >> 48));
        lane[k] = r0 * (wxy[k] * gz) + r1 * (wxy[k] * fz);
    }
    return (lane[0] + lane[1]) + (lane[2] + lane[3]);
}

/* Fibre direction at the point: dominant eigenvector of the structure
 * tensor sum_k w_k d_k d_k^T by eight power iterations from (1,1,1)/sqrt(3)
 * (SampleVolumePoint, ray-tracing.cc:110-182).  The corner weights are the
 * reference's: the lower corner of an axis weighs f, the upper 1 - f.  A
 * zero result (zero tensor) falls back to +z. */
template <bool QUANT>
TR_HD V3 cell_orientation(const VoxelCell& c, const uint64_t* voxels, const VolumeRec& V)
{
    float lf[3] = {c.f[0], c.f[1], c.f[2]};
    float uf[3] = {1.0f - c.f[0], 1.0f - c.f[1], 1.0f - c.f[2]};
    float t00 = 0.0f, t01 = 0.0f, t02 = 0.0f, t11 = 0.0f, t12 = 0.0f, t22 = 0.0f;
    const float* tab = reinterpret_cast<const float*>(voxels + V.table_base);
    const bool quant = QUANT;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float w = ((k & 4) ? uf[2] : lf[2]) * ((k & 2) ? uf[1] : lf[1]) * ((k & 1) ? uf[0] : lf[0]);
        uint64_t s = c.s[k];
        float d0, d1, d2;
        if (quant) {
            /* d = (cos phi sin theta, sin phi sin theta, cos theta) from the
             * quantized angles (ray-tracing.cc:133-144) */
            uint32_t th = (uint32_t)(s & 0xffu), ph = (uint32_t)((s >> 8) & 0xffu);
            float sin_t = tab[768 + th];
            d0 = tab[ph] * sin_t;
            d1 = tab[256 + ph] * sin_t;
            d2 = tab[512 + th];
        } else {
            d0 = h2f((uint32_t)(s & 0xffffu));
            d1 = h2f((uint32_t)((s >> 16) & 0xffffu));
            d2 = h2f((uint32_t)((s >> 32) & 0xffffu));
        }
        t00 += w * d0 * d0;
        t01 += w * d0 * d1;
        t02 += w * d0 * d2;
        t11 += w * d1 * d1;
        t12 += w * d1 * d2;
        t22 += w * d2 * d2;
    }
    V3 v = mk(0.0f, 0.0f, 0.0f);
    if (t00 != 0.0f || t01 != 0.0f || t02 != 0.0f || t11 != 0.0f || t12 != 0.0f || t22 != 0.0f) {
        const float r3 = 0.5773502691f;
        v = mk(r3, r3, r3);
        for (int i = 0; i < 8; ++i) {
            v = mk(t00 * v.x + t01 * v.y + t02 * v.z, t01 * v.x + t11 * v.y + t12 * v.z,
                   t02 * v.x + t12 * v.y + t22 * v.z);
            float l = length(v);
            if (l != 0.0f) {
                float rl = 1.0f / l;
                v = mk(v.x * rl, v.y * rl, v.z * rl);
            }
        }
    }
    if (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f) v = mk(0.0f, 0.0f, 1.0f);
    return v;
}

/* Position along a ray, the same expression everywhere. */
TR_HD V3 ray_point(V3 org, V3 dir, float t) { return add(org, smul(t, dir)); }

/* The fibre frame at a point and sigma(dir) of the SGGX micro-flakes.  The
 * voxel format is a template argument so that each walk loop carries only
 * its own decoding (the registers of the two do not add up). */
template <bool QUANT> TR_HD M3 volume_frame(const VoxelCell& c, const uint64_t* voxels, const VolumeRec& V)
{
    return make_basis_tangent(cell_orientation<QUANT>(c, voxels, V));
}
TR_HD M3 volume_frame_any(const VoxelCell& c, const uint64_t* voxels, const VolumeRec& V)
{
    return V.format == kVoxelQuantized ? volume_frame<true>(c, voxels, V) : volume_frame<false>(c, voxels, V);
}

/* One tentative collision: 0 = rejected, 1 = accepted (scatter / absorb),
 * 2 = zero density (the reference's sampleScattering returns false).
 * Brick volumes follow sampleScattering's own arithmetic (ray-tracing.cc:
 * 316-337 and the loops' TotalDensity *= RcpMaxExtinction): u8 density / 255
 * against the threshold, then (DensityScale / 255) * density * sigma *
 * RcpMaxExtinction. */
/* orient (if given) receives the cell's fibre orientation when the
 * collision is accepted (return 1): the frame of volume_frame is
 * make_basis_tangent of it. */
template <bool QUANT>
TR_HD int accept_cell(const VoxelCell& c, const uint64_t* voxels, const VolumeRec& V, V3 dir, float threshold,
                      V3* orient = nullptr)
{
    float rho = cell_density(c);
    if (rho == 0.0f) return 2;
    const bool quant = QUANT;
    float dens = quant ? rho / 255.0f : rho * V.rcp_rho_max;
    if (dens < threshold) return 0;
    const V3 o = cell_orientation<QUANT>(c, voxels, V);
    M3 basis = make_basis_tangent(o);
    float sigma = length(mulv(transform_rotation_inverse(basis, dir), ld3(V.stddev)));
    const bool acc = quant ? (kDensityScale / 255.0f) * rho * sigma * V.rcp_majorant >= threshold
                           : dens * sigma * V.rcp_smax >= threshold;
    if (acc && orient) *orient = o;
    return acc ? 1 : 0;
}
template <bool QUANT>
TR_HD int volume_accept(const uint64_t* voxels, const VolumeRec& V, V3 pos, V3 dir, float threshold)
{
    return accept_cell<QUANT>(fetch_cell(voxels, V, pos), voxels, V, dir, threshold);
}

/* Tracking budget: a walk that makes this many tentative collisions in one
 * vertex scatters at the last one (the step length is 1 / majorant, so this
 * is only reached by rays that cross more than 2^20 mean free paths). */
constexpr uint32_t kMaxVolumeSteps = 1u << 20;
/* Shadow rays: a Woodcock walk beyond 2^17 steps counts as absorbed, and a
 * shadow ray crosses at most 256 volume boxes. */
constexpr uint32_t kMaxShadowSteps = 1u << 17;
constexpr uint32_t kMaxVolumeCrossings = 256u;

/* Philox streams of the volume walks: delta tracking of the path vertex
 * (sub = step pair) and ratio tracking of shadow rays (sub = light << 24 |
 * light sample << 16 | step pair).  Each draw covers two steps: words
 * (a, b) then (c, d) = (free-flight, acceptance). */
enum { STREAM_VOLUME = 4, STREAM_SHADOW_VOLUME = 5 };

/* Delta tracking from the volume entry t0 to its exit t1, one pair of
 * tentative collisions per call.  Collision k draws its free flight and
 * acceptance numbers from words (a, b) (k even) or (c, d) (k odd) of Philox
 * draw k >> 1.  A free flight does not depend on the previous collision's
 * outcome, so both cells of a pair are fetched together (one memory round
 * trip for two collisions; the second fetch is wasted when the first
 * collision is accepted, and clamped into the grid when it lies beyond the
 * exit), then tested in order, with the results of one at a time.  *tc is
 * the walk's position; returns 1 when a collision is accepted at *tc (or
 * the step budget is spent), -1 when the walk left the volume (*tc beyond
 * the exit), 0 when it goes on.  `step` counts tentative collisions of
 * this path vertex across volumes. */
template <bool QUANT>
TR_HD int delta_track_pair(const uint64_t* voxels, const VolumeRec& V, V3 org, V3 dir, float* tc, float t1,
                           const Rng& rng, uint32_t vertex, uint32_t* step, uint32_t* nsteps, V3* orient = nullptr)
{
    const uint32_t k = *step;
    const U4 r0 = rng.draw(vertex, STREAM_VOLUME, k >> 1);
    const uint32_t xa1 = (k & 1u) ? r0.c : r0.a, xb1 = (k & 1u) ? r0.d : r0.b;
    *step = k + 1u;
    if (nsteps) ++*nsteps;
    /* d = logf(xi) * RcpMaxExtinction; pos -= dir * d */
    const float tc1 = *tc - tr_logf(u01(xa1)) * V.rcp_majorant;
    *tc = tc1;
    if (tc1 >= t1) return -1;
    if (k + 1u >= kMaxVolumeSteps) return 1;
    const U4 r1 = (k & 1u) ? rng.draw(vertex, STREAM_VOLUME, (k + 1u) >> 1) : r0;
    const uint32_t xa2 = (k & 1u) ? r1.a : r1.c, xb2 = (k & 1u) ? r1.b : r1.d;
    const float tc2 = tc1 - tr_logf(u01(xa2)) * V.rcp_majorant;
    const VoxelCell c1 = fetch_cell(voxels, V, ray_point(org, dir, tc1));
    const VoxelCell c2 = fetch_cell(voxels, V, ray_point(org, dir, tc2));
    if (accept_cell<QUANT>(c1, voxels, V, dir, u01(xb1), orient) == 1) return 1;
    *step = k + 2u;
    if (nsteps) ++*nsteps;
    *tc = tc2;
    if (tc2 >= t1) return -1;
    if (k + 2u >= kMaxVolumeSteps) return 1;
    if (accept_cell<QUANT>(c2, voxels, V, dir, u01(xb2), orient) == 1) return 1;
    return 0;
}
TR_HD int delta_track_pair_any(const uint64_t* voxels, const VolumeRec& V, V3 org, V3 dir, float* tc, float t1,
                               const Rng& rng, uint32_t vertex, uint32_t* step, uint32_t* nsteps, V3* orient = nullptr)
{
    if (V.format == kVoxelQuantized)
        return delta_track_pair<true>(voxels, V, org, dir, tc, t1, rng, vertex, step, nsteps, orient);
    return delta_track_pair<false>(voxels, V, org, dir, tc, t1, rng, vertex, step, nsteps, orient);
}

/* The whole walk: true and the collision distance in *t when a collision
 * is accepted, false and the distance beyond the exit where the walk left
 * the volume otherwise. */
template <bool QUANT>
TR_HD bool delta_track(const uint64_t* voxels, const VolumeRec& V, V3 org, V3 dir, float t0, float t1,
                       const Rng& rng, uint32_t vertex, uint32_t* step, float* t, uint32_t* nsteps)
{
    float tc = t0;
    for (;;) {
        const int res = delta_track_pair<QUANT>(voxels, V, org, dir, &tc, t1, rng, vertex, step, nsteps);
        if (res != 0) {
            *t = tc;
            return res > 0;
        }
    }
}
TR_HD bool delta_track_any(const uint64_t* voxels, const VolumeRec& V, V3 org, V3 dir, float t0, float t1,
                           const Rng& rng, uint32_t vertex, uint32_t* step, float* t, uint32_t* nsteps)
{
    if (V.format == kVoxelQuantized) return delta_track<true>(voxels, V, org, dir, t0, t1, rng, vertex, step, t, nsteps);
    return delta_track<false>(voxels, V, org, dir, t0, t1, rng, vertex, step, t, nsteps);
}

/* The two Woodcock repetitions of a shadow ray through one volume crossing
 * (ray-tracing.cc:1889-1916).  A repetition succeeds when its walk leaves
 * the volume (or meets zero density) and records where in *step_through;
 * one that accepts a collision is absorbed.  Each repetition starts at
 * max(*step_through, t0): the reference starts at *step_through, measured
 * from the shadow-ray origin, which skips a volume lying further than one
 * free path from the shading point.  Returns the number of successes. */
template <bool QUANT>
TR_HD uint32_t woodcock(const uint64_t* voxels, const VolumeRec& V, V3 org, V3 dir, float t0, float t1,
                        const Rng& rng, uint32_t vertex, uint32_t sub_hi, uint32_t* step, float* step_through,
                        uint32_t* nsteps)
{
    uint32_t success = 0;
    U4 r;
    bool have = false;
    for (int rep = 0; rep < 2; ++rep) {
        float sd = *step_through > t0 ? *step_through : t0;
        for (;;) {
            uint32_t k = *step;
            if (k >= kMaxShadowSteps) break;
            if (!have || (k & 1u) == 0u) r = rng.draw(vertex, STREAM_SHADOW_VOLUME, sub_hi | (k >> 1));
            have = true;
            uint32_t xa = (k & 1u) ? r.c : r.a, xb = (k & 1u) ? r.d : r.b;
            *step = k + 1u;
            if (nsteps) ++*nsteps;
            sd = sd - tr_logf(u01(xa)) * V.rcp_majorant;
            if (sd >= t1) {
                *step_through = sd;
                ++success;
                break;
            }
            int acc = volume_accept<QUANT>(voxels, V, ray_point(org, dir, sd), dir, u01(xb));
            if (acc == 2) {
                *step_through = sd;
                ++success;
                break;
            }
            if (acc == 1) break;
        }
    }
    return success;
}
TR_HD uint32_t woodcock_any(const uint64_t* voxels, const VolumeRec& V, V3 org, V3 dir, float t0, float t1,
                            const Rng& rng, uint32_t vertex, uint32_t sub_hi, uint32_t* step, float* step_through,
                            uint32_t* nsteps)
{
    if (V.format == kVoxelQuantized)
        return woodcock<true>(voxels, V, org, dir, t0, t1, rng, vertex, sub_hi, step, step_through, nsteps);
    return woodcock<false>(voxels, V, org, dir, t0, t1, rng, vertex, sub_hi, step, step_through, nsteps);
}

}  // namespace trk

#endif
