/*
 * orc_philox.h - Philox-4x32-10 counter-based generator of the oracle.
 *
 * ORACLE: test infrastructure only.  Written from the description in
 * J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11, section 3.3: 4x32 state, two
 * 32x32->64 multiplications per round with M0 = 0xD2511F53 and
 * M1 = 0xCD9E8D57, key bumped by the Weyl constants 0x9E3779B9 and
 * 0xBB67AE85 between the 10 rounds.
 *
 * The reference renderer has no counterpart: it draws from a 15-bit LCG
 * (include/tempest/math/functions.hh:214-218) seeded per thread
 * (src/graphics/ray-tracing/ray-tracing.cc:1987-1989).  The GPU kernels
 * and this oracle both key Philox by (pixel, camera sample, split/vertex/
 * stream, sub-index), which makes every path reproducible on either side.
 */
#ifndef ORC_PHILOX_H
#define ORC_PHILOX_H

#include <stdint.h>

static inline void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                     uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* A uniform float in the open interval (0, 1): the top 23 bits of x as an
 * odd multiple of 2^-24.  Exact 0 and 1 never occur, so log(u), u/(1-u)
 * and sqrt(1-u) stay finite. */
static inline float orc_u01(uint32_t x)
{
    return (float)((x >> 8) | 1u) * 5.9604644775390625e-08f;
}

#endif
