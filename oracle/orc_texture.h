/*
 * orc_texture.h - the oracle's texture maps and global cube map.
 *
 * ORACLE: test infrastructure only.  Restates
 *   Texture::sample<Extractor>   include/tempest/graphics/texture.hh:470-520
 *     (bilinear: texel centres at +0.5, wrapped addresses; nearest)
 *   SpectrumExtractor / RGExtractor / RGBAExtractor  texture.hh:126-330
 *     (RGBA8 spectra through SRGBToSpectrum = ConvertSLuminanceToLinear,
 *     spectrum.hh:156-162, of c * (1/255), spectrum.hh:815-819)
 *   FastFloorToInt64              include/tempest/math/functions.hh:82-88
 *   Modulo                        functions.hh:547-550
 *   SampleCubeMapTexel            src/graphics/cube-map.cc:73-105
 *   CartesianToCubeMapCoordinates include/tempest/graphics/cube-map.hh:35-77
 * All texture arithmetic is FP32, in both precision modes of the oracle.
 */
#ifndef ORC_TEXTURE_H
#define ORC_TEXTURE_H

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

#include "../include/tempest_rt.h"

namespace orc {

struct OrcTexture {
    uint32_t w = 0, h = 0, format = 0, sampling = 0;
    std::vector<uint8_t> data;
};

struct OrcTexSet {
    std::vector<OrcTexture> tex;
    float srgb[256];
    uint32_t cube_side = 0;
    std::vector<uint8_t> cube;  /* six faces, rows flipped as CubeMap stores them */
    OrcTexSet()
    {
        for (int i = 0; i < 256; ++i) {
            float c = (float)i * (1.0f / 255.0f);
            srgb[i] = c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
        }
    }
};

/* SampleSpectrum, SampleRG, SampleRGBA and sampleRGB (8-bit channels as
 * byte / 255, texture.hh:203-240) */
enum { ORC_TEX_SPECTRUM = 0, ORC_TEX_RG = 1, ORC_TEX_RGBA = 2, ORC_TEX_RGB = 3 };

inline uint32_t orc_texel_size(uint32_t fmt)
{
    static const uint32_t sizes[7] = {1, 2, 4, 4, 8, 12, 16};
    return fmt < 7 ? sizes[fmt] : 0;
}

struct Rgba {
    float v[4];
};

inline Rgba orc_extract(const OrcTexSet& ts, const OrcTexture& t, size_t idx, int kind)
{
    const uint8_t* p = t.data.data() + idx * orc_texel_size(t.format);
    Rgba r = {{0.0f, 0.0f, 0.0f, 0.0f}};
    switch (t.format) {
    case TR_TEX_RGBA8_UNORM:
        if (kind == ORC_TEX_SPECTRUM) {
            for (int c = 0; c < 3; ++c) r.v[c] = ts.srgb[p[c]];
        } else {
            for (int c = 0; c < 4; ++c) r.v[c] = p[c] / 255.0f;
        }
        break;
    case TR_TEX_RG8_UNORM:
        r.v[0] = p[0] / 255.0f;
        r.v[1] = p[1] / 255.0f;
        break;
    default: {
        uint32_t n = orc_texel_size(t.format) / 4;
        float f[4];
        memcpy(f, p, n * 4);
        for (uint32_t c = 0; c < n; ++c) r.v[c] = f[c];
    }
    }
    return r;
}

inline int64_t orc_fast_floor64(float f)
{
    if (!(f > -9.0e18f && f < 9.0e18f)) return 0;
    int64_t k = (int64_t)f;
    uint32_t bits;
    memcpy(&bits, &f, 4);
    return k - (int64_t)(bits >> 31);
}

inline int64_t orc_modulo(int64_t a, int64_t b) { return ((a % b) + b) % b; }

inline Rgba orc_bilerp(const Rgba& c00, const Rgba& c01, const Rgba& c10, const Rgba& c11, float fx0, float fx1,
                       float fy0, float fy1)
{
    Rgba r;
    for (int c = 0; c < 4; ++c)
        r.v[c] = (fx0 * c00.v[c] + fx1 * c01.v[c]) * fy0 + (fx0 * c10.v[c] + fx1 * c11.v[c]) * fy1;
    return r;
}

/* Texture::sample of the texture with 1-based handle `h` at (u, v). */
inline Rgba orc_tex_sample(const OrcTexSet& ts, uint32_t h, float u, float v, int kind)
{
    const OrcTexture& t = ts.tex[h - 1];
    int64_t w = t.w, hh = t.h;
    if (t.sampling == TR_TEX_NEAREST) {
        size_t x = (size_t)orc_modulo(orc_fast_floor64((float)w * u), w);
        size_t y = (size_t)orc_modulo(orc_fast_floor64((float)hh * v), hh);
        return orc_extract(ts, t, y * (size_t)w + x, kind);
    }
    float tx = (float)w * u - 0.5f, ty = (float)hh * v - 0.5f;
    int64_t xt = orc_fast_floor64(tx), yt = orc_fast_floor64(ty);
    size_t x0 = (size_t)orc_modulo(xt, w), y0 = (size_t)orc_modulo(yt, hh);
    size_t x1 = (x0 + 1) % (size_t)w, y1 = (y0 + 1) % (size_t)hh;
    const size_t sw = (size_t)w;
    Rgba c00 = orc_extract(ts, t, y0 * sw + x0, kind), c01 = orc_extract(ts, t, y0 * sw + x1, kind);
    Rgba c10 = orc_extract(ts, t, y1 * sw + x0, kind), c11 = orc_extract(ts, t, y1 * sw + x1, kind);
    float fx1 = tx - (float)xt, fx0 = 1.0f - fx1;
    float fy1 = ty - (float)yt, fy0 = 1.0f - fy1;
    return orc_bilerp(c00, c01, c10, c11, fx0, fx1, fy0, fy1);
}

/* CubeMap::sampleSpectrum(dir): returns RGB in out[3]. */
inline void orc_cube_sample(const OrcTexSet& ts, const float dir[3], float out[3])
{
    float ax = fabsf(dir[0]), ay = fabsf(dir[1]), az = fabsf(dir[2]);
    float s, t, m;
    size_t face;
    auto sgn = [](float x) { return x >= 0.0f ? 1.0f : -1.0f; };
    if (ax > ay) {
        if (ax > az) {
            s = -dir[2] * sgn(dir[0]); t = -dir[1]; m = ax; face = signbit(dir[0]) ? 1 : 0;
        } else {
            s = dir[0] * sgn(dir[2]); t = -dir[1]; m = az; face = 4 + (signbit(dir[2]) ? 1 : 0);
        }
    } else {
        if (ay > az) {
            s = dir[0]; t = dir[2] * sgn(dir[1]); m = ay; face = 2 + (signbit(dir[1]) ? 1 : 0);
        } else {
            s = dir[0] * sgn(dir[2]); t = -dir[1]; m = az; face = 4 + (signbit(dir[2]) ? 1 : 0);
        }
    }
    float tcx = s / m * 0.5f + 0.5f, tcy = t / m * 0.5f + 0.5f;
    int side = (int)ts.cube_side;
    float sf = (float)side * tcx - 0.5f, tf = (float)side * tcy - 0.5f;
    int sf0 = (int)sf, tf0 = (int)tf;
    sf0 = sf0 < 0 ? 0 : (sf0 > side - 1 ? side - 1 : sf0);
    tf0 = tf0 < 0 ? 0 : (tf0 > side - 1 ? side - 1 : tf0);
    int sf1 = sf0 + 1 < side - 1 ? sf0 + 1 : side - 1;
    int tf1 = tf0 + 1 < side - 1 ? tf0 + 1 : side - 1;
    const uint8_t* base = ts.cube.data() + face * (size_t)side * side * 4;
    auto px = [&](int x, int y) {
        const uint8_t* p = base + ((size_t)y * side + x) * 4;
        Rgba r = {{ts.srgb[p[0]], ts.srgb[p[1]], ts.srgb[p[2]], 0.0f}};
        return r;
    };
    float fx1 = sf - (float)sf0, fx0 = 1.0f - fx1, fy1 = tf - (float)tf0, fy0 = 1.0f - fy1;
    Rgba r = orc_bilerp(px(sf0, tf0), px(sf1, tf0), px(sf0, tf1), px(sf1, tf1), fx0, fx1, fy0, fy1);
    out[0] = r.v[0];
    out[1] = r.v[1];
    out[2] = r.v[2];
}

}  // namespace orc

#endif
