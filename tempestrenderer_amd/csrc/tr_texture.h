/*
 * tr_texture.h - texture maps and the global cube map, shared by the
 * kernels and the host.
 *
 * Texture::sample (graphics/texture.hh:470-520): bilinear with texel
 * centres at +0.5 and wrapping addresses, or nearest; the extractors of
 * texture.hh:126-330 per use (spectrum: RGBA8 through SRGBToSpectrum, the
 * sRGB decode, which comes from a 256-entry table built on the host with
 * the reference's ConvertSLuminanceToLinear, spectrum.hh:156-162; RG and
 * RGBA: 8-bit channels divided by 255; float formats as stored).  The
 * integer floor is FastFloorToInt64 (math/functions.hh:82-88): truncation
 * minus the sign bit, so -0.0 and negative whole numbers step one further
 * down, as in the reference.
 *
 * CubeMap::sampleSpectrum (src/graphics/cube-map.cc:73-110) with
 * CartesianToCubeMapCoordinates (graphics/cube-map.hh:35-77): face by the
 * major axis, truncated (not floored) texel index, clamped bilinear.
 */
#ifndef TR_TEXTURE_H
#define TR_TEXTURE_H

#include "tr_math.h"
#include "tr_scene.h"

namespace trk {

/* Extractors: SampleSpectrum (8-bit channels through the sRGB decode),
 * SampleRG, SampleRGBA and sampleRGB (texture.hh:557-560; 8-bit channels as
 * byte / 255, texture.hh:203-240). */
enum { TEX_SPECTRUM = 0, TEX_RG = 1, TEX_RGBA = 2, TEX_RGB = 3 };

struct V4f {
    float x, y, z, w;
};

TR_HD uint32_t texel_bytes(uint32_t fmt)
{
    switch (fmt) {
    case TR_TEX_R8_UNORM: return 1;
    case TR_TEX_RG8_UNORM: return 2;
    case TR_TEX_RGBA8_UNORM: return 4;
    case TR_TEX_R32F: return 4;
    case TR_TEX_RG32F: return 8;
    case TR_TEX_RGB32F: return 12;
    case TR_TEX_RGBA32F: return 16;
    default: return 0;
    }
}

/* Whether an extractor reads a format (the others assert in the reference). */
TR_HD bool texel_kind_ok(uint32_t fmt, int kind)
{
    if (kind == TEX_SPECTRUM) return fmt == TR_TEX_RGBA8_UNORM || fmt == TR_TEX_RGB32F || fmt == TR_TEX_RGBA32F;
    if (kind == TEX_RG)
        return fmt == TR_TEX_RG8_UNORM || fmt == TR_TEX_RG32F || fmt == TR_TEX_RGB32F || fmt == TR_TEX_RGBA32F;
    if (kind == TEX_RGB) return fmt == TR_TEX_RGBA8_UNORM || fmt == TR_TEX_RGB32F || fmt == TR_TEX_RGBA32F;
    return fmt == TR_TEX_RGBA8_UNORM || fmt == TR_TEX_RGBA32F;
}

TR_HD V4f texel(const TexView& tv, const TexRec& t, uint64_t idx, int kind)
{
    const uint8_t* p = tv.texels + t.offset + idx * texel_bytes(t.format);
    V4f r = {0.0f, 0.0f, 0.0f, 0.0f};
    if (t.format == TR_TEX_RGBA8_UNORM) {
        if (kind == TEX_SPECTRUM) {
            r.x = tv.srgb[p[0]];
            r.y = tv.srgb[p[1]];
            r.z = tv.srgb[p[2]];
        } else {
            r.x = p[0] / 255.0f;
            r.y = p[1] / 255.0f;
            r.z = p[2] / 255.0f;
            r.w = p[3] / 255.0f;
        }
    } else if (t.format == TR_TEX_RG8_UNORM) {
        r.x = p[0] / 255.0f;
        r.y = p[1] / 255.0f;
    } else {
        const float* f = reinterpret_cast<const float*>(p);
        const uint32_t n = texel_bytes(t.format) / 4u;
        r.x = f[0];
        if (n > 1) r.y = f[1];
        if (n > 2) r.z = f[2];
        if (n > 3) r.w = f[3];
    }
    return r;
}

/* FastFloorToInt64, with arguments beyond the int64 range taken as 0. */
TR_HD int64_t floor_i64(float f)
{
    if (!(f > -9.0e18f && f < 9.0e18f)) return 0;
    int64_t k = (int64_t)f;
    return k - (int64_t)(tr_f2u(f) >> 31);
}

/* Modulo (math/functions.hh:547-550) */
TR_HD int64_t wrap_i64(int64_t v, int64_t n) { return ((v % n) + n) % n; }

TR_HD V4f lerp4(V4f c00, V4f c01, V4f c10, V4f c11, float fx0, float fx1, float fy0, float fy1)
{
    /* (fx0 * c00 + fx1 * c01) * fy0 + (fx0 * c10 + fx1 * c11) * fy1 */
    V4f r;
    r.x = (fx0 * c00.x + fx1 * c01.x) * fy0 + (fx0 * c10.x + fx1 * c11.x) * fy1;
    r.y = (fx0 * c00.y + fx1 * c01.y) * fy0 + (fx0 * c10.y + fx1 * c11.y) * fy1;
    r.z = (fx0 * c00.z + fx1 * c01.z) * fy0 + (fx0 * c10.z + fx1 * c11.z) * fy1;
    r.w = (fx0 * c00.w + fx1 * c01.w) * fy0 + (fx0 * c10.w + fx1 * c11.w) * fy1;
    return r;
}

/* Texture::sample<Extractor>(tc) of texture `handle` (1-based). */
TR_HD V4f tex_sample(const TexView& tv, uint32_t handle, float u, float v, int kind)
{
    const TexRec& t = tv.recs[handle - 1u];
    const int64_t w = t.width, h = t.height;
    if (t.sampling == TR_TEX_NEAREST) {
        int64_t x = wrap_i64(floor_i64((float)w * u), w), y = wrap_i64(floor_i64((float)h * v), h);
        return texel(tv, t, (uint64_t)(y * w + x), kind);
    }
    float fx = (float)w * u - 0.5f, fy = (float)h * v - 0.5f;
    int64_t xt = floor_i64(fx), yt = floor_i64(fy);
    int64_t x0 = wrap_i64(xt, w), y0 = wrap_i64(yt, h);
    int64_t x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    V4f c00 = texel(tv, t, (uint64_t)(y0 * w + x0), kind), c01 = texel(tv, t, (uint64_t)(y0 * w + x1), kind);
    V4f c10 = texel(tv, t, (uint64_t)(y1 * w + x0), kind), c11 = texel(tv, t, (uint64_t)(y1 * w + x1), kind);
    float fx1 = fx - (float)xt, fx0 = 1.0f - fx1, fy1 = fy - (float)yt, fy0 = 1.0f - fy1;
    return lerp4(c00, c01, c10, c11, fx0, fx1, fy0, fy1);
}

TR_HD float sign_of(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

/* CubeMap::sampleSpectrum(dir) */
TR_HD V3 cube_sample(const TexView& tv, V3 dir)
{
    float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    float s, t, m;
    uint32_t face;
    if (ax > ay) {
        if (ax > az) {
            s = -dir.z * sign_of(dir.x);
            t = -dir.y;
            m = ax;
            face = tr_f2u(dir.x) >> 31;
        } else {
            s = dir.x * sign_of(dir.z);
            t = -dir.y;
            m = az;
            face = 4u + (tr_f2u(dir.z) >> 31);
        }
    } else if (ay > az) {
        s = dir.x;
        t = dir.z * sign_of(dir.y);
        m = ay;
        face = 2u + (tr_f2u(dir.y) >> 31);
    } else {
        s = dir.x * sign_of(dir.z);
        t = -dir.y;
        m = az;
        face = 4u + (tr_f2u(dir.z) >> 31);
    }
    const float tcx = s / m * 0.5f + 0.5f, tcy = t / m * 0.5f + 0.5f;
    const int side = (int)tv.cube_side;
    float sf = (float)side * tcx - 0.5f, tf = (float)side * tcy - 0.5f;
    int sf0 = (int)sf, tf0 = (int)tf;
    sf0 = sf0 < 0 ? 0 : (sf0 > side - 1 ? side - 1 : sf0);
    tf0 = tf0 < 0 ? 0 : (tf0 > side - 1 ? side - 1 : tf0);
    int sf1 = sf0 + 1 < side - 1 ? sf0 + 1 : side - 1, tf1 = tf0 + 1 < side - 1 ? tf0 + 1 : side - 1;
    const uint8_t* base = tv.texels + tv.cube_offset + (uint64_t)face * side * side * 4u;
    auto fetch = [&](int x, int y) {
        const uint8_t* p = base + ((uint64_t)y * side + x) * 4u;
        V4f r = {tv.srgb[p[0]], tv.srgb[p[1]], tv.srgb[p[2]], 0.0f};
        return r;
    };
    float fx1 = sf - (float)sf0, fx0 = 1.0f - fx1, fy1 = tf - (float)tf0, fy0 = 1.0f - fy1;
    V4f r = lerp4(fetch(sf0, tf0), fetch(sf1, tf0), fetch(sf0, tf1), fetch(sf1, tf1), fx0, fx1, fy0, fy1);
    return mk(r.x, r.y, r.z);
}

}  // namespace trk

#endif
