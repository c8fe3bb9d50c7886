/*
 * tr_image.cpp - image files of an RGBA8 backbuffer.
 *
 * SaveImage (src/image/image.cc:60-80) picks the writer by the lower-cased
 * text after the last '.' of the file name (Path::extension,
 * src/utils/file-system.cc:127-135):
 *  - "tga": SaveTGAImage (src/image/tga-image.cc:247-309), an 18-byte
 *    header (no ID, no colour map, image type 2 = uncompressed true colour,
 *    origin 0/0, width and height as little-endian 16-bit, 32 bits per
 *    pixel, descriptor 0 = rows from the bottom) followed by the pixels
 *    with red and blue swapped (B, G, R, A bytes), rows in memory order;
 *  - "png": SavePNGImage (src/image/png-image.cc:249-280) hands the rows
 *    to stb_image_write last row first (PNG stores the top row first and
 *    the backbuffer's row 0 is the bottom of the picture); this writer
 *    stores the same rows in the same order as 8-bit RGBA (colour type 6)
 *    with zlib's deflate, so a decoder reads the same pixels while the
 *    compressed bytes differ from stb's;
 *  - "exr": the reference's SaveEXRImage (src/image/exr-image.cc:201-300)
 *    has no RGBA8 case and fails, so does this one;
 *  - anything else fails as unsupported.
 */
#include <zlib.h>

#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tr_host.h"

using namespace trk;

namespace {

inline void put_le16(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)(v & 0xFF);
    p[1] = (uint8_t)((v >> 8) & 0xFF);
}

inline void put_be32(std::vector<uint8_t>& o, uint32_t v)
{
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void encode_tga(uint32_t w, uint32_t h, const uint32_t* rgba, std::vector<uint8_t>& o)
{
    const size_t n = (size_t)w * h;
    o.assign(18 + n * 4, 0);
    uint8_t* hd = o.data();
    hd[2] = 2;  /* uncompressed true colour */
    put_le16(hd + 12, w);
    put_le16(hd + 14, h);
    hd[16] = 32;
    uint8_t* px = hd + 18;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t c = rgba[i];  /* R in the low byte */
        px[i * 4 + 0] = (uint8_t)(c >> 16);
        px[i * 4 + 1] = (uint8_t)(c >> 8);
        px[i * 4 + 2] = (uint8_t)c;
        px[i * 4 + 3] = (uint8_t)(c >> 24);
    }
}

void png_chunk(std::vector<uint8_t>& o, const char type[4], const uint8_t* data, size_t len)
{
    put_be32(o, (uint32_t)len);
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    if (len) o.insert(o.end(), data, data + len);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, o.data() + start, (uInt)(len + 4));
    put_be32(o, (uint32_t)crc);
}

tr_status encode_png(uint32_t w, uint32_t h, const uint32_t* rgba, std::vector<uint8_t>& o)
{
    const size_t row = (size_t)w * 4 + 1;
    std::vector<uint8_t> raw(row * h);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* dst = raw.data() + row * y;
        const uint32_t* src = rgba + (size_t)(h - 1 - y) * w;
        dst[0] = 0;  /* filter: none */
        for (uint32_t x = 0; x < w; ++x) {
            const uint32_t c = src[x];
            dst[1 + x * 4 + 0] = (uint8_t)c;
            dst[1 + x * 4 + 1] = (uint8_t)(c >> 8);
            dst[1 + x * 4 + 2] = (uint8_t)(c >> 16);
            dst[1 + x * 4 + 3] = (uint8_t)(c >> 24);
        }
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
        return fail(TR_ERROR_OUT_OF_MEMORY, "PNG deflate failed");
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    o.assign(sig, sig + 8);
    uint8_t ihdr[13];
    const uint32_t wh[2] = {w, h};
    for (int k = 0; k < 2; ++k)
        for (int b = 0; b < 4; ++b) ihdr[k * 4 + b] = (uint8_t)(wh[k] >> (24 - 8 * b));
    ihdr[8] = 8;   /* bits per channel */
    ihdr[9] = 6;   /* RGBA */
    ihdr[10] = 0;  /* deflate */
    ihdr[11] = 0;  /* adaptive filtering, every row "none" */
    ihdr[12] = 0;  /* no interlace */
    png_chunk(o, "IHDR", ihdr, 13);
    png_chunk(o, "IDAT", z.data(), zlen);
    png_chunk(o, "IEND", nullptr, 0);
    return TR_OK;
}

tr_status encode(uint32_t format, uint32_t width, uint32_t height, const uint32_t* rgba8, std::vector<uint8_t>& o)
{
    if (!rgba8) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (width == 0 || height == 0) return fail(TR_ERROR_INVALID_ARG, "empty image");
    if (format == TR_IMAGE_TGA) {
        if (width > 0xFFFF || height > 0xFFFF) return fail(TR_ERROR_INVALID_ARG, "TGA sides are 16-bit");
        encode_tga(width, height, rgba8, o);
        return TR_OK;
    }
    if (format == TR_IMAGE_PNG) {
        if (width > 0x7FFFFFFF || height > 0x7FFFFFFF) return fail(TR_ERROR_INVALID_ARG, "PNG sides are 31-bit");
        return encode_png(width, height, rgba8, o);
    }
    return fail(TR_ERROR_UNSUPPORTED, "unknown image format");
}

std::string lower_extension(const char* path)
{
    std::string p(path);
    for (int i = (int)p.size() - 1; i >= 0 && p[i] != '/'; --i)
        if (p[i] == '.') {
            std::string e = p.substr(i + 1);
            for (auto& c : e) c = (char)tolower((unsigned char)c);
            return e;
        }
    return std::string();
}

}  // namespace

extern "C" {

tr_status tr_image_encode(uint32_t format, uint32_t width, uint32_t height, const uint32_t* rgba8, uint8_t* out,
                          size_t capacity, size_t* size)
{
    if (!size) return fail(TR_ERROR_INVALID_ARG, "null argument");
    std::vector<uint8_t> o;
    tr_status st = encode(format, width, height, rgba8, o);
    if (st != TR_OK) return st;
    *size = o.size();
    if (!out) return TR_OK;
    if (capacity < o.size()) return fail(TR_ERROR_INVALID_ARG, "output buffer too small");
    memcpy(out, o.data(), o.size());
    return TR_OK;
}

tr_status tr_image_save(const char* path, uint32_t width, uint32_t height, const uint32_t* rgba8)
{
    if (!path || !rgba8) return fail(TR_ERROR_INVALID_ARG, "null argument");
    const std::string ext = lower_extension(path);
    uint32_t format;
    if (ext == "tga") format = TR_IMAGE_TGA;
    else if (ext == "png") format = TR_IMAGE_PNG;
    else if (ext == "exr") return fail(TR_ERROR_UNSUPPORTED, "EXR holds no RGBA8 images: " + std::string(path));
    else return fail(TR_ERROR_UNSUPPORTED, "unsupported file format: " + ext);
    std::vector<uint8_t> o;
    tr_status st = encode(format, width, height, rgba8, o);
    if (st != TR_OK) return st;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(TR_ERROR_IO, "cannot open for writing: " + std::string(path));
    const bool ok = fwrite(o.data(), 1, o.size(), f) == o.size();
    if (fclose(f) != 0 || !ok) return fail(TR_ERROR_IO, "write failed: " + std::string(path));
    return TR_OK;
}

}  // extern "C"
