/*
 * tr_image.cpp - image files: the backbuffer writers and the texture
 * readers.
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
 *
 * LoadImage (src/image/image.cc:36-58) reads textures the same way:
 *  - "tga": LoadTGAImage (src/image/tga-image.cc:98-242), uncompressed or
 *    run-length true colour of 24 or 32 bits into RGBA8 (B and R swapped
 *    back), rows flipped when the descriptor's top bit is set and columns
 *    when its right bit is; 24-bit pixels get alpha 255 (the reference
 *    leaves that byte uninitialised) and the flips use the image's own
 *    width and height (the reference's index both by the height, right only
 *    for square images); grey-scale TGA files, whose red/blue swap runs
 *    across pixels and past the buffer in the reference, are rejected;
 *  - "png": LoadPNGImage (src/image/png-image.cc:125-246) with its libpng
 *    transforms: grey below 8 bits expanded, palettes to RGB, tRNS to
 *    alpha, RGB given alpha 255; 8-bit grey becomes an R8 texture, grey with
 *    alpha RGBA8 (grey in R, G, B); rows top-down as stored.  16-bit and
 *    interlaced files are rejected (they would need RGBA16 textures and
 *    the Adam7 passes).
 */
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
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

/* ---------------- readers ---------------- */

struct ImgError {
    std::string msg;
};

bool read_all(const char* path, std::vector<uint8_t>* out)
{
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t r;
    out->clear();
    while ((r = fread(buf, 1, sizeof(buf), f)) > 0) out->insert(out->end(), buf, buf + r);
    fclose(f);
    return true;
}

inline uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
inline uint32_t le16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }

void load_tga(const std::vector<uint8_t>& f, uint32_t* w_out, uint32_t* h_out, std::vector<uint8_t>* px)
{
    if (f.size() < 18) throw ImgError{"TGA header truncated"};
    const uint32_t id_len = f[0], cmap = f[1], type = f[2];
    const uint32_t w = le16(&f[12]), h = le16(&f[14]), depth = f[16], desc = f[17];
    if (type == 0) throw ImgError{"TGA without image data"};
    if (type != 2 && type != 10) throw ImgError{"unsupported TGA image type (true colour only)"};
    if (cmap || (depth != 24 && depth != 32)) throw ImgError{"unsupported TGA pixel depth"};
    if (w == 0 || h == 0) throw ImgError{"empty TGA image"};
    const size_t bpp = depth / 8, n = (size_t)w * h;
    size_t pos = 18 + id_len;
    px->assign(n * 4, 0);
    auto take = [&](uint8_t* dst) {
        if (pos + bpp > f.size()) throw ImgError{"TGA data truncated"};
        memcpy(dst, &f[pos], bpp);
        if (bpp == 3) dst[3] = 255;
        pos += bpp;
    };
    if (type == 10) {
        size_t i = 0;
        while (i < n) {
            if (pos >= f.size()) throw ImgError{"TGA data truncated"};
            const uint32_t hdr = f[pos++];
            const size_t count = (hdr & 0x7f) + 1;
            if (i + count > n) throw ImgError{"TGA run beyond the image"};
            if (hdr & 0x80) {
                take(&(*px)[i * 4]);
                for (size_t k = 1; k < count; ++k) memcpy(&(*px)[(i + k) * 4], &(*px)[i * 4], 4);
            } else {
                for (size_t k = 0; k < count; ++k) take(&(*px)[(i + k) * 4]);
            }
            i += count;
        }
    } else {
        for (size_t i = 0; i < n; ++i) take(&(*px)[i * 4]);
    }
    for (size_t i = 0; i < n; ++i) std::swap((*px)[i * 4], (*px)[i * 4 + 2]);
    if (desc & 0x10)  /* right-to-left columns */
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0, rx = w - 1; x < w / 2; ++x, --rx)
                for (int k = 0; k < 4; ++k) std::swap((*px)[((size_t)y * w + x) * 4 + k], (*px)[((size_t)y * w + rx) * 4 + k]);
    if (desc & 0x20)  /* top-to-bottom rows */
        for (uint32_t y = 0, ry = h - 1; y < h / 2; ++y, --ry)
            for (uint32_t x = 0; x < w; ++x)
                for (int k = 0; k < 4; ++k) std::swap((*px)[((size_t)y * w + x) * 4 + k], (*px)[((size_t)ry * w + x) * 4 + k]);
    *w_out = w;
    *h_out = h;
}

inline uint8_t paeth(int a, int b, int c)
{
    int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    return (uint8_t)((pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c));
}

void load_png(const std::vector<uint8_t>& f, uint32_t* w_out, uint32_t* h_out, uint32_t* fmt_out,
              std::vector<uint8_t>* px)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (f.size() < 8 || memcmp(f.data(), sig, 8) != 0) throw ImgError{"not a PNG file"};
    size_t pos = 8;
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool have_ihdr = false, have_iend = false;
    while (pos + 12 <= f.size() && !have_iend) {
        const uint32_t len = be32(&f[pos]);
        if (pos + 12 + (size_t)len > f.size()) throw ImgError{"PNG chunk truncated"};
        const uint8_t* type = &f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        uLong crc = crc32(0L, Z_NULL, 0);
        crc = crc32(crc, type, (uInt)(len + 4));
        if ((uint32_t)crc != be32(d + len)) throw ImgError{"PNG chunk CRC mismatch"};
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13) throw ImgError{"bad PNG header"};
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (d[10] != 0 || d[11] != 0) throw ImgError{"unknown PNG compression or filter method"};
            interlace = d[12];
            have_ihdr = true;
        } else if (!memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            have_iend = true;
        } else if (!(type[0] & 0x20)) {
            throw ImgError{"unknown critical PNG chunk"};
        }
        pos += 12 + len;
    }
    if (!have_ihdr || !have_iend) throw ImgError{"PNG without header or end"};
    if (w == 0 || h == 0 || w > 65535 || h > 65535) throw ImgError{"PNG size outside 1..65535"};
    if (interlace) throw ImgError{"interlaced PNG files are not read"};
    if (depth == 16) throw ImgError{"16-bit PNG files are not read"};
    uint32_t ch;
    switch (ctype) {
    case 0: ch = 1; if (depth != 1 && depth != 2 && depth != 4 && depth != 8) throw ImgError{"bad PNG grey depth"}; break;
    case 2: ch = 3; if (depth != 8) throw ImgError{"bad PNG RGB depth"}; break;
    case 3: ch = 1; if (depth != 1 && depth != 2 && depth != 4 && depth != 8) throw ImgError{"bad PNG palette depth"};
            if (plte.empty() || plte.size() % 3) throw ImgError{"PNG palette missing"}; break;
    case 4: ch = 2; if (depth != 8) throw ImgError{"bad PNG grey-alpha depth"}; break;
    case 6: ch = 4; if (depth != 8) throw ImgError{"bad PNG RGBA depth"}; break;
    default: throw ImgError{"unknown PNG colour type"};
    }
    const size_t rowbytes = ((size_t)w * ch * depth + 7) / 8, bpp = std::max<size_t>(1, ch * depth / 8);
    std::vector<uint8_t> raw((rowbytes + 1) * h);
    uLongf rlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rlen, idat.data(), (uLong)idat.size()) != Z_OK || rlen != raw.size())
        throw ImgError{"PNG image data does not inflate to the image size"};
    std::vector<uint8_t> img(rowbytes * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (rowbytes + 1)];
        const uint8_t* src = &raw[y * (rowbytes + 1) + 1];
        uint8_t* dst = &img[y * rowbytes];
        const uint8_t* up = y ? &img[(y - 1) * rowbytes] : nullptr;
        for (size_t i = 0; i < rowbytes; ++i) {
            int a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            uint8_t v = src[i];
            switch (ft) {
            case 0: break;
            case 1: v = (uint8_t)(v + a); break;
            case 2: v = (uint8_t)(v + b); break;
            case 3: v = (uint8_t)(v + ((a + b) >> 1)); break;
            case 4: v = (uint8_t)(v + paeth(a, b, c)); break;
            default: throw ImgError{"unknown PNG row filter"};
            }
            dst[i] = v;
        }
    }
    auto sample = [&](uint32_t y, uint32_t x) -> uint32_t {  /* one sub-byte or byte sample */
        if (depth == 8) return img[y * rowbytes + x];
        const size_t bit = (size_t)x * depth;
        const uint32_t byte = img[y * rowbytes + bit / 8];
        return (byte >> (8 - depth - (bit % 8))) & ((1u << depth) - 1);
    };
    const size_t n = (size_t)w * h;
    if (ctype == 0 && trns.empty()) {
        px->resize(n);
        const uint32_t scale = depth == 8 ? 1 : 255 / ((1u << depth) - 1);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) (*px)[(size_t)y * w + x] = (uint8_t)(sample(y, x) * scale);
        *fmt_out = TR_TEX_R8_UNORM;
    } else {
        px->resize(n * 4);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                uint8_t* o = &(*px)[((size_t)y * w + x) * 4];
                const uint8_t* p = &img[y * rowbytes + (size_t)x * bpp];
                if (ctype == 0) {  /* grey + tRNS: grey-alpha */
                    const uint32_t g = sample(y, x), scale = depth == 8 ? 1 : 255 / ((1u << depth) - 1);
                    const uint32_t key = trns.size() >= 2 ? (((uint32_t)trns[0] << 8) | trns[1]) : 0xFFFFFFFFu;
                    o[0] = o[1] = o[2] = (uint8_t)(g * scale);
                    o[3] = g == key ? 0 : 255;
                } else if (ctype == 2) {
                    o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
                    o[3] = (trns.size() >= 6 && p[0] == trns[1] && p[1] == trns[3] && p[2] == trns[5]) ? 0 : 255;
                } else if (ctype == 3) {
                    const uint32_t idx = sample(y, x);
                    if ((size_t)idx * 3 + 2 >= plte.size()) throw ImgError{"PNG palette index out of range"};
                    o[0] = plte[idx * 3]; o[1] = plte[idx * 3 + 1]; o[2] = plte[idx * 3 + 2];
                    o[3] = idx < trns.size() ? trns[idx] : 255;
                } else if (ctype == 4) {
                    o[0] = o[1] = o[2] = p[0];
                    o[3] = p[1];
                } else {
                    memcpy(o, p, 4);
                }
            }
        *fmt_out = TR_TEX_RGBA8_UNORM;
    }
    *w_out = w;
    *h_out = h;
}

}  // namespace

extern "C" {

tr_status tr_image_load(const char* path, tr_image* out)
{
    if (!path || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    const std::string ext = lower_extension(path);
    if (ext != "tga" && ext != "png") return fail(TR_ERROR_UNSUPPORTED, "unsupported image file format: " + ext);
    std::vector<uint8_t> file, px;
    if (!read_all(path, &file)) return fail(TR_ERROR_IO, "cannot read image file: " + std::string(path));
    uint32_t w = 0, h = 0, fmt = TR_TEX_RGBA8_UNORM;
    try {
        if (ext == "tga") load_tga(file, &w, &h, &px);
        else load_png(file, &w, &h, &fmt, &px);
    } catch (const ImgError& e) {
        return fail(TR_ERROR_UNSUPPORTED, e.msg + ": " + std::string(path));
    } catch (const std::bad_alloc&) {
        return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory decoding " + std::string(path));
    }
    out->data = (uint8_t*)malloc(px.size());
    if (!out->data) return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory for an image");
    memcpy(out->data, px.data(), px.size());
    out->width = w;
    out->height = h;
    out->format = fmt;
    return TR_OK;
}

void tr_image_free(tr_image* image)
{
    if (!image) return;
    free(image->data);
    memset(image, 0, sizeof(*image));
}


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
