/*
 * tr_volfile.cpp - the reference's "VOL\x3" volume grid files.
 *
 * ParseVolumeElement, ParseVolumeHierarchy and ParseVolume
 * (src/volume/volume.cc:93-453).  A grid file is: the FourCC 'V' 'O' 'L'
 * 3 (little-endian u32), GridDataType (u32: 1 Float32, 2 Float16, 3 UInt8,
 * 4 QuantizedDirections), dimensions (3 x i32), channel count (i32), the
 * bounding box (2 x 3 floats), then the samples, x fastest.  A brick holds
 * its u8 densities followed by (theta, phi) byte pairs; Float32 densities
 * become (uint8)(255 v), Float32 directions (x, y[, z]) are quantized to
 * theta * 255 / pi + 0.5 and phi * 255 / (2 pi) + 0.5, and UInt8 direction
 * grids leave the pairs zero, as in the reference.
 *
 * Where the reference goes on after a failure with undefined results (a
 * missing brick file, a short read, a density outside [0, 1], a channel
 * count or brick size that does not match, the description file's final
 * partial read) this reader fails with an error instead; non-finite
 * directions quantize to 0.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tr_host.h"
#include "tr_math.h"

using namespace trk;

namespace {

enum { GRID_FLOAT32 = 1, GRID_FLOAT16 = 2, GRID_UINT8 = 3, GRID_QUANTIZED = 4 };
constexpr uint32_t kVolMagic = 'V' | ('O' << 8) | ('L' << 16) | (3u << 24);

struct VolError {
    std::string msg;
};

size_t grid_type_size(uint32_t t)
{
    switch (t) {
    case GRID_FLOAT32: return 4;
    case GRID_FLOAT16: return 2;
    case GRID_UINT8: return 1;
    case GRID_QUANTIZED: return 2;
    default: return 0;
    }
}

struct Brick {
    int32_t dims[3] = {0, 0, 0};
    std::vector<uint8_t> data;  /* n densities, then n (theta, phi) pairs */
};

inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ParseVolumeElement (volume.cc:93-359) for one of the two grids of a
 * brick: param_type GRID_UINT8 (density, at offset 0) or GRID_QUANTIZED
 * (directions, after the densities). */
void parse_element(const std::string& file, uint32_t param_type, Brick* brick, float lo[3], float hi[3])
{
    FILE* f = fopen(file.c_str(), "rb");
    if (!f) throw VolError{"cannot read volume grid file: " + file};
    struct Closer {
        FILE* f;
        ~Closer() { fclose(f); }
    } closer{f};
    uint32_t magic = 0, fmt = 0;
    int32_t dims[3], channels = 0;
    if (fread(&magic, 4, 1, f) != 1 || magic != kVolMagic) throw VolError{"invalid volume grid file format: " + file};
    if (fread(&fmt, 4, 1, f) != 1 || fread(dims, 4, 3, f) != 3 || fread(&channels, 4, 1, f) != 1 ||
        fread(lo, 4, 3, f) != 3 || fread(hi, 4, 3, f) != 3)
        throw VolError{"invalid volume bounding box: " + file};
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) throw VolError{"invalid volume dimensions: " + file};
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (n > (1ull << 31)) throw VolError{"volume grid larger than 2^31 samples: " + file};
    if (brick->data.empty()) {
        memcpy(brick->dims, dims, sizeof(dims));
        brick->data.assign(3 * n, 0);
    } else if (memcmp(brick->dims, dims, sizeof(dims)) != 0) {
        throw VolError{"density and orientation grids differ in size: " + file};
    }
    if (channels < 1) throw VolError{"invalid channel count: " + file};
    const bool dirs = param_type == GRID_QUANTIZED;
    uint8_t* out = brick->data.data() + (dirs ? n : 0);
    if (fmt == param_type) {
        if (fread(out, 1, (dirs ? 2 : 1) * n, f) != (dirs ? 2 : 1) * n) throw VolError{"incomplete volume data: " + file};
        return;
    }
    const size_t es = grid_type_size(fmt);
    if (es == 0) throw VolError{"unsupported grid data type: " + file};
    const size_t actual = (fmt != GRID_QUANTIZED ? (size_t)channels : 1) * es * n;
    std::vector<uint8_t> in(actual);
    if (fread(in.data(), 1, actual, f) != actual) throw VolError{"incomplete volume data: " + file};
    if (!dirs) {
        if (fmt != GRID_FLOAT32) throw VolError{"unsupported density grid data type: " + file};
        if (channels != 1) throw VolError{"density grids have one channel: " + file};
        const float* v = reinterpret_cast<const float*>(in.data());
        for (size_t i = 0; i < n; ++i) {
            if (!(v[i] >= 0.0f && v[i] <= 1.0f)) throw VolError{"density outside [0, 1]: " + file};
            out[i] = (uint8_t)(255.0f * v[i]);
        }
        return;
    }
    if (fmt == GRID_UINT8) return;  /* the reference's empty conversion case */
    if (fmt != GRID_FLOAT32 || channels < 2) throw VolError{"unsupported direction grid data type: " + file};
    const float* v = reinterpret_cast<const float*>(in.data());
    for (size_t i = 0; i < n; ++i) {
        float x = v[i * channels + 0], y = v[i * channels + 1];
        float sin_theta2 = x * x + y * y;
        float cos_theta, sin_theta;
        if (channels < 3) {
            sin_theta = sqrtf(sin_theta2);
            cos_theta = sqrtf(1.0f - sin_theta2);
        } else {
            cos_theta = v[i * channels + 2];
            sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        }
        float cos_phi, sin_phi;
        if (sin_theta < 1e-3f) {
            cos_phi = 1.0f;
            sin_phi = 0.0f;
        } else {
            cos_phi = clampf(x / sin_theta, -1.0f, 1.0f);
            sin_phi = copysignf(sqrtf(1.0f - cos_phi * cos_phi), y);
        }
        float theta = atan2f(sin_theta, cos_theta);
        float phi = atan2f(sin_phi, cos_phi);
        if (phi < 0.0f) phi += 2.0f * kPi;
        float tq = theta * 255.0f / kPi + 0.5f, pq = phi * 255.0f / (2.0f * kPi) + 0.5f;
        out[2 * i + 0] = std::isfinite(tq) ? (uint8_t)clampf(tq, 0.0f, 255.0f) : 0;
        out[2 * i + 1] = std::isfinite(pq) ? (uint8_t)clampf(pq, 0.0f, 255.0f) : 0;
    }
}

std::string brick_name(const std::string& name, int32_t x, int32_t y, int32_t z)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "_%03d_%03d_%03d", x, y, z);
    return name + buf;
}

tr_status fill_root(const float lo[3], const float hi[3], const int32_t dims[3], std::vector<Brick>& bricks,
                    tr_volume_root* out)
{
    const size_t nb = bricks.size();
    out->bricks = (tr_volume_brick*)calloc(nb ? nb : 1, sizeof(tr_volume_brick));
    if (!out->bricks) return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory for the volume root");
    memcpy(out->min_corner, lo, 12);
    memcpy(out->max_corner, hi, 12);
    memcpy(out->dims, dims, 12);
    for (size_t b = 0; b < nb; ++b) {
        if (bricks[b].data.empty()) continue;
        uint8_t* p = (uint8_t*)malloc(bricks[b].data.size());
        if (!p) {
            tr_volume_root_free(out);
            return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory for a volume brick");
        }
        memcpy(p, bricks[b].data.data(), bricks[b].data.size());
        memcpy(out->bricks[b].dims, bricks[b].dims, 12);
        out->bricks[b].data = p;
    }
    return TR_OK;
}

}  // namespace

extern "C" {

tr_status tr_parse_volume_hierarchy(const char* name, tr_volume_root* out)
{
    if (!name || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    const std::string base(name);
    const std::string descr = base + "_description.vol";
    FILE* f = fopen(descr.c_str(), "rb");
    if (!f) return fail(TR_ERROR_INVALID_ARG, "cannot read volume description file: " + descr);
    float lo[3], hi[3];
    int32_t dims[3];
    std::vector<int32_t> blocks;
    bool ok = fread(lo, 4, 3, f) == 3 && fread(hi, 4, 3, f) == 3 && fread(dims, 4, 3, f) == 3;
    if (ok) {
        int32_t b[3];
        size_t r;
        while ((r = fread(b, 4, 3, f)) == 3) blocks.insert(blocks.end(), b, b + 3);
        ok = r == 0 && feof(f);
    }
    fclose(f);
    if (!ok) return fail(TR_ERROR_INVALID_ARG, "truncated volume description file: " + descr);
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || (uint64_t)dims[0] * dims[1] * dims[2] > (1u << 20))
        return fail(TR_ERROR_INVALID_ARG, "invalid brick grid dimensions: " + descr);
    std::vector<Brick> bricks((size_t)dims[0] * dims[1] * dims[2]);
    try {
        for (size_t k = 0; k < blocks.size(); k += 3) {
            const int32_t x = blocks[k], y = blocks[k + 1], z = blocks[k + 2];
            if (x < 0 || y < 0 || z < 0 || x >= dims[0] || y >= dims[1] || z >= dims[2])
                throw VolError{"brick outside the grid in " + descr};
            Brick& br = bricks[((size_t)z * dims[1] + y) * dims[0] + x];
            const std::string stem = brick_name(base, x, y, z);
            float a[3], c[3];
            parse_element(stem + "-density.vol", GRID_UINT8, &br, a, c);
            parse_element(stem + "-orientation.vol", GRID_QUANTIZED, &br, a, c);
        }
    } catch (const VolError& e) {
        return fail(TR_ERROR_INVALID_ARG, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory reading a volume");
    }
    return fill_root(lo, hi, dims, bricks, out);
}

tr_status tr_parse_volume(const char* density_file, const char* orientation_file, tr_volume_root* out)
{
    if (!density_file || !orientation_file || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    std::vector<Brick> bricks(1);
    float lo[3], hi[3], a[3], c[3];
    const int32_t dims[3] = {1, 1, 1};
    try {
        parse_element(density_file, GRID_UINT8, &bricks[0], lo, hi);
        parse_element(orientation_file, GRID_QUANTIZED, &bricks[0], a, c);
    } catch (const VolError& e) {
        return fail(TR_ERROR_INVALID_ARG, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory reading a volume");
    }
    return fill_root(lo, hi, dims, bricks, out);
}

void tr_volume_root_free(tr_volume_root* root)
{
    if (!root) return;
    if (root->bricks) {
        const uint64_t nb = (uint64_t)std::max(root->dims[0], 0) * std::max(root->dims[1], 0) *
                            std::max(root->dims[2], 0);
        for (uint64_t b = 0; b < nb; ++b) free(root->bricks[b].data);
        free(root->bricks);
    }
    memset(root, 0, sizeof(*root));
}

}  // extern "C"
