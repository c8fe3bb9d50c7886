/*
 * tr_scene.h - the scene as it lives in HBM, and the 8-wide BVH format.
 *
 * Layout (all arrays 256-byte aligned, read-only during a frame):
 *   Node8[]     128 B per node: quantised child boxes of an 8-wide BVH.
 *   PrimRec[]   48 B per primitive, in BVH leaf order: the three world-space
 *               vertices of a triangle (or centre and radius of a sphere),
 *               its primitive id, geometry slot and flags.  Traversal reads
 *               nothing else.
 *   GeomRec[]   one per geometry slot (submesh, sphere, sphere light, volume).
 *   VolumeRec[] and the packed binary16 samples of grid volumes.
 *   tr_material[], LightRec[], vertex bytes and int32 index triplets for
 *               shading normals / tangents and mesh-light sampling.
 */
#ifndef TR_SCENE_H
#define TR_SCENE_H

#include <stdint.h>
#include <string.h>
#include "../../include/tempest_rt.h"

namespace trk {

/* 8-wide BVH node, one 128-byte cache line.
 * Child slot s (0..7) has the box
 *   lo = origin + h_lo[s] * 2^exp,  hi = origin + h_hi[s] * 2^exp  (per axis)
 * where h_lo and h_hi are binary16 values (signed offsets from the node's
 * centre, so the coarsest code spacing is 2^-12 of the node's extent) and
 * the sum is computed in FP32 exactly as written; the builder rounds
 * outward and checks that the decoded box contains the child.  The
 * traversal converts and scales a code in one mixed-precision fused
 * multiply-add.  An empty slot has an inverted box (lo = +65504,
 * hi = -65504), which no ray enters.  imask bit s marks an internal child:
 * its node index is child_base + popcount(imask & ((1<<s)-1)).  A leaf
 * child has meta[s] = count << 5 | offset (count 1..7): primitives
 * prim_base + offset .. + count - 1.  Slots are assigned so that slot s
 * lies towards the (+x if bit0, +y if bit1, +z if bit2) side of the node;
 * a ray visits slots in the order i ^ octant(ray direction), i = 0..7,
 * i.e. roughly front to back. */
struct alignas(128) Node8 {
    float origin[3];
    int8_t exp[3];
    uint8_t imask;
    uint32_t child_base;
    uint32_t prim_base;
    uint8_t meta[8];
    uint16_t qlo[3][8];   /* binary16 bit patterns */
    uint16_t qhi[3][8];
};
static_assert(sizeof(Node8) == 128, "Node8 must be one 128-byte line");

/* Empty-slot codes: +65504 and -65504. */
constexpr uint16_t kHalfMax = 0x7BFFu, kHalfMinusMax = 0xFBFFu;

/* The value of a finite binary16 bit pattern (exact in FP32). */
inline float half_value(uint16_t h)
{
    const uint32_t e = (h >> 10) & 31u, m = h & 1023u;
    float v;
    if (e == 0) v = (float)m * 5.9604644775390625e-08f; /* m * 2^-24 */
    else {
        uint32_t bits = ((e + 112u) << 23) | (m << 13);
        memcpy(&v, &bits, 4);
    }
    return (h & 0x8000u) ? -v : v;
}

/* Primitive kinds.  A grid volume is one PRIM_VOLUME primitive: its box
 * lo = v0, hi = v1.  A user shape (PRIM_SHAPE, the shape kind in bits 8-15
 * of flags) keeps its ShapeRec index in the bits of v0[0] and its box in
 * v1 (lo) and v2 (hi). */
enum { PRIM_SPHERE = 1u, PRIM_TWO_SIDED = 2u, PRIM_VOLUME = 4u, PRIM_SHAPE = 8u };

/* User shapes of the scene API (ray-tracing.cc:1315-1446).  Rects and
 * blockers share the rect test. */
enum { SHAPE_RECT = 1u, SHAPE_DISK = 2u, SHAPE_CYLINDER = 3u, SHAPE_OBLIQUE_CYLINDER = 4u, SHAPE_ELLIPSOID = 5u };

/* A user shape's parameters, precomputed once on the host in FP32:
 *   rect:      f[0-2] centre, f[3-11] columns T, B, N of ToMatrix3 of its
 *              quaternion, f[12-13] half sizes, f[14-15] texture-coordinate
 *              start, f[16-17] multiplier
 *   disk:      f[0-2] centre, f[3-5] normal, f[6] inner, f[7] outer radius
 *   cylinder:  f[0-2] centre, f[3] radius, f[4] half height
 *   oblique:   as cylinder, f[5-6] bending = Tilt * TiltDirection / HalfHeight
 *   ellipsoid: f[0-2] centre, f[3-5] scale, f[6-9] orientation (x, y, z, w),
 *              f[10-13] its inverse (the conjugate), f[14-16] ToNormal of it
 * tangent_map: the rect's tangent-space texture handle (0 = none). */
struct alignas(16) ShapeRec {
    uint32_t kind;
    uint32_t geom;
    uint32_t tangent_map;
    uint32_t pad;
    float f[20];
};
static_assert(sizeof(ShapeRec) == 96, "ShapeRec layout");

struct alignas(16) PrimRec {
    float v0[3];
    uint32_t prim;
    float v1[3];
    uint32_t geom;
    float v2[3];
    uint32_t flags;
};
static_assert(sizeof(PrimRec) == 48, "PrimRec must be 48 bytes");

/* A primitive's bounding box (host side: BVH build, scene box, checks). */
inline void prim_box(const PrimRec& p, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a) {
        if (p.flags & PRIM_SPHERE) { lo[a] = p.v0[a] - p.v1[0]; hi[a] = p.v0[a] + p.v1[0]; }
        else if (p.flags & PRIM_VOLUME) { lo[a] = p.v0[a]; hi[a] = p.v1[a]; }
        else if (p.flags & PRIM_SHAPE) { lo[a] = p.v1[a]; hi[a] = p.v2[a]; }
        else {
            lo[a] = p.v0[a] < p.v1[a] ? (p.v0[a] < p.v2[a] ? p.v0[a] : p.v2[a]) : (p.v1[a] < p.v2[a] ? p.v1[a] : p.v2[a]);
            hi[a] = p.v0[a] > p.v1[a] ? (p.v0[a] > p.v2[a] ? p.v0[a] : p.v2[a]) : (p.v1[a] > p.v2[a] ? p.v1[a] : p.v2[a]);
        }
    }
}

/* GEOM_SHAPE: a user shape, its ShapeRec at index_base; a blocker
 * (addBlocker) has material kNoMaterial and ends every path that hits it. */
enum { GEOM_MESH = 0, GEOM_SPHERE = 1, GEOM_SPHERE_LIGHT = 2, GEOM_VOLUME = 3, GEOM_SHAPE = 4 };
constexpr uint32_t kNoMaterial = 0xffffffffu;

struct GeomRec {
    uint32_t type;
    uint32_t material;
    uint32_t two_sided;
    uint32_t stride;
    uint64_t vbase;       /* byte offset of the submesh's vertex 0 in the vertex pool */
    uint32_t index_base;  /* offset of the submesh's first index in the index pool */
    uint32_t tri_count;   /* volumes: index into SceneView::volumes */
    float sphere[4];
    uint32_t shade_class; /* shading class of its hits (shade_class_of), set at commit */
    uint32_t pad;
};

/* Shading classes: the order in which a shading block runs its paths, so
 * that the lanes of a wave run the same BSDF code.  0 is a miss; an entry
 * past the end of the queue takes kShadeClasses - 1. */
constexpr uint32_t kShadeClasses = 9;
enum { SC_MISS = 0, SC_NO_BSDF = 1, SC_PHONG = 2, SC_GGX = 3, SC_BECKMANN = 4, SC_ANISO = 5, SC_SGGX = 6,
       SC_WALK = 7, SC_NONE = 8 };

/* The class of a hit on geometry g with material m (nullptr: none). */
inline uint32_t shade_class_of(const GeomRec& g, const tr_material* m)
{
    if (!m) return SC_NO_BSDF;
    if (g.type == GEOM_VOLUME) return SC_WALK;
    switch (m->model) {
    case TR_MODEL_EMISSIVE: return SC_NO_BSDF;
    case TR_MODEL_BLINN_PHONG: return SC_PHONG;
    case TR_MODEL_GGX_MICROFACET:
    case TR_MODEL_GGX_MICROFACET_CONDUCTOR:
    case TR_MODEL_GGX_MICROFACET_DIELECTRIC: return SC_GGX;
    case TR_MODEL_BECKMANN_MICROFACET:
    case TR_MODEL_BECKMANN_MICROFACET_CONDUCTOR:
    case TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC: return SC_BECKMANN;
    case TR_MODEL_SGGX_SURFACE: return SC_SGGX;
    case TR_MODEL_SGGX_PSEUDO_VOLUME: return SC_WALK;
    default: return SC_ANISO;
    }
}

/* A grid volume as the tracking code reads it.  Samples are packed per
 * grid node as four binary16 values (direction x, y, z, density): one
 * 8-byte load per trilinear corner.  */
struct VolumeRec {
    float lo[3];
    float rcp_rho_max;    /* 1 / max density */
    float hi[3];
    float rcp_majorant;   /* 1 / (max density * max(stddev)) */
    float scale[3];       /* (dims - 1) / (hi - lo) */
    float rcp_smax;       /* 1 / max(stddev) */
    float stddev[3];      /* SGGX standard deviation of the fibre frame (tangent, binormal, normal) */
    uint32_t format;      /* kVoxelHalf or kVoxelQuantized */
    uint32_t dims[3];
    uint32_t pad1;
    uint64_t voxel_base;  /* first sample in SceneView::voxels */
    uint64_t table_base;  /* kVoxelQuantized: the direction tables in SceneView::voxels */
};
/* Voxel words.  kVoxelHalf: fibre x, y, z and density as binary16 (bits 0,
 * 16, 32, 48).  kVoxelQuantized (the reference's brick volumes): the
 * quantized theta and phi bytes at bits 0 and 8 and the u8 density as an
 * exact binary16 at bit 48; the direction tables cos(phi), sin(phi),
 * cos(theta), sin(theta) (256 floats each) sit at table_base. */
enum { kVoxelHalf = 0, kVoxelQuantized = 1 };
/* DensityScale (ray-tracing.cc:60): the brick volumes' majorant. */
constexpr float kDensityScale = 100.0f;
static_assert(sizeof(VolumeRec) == 96, "VolumeRec layout");

struct LightRec {
    tr_light src;
    uint32_t geom;        /* mesh and sphere lights: geometry slot */
    uint32_t area;        /* IsAreaLight: sphere lights only (ray-tracing-common.hh:546-551) */
};

/* A bound texture: TR_TEX_* format and filter, texels at `offset` bytes
 * into the texel pool (16-byte aligned). */
struct TexRec {
    uint32_t width, height, format, sampling;
    uint64_t offset;
    uint64_t pad;
};

/* The scene's textures, the sRGB decode table of 8-bit values and the
 * global cube map (six RGBA8 faces of cube_side^2 texels at cube_offset,
 * cube_side 0 = none). */
struct TexView {
    const TexRec* recs;
    const uint8_t* texels;
    const float* srgb;
    uint32_t count;
    uint32_t cube_side;
    uint64_t cube_offset;
    uint32_t pseudo_volume;  /* a material of the scene is the SGGX pseudo-volume */
};

/* Device pointers of a committed scene. */
struct SceneView {
    const Node8* nodes;
    const PrimRec* prims;
    const GeomRec* geoms;
    const tr_material* mats;
    const LightRec* lights;
    const uint8_t* vpool;
    const int32_t* ipool;
    const VolumeRec* volumes;
    const uint64_t* voxels; /* packed binary16 x4 samples of every grid volume */
    const ShapeRec* shapes;
    uint32_t n_lights;
    uint32_t n_prims;
    uint32_t n_volumes;
    TexView tex;
};

}  // namespace trk

#endif
