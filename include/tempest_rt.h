/*
 * tempest_rt.h - C ABI of the MI355X path tracer.
 *
 * The entry points mirror the ray-trace scene API of the reference renderer,
 * the duck-typed twins RayTracerScene
 * (include/tempest/graphics/ray-tracing/ray-tracing.hh:241-395) and
 * RayTracerCuda (include/tempest/compute/ray-tracing-cuda.hh:80-172).  Each
 * declaration below names the member it replaces.  Plain C types only: no
 * HIP, no torch.  Every call returns a tr_status and never aborts;
 * tr_last_error_message() describes the last failure of the calling thread.
 *
 * Ownership differs from the CPU reference on purpose: every add* call
 * deep-copies vertices, indices, materials and lights, so the caller keeps
 * its buffers (the reference retains index and material pointers,
 * ray-tracing.cc:1532-1535, and takes ownership of LightSource*).
 */
#ifndef TEMPEST_RT_H
#define TEMPEST_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tr_status {
    TR_OK = 0,
    TR_ERROR_INVALID_ARG = 1,
    TR_ERROR_OUT_OF_MEMORY = 2,
    TR_ERROR_DEVICE = 3,
    TR_ERROR_NOT_COMMITTED = 4,
    TR_ERROR_UNSUPPORTED = 5,
    TR_ERROR_IO = 6
} tr_status;

/* IlluminationModel, numeric values kept (ray-tracing-common.hh:68-105). */
enum {
    TR_MODEL_EMISSIVE = 0,
    TR_MODEL_BLINN_PHONG = 1,
    TR_MODEL_KAJIYA_KAY = 2,
    TR_MODEL_ASHIKHMIN_SHIRLEY = 3,
    TR_MODEL_MICRO_FLAKE = 4,
    TR_MODEL_PATCHWORK = 8,
    TR_MODEL_MIRROR = 9,
    TR_MODEL_GGX_MICROFACET = 11,
    TR_MODEL_GGX_MICROFACET_CONDUCTOR = 12,
    TR_MODEL_GGX_MICROFACET_DIELECTRIC = 13,
    TR_MODEL_MIX = 14,
    TR_MODEL_GGX_MICROFACET_ANISOTROPIC = 15,
    TR_MODEL_GGX_MICROFACET_CONDUCTOR_ANISOTROPIC = 16,
    TR_MODEL_GGX_MICROFACET_DIELECTRIC_ANISOTROPIC = 17,
    TR_MODEL_STOCHASTIC_ROTATOR = 20,
    TR_MODEL_SGGX_MICRO_FLAKE = 22,
    TR_MODEL_SPATIALLY_VARYING_EMISSIVE = 23,
    TR_MODEL_SGGX_SURFACE = 24,
    TR_MODEL_SGGX_PSEUDO_VOLUME = 25,
    TR_MODEL_BTF = 27,
    TR_MODEL_BECKMANN_MICROFACET = 28,
    TR_MODEL_BECKMANN_MICROFACET_CONDUCTOR = 29,
    TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC = 30,
    TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC = 31,
    TR_MODEL_BECKMANN_MICROFACET_CONDUCTOR_ANISOTROPIC = 32,
    TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC_ANISOTROPIC = 33,
    TR_MODEL_COUNT = 34
};

/* LightSourceType (ray-tracing-common.hh:478-485). */
enum {
    TR_LIGHT_DIRECTIONAL = 0,
    TR_LIGHT_SPHERE_AREA = 1,
    TR_LIGHT_MESH_AREA = 2,
    TR_LIGHT_POINT = 3
};

/* PicturePostProcess (ray-tracing-common.hh:59-66). */
enum {
    TR_POST_LINEAR = 0,
    TR_POST_SRGB = 1,
    TR_POST_AUTO_EXPOSURE_SRGB = 2,
    TR_POST_AUTO_EXPOSURE_HDR = 3,
    TR_POST_ACES = 4
};

/* Estimator of the path loop.  REFERENCE reproduces the CPU reference's
 * balance-heuristic weighting (ray-tracing.cc:1791-1930, 2182-2332),
 * including its quirks; it is the mode of the per-path parity contract.
 * UNBIASED is multi-sample MIS with the balance heuristic between each
 * area light (sphere and mesh lights, SamplesLocalAreaLight samples each)
 * and the BSDF (SamplesGlobalIllumination at the camera vertex, 1 later):
 * the weight applies to the NEE term and to the emitter hit only, not to
 * the continuing throughput, the denominator holds only the strategies
 * that can produce the same path, and a mesh light's density is the true
 * solid-angle density dist^2 / (cos * area * triangles).  Scenes with
 * volumes take the REFERENCE estimator only. */
enum {
    TR_ESTIMATOR_REFERENCE = 0,
    TR_ESTIMATOR_UNBIASED = 1
};

/* One flat record for the material structs of ray-tracing-common.hh.
 * RTMicrofacetMaterial (:159-174): specular, specular_power (the union
 * member StandardDeviation for SGGX surfaces), diffuse (radiance for
 * Emissive), fresnel {eta, k}, normalization.  RTSGGXSurface (:397-406):
 * sggx_basis quaternion (x, y, z, w), sggx_depth, sggx_sample_count.
 * RTVolumeMaterial / RTSGGXMicroFlakeMaterial (:379-395): albedo,
 * sggx_stddev.  The texture maps DiffuseMap, SpecularMap (:171-172),
 * StandardDeviationMap and BasisMap (:404-405) are handles from
 * tr_scene_bind_texture of the scene the material is added to (0 = none);
 * Blinn-Phong reads DiffuseMap (illumination-models.hh:618-698), the SGGX
 * surface all four (ggx-models-impl.hh:476-619).  tr_material_init fills
 * the reference's defaults. */
typedef struct tr_material {
    uint32_t model;
    float specular[3];
    float specular_power[2];
    float diffuse[3];
    float fresnel[2];
    float normalization;
    float sggx_basis[4];
    uint32_t sggx_depth;
    uint32_t sggx_sample_count;
    float albedo[3];
    float sggx_stddev[3];
    uint32_t diffuse_map;
    uint32_t specular_map;
    uint32_t stddev_map;
    uint32_t basis_map;
} tr_material;

/* LightSource and its subclasses (ray-tracing-common.hh:486-544). */
typedef struct tr_light {
    uint32_t type;
    float direction[3];      /* DirectionalLight::Direction (towards the light) */
    float position[3];       /* PointLight::Position */
    float radiance[3];       /* Radiance of every type */
    float sphere_center[3];  /* SphereAreaLight::SphereShape */
    float sphere_radius;
    uint32_t geometry_id;    /* set by tr_scene_add_sphere_light_source */
} tr_light;

/* RTSubmesh (ray-tracing-common.hh:240-247): vertex_count is the number of
 * indices of the submesh, vertex_offset a byte offset into the vertex
 * buffer, base_index an offset into the index array, stride the vertex
 * size in bytes.  Positions sit at byte 0 of a vertex, the normal at
 * stride-12 (ray-tracing.cc:377); strides 44 (PTTNFormat) and 48
 * (PTTcNFormat) carry a tangent at byte 20. */
typedef struct tr_submesh {
    const tr_material* material;
    uint32_t vertex_count;
    uint32_t vertex_offset;
    uint32_t base_index;
    uint32_t stride;
} tr_submesh;

/* MeshOptions (ray-tracing-common.hh:408-415) without the host callbacks,
 * which a GPU backend cannot call. */
typedef struct tr_mesh_options {
    int32_t two_sided;
} tr_mesh_options;

/* RTSettings (ray-tracing-common.hh:141-144) plus the device path pool. */
typedef struct tr_settings {
    uint32_t data_pool_size;   /* accepted for compatibility, unused */
    uint32_t path_pool_size;   /* paths resident per batch; 0 = automatic: the frame's paths, at most
                                  2^27 and a quarter of the device memory free when the pool is made */
    uint32_t trace_waves_per_cu; /* persistent traversal waves per CU, 1..32; 0 = automatic (24 for
                                    surface scenes, 16 with volumes) */
    uint32_t shade_order;      /* 0: shading runs paths in queue order (slot order, coalesced);
                                  1: each shading block first sorts its paths by material class.
                                  Results are the same. */
} tr_settings;

/* A ray of a fixed ray set and its closest hit. */
typedef struct tr_ray {
    float org[3];
    float tnear;
    float dir[3];
    float tfar;
} tr_ray;

typedef struct tr_hit {
    float t;
    float u;        /* barycentric weight of vertex 1 */
    float v;        /* barycentric weight of vertex 2 */
    uint32_t geom;  /* scene geometry slot, TR_INVALID_GEOMETRY on a miss */
    uint32_t prim;  /* triangle index within the submesh, 0 for spheres */
} tr_hit;

#define TR_INVALID_GEOMETRY 0xffffffffu

typedef struct tr_scene tr_scene;

/* Counters and timings of the last drawOnce / render call. */
typedef struct tr_stats {
    uint64_t paths;              /* camera samples traced */
    uint64_t segments;           /* extension rays traced */
    uint64_t shadow_rays;        /* next-event shadow rays traced */
    double frame_ms;             /* host wall time of the call */
    double trace_ms;             /* timing mode: summed HIP-event time of closest-hit kernels */
    double shade_ms;             /* timing mode: summed HIP-event time of shading kernels */
    double shadow_ms;            /* timing mode: summed HIP-event time of shadow kernels */
    double other_ms;             /* timing mode: ray generation and accumulation */
    uint64_t trace_launches;     /* closest-hit launches behind trace_ms */
    uint64_t node_visits;        /* BVH8 nodes fetched (only if counting on) */
    uint64_t prim_tests;         /* primitives tested (only if counting on) */
    uint64_t counted_rays;       /* rays behind node_visits / prim_tests */
    uint64_t volume_steps;       /* tentative collisions of delta / ratio tracking (only if counting on) */
    double cache_ms;             /* timing mode: radiance-cache inference, targets and training */
    double cache_loss;           /* loss of the last training step of the call (cache mode) */
    double exposure;             /* auto-exposure modes: the factor applied to the frame's mean image */
} tr_stats;

/* Scene-construction information. */
typedef struct tr_scene_info {
    uint64_t primitives;
    uint64_t bvh_nodes;
    uint64_t bvh_bytes;
    uint64_t prim_bytes;
    uint32_t bvh_depth;
    uint32_t geometries;
    uint32_t lights;
    uint32_t materials;
    double build_ms;
} tr_scene_info;

const char* tr_last_error_message(void);

/* The reference's pseudo-volume benchmark kernel ComputeImage
 * (tests/benchmark/pseudo-volume-performance-test.cu:16-41): every pixel of
 * a width x height image evaluates SGGXMicroFlakePseudoVolumeBRDF of `m`
 * (model 25, constant StandardDeviation and basis) at wi = (0, sin 45,
 * cos 45), wo = (0, -sin 45, cos 45) after SGGXSurfaceCache and stores
 * ToColor of it.  One warm-up and `repeats` timed launches on the current
 * device; rgba receives the image (host memory), probe pixel (0, 0)'s BRDF
 * and directional density, ms_per_image the mean launch time. */
tr_status tr_bench_pseudo_volume_image(const tr_material* m, uint32_t width, uint32_t height, uint32_t repeats,
                                       uint32_t* rgba, float probe[4], float* ms_per_image);

/* RTMaterial::setup (ray-tracing.cc:80-87): applies MaterialSetupLookup,
 * i.e. DefaultMicrofacetMaterialSetup (illumination-models.hh:599-614). */
tr_status tr_material_init(tr_material* material, uint32_t model);
tr_status tr_material_setup(tr_material* material);

/* RayTracerScene(width, height, view_proj_inv, settings) (ray-tracing.hh:302). */
tr_status tr_scene_create(uint32_t width, uint32_t height, const float view_proj_inv[16],
                          const tr_settings* settings, tr_scene** out_scene);
/* ~RayTracerScene. */
void tr_scene_destroy(tr_scene* scene);

/* Setters (ray-tracing.hh:305-321). */
tr_status tr_scene_set_chunk_size(tr_scene* scene, uint32_t chunk_size);
tr_status tr_scene_set_samples_local_area_light(tr_scene* scene, uint32_t rays);
tr_status tr_scene_set_samples_global_illumination(tr_scene* scene, uint32_t rays);
tr_status tr_scene_set_samples_camera(tr_scene* scene, uint32_t rays);
tr_status tr_scene_set_max_ray_depth(tr_scene* scene, uint32_t depth);
tr_status tr_scene_set_russian_roulette(tr_scene* scene, float rval);
tr_status tr_scene_set_background_spectrum(tr_scene* scene, const float rgb[3]);

/* Textures (graphics/texture.hh:373-600): texels row by row, x fastest;
 * Texture::sample's bilinear (wrapping, texel centres at +0.5) or nearest
 * filter.  Extractors: spectrum maps (DiffuseMap, SpecularMap) take RGBA8
 * as sRGB (SRGBToSpectrum) or RGB32F / RGBA32F as linear; the
 * standard-deviation map takes RG8 / RG32F / RGB32F / RGBA32F; the basis map
 * RGBA8 / RGBA32F. */
enum {
    TR_TEX_R8_UNORM = 0,
    TR_TEX_RG8_UNORM = 1,
    TR_TEX_RGBA8_UNORM = 2,
    TR_TEX_R32F = 3,
    TR_TEX_RG32F = 4,
    TR_TEX_RGB32F = 5,
    TR_TEX_RGBA32F = 6
};
enum { TR_TEX_BILINEAR = 0, TR_TEX_NEAREST = 1 };
typedef struct tr_texture_desc {
    uint32_t width, height;   /* 1..65535, TextureDescription's uint16 */
    uint32_t format;          /* TR_TEX_* */
    uint32_t sampling;        /* TR_TEX_BILINEAR or TR_TEX_NEAREST */
} tr_texture_desc;
/* bindTexture (ray-tracing.hh:340-354): copies the texels; *out_handle
 * (1, 2, ...) goes into tr_material map fields. */
tr_status tr_scene_bind_texture(tr_scene* scene, const tr_texture_desc* desc, const void* texels,
                                uint32_t* out_handle);
/* setGlobalCubeMap (ray-tracing.hh:316; CubeMap, src/graphics/cube-map.cc):
 * six square RGBA8 faces (+x, -x, +y, -y, +z, -z) of side > 16, each as
 * the reference's Texture holds it (the faces are flipped vertically on
 * copy, cube-map.cc:58-66).  Rays that miss look up the map (clamped
 * bilinear, sRGB) instead of the background spectrum.  faces = NULL
 * removes it. */
tr_status tr_scene_set_global_cube_map(tr_scene* scene, uint32_t side, const uint32_t* const faces[6]);
tr_status tr_scene_set_picture_post_process(tr_scene* scene, uint32_t mode);
/* setTransparentBackground (ray-tracing.hh:319; ray-tracing.cc:2364-2365):
 * a pixel none of whose camera rays of a frame hit geometry keeps its
 * backbuffer value (0 before any write) and stays out of the
 * auto-exposure log sum; accumulation is unchanged. */
tr_status tr_scene_set_transparent_background(tr_scene* scene, int32_t enabled);
/* RenderMode (ray-tracing-common.hh:132-139; ray-tracing.cc:2078-2089):
 * 0 Normal; 1 DebugTangents, 2 DebugBinormals, 3 DebugNormals (the frame
 * vector * 0.5 + 0.5 of the first hit), 4 DebugLighting (max(0, N . -dir))
 * at the camera vertex, which ends the path.  Debug modes reject volumes,
 * the radiance cache and render requests. */
tr_status tr_scene_set_render_mode(tr_scene* scene, uint32_t mode);
/* repaint (ray-tracing.hh:323): restart progressive accumulation. */
tr_status tr_scene_repaint(tr_scene* scene);

/* addLightSource / updateLightSource / clearLightSources (ray-tracing.hh:326-331). */
tr_status tr_scene_add_light_source(tr_scene* scene, const tr_light* light, uint32_t* out_index);
tr_status tr_scene_update_light_source(tr_scene* scene, uint32_t index, const tr_light* light);
tr_status tr_scene_clear_light_sources(tr_scene* scene);
/* addSphereLightSource (ray-tracing.cc:1285-1301): light + emissive sphere. */
tr_status tr_scene_add_sphere_light_source(tr_scene* scene, const tr_light* light,
                                           uint32_t* out_geom_id);

/* addTriangleMesh (ray-tracing.hh:356-362, ray-tracing.cc:1506-1582).
 * geom_ids[i] = (instance << 32) | submesh.  Emissive submeshes register a
 * MeshLight, as the reference does (ray-tracing.cc:1554-1564). */
tr_status tr_scene_add_triangle_mesh(tr_scene* scene, const float world[16],
                                     size_t submesh_count, const tr_submesh* submeshes,
                                     size_t index_count, const int32_t* tris,
                                     size_t vert_size, const void* verts,
                                     const tr_mesh_options* mesh_opts, uint64_t* geom_ids);
/* addSphere (ray-tracing.cc:1341-1355). */
tr_status tr_scene_add_sphere(tr_scene* scene, const float center[3], float radius,
                              const tr_material* material, uint64_t* out_geom_id);
/* User shapes.  Each returns its geometry slot.  Hits are limited to the
 * ray's (tnear, tfar), which the reference's Embree callbacks do not check.
 *
 * addRect (ray-tracing.cc:1402-1428; IntersectRect3, intersect.hh:105-131):
 * a one-sided rectangle at pos with half sizes size[0] along tan and
 * size[1] along norm x tan, facing norm.  tc_box = {min u, min v, max u,
 * max v} maps the hit's (u, v) to TexCoord (NULL = the unit square);
 * tangent_map is a handle from tr_scene_bind_texture (0 = none) whose
 * sampleRGB is the tangent-space direction of the shading frame. */
tr_status tr_scene_add_rect(tr_scene* scene, const float pos[3], const float tan[3], const float norm[3],
                            const float size[2], const tr_material* material, const float tc_box[4],
                            uint32_t tangent_map, uint64_t* out_geom_id);
/* addDisk (ray-tracing.cc:1430-1446; IntersectDisk3, intersect.hh:135-161):
 * a one-sided annulus facing norm. */
tr_status tr_scene_add_disk(tr_scene* scene, const float pos[3], const float norm[3], float inner_radius,
                            float outer_radius, const tr_material* material, uint64_t* out_geom_id);
/* addCylinder (ray-tracing.cc:1356-1369; intersection :892-960): a capped
 * cylinder along z. */
tr_status tr_scene_add_cylinder(tr_scene* scene, const float center[3], float radius, float half_height,
                                const tr_material* material, uint64_t* out_geom_id);
/* addObliqueCylinder (ray-tracing.cc:1371-1384; intersection :992-1046): a
 * capped cylinder whose axis leans by tilt * tilt_direction over its half
 * height.  Shaded with the frame of its hit normal (the reference leaves
 * its frame zero, ray-tracing.cc:1210-1237). */
tr_status tr_scene_add_oblique_cylinder(tr_scene* scene, const float center[3], float radius, float half_height,
                                        const float tilt_direction[2], float tilt, const tr_material* material,
                                        uint64_t* out_geom_id);
/* addEllipsoid (ray-tracing.cc:1315-1339; IntersectEllipsoid,
 * intersect.hh:192-221): orientation is a quaternion (x, y, z, w). */
tr_status tr_scene_add_ellipsoid(tr_scene* scene, const float center[3], const float scale[3],
                                 const float orientation[4], const tr_material* material, uint64_t* out_geom_id);
/* addBlocker (ray-tracing.cc:1386-1400): a rect without material; paths
 * and shadow rays that hit it end there. */
tr_status tr_scene_add_blocker(tr_scene* scene, const float pos[3], const float tan[3], const float norm[3],
                               const float size[2], uint64_t* out_geom_id);

/* A heterogeneous participating medium sampled on a regular grid.  It
 * replaces addHierarchicalVolume (ray-tracing.hh:381, ray-tracing.cc:
 * 1448-1504), whose bricks hold u8 density and u8 (theta, phi) fibre
 * angles: a caller flattens its bricks into one grid (INTEGRATION shows the
 * conversion).  Samples sit on the grid nodes: sample (i, j, k) is at
 * min_corner + (i, j, k) * (max_corner - min_corner) / (dims - 1), x
 * fastest, and values between them are trilinear, as RTVolumeSet::
 * sampleScattering interpolates (ray-tracing.cc:283-325).
 *   density    IEEE binary16 per sample: extinction per unit length of a
 *              flake whose projected area sigma(w) is 1 (>= 0, finite).
 *   direction  three binary16 per sample: the fibre (SGGX tangent)
 *              direction; its length does not matter.  The orientation at
 *              a point is the dominant eigenvector of the trilinearly
 *              weighted structure tensor (SampleVolumePoint, :110-182).
 *   max_density  the majorant; 0 takes the grid's largest density, a value
 *              below it is rejected.
 * The material must be SGGXMicroFlake (albedo, sggx_stddev).  Extinction
 * at x towards w is density(x) * sigma(w), tracked with the global majorant
 * max_density * max(sggx_stddev). */
typedef struct tr_grid_volume {
    uint32_t dims[3];        /* >= 2 each */
    float min_corner[3];
    float max_corner[3];
    const uint16_t* density;
    const uint16_t* direction;
    float max_density;
} tr_grid_volume;

tr_status tr_scene_add_grid_volume(tr_scene* scene, const tr_grid_volume* volume,
                                   const tr_material* material, uint64_t* out_geom_id);

/* VolumeRoot / Volume (volume/volume.hh:65-83): a grid of dims bricks
 * spanning [min_corner, max_corner]; brick (x, y, z) is
 * bricks[(z * dims[1] + y) * dims[0] + x].  A brick's data is its u8
 * densities (x fastest) followed by one (theta, phi) byte pair per voxel
 * (the QuantizedDirections grid); data = NULL is an empty brick. */
typedef struct tr_volume_brick {
    int32_t dims[3];
    uint8_t* data;
} tr_volume_brick;
typedef struct tr_volume_root {
    float min_corner[3];
    float max_corner[3];
    int32_t dims[3];
    tr_volume_brick* bricks;
} tr_volume_root;

/* ParseVolumeHierarchy (src/volume/volume.cc:361-420): reads
 * <name>_description.vol and the <name>_XXX_YYY_ZZZ-density.vol /
 * -orientation.vol pair of every brick it lists ("VOL\x3" grids;
 * src/volume/volume.cc:93-359 converts Float32 and UInt8 inputs).
 * Release with tr_volume_root_free. */
tr_status tr_parse_volume_hierarchy(const char* name, tr_volume_root* root);
/* ParseVolume (src/volume/volume.cc:422-453): one brick from a density and
 * an orientation file; the bounds come from the density file. */
tr_status tr_parse_volume(const char* density_file, const char* orientation_file, tr_volume_root* root);
void tr_volume_root_free(tr_volume_root* root);
/* addHierarchicalVolume (ray-tracing.cc:1448-1504): every non-empty brick
 * becomes a volume box sampled as sampleScattering samples bricks (u8
 * density trilinear, quantized fibre angles through cos/sin tables,
 * majorant DensityScale = 100).  The material must be SGGXMicroFlake.
 * *out_geom_id is the first brick's geometry slot (UINT64_MAX if every brick
 * is empty). */
tr_status tr_scene_add_hierarchical_volume(tr_scene* scene, const tr_volume_root* root,
                                           const tr_material* material, uint64_t* out_geom_id);

/* commitScene (ray-tracing.cc:1633-1636): packs the scene into HBM and
 * builds the 8-wide BVH. */
tr_status tr_scene_commit(tr_scene* scene);
tr_status tr_scene_get_info(const tr_scene* scene, tr_scene_info* info);

/* drawOnce (ray-tracing.cc:1698-1703): one frame of SamplesCamera camera
 * samples per pixel, synchronous, accumulated progressively. */
tr_status tr_scene_draw_once(tr_scene* scene);
/* draw(width, height, view_proj_inv) (ray-tracing.cc:1705-1773): starts the
 * next frame; a changed camera or size resets accumulation. */
tr_status tr_scene_draw(tr_scene* scene, uint32_t width, uint32_t height,
                        const float view_proj_inv[16]);
/* Completes the frame started by tr_scene_draw. */
tr_status tr_scene_frame_wait(tr_scene* scene);
/* The RGBA8 backbuffer of the last frame (FrameData::Backbuffer). */
tr_status tr_scene_read_backbuffer(tr_scene* scene, uint32_t* rgba8, size_t pixel_count);
/* The FP64 accumulation: per pixel the RGB sum over all camera samples of
 * all frames since the last reset, and the number of samples in it. */
tr_status tr_scene_read_accum(tr_scene* scene, double* rgb_sum, size_t pixel_count,
                              uint64_t* samples_per_pixel);

/* A decoded image: rows in the order the reference's loaders leave them,
 * format TR_TEX_RGBA8_UNORM or TR_TEX_R8_UNORM. */
typedef struct tr_image {
    uint32_t width, height, format;
    uint8_t* data;
} tr_image;

/* RTMeshBlob (mesh/obj-loader.hh:47-59): submeshes whose material
 * pointers point into `materials`, interleaved vertices (byte offsets in
 * the submeshes), group-local int32 index triplets, the decoded map_Kd
 * images (TextureData).  `stride` is kept for the reference's field, which
 * its loader never sets (0).  `ignored_maps` counts map_Kd files that did
 * not load. */
typedef struct tr_mesh_blob {
    uint32_t submesh_count;
    tr_submesh* submeshes;
    uint8_t* vertex_data;
    size_t vertex_bytes;
    int32_t* index_data;
    size_t index_count;
    uint32_t stride;
    tr_material* materials;
    uint32_t material_count;
    uint32_t ignored_maps;
    tr_image* textures;        /* the map_Kd images; materials' diffuse_map = index + 1 */
    uint32_t texture_count;
} tr_mesh_blob;
/* TEMPEST_OBJ_LOADER_* flags (mesh/obj-loader.hh:36-40) */
enum { TR_OBJ_GENERATE_TANGENTS = 1, TR_OBJ_GENERATE_CONSISTENT_NORMALS = 2 };
/* LoadObjFileStaticRTGeometry(filename, nullptr, blob, flags)
 * (src/mesh/obj-loader.cc:635-840): parses the OBJ file and the MTL files
 * it names; any diagnostic fails the load (TR_ERROR_INVALID_ARG, the
 * message names file and line).  Release with tr_mesh_blob_free. */
tr_status tr_load_obj_rt_geometry(const char* filename, uint32_t flags, tr_mesh_blob* blob);
/* RTMeshBlob::~RTMeshBlob (src/mesh/obj-loader.cc:51-59) */
void tr_mesh_blob_free(tr_mesh_blob* blob);
/* Binds the blob's textures to a scene (bilinear, as the reference's
 * loaded textures sample) and rewrites its materials' diffuse_map from
 * blob indices to that scene's handles; call once, before adding the
 * blob's meshes. */
tr_status tr_scene_bind_mesh_blob_textures(tr_scene* scene, tr_mesh_blob* blob);

/* Image files of an RGBA8 picture whose row 0 is the bottom row (the
 * backbuffer's layout). */
enum { TR_IMAGE_TGA = 0, TR_IMAGE_PNG = 1 };
/* The file bytes of one format: SaveTGAImage (src/image/tga-image.cc:
 * 247-309) byte for byte; PNG holds the pixels SavePNGImage
 * (src/image/png-image.cc:249-280) stores, deflated by zlib.  out = NULL
 * asks for the size only. */
tr_status tr_image_encode(uint32_t format, uint32_t width, uint32_t height, const uint32_t* rgba8,
                          uint8_t* out, size_t capacity, size_t* size);
/* SaveImage (src/image/image.cc:60-80): the format from the file's
 * lower-cased extension, "tga" or "png"; "exr" and others fail with
 * TR_ERROR_UNSUPPORTED (the reference's EXR writer has no RGBA8 case). */
tr_status tr_image_save(const char* path, uint32_t width, uint32_t height, const uint32_t* rgba8);
/* RayTracingSystem::saveImage (ray-tracing-system.hh:125-130): the last
 * frame's backbuffer through tr_image_save. */
tr_status tr_scene_save_image(tr_scene* scene, const char* path);

/* LoadImage (src/image/image.cc:36-58): "tga" (LoadTGAImage,
 * tga-image.cc:98-242: true colour, 24 or 32 bits, plain or run-length;
 * rows bottom-up as stored unless the descriptor says top-down) or "png"
 * (LoadPNGImage, png-image.cc:125-246: 8-bit grey, grey-alpha, RGB, RGBA
 * and palettes of 1-8 bits, tRNS alpha, non-interlaced; rows top-down as
 * stored) by the lower-cased extension.  Release with tr_image_free. */
tr_status tr_image_load(const char* path, tr_image* image);
void tr_image_free(tr_image* image);

/* Philox key of the counter-based RNG (new; the reference has no seed). */
tr_status tr_scene_set_seed(tr_scene* scene, uint64_t seed);
tr_status tr_scene_set_estimator(tr_scene* scene, uint32_t mode);
/* Frame index used in the RNG counter of the next draw. */
tr_status tr_scene_set_frame_index(tr_scene* scene, uint32_t frame);

/* Closest hit of a fixed ray set, as rtcIntersect / IntersectBVH
 * (ray-tracing.cc:2056; ray-tracing.cu:821-846) answer it. */
tr_status tr_scene_trace_rays(tr_scene* scene, size_t ray_count, const tr_ray* rays,
                              tr_hit* hits);
/* tr_scene_trace_rays without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Rays and hits are DEVICE
 * pointers the caller owns; the traversal's spill stack lives in a 16 MiB
 * file-scope device array that each GPU holds until the process ends, so
 * the BVH may be at most 24 levels deep (TR_ERROR_UNSUPPORTED otherwise)
 * and the persistent grid at most 2048 waves.  `stream` (a hipStream_t)
 * orders the work and the call returns without waiting; NULL uses the
 * scene's stream and waits.  Calls that overlap on one GPU share that
 * array and race. */
tr_status tr_scene_trace_rays_noalloc(tr_scene* scene, size_t ray_count, const tr_ray* device_rays,
                                      tr_hit* device_hits, void* stream);
/* SampleData as rayQuery fills it (ray-tracing-common.hh:184-198, the
 * fields sample() writes): the material pointer the geometry was added
 * with (NULL for blockers and sphere lights, whose material is internal),
 * the shading frame, TexCoord, and the hit distance along the query ray. */
typedef struct tr_sample_data {
    const tr_material* material;
    float normal[3];
    float tangent[3];
    float binormal[3];
    float texcoord[2];
    float t;
} tr_sample_data;

/* rayQuery (ray-tracing.hh:372, ray-tracing.cc:1932-1979): the closest hit
 * of the camera ray through texture coordinate tc of the frame (tc.y
 * downwards).  *out_geom_id is the id the geometry's add call returned, or
 * TR_INVALID_GEOMETRY on a miss (sample_data is then left untouched). */
tr_status tr_scene_ray_query(tr_scene* scene, const float tc[2], tr_sample_data* sample_data,
                             uint64_t* out_geom_id);
/* Renders every camera sample of the listed pixels and returns, per
 * camera sample, its radiance (3 floats) and a hash of its hit sequence. */
tr_status tr_scene_render_paths(tr_scene* scene, size_t pixel_count, const uint32_t* pixels,
                                float* out_rgb, uint32_t* out_hit_hash);
tr_status tr_scene_get_stats(const tr_scene* scene, tr_stats* stats);

/* Batched render requests.  Replaces the reference's serial request loop
 * of samples/virtual-gonioreflectometer.cc:161-189 (per request:
 * updateLightSource of the directional light, drawOnce, copy of the
 * image) by one call: every request is a frame of its own camera, image
 * size and camera sample count, with light 0 (a directional light) turned
 * to the request's direction.  All requests' paths run through one path
 * pool in batches that end at pixel boundaries; the bounce loop of a batch
 * is a captured HIP graph replayed once per batch.  A request's pixel p,
 * camera sample c draws the random numbers of frame pixel
 * (first pixel of the request over all requests) + p, sample c, so the
 * results do not depend on the batching. */
typedef struct tr_request {
    float view_proj_inv[16];
    uint32_t width, height;    /* >= 2 each */
    uint32_t samples_camera;   /* 1 .. 65536 */
    float light_direction[3];  /* DirectionalLight::Direction of light 0 */
} tr_request;

typedef struct tr_request_options {
    int32_t use_graph;           /* 1: the bounce loops of all batches of a configuration in one
                                  * captured graph, launched once; 0: plain launches */
    int32_t autotune;            /* 1: time candidate pool sizes and traversal grids on the
                                  * first batches and keep the fastest for the rest */
    uint32_t pool_slots;         /* paths per batch without autotuning (0: 2^24) */
    uint32_t trace_waves_per_cu; /* persistent traversal waves per CU without autotuning (0: 16) */
} tr_request_options;

typedef struct tr_request_stats {
    uint64_t paths, pixels;
    uint32_t batches;
    uint32_t pool_slots, trace_waves_per_cu; /* configuration of the bulk of the batches */
    uint32_t candidates;                     /* configurations timed by the autotuner */
    double ms;                               /* wall time of the call */
    double tune_ms;                          /* part of it spent on the timed trial batches */
} tr_request_stats;

/* Renders `count` requests.  out_mean (count x 3, may be null): mean
 * radiance of each request over its pixels and samples; out_image (may be
 * null): per-pixel mean radiance, all requests in order, width x height x
 * 3 floats each.  Needs a committed scene whose light 0 is directional,
 * no grid volumes and no radiance cache.  options and stats may be null. */
tr_status tr_scene_render_requests(tr_scene* scene, size_t count, const tr_request* requests,
                                   const tr_request_options* options, float* out_mean, float* out_image,
                                   tr_request_stats* stats);
/* Radiance cache (new; no reference counterpart).  With `enabled`, a path
 * ends at the vertex after its first rough vertex (microfacet alpha >=
 * rough_alpha, a lobe without specular part, or an SGGX surface): the
 * network's prediction of the radiance scattered there, times the path
 * throughput, replaces next-event estimation and the rest of the path.
 * About train_paths paths per frame (every k-th path) are training paths:
 * they run to MaxRayDepth, record the network inputs at their query vertex
 * and, as target, the radiance they gathered after it divided by their
 * throughput there; with `train` the frame ends with Adam steps of `batch`
 * samples over these records.  Needs the unbiased estimator,
 * SamplesGlobalIllumination 1 and no volumes. */
typedef struct tr_cache_settings {
    int32_t enabled;
    int32_t train;
    uint32_t train_paths;
    uint32_t batch;
    float rough_alpha;
    /* 1: the query vertex keeps its next-event estimate and the network
     * predicts the radiance beyond it; 0: the network predicts the whole
     * radiance leaving the query vertex (no shadow rays there) */
    int32_t query_nee;
    uint64_t seed;           /* network initialisation */
} tr_cache_settings;

tr_status tr_scene_set_radiance_cache(tr_scene* scene, const tr_cache_settings* settings);
/* The scene's network (owned by the scene). */
tr_status tr_scene_get_radiance_cache(tr_scene* scene, struct tr_mlp** out_mlp);
/* Training records of the last frame: inputs count x 10, targets count x 3,
 * weights count (record r belongs to path p of the frame with
 * (p + frame) = r * stride, stride = ceil(paths / train_paths)). */
tr_status tr_scene_read_cache_records(tr_scene* scene, size_t count, float* inputs, float* targets,
                                      float* valid);

/* Turns on per-launch node/primitive counting (slower; for measurement). */
tr_status tr_scene_set_counting(tr_scene* scene, int32_t enabled);
/* Turns on per-kernel timing: every launch carries a start and a stop
 * event in its dispatch (hipExtLaunchKernelGGL) and is waited for before
 * the next one (slower; for measurement).  The *_ms fields of tr_stats are
 * filled only in this mode. */
tr_status tr_scene_set_timing(tr_scene* scene, int32_t enabled);

/* ---- radiance-cache network (new; no reference counterpart) ----
 * The MLP that caches outgoing radiance for the path tracer's cache mode:
 * 10 inputs in [0, 1] (position in the scene box, (w_o + 1) / 2,
 * (n + 1) / 2, roughness), 12-band sin/cos encoding (240 features),
 * 4 hidden ReLU layers of 64, 3 linear outputs; 28,099 FP32 parameters in
 * the order W0 (64 x 240, row major), b0, W1, b1, W2, b2, W3, b3, W4
 * (3 x 64), b4.  Products run in BF16 on the matrix cores with FP32
 * accumulation; training minimises the relative L2 loss
 * mean_samples sum_rgb (y - t)^2 / (y^2 + 0.01) (y^2 held constant) with
 * Adam (lr 1e-2, betas 0.9 / 0.99, eps 1e-8) on FP32 master weights and a
 * loss scale of 128.  The host-pointer entry points below copy their
 * arrays (inputs n x 10, outputs and targets n x 3, valid n weights 0/1,
 * row major) and are synchronous. */
typedef struct tr_mlp tr_mlp;
enum { TR_MLP_INPUTS = 10, TR_MLP_OUTPUTS = 3, TR_MLP_PARAMS = 28099 };

tr_status tr_mlp_create(uint64_t seed, tr_mlp** out_mlp);
void tr_mlp_destroy(tr_mlp* mlp);
tr_status tr_mlp_get_params(tr_mlp* mlp, float* params, size_t count);
/* Replaces the parameters and resets the Adam state. */
tr_status tr_mlp_set_params(tr_mlp* mlp, const float* params, size_t count);
tr_status tr_mlp_infer(tr_mlp* mlp, size_t n, const float* inputs, float* outputs);
/* Gradient of the loss at the current parameters (no update) and the loss. */
tr_status tr_mlp_gradient(tr_mlp* mlp, size_t n, const float* inputs, const float* targets,
                          const float* valid, float* grad, float* loss);
/* One Adam step on the n samples; returns the loss before the step. */
tr_status tr_mlp_train_step(tr_mlp* mlp, size_t n, const float* inputs, const float* targets,
                            const float* valid, float* loss);

#ifdef __cplusplus
}
#endif

#endif /* TEMPEST_RT_H */
