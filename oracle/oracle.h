/*
 * oracle.h - C API of the CPU oracle.
 *
 * ORACLE: test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as
 * the timed CPU baseline.  It is a CPU restatement of the reference CPU
 * path tracer (src/graphics/ray-tracing/ray-tracing.cc:245-340,
 * 1791-2379) that reads the same scene description as the C ABI in
 * include/tempest_rt.h, so a test can build one scene twice.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include "../include/tempest_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

orc_scene* orc_scene_create(uint32_t width, uint32_t height, const float view_proj_inv[16]);
void orc_scene_destroy(orc_scene* scene);

void orc_set_samples_camera(orc_scene* scene, uint32_t n);
void orc_set_samples_global_illumination(orc_scene* scene, uint32_t n);
void orc_set_samples_local_area_light(orc_scene* scene, uint32_t n);
void orc_set_max_ray_depth(orc_scene* scene, uint32_t depth);
void orc_set_russian_roulette(orc_scene* scene, float rr);
void orc_set_background_spectrum(orc_scene* scene, const float rgb[3]);
void orc_set_seed(orc_scene* scene, uint64_t seed);
void orc_set_estimator(orc_scene* scene, uint32_t mode);

int orc_add_triangle_mesh(orc_scene* scene, const float world[16], size_t submesh_count,
                          const tr_submesh* submeshes, size_t index_count, const int32_t* tris,
                          size_t vert_size, const void* verts, const tr_mesh_options* opts,
                          uint64_t* geom_ids);
int orc_add_sphere(orc_scene* scene, const float center[3], float radius,
                   const tr_material* material, uint64_t* out_geom_id);
int orc_add_grid_volume(orc_scene* scene, const tr_grid_volume* volume, const tr_material* material,
                        uint64_t* out_geom_id);
int orc_add_light_source(orc_scene* scene, const tr_light* light, uint32_t* out_index);
int orc_add_sphere_light_source(orc_scene* scene, const tr_light* light, uint32_t* out_geom_id);
int orc_commit(orc_scene* scene);

/* Closest hit of each ray; brute_force != 0 tests every primitive. */
void orc_trace_rays(const orc_scene* scene, size_t n, const tr_ray* rays, tr_hit* hits,
                    int brute_force, int nthreads);

/* Per camera sample of the listed pixels: radiance (FP32 path) and hash of
 * the hit sequence.  Output index = pixel_list_index * samples_camera + s. */
void orc_render_paths(const orc_scene* scene, size_t pixel_count, const uint32_t* pixels,
                      uint32_t frame, float* out_rgb, uint32_t* out_hash, int nthreads);

/* Renders one camera sample and prints every vertex to stderr. */
void orc_trace_path(const orc_scene* scene, uint32_t pixel, uint32_t camera_sample);

/* Rows [row_begin, row_end) of one frame: per pixel the sum and the sum of
 * squares of the camera-sample radiance.  use_fp64 selects the FP64
 * instantiation of the integrator.  Returns the wall time in seconds. */
double orc_render_rows(const orc_scene* scene, uint32_t row_begin, uint32_t row_end,
                       uint32_t frame, int use_fp64, double* out_sum, double* out_sumsq,
                       int nthreads);

#ifdef __cplusplus
}
#endif

#endif
