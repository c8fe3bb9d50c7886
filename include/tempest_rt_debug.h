/*
 * tempest_rt_debug.h - host-side checks of the scene packing and the
 * 8-wide BVH, for tests.  They have no counterpart in the reference and
 * render nothing: tr_debug_trace_rays_host walks the same Node8/PrimRec
 * arrays the GPU kernels read, with the same traversal code compiled for
 * the host, so the builder can be checked on a machine without a GPU.
 */
#ifndef TEMPEST_RT_DEBUG_H
#define TEMPEST_RT_DEBUG_H

#include "tempest_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Builds the BVH on the host without touching a device. */
tr_status tr_debug_build_scene_host(tr_scene* scene);
/* Closest hits through the host copy of the BVH; optional visit counters. */
tr_status tr_debug_trace_rays_host(const tr_scene* scene, size_t ray_count, const tr_ray* rays, tr_hit* hits,
                                   uint64_t* node_visits, uint64_t* prim_tests);
/* Counts containment violations of the quantised boxes and primitives that
 * are not referenced exactly once. */
tr_status tr_debug_bvh_check(const tr_scene* scene, uint64_t* violations);

#ifdef __cplusplus
}
#endif

#endif
