/*
 * tr_kernels.h - parameters of the wavefront kernels and their launchers.
 *
 * One frame is processed in batches of path slots.  A slot is one
 * (pixel, camera sample, GI split) triple; slots are numbered pixel-major,
 * so a batch holds whole pixels and the accumulation of a pixel happens in
 * one thread, in a fixed order.  Per batch: ray generation, then for each
 * vertex b = 0..MaxRayDepth: closest-hit trace of the live queue, shading
 * (emission, next-event shadow rays, BSDF sample, queue compaction) and the
 * shadow-ray trace; finally accumulation into FP64 per-pixel sums.
 */
#ifndef TR_KERNELS_H
#define TR_KERNELS_H

#include <hip/hip_runtime.h>
#include "tr_mlp.h"
#include "tr_scene.h"

namespace trk {

/* Batched render requests (tr_scene_render_requests): camera, image size,
 * camera samples and the direction of light 0 per request.  The paths of
 * all requests form one index space (request-major, then pixel, camera
 * sample, GI split), cut into pool batches at pixel boundaries; the
 * kernels read the current batch through a device-side cursor, so one
 * captured graph serves every batch. */
struct ReqRec {
    float vpi[16];
    uint32_t width, height, spp;
    float light_dir[3];       /* DirectionalLight::Direction of light 0 */
};
struct ReqBatch {
    uint64_t path_base;       /* first path of the batch */
    uint64_t pix_base;        /* first pixel (over all requests) */
    uint32_t n_slots;         /* paths of the batch */
    uint32_t n_pix;           /* pixels of the batch */
};
struct ReqView {
    const ReqRec* reqs;
    const uint64_t* path_off; /* [n + 1]: first path of each request */
    const uint64_t* pix_off;  /* [n + 1]: first pixel of each request */
    const ReqBatch* batches;
    const uint32_t* cursor;   /* index of the current batch */
    uint32_t n;               /* requests; 0 outside request mode */
};

struct FrameParams {
    SceneView sc;
    uint32_t width, height;
    uint32_t spp;        /* SamplesCamera */
    uint32_t n_gi;       /* SamplesGlobalIllumination */
    uint32_t n_light;    /* SamplesLocalAreaLight */
    uint32_t max_depth;  /* MaxRayDepth */
    uint32_t rays_sq;    /* sqrt(SamplesGlobalIllumination) */
    uint32_t max_shadow; /* shadow entries per slot */
    float rr;            /* RussianRoulette */
    float background[3];
    uint32_t key0, key1;
    uint32_t frame;
    uint32_t estimator;  /* TR_ESTIMATOR_REFERENCE or TR_ESTIMATOR_UNBIASED */
    float vpi[16];
    /* radiance cache (cache mode only) */
    uint32_t cache;          /* 1: paths end at the vertex after their first rough one */
    uint32_t train_stride;   /* path p of a frame trains iff (p + frame) % train_stride == 0 */
    float rough_alpha;       /* roughness threshold */
    uint32_t cache_nee;      /* 1: next-event estimate at the query vertex, network predicts the rest */
    float box_lo[3];         /* scene box for the position input */
    float box_rcp[3];
    ReqView req;             /* request mode (req.n > 0) */
    uint32_t render_mode;    /* RenderMode (ray-tracing-common.hh:132-139): 0 Normal, 1-3 debug tangents /
                                binormals / normals, 4 debug lighting */
    uint32_t shade_sort;     /* 1: shading blocks sort their paths by material class */
    uint32_t camera_fused;   /* 1: the camera vertex's kernels (k_trace_packet, k_shade) compute the camera
                                rays from the slot and take queue position = slot; k_raygen only sets the
                                queue size (camera_rays_in_kernels) */
};

/* Path flags of cache mode: training path, next vertex is the query
 * vertex, record written, record waits for its radiance-before value
 * (captured when the path reaches the vertex after the query vertex). */
enum { PF_TRAIN = 1u, PF_ROUGH = 2u, PF_RECORDED = 4u, PF_BEFORE = 8u };

/* Training records of one frame, indexed by training path: the network
 * inputs at the path's query vertex, the throughput there, the path
 * radiance gathered up to and including that vertex's next-event
 * estimate, and the path's slot. */
struct CacheRecords {
    MlpSampleSoA smp;     /* inputs, targets, valid (capacity = smp.stride) */
    float* thr;           /* [3][stride] */
    float* before;        /* [3][stride] */
    uint32_t* slot;       /* [stride]: batch-local slot, or ~0 if none */
    MlpQueries queries;   /* the shade's queries at their queue-segment positions (stride seg_cap * kSegment) */
};

/* Structure-of-arrays path state; every array has `capacity` entries
 * (shadow arrays: max_shadow * capacity). */
struct PathPool {
    float4* ray_o;      /* origin, .w unused */
    float4* ray_d;      /* direction, .w unused */
    float4* hit;        /* t, u, v, geom (bits) */
    uint32_t* hit_prim;
    float4* thr;        /* throughput, .w = pending balance weight */
    float4* local;      /* radiance of the camera vertex */
    float4* global;     /* radiance of later vertices */
    uint32_t* hash;     /* FNV-1a of the hit sequence */
    uint32_t* queue[2];
    uint32_t* shadow_queue;
    float4* sh_o;       /* origin, .w = tfar */
    float4* sh_d;       /* direction, .w = light geometry (bits) */
    float4* sh_c;       /* contribution, .w = light index (bits) */
    uint32_t* sh_n;     /* entries per slot */
    float* sh_tmax;     /* largest path-throughput component at the shading point */
    uint32_t* query_queue; /* cache mode: positions of the queries in CacheRecords::queries */
    uint32_t* counters; /* [0],[1] queue sizes, [2] shadow queue size, [3] cache queries */
    /* Segment-local queues written by the shading kernel: block b appends
     * queue k's entries of its threads, in thread order, at
     * seg[k * seg_cap * kSegment + b * kSegment ...] and their number at
     * seg_n[k * seg_cap + b]; k_compact concatenates the segments into the
     * dense queues (k = 0 next path queue, 1 shadow queue, 2 query queue). */
    uint32_t* seg;
    uint32_t* seg_n;
    uint32_t seg_cap;   /* segments per queue */
    uint32_t* group_sum; /* k_compact: [3][kCompactGroups] segment-group sums */
    uint32_t* pflags;   /* cache mode path flags */
    uint32_t* slot_req; /* request mode: request of each slot */
    uint2* spill;       /* traversal stack spill, per persistent thread */
    uint32_t capacity;
};

/* Output of one batch. */
struct BatchOut {
    double* accum;        /* per pixel RGB sums (frame accumulation), may be null */
    float* path_rgb;      /* per camera sample radiance, may be null */
    uint32_t* path_hash;  /* per camera sample hit hash, may be null */
    const uint32_t* pixel_list; /* explicit pixel list, or null for pixel-major order */
    uint64_t* count_out;  /* node/prim counters when counting, may be null */
    uint8_t* hitmask;     /* per pixel: 1 if a camera ray of the frame hit geometry (TransparentBackground), may be null */
};

/* Timing mode: the launch records `start` and `stop` as part of the kernel
 * dispatch itself (hipExtLaunchKernelGGL), so the pair measures exactly the
 * kernel's execution, as a kernel-trace profiler does. */
struct LaunchEvents {
    hipEvent_t start, stop;
};

constexpr int kTraceBlock = 64;       /* one wave per block */
constexpr uint32_t kSegment = 256;    /* shading block = queue segment */
constexpr uint32_t kCompactGroups = 4096; /* k_compact blocks per queue, at most (a multiple of kSegment) */
constexpr int kStackLds = 10;         /* stack entries per lane kept in LDS (5 KB per wave) */
constexpr int kStackSpill = 48;       /* further entries per lane in global memory */
/* Waves per SIMD the refill traversal is compiled for (its register budget,
 * 72 VGPRs; 6 waves with 80 VGPRs were 1.2 % slower on the heightfield, 8
 * with 64 spill); its default persistent grid is 4 * kRefillWavesPerSimd
 * waves per CU, the most that stay resident (LDS: 28 x 5 KB per CU). */
constexpr uint32_t kRefillWavesPerSimd = 7;

/* Whether a frame's camera vertex can compute its rays in the kernels that
 * use them: plain frames (no requests, no cache, not counting) of scenes
 * whose camera rays k_trace_packet traces. */
bool camera_rays_in_kernels(const FrameParams& p, bool counting);
void launch_raygen(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out, uint64_t slot_base,
                   uint32_t n_slots, const CacheRecords& rec, const LaunchEvents* ev = nullptr);
/* grid: persistent grid of the volume-tracking kernel; grid_refill: of the
 * refill kernel (scenes without volumes). */
void launch_trace(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out, uint64_t slot_base,
                  uint32_t vertex, int cur, int nxt, uint32_t grid, uint32_t grid_refill, uint64_t* counts,
                  const LaunchEvents* ev = nullptr);
void launch_shade(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out, uint64_t slot_base,
                  uint32_t vertex, int cur, int nxt, uint32_t n_slots, const CacheRecords& rec,
                  const LaunchEvents* ev = nullptr);
/* Concatenates the segment-local queues of the last shade into the next
 * path queue, the shadow queue and (cache mode) the query queue, in
 * segment order, and writes their sizes to the counters. */
void launch_compact(hipStream_t s, const PathPool& pool, uint32_t n_slots, int cur, int nxt, bool queries,
                    const LaunchEvents* ev = nullptr);
/* Cache mode: network predictions for the queries of the last shade, added
 * to their paths' radiance. */
void launch_cache_infer(hipStream_t s, const MlpWeights& w, const PathPool& pool, const CacheRecords& rec,
                        uint32_t grid, const LaunchEvents* ev = nullptr);
/* Cache mode: targets of the training records [r0, r1) of a batch. */
void launch_cache_targets(hipStream_t s, const PathPool& pool, const CacheRecords& rec, uint32_t r0, uint32_t r1,
                          const LaunchEvents* ev = nullptr);
/* wide: the persistent grid holds more than 16 waves per CU */
void launch_shadow(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out,
                   uint64_t slot_base, uint32_t vertex, uint32_t grid, uint64_t* counts, bool wide,
                   const LaunchEvents* ev = nullptr);
void launch_accumulate(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out,
                       uint64_t slot_base, uint32_t n_pixels, const LaunchEvents* ev = nullptr);
/* Request mode: per-pixel sums of the current batch into accum (one
 * thread per pixel, up to max_pix), then the cursor moves to the next
 * batch. */
void launch_accumulate_requests(hipStream_t s, const FrameParams& p, const PathPool& pool, double* accum,
                                uint32_t max_pix, uint32_t* cursor, const LaunchEvents* ev = nullptr);
/* Request mode outputs: mean RGB of each request over its pixels and
 * samples; optionally the per-pixel mean image (float RGB, all requests
 * in order). */
void launch_request_outputs(hipStream_t s, const ReqView& rv, uint32_t n_gi, const double* accum, float* mean,
                            float* image, uint64_t n_pix_total);
/* The allocation-free ray-set trace: at most kNoAllocWaves persistent
 * waves, kNoAllocSpill spill entries per lane in a file-scope device array
 * (its per-device address is looked up once per device); BVHs up to
 * kStackLds + kNoAllocSpill - 2 levels deep. */
constexpr uint32_t kNoAllocWaves = 2048;
constexpr uint32_t kNoAllocSpill = 16;
hipError_t launch_trace_rays_noalloc(hipStream_t s, const SceneView& sc, const tr_ray* rays, tr_hit* hits, uint32_t n,
                                     uint32_t grid);
void launch_trace_rays(hipStream_t s, const SceneView& sc, const tr_ray* rays, tr_hit* hits, uint32_t n,
                       uint2* spill, uint32_t grid);
/* The backbuffer of the frame's mean radiance through the picture
 * post-process; the auto-exposure modes use `partial` (3 doubles per 256
 * pixels) and write their exposure to `exposure`. */
void launch_pseudo_volume_image(hipStream_t s, const tr_material* m, uint32_t width, uint32_t height, uint32_t* rgba,
                                float* probe);
/* rayQuery: the closest hit of the ray through a screen point and its
 * geometry's frame (zero for blockers and volumes). */
struct Mat16 {
    float m[16];
};
struct RayQueryOut {
    float t, u, v;
    uint32_t geom, prim;
    float normal[3], tangent[3], binormal[3], tc[2];
};
void launch_ray_query(hipStream_t s, const SceneView& sc, const float vpi[16], float tx, float ty, RayQueryOut* out);
/* hitmask (TransparentBackground, may be null): pixels whose camera rays
 * all missed keep their backbuffer value and stay out of the exposure's
 * log sum (ray-tracing.cc:2364-2365). */
void launch_resolve(hipStream_t s, const double* accum, double rcp_count, uint32_t* rgba,
                    uint32_t n_pixels, uint32_t post, double* partial, float* exposure,
                    const uint8_t* hitmask = nullptr);

}  // namespace trk

#endif
