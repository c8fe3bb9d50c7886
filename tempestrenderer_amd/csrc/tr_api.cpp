/*
 * tr_api.cpp - the C ABI of include/tempest_rt.h: scene construction on
 * the host, packing into HBM, and orchestration of the wavefront kernels.
 *
 * Mirrors RayTracerScene (include/tempest/graphics/ray-tracing/
 * ray-tracing.hh:241-395, src/graphics/ray-tracing/ray-tracing.cc) and
 * RayTracerCuda (include/tempest/compute/ray-tracing-cuda.hh:80-172):
 * the same setters and defaults (SamplesGlobalIllumination 64,
 * SamplesLocalAreaLight 4, SamplesCamera 1, MaxRayDepth 1, RussianRoulette
 * 1, ray-tracing.hh:255-261), the same geometry-id scheme for meshes and
 * the same emissive-submesh -> MeshLight registration.
 */
#include "../../include/tempest_rt.h"
#include "../../include/tempest_rt_debug.h"
#include "tr_bvh8.h"
#include "tr_cache.h"
#include "tr_host.h"
#include "tr_kernels.h"
#include "tr_shading.h"
#include "tr_traverse.h"
#include "tr_volume.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace trk;

namespace trk {
thread_local std::string g_error;
}

namespace {

enum { EV_TRACE = 0, EV_SHADE = 1, EV_SHADOW = 2, EV_OTHER = 3, EV_CACHE = 4 };

}  // namespace

struct tr_scene {
    uint32_t width = 0, height = 0;
    float vpi[16];
    tr_settings settings;
    uint32_t spp = 1, n_gi = 64, n_light = 4, max_depth = 1, chunk = 16;
    float rr = 1.0f;
    float background[3] = {0.0f, 0.0f, 0.0f};
    uint32_t post = TR_POST_SRGB;
    bool expo_used = false; /* the last frame ran the auto-exposure passes */
    int32_t transparent = 0;
    uint32_t render_mode = 0;
    uint64_t seed = 1;
    uint32_t estimator = TR_ESTIMATOR_REFERENCE;
    uint32_t frame_index = 0;
    bool counting = false;
    bool timing = false;
    bool cache_trained = false;
    uint32_t cache_steps = 0;
    float cache_step_w[64];

    /* radiance cache */
    tr_cache_settings cache;
    tr_mlp* net = nullptr;
    float box_lo[3] = {0, 0, 0}, box_hi[3] = {1, 1, 1};
    uint32_t rec_cap = 0;
    DevBuf<float> rec_thr, rec_before;
    DevBuf<uint32_t> rec_slot, pflags;
    DevBuf<float> q_in, q_thr;   /* cache queries, [kMlpIn | 3][seg_cap * kSegment] */
    DevBuf<uint32_t> q_slot;

    /* host copy of the scene */
    std::vector<tr_material> mats;
    std::vector<const tr_material*> mat_keys;
    std::vector<uint8_t> vpool;
    std::vector<int32_t> ipool;
    std::vector<GeomRec> geoms;
    std::vector<PrimRec> prims_in;
    std::vector<ShapeRec> shapes;   /* user shapes (addRect, addDisk, ...) */
    std::vector<uint64_t> ext_ids;  /* geometry slot -> the id its add call returned (meshes: instance << 32 | submesh) */
    DevBuf<RayQueryOut> query_out;
    std::vector<LightRec> lights;
    std::vector<VolumeRec> vols;
    std::vector<uint64_t> voxels;   /* voxel words of all volumes (kVoxelHalf / kVoxelQuantized) */
    uint64_t dir_table = UINT64_MAX; /* brick volumes' direction tables in voxels */
    uint32_t instances = 0;
    bool committed = false;
    bool lights_dirty = false;

    /* BVH (host copy kept for the debug traversal) */
    std::vector<Node8> nodes;
    std::vector<PrimRec> prims;
    Bvh8Stats bvh_stats;
    double build_ms = 0.0;

    /* device */
    bool device_ready = false;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    DevBuf<Node8> d_nodes;
    DevBuf<PrimRec> d_prims;
    DevBuf<GeomRec> d_geoms;
    DevBuf<tr_material> d_mats;
    DevBuf<LightRec> d_lights;
    DevBuf<uint8_t> d_vpool;
    DevBuf<int32_t> d_ipool;
    DevBuf<VolumeRec> d_vols;
    DevBuf<ShapeRec> d_shapes;
    DevBuf<uint64_t> d_voxels;
    std::vector<TexRec> tex_recs;    /* bound textures */
    std::vector<uint8_t> texels;     /* their texels, then the cube map's faces */
    std::vector<uint8_t> cube;       /* six flipped RGBA8 faces */
    uint32_t cube_side = 0;
    DevBuf<TexRec> d_tex;
    DevBuf<uint8_t> d_texels;
    DevBuf<float> d_srgb;
    std::vector<float> srgb;         /* host copy of the sRGB decode table */
    uint64_t cube_offset = 0;

    /* path pool */
    uint32_t capacity = 0;
    uint32_t max_shadow = 0;
    DevBuf<float4> ray_o, ray_d, hit, thr, local, global, sh_o, sh_d, sh_c;
    DevBuf<uint32_t> hit_prim, hash, q0, q1, shq, qq, seg, seg_n, sh_n, counters, slot_req, gsum;
    DevBuf<float> sh_tmax;
    DevBuf<uint2> spill;
    DevBuf<uint64_t> counts;

    /* frame accumulation */
    DevBuf<double> accum;
    DevBuf<uint32_t> rgba;
    DevBuf<uint8_t> hitmask;      /* TransparentBackground: camera rays of the frame hit geometry, per pixel */
    DevBuf<double> expo_partial;  /* auto exposure: 3 log sums per 256 pixels */
    DevBuf<float> exposure;       /* auto exposure: the frame's factor */
    std::vector<uint32_t> rgba_host;
    uint64_t accum_samples = 0;
    bool frame_pending = false;

    /* events */
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    tr_stats stats;

    tr_scene() { memset(&cache, 0, sizeof(cache)); }
    ~tr_scene();
};

namespace {

void release_events(tr_scene* s)
{
    for (auto& e : s->ev_pool) (void)hipEventDestroy(e);
    s->ev_pool.clear();
    s->ev_used.clear();
}

hipEvent_t get_event(tr_scene* s)
{
    if (!s->ev_pool.empty()) {
        hipEvent_t e = s->ev_pool.back();
        s->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

tr_scene::~tr_scene()
{
    if (frame_pending && stream) (void)hipStreamSynchronize(stream);
    for (auto& u : ev_used) { ev_pool.push_back(u.second.first); ev_pool.push_back(u.second.second); }
    ev_used.clear();
    release_events(this);
    d_nodes.release(); d_prims.release(); d_geoms.release(); d_mats.release(); d_lights.release();
    d_vpool.release(); d_ipool.release(); d_vols.release(); d_voxels.release(); sh_tmax.release(); d_shapes.release();
    query_out.release();
    ray_o.release(); ray_d.release(); hit.release(); thr.release(); local.release(); global.release();
    sh_o.release(); sh_d.release(); sh_c.release(); hit_prim.release(); hash.release(); q0.release(); q1.release();
    shq.release(); qq.release(); seg.release(); seg_n.release(); sh_n.release(); counters.release(); spill.release(); counts.release(); accum.release(); rgba.release(); hitmask.release(); expo_partial.release(); exposure.release();
    if (net) tr_mlp_destroy(net);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

uint32_t intern_material(tr_scene* s, const tr_material* m)
{
    for (size_t i = 0; i < s->mat_keys.size(); ++i)
        if (s->mat_keys[i] == m && memcmp(&s->mats[i], m, sizeof(*m)) == 0) return (uint32_t)i;
    s->mat_keys.push_back(m);
    s->mats.push_back(*m);
    return (uint32_t)(s->mats.size() - 1);
}

/* Surface models of meshes and spheres.  Volume models belong to grid
 * volumes only: the reference treats a hit on a volume-model geometry as a
 * volume set (ray-tracing.cc:2108-2110). */
bool model_supported(uint32_t m)
{
    return m == TR_MODEL_EMISSIVE || m == TR_MODEL_BLINN_PHONG || (m >= TR_MODEL_GGX_MICROFACET && m <= TR_MODEL_GGX_MICROFACET_DIELECTRIC) ||
           (m >= TR_MODEL_BECKMANN_MICROFACET && m <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC) ||
           (m >= TR_MODEL_GGX_MICROFACET_ANISOTROPIC && m <= TR_MODEL_GGX_MICROFACET_DIELECTRIC_ANISOTROPIC) ||
           (m >= TR_MODEL_BECKMANN_MICROFACET_ANISOTROPIC && m <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC_ANISOTROPIC) ||
           m == TR_MODEL_SGGX_SURFACE || m == TR_MODEL_SGGX_PSEUDO_VOLUME;
}

tr_status ensure_device(tr_scene* s)
{
    if (s->device_ready) return TR_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(TR_ERROR_DEVICE, "no HIP device available");
    int dev = 0;
    TR_HIP(hipGetDevice(&dev));
    TR_HIP(hipDeviceGetAttribute(&s->cu_count, hipDeviceAttributeMultiprocessorCount, dev));
    TR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->device_ready = true;
    return TR_OK;
}

template <class T> tr_status upload(tr_scene* s, DevBuf<T>& buf, const std::vector<T>& v)
{
    TR_HIP(buf.reserve(v.size()));
    if (!v.empty()) TR_HIP(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s->stream));
    return TR_OK;
}

SceneView scene_view(const tr_scene* s)
{
    SceneView v;
    v.nodes = s->d_nodes.ptr;
    v.prims = s->d_prims.ptr;
    v.geoms = s->d_geoms.ptr;
    v.mats = s->d_mats.ptr;
    v.lights = s->d_lights.ptr;
    v.vpool = s->d_vpool.ptr;
    v.ipool = s->d_ipool.ptr;
    v.volumes = s->d_vols.ptr;
    v.voxels = s->d_voxels.ptr;
    v.shapes = s->shapes.empty() ? nullptr : s->d_shapes.ptr;   /* null selects the shape-free traversal */
    v.n_lights = (uint32_t)s->lights.size();
    v.n_prims = (uint32_t)s->prims.size();
    v.n_volumes = (uint32_t)s->vols.size();
    v.tex.recs = s->d_tex.ptr;
    v.tex.texels = s->d_texels.ptr;
    v.tex.srgb = s->d_srgb.ptr;
    v.tex.count = (uint32_t)s->tex_recs.size();
    v.tex.cube_side = s->cube_side;
    v.tex.cube_offset = s->cube_offset;
    v.tex.pseudo_volume = 0;
    for (const tr_material& m : s->mats) v.tex.pseudo_volume |= m.model == TR_MODEL_SGGX_PSEUDO_VOLUME;
    return v;
}

/* Persistent traversal grids: trace_waves_per_cu if set, else 16 waves per
 * CU for the volume-tracking, shadow and ray-set kernels (4 resident per
 * SIMD at their register counts) and 4 * kRefillWavesPerSimd for the
 * refill kernel.  The spill area is sized for the larger. */
uint32_t persistent_grid(const tr_scene* s, bool refill = false)
{
    const uint32_t dflt = refill ? 4u * kRefillWavesPerSimd : 16u;
    const uint32_t waves = s->settings.trace_waves_per_cu ? s->settings.trace_waves_per_cu : dflt;
    return (uint32_t)std::max(1, s->cu_count) * waves;
}

uint32_t spill_grid(const tr_scene* s) { return std::max(persistent_grid(s), persistent_grid(s, true)); }

tr_status sync_lights(tr_scene* s)
{
    if (!s->lights_dirty) return TR_OK;
    tr_status st = upload(s, s->d_lights, s->lights);
    if (st != TR_OK) return st;
    s->lights_dirty = false;
    return TR_OK;
}

uint32_t shadow_entries(const tr_scene* s)
{
    uint32_t n = 0;
    for (const LightRec& L : s->lights)
        n += multi_sample_light(L, s->estimator == TR_ESTIMATOR_UNBIASED) ? std::max(1u, s->n_light) : 1u;
    return n;
}

/* Device bytes per path slot: six float4 and six uint32 arrays, three
 * queue segments, three more words, and the shadow entries. */
uint64_t slot_bytes(uint32_t max_shadow) { return 96u + 24u + 12u + 12u + 48ull * max_shadow; }

/* The default path pool: the whole frame's paths resident at once, up to
 * 2^27 slots (25 GB at one shadow entry per slot) and a quarter of the
 * device memory free now.  Fewer, larger batches leave the persistent
 * traversal less tail per launch and fewer nearly empty late-bounce
 * launches: heightfield frame 2^22 / 2^23 / 2^24 / 2^25 / 2^26 / 2^27
 * slots: 2,112 / 2,614 / 3,029 / 3,253 / 3,340 / 3,450 Msamples/s. */
uint32_t default_pool_slots(uint32_t max_shadow)
{
    uint32_t slots = 1u << 27;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) {
        const uint64_t fit = (uint64_t)free_b / 4u / slot_bytes(max_shadow);
        while (slots > (1u << 20) && slots > fit) slots >>= 1;
    }
    return slots;
}

tr_status ensure_pool(tr_scene* s, uint64_t slots_needed)
{
    uint32_t want = s->settings.path_pool_size ? s->settings.path_pool_size
                                               : default_pool_slots(std::max(1u, shadow_entries(s)));
    uint64_t per_pixel = (uint64_t)s->spp * s->n_gi;
    if (per_pixel > want) want = (uint32_t)per_pixel;
    if (slots_needed < want) want = (uint32_t)std::max<uint64_t>(slots_needed, per_pixel);
    want = (uint32_t)(want / per_pixel * per_pixel);
    uint32_t ms = std::max(1u, shadow_entries(s));
    if (want <= s->capacity && ms <= s->max_shadow) return TR_OK;
    uint32_t cap = std::max(want, s->capacity);
    uint32_t msh = std::max(ms, s->max_shadow);
    TR_HIP(s->ray_o.reserve(cap));
    TR_HIP(s->ray_d.reserve(cap));
    TR_HIP(s->hit.reserve(cap));
    TR_HIP(s->thr.reserve(cap));
    TR_HIP(s->local.reserve(cap));
    TR_HIP(s->global.reserve(cap));
    TR_HIP(s->hit_prim.reserve(cap));
    TR_HIP(s->hash.reserve(cap));
    TR_HIP(s->q0.reserve(cap));
    TR_HIP(s->q1.reserve(cap));
    TR_HIP(s->shq.reserve(cap));
    TR_HIP(s->qq.reserve(cap));
    uint32_t seg_cap = (cap + kSegment - 1) / kSegment;
    TR_HIP(s->seg.reserve((size_t)3 * seg_cap * kSegment));
    TR_HIP(s->seg_n.reserve((size_t)3 * seg_cap));
    TR_HIP(s->sh_n.reserve(cap));
    TR_HIP(s->sh_tmax.reserve(cap));
    TR_HIP(s->pflags.reserve(cap));
    TR_HIP(s->sh_o.reserve((size_t)cap * msh));
    TR_HIP(s->sh_d.reserve((size_t)cap * msh));
    TR_HIP(s->sh_c.reserve((size_t)cap * msh));
    TR_HIP(s->counters.reserve(8));
    TR_HIP(s->gsum.reserve(3 * kCompactGroups));
    TR_HIP(s->spill.reserve((size_t)spill_grid(s) * kTraceBlock * kStackSpill));
    TR_HIP(s->counts.reserve(8));
    s->capacity = cap;
    s->max_shadow = msh;
    return TR_OK;
}

PathPool path_pool(const tr_scene* s)
{
    PathPool p;
    memset(&p, 0, sizeof(p));
    p.ray_o = s->ray_o.ptr; p.ray_d = s->ray_d.ptr; p.hit = s->hit.ptr; p.hit_prim = s->hit_prim.ptr;
    p.thr = s->thr.ptr; p.local = s->local.ptr; p.global = s->global.ptr; p.hash = s->hash.ptr;
    p.queue[0] = s->q0.ptr; p.queue[1] = s->q1.ptr; p.shadow_queue = s->shq.ptr;
    p.sh_o = s->sh_o.ptr; p.sh_d = s->sh_d.ptr; p.sh_c = s->sh_c.ptr; p.sh_n = s->sh_n.ptr; p.sh_tmax = s->sh_tmax.ptr;
    p.counters = s->counters.ptr; p.spill = s->spill.ptr; p.capacity = s->capacity; p.pflags = s->pflags.ptr;
    p.query_queue = s->qq.ptr; p.seg = s->seg.ptr; p.seg_n = s->seg_n.ptr; p.slot_req = s->slot_req.ptr;
    p.group_sum = s->gsum.ptr;
    p.seg_cap = (s->capacity + kSegment - 1) / kSegment;
    return p;
}

FrameParams frame_params(const tr_scene* s)
{
    FrameParams p;
    memset(&p, 0, sizeof(p));
    p.sc = scene_view(s);
    p.width = s->width;
    p.height = s->height;
    p.spp = s->spp;
    p.n_gi = s->n_gi;
    p.n_light = s->n_light;
    p.max_depth = s->max_depth;
    p.rays_sq = (uint32_t)sqrtf((float)s->n_gi);
    if (p.rays_sq == 0) p.rays_sq = 1;
    p.max_shadow = s->max_shadow;
    p.rr = s->rr;
    memcpy(p.background, s->background, sizeof(p.background));
    p.key0 = (uint32_t)s->seed;
    p.key1 = (uint32_t)(s->seed >> 32);
    p.frame = s->frame_index;
    p.estimator = s->estimator;
    p.cache = s->cache.enabled ? 1u : 0u;
    p.train_stride = 1;
    p.rough_alpha = s->cache.rough_alpha;
    p.cache_nee = s->cache.query_nee ? 1u : 0u;
    for (int a = 0; a < 3; ++a) {
        p.box_lo[a] = s->box_lo[a];
        float ext = s->box_hi[a] - s->box_lo[a];
        p.box_rcp[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    memcpy(p.vpi, s->vpi, sizeof(p.vpi));
    p.render_mode = s->render_mode;
    p.shade_sort = s->settings.shade_order;
    return p;
}

void record(tr_scene* s, int cat, hipEvent_t a, hipEvent_t b) { s->ev_used.push_back({cat, {a, b}}); }

/* Enqueues every batch of a frame (or of a pixel list) on the scene stream. */
tr_status enqueue_frame(tr_scene* s, uint64_t n_pixels, const BatchOut& out)
{
    if (!s->committed) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    if (s->spp == 0 || s->n_gi == 0) return fail(TR_ERROR_INVALID_ARG, "SamplesCamera and SamplesGlobalIllumination must be positive");
    uint32_t rsq = (uint32_t)sqrtf((float)s->n_gi);
    if (rsq * rsq != s->n_gi) return fail(TR_ERROR_INVALID_ARG, "SamplesGlobalIllumination must be a perfect square");
    if (s->bvh_stats.depth + 2 > (uint32_t)(kStackLds + kStackSpill))
        return fail(TR_ERROR_UNSUPPORTED, "BVH deeper than the traversal stack");
    if (s->n_light > 65535) return fail(TR_ERROR_UNSUPPORTED, "SamplesLocalAreaLight above 65535");
    if (!s->vols.empty() && s->estimator == TR_ESTIMATOR_UNBIASED)
        return fail(TR_ERROR_UNSUPPORTED, "the unbiased estimator does not take volumes: the reference's micro-flake "
                                          "sampler has no closed-form density");
    if (!s->vols.empty() && (s->lights.size() > 256 || s->n_light > 256))
        return fail(TR_ERROR_UNSUPPORTED, "scenes with volumes take at most 256 lights and 256 samples per area light");
    if (s->render_mode != 0 && !s->vols.empty())
        return fail(TR_ERROR_UNSUPPORTED, "debug render modes do not take volumes");
    if (s->render_mode != 0 && s->cache.enabled)
        return fail(TR_ERROR_UNSUPPORTED, "debug render modes do not take the radiance cache");
    if (s->cache.enabled) {
        if (s->estimator != TR_ESTIMATOR_UNBIASED)
            return fail(TR_ERROR_UNSUPPORTED, "the radiance cache needs the unbiased estimator");
        if (s->n_gi != 1) return fail(TR_ERROR_UNSUPPORTED, "the radiance cache needs SamplesGlobalIllumination = 1");
        if (!s->vols.empty()) return fail(TR_ERROR_UNSUPPORTED, "the radiance cache does not take volumes");
    }
    tr_status st = sync_lights(s);
    if (st != TR_OK) return st;
    uint64_t per_pixel = (uint64_t)s->spp * s->n_gi;
    st = ensure_pool(s, n_pixels * per_pixel);
    if (st != TR_OK) return st;
    const size_t q_stride = (size_t)(s->capacity + kSegment - 1) / kSegment * kSegment;
    if (s->cache.enabled && s->q_slot.count < q_stride) {
        TR_HIP(s->q_in.reserve(q_stride * kMlpIn));
        TR_HIP(s->q_thr.reserve(q_stride * 3));
        TR_HIP(s->q_slot.reserve(q_stride));
    }
    FrameParams p = frame_params(s);
    PathPool pool = path_pool(s);
    /* cache mode: every train_stride-th path of the frame trains */
    uint64_t total_paths = n_pixels * per_pixel;
    CacheRecords rec;
    memset(&rec, 0, sizeof(rec));
    uint32_t n_records = 0;
    if (s->cache.enabled) {
        uint32_t want = std::max(1u, s->cache.train_paths);
        p.train_stride = (uint32_t)std::max<uint64_t>(1, (total_paths + want - 1) / want);
        n_records = (uint32_t)((total_paths + p.frame + p.train_stride - 1) / p.train_stride);
        MlpState& ms = s->net->st;
        if ((st = ms.ensure_train(n_records)) != TR_OK) return st;
        if (s->rec_cap < ms.train_cap) {
            TR_HIP(s->rec_thr.reserve((size_t)3 * ms.train_cap));
            TR_HIP(s->rec_before.reserve((size_t)3 * ms.train_cap));
            TR_HIP(s->rec_slot.reserve(ms.train_cap));
            s->rec_cap = ms.train_cap;
        }
        rec.smp = ms.samples();
        rec.thr = s->rec_thr.ptr;
        rec.before = s->rec_before.ptr;
        rec.slot = s->rec_slot.ptr;
        rec.queries.in = s->q_in.ptr;
        rec.queries.thr = s->q_thr.ptr;
        rec.queries.slot = s->q_slot.ptr;
        rec.queries.stride = (uint32_t)q_stride;
        /* records beyond the frame's paths stay at weight 0 */
        TR_HIP(hipMemsetAsync(ms.s_valid.ptr, 0, (size_t)ms.train_cap * sizeof(float), s->stream));
    }
    uint32_t ppb = (uint32_t)std::max<uint64_t>(1, s->capacity / per_pixel);
    uint32_t grid = persistent_grid(s);
    uint32_t grid_refill = persistent_grid(s, true);
    uint64_t* counts = nullptr;
    if (s->counting) {
        counts = s->counts.ptr;
        TR_HIP(hipMemsetAsync(counts, 0, 8 * sizeof(uint64_t), s->stream));
    }
    p.camera_fused = camera_rays_in_kernels(p, counts != nullptr) ? 1u : 0u;
    /* In timing mode every launch carries its own start / stop events in
     * the dispatch (hipExtLaunchKernelGGL) and is waited for, so the pair
     * measures the kernel alone, without launch gaps. */
    auto stage = [&](int cat, auto&& launch) -> tr_status {
        if (!s->timing) {
            launch((const LaunchEvents*)nullptr);
            TR_HIP(hipGetLastError());
            return TR_OK;
        }
        LaunchEvents ev;
        ev.start = get_event(s);
        ev.stop = get_event(s);
        if (!ev.start || !ev.stop) return fail(TR_ERROR_DEVICE, "event creation failed");
        launch(&ev);
        TR_HIP(hipGetLastError());
        TR_HIP(hipEventSynchronize(ev.stop));
        record(s, cat, ev.start, ev.stop);
        return TR_OK;
    };
    /* A stage of several launches: in timing mode bracketed by events
     * recorded on the stream, after the previous stage has completed. */
    auto stage_multi = [&](int cat, auto&& launch) -> tr_status {
        if (!s->timing) {
            launch();
            TR_HIP(hipGetLastError());
            return TR_OK;
        }
        hipEvent_t a = get_event(s), b = get_event(s);
        if (!a || !b) return fail(TR_ERROR_DEVICE, "event creation failed");
        TR_HIP(hipEventRecord(a, s->stream));
        launch();
        TR_HIP(hipGetLastError());
        TR_HIP(hipEventRecord(b, s->stream));
        TR_HIP(hipEventSynchronize(b));
        record(s, cat, a, b);
        return TR_OK;
    };
    for (uint64_t pix0 = 0; pix0 < n_pixels; pix0 += ppb) {
        uint32_t npx = (uint32_t)std::min<uint64_t>(ppb, n_pixels - pix0);
        uint32_t nslots = (uint32_t)(npx * per_pixel);
        uint64_t slot_base = pix0 * per_pixel;
        tr_status st2;
        if ((st2 = stage(EV_OTHER, [&](const LaunchEvents* ev) {
                 launch_raygen(s->stream, p, pool, out, slot_base, nslots, rec, ev);
             })) != TR_OK)
            return st2;
        int cur = 0, nxt = 1;
        for (uint32_t v = 0; v <= s->max_depth; ++v) {
            if ((st2 = stage(EV_TRACE, [&](const LaunchEvents* ev) {
                     launch_trace(s->stream, p, pool, out, slot_base, v, cur, nxt, grid, grid_refill, counts, ev);
                 })) != TR_OK)
                return st2;
            if ((st2 = stage(EV_SHADE, [&](const LaunchEvents* ev) {
                     launch_shade(s->stream, p, pool, out, slot_base, v, cur, nxt, nslots, rec, ev);
                 })) != TR_OK)
                return st2;
            if ((st2 = stage(EV_SHADE, [&](const LaunchEvents* ev) {
                     launch_compact(s->stream, pool, nslots, cur, nxt, s->cache.enabled, ev);
                 })) != TR_OK)
                return st2;
            if (s->cache.enabled && v > 0) {
                MlpWeights w = s->net->st.weights();
                if ((st2 = stage_multi(EV_CACHE, [&]() {
                         launch_cache_infer(s->stream, w, pool, rec, (uint32_t)std::max(1, s->cu_count));
                     })) != TR_OK)
                    return st2;
            }
            if (p.sc.n_lights > 0) {
                if ((st2 = stage(EV_SHADOW, [&](const LaunchEvents* ev) {
                         launch_shadow(s->stream, p, pool, out, slot_base, v, grid, counts,
                                       grid > 16u * (uint32_t)std::max(1, s->cu_count), ev);
                     })) != TR_OK)
                    return st2;
            }
            std::swap(cur, nxt);
        }
        if (s->cache.enabled) {
            /* records of the paths of this batch: path q = slot_base + i */
            uint64_t q0 = slot_base + p.frame, q1 = slot_base + nslots + p.frame;
            uint32_t r0 = (uint32_t)((q0 + p.train_stride - 1) / p.train_stride);
            uint32_t r1 = (uint32_t)((q1 + p.train_stride - 1) / p.train_stride);
            if ((st2 = stage(EV_CACHE, [&](const LaunchEvents* ev) {
                     launch_cache_targets(s->stream, pool, rec, r0, r1, ev);
                 })) != TR_OK)
                return st2;
        }
        if ((st2 = stage(EV_OTHER, [&](const LaunchEvents* ev) {
                 launch_accumulate(s->stream, p, pool, out, slot_base, npx, ev);
             })) != TR_OK)
            return st2;
    }
    if (s->cache.enabled && s->cache.train) {
        /* online training on this frame's records, in batches */
        MlpState& ms = s->net->st;
        uint32_t batch = std::max(MlpState::kSlice, s->cache.batch / MlpState::kSlice * MlpState::kSlice);
        /* records below ceil(frame / stride) belong to no path of the frame */
        uint32_t first = (uint32_t)((p.frame + p.train_stride - 1) / p.train_stride) / MlpState::kSlice * MlpState::kSlice;
        s->cache_steps = 0;
        for (uint32_t off = first, nb = 0; off < n_records; off += nb) {
            nb = std::min(batch, n_records - off);
            /* a remainder below a quarter batch joins the step before it
             * instead of paying for a step of its own */
            if (n_records - off - nb > 0 && n_records - off - nb < batch / 4) nb = n_records - off;
            tr_status st3 = TR_OK;
            uint32_t slot = s->cache_steps;
            tr_status st2 = stage_multi(EV_CACHE, [&]() { st3 = ms.enqueue_step(nb, nb, s->stream, true, off, slot); });
            if (st2 != TR_OK) return st2;
            if (st3 != TR_OK) return st3;
            if (s->cache_steps < MlpState::kMaxLossSlots) s->cache_step_w[s->cache_steps++] = (float)nb;
        }
        s->cache_trained = true;
    }
    return TR_OK;
}

/* Collects event timings after the stream has been synchronised. */
void collect_stats(tr_scene* s, uint64_t paths)
{
    double ms[5] = {0, 0, 0, 0, 0};
    uint64_t traces = 0;
    for (auto& u : s->ev_used) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, u.second.first, u.second.second) == hipSuccess) ms[u.first] += t;
        if (u.first == EV_TRACE) ++traces;
        s->ev_pool.push_back(u.second.first);
        s->ev_pool.push_back(u.second.second);
    }
    s->ev_used.clear();
    s->stats.paths = paths;
    s->stats.trace_ms = ms[EV_TRACE];
    s->stats.shade_ms = ms[EV_SHADE];
    s->stats.shadow_ms = ms[EV_SHADOW];
    s->stats.other_ms = ms[EV_OTHER];
    s->stats.cache_ms = ms[EV_CACHE];
    s->stats.cache_loss = 0.0;
    if (s->cache_trained && s->net) s->stats.cache_loss = s->net->st.read_loss(s->cache_steps, s->cache_step_w);
    s->cache_trained = false;
    s->stats.exposure = 0.0;
    if (s->expo_used) {
        float e = 0.0f;
        if (hipMemcpy(&e, s->exposure.ptr, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess) s->stats.exposure = e;
    }
    s->stats.trace_launches = traces;
    s->stats.node_visits = s->stats.prim_tests = s->stats.counted_rays = 0;
    s->stats.segments = s->stats.shadow_rays = s->stats.volume_steps = 0;
    if (s->counting) {
        /* [0..2] closest-hit kernel: nodes, primitives, rays;
         * [3] shadow rays, [4..5] shadow kernel: nodes, primitives;
         * [6] tentative volume collisions of both kernels */
        uint64_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (hipMemcpy(c, s->counts.ptr, sizeof(c), hipMemcpyDeviceToHost) == hipSuccess) {
            s->stats.node_visits = c[0];
            s->stats.prim_tests = c[1];
            s->stats.counted_rays = c[2];
            s->stats.segments = c[2];
            s->stats.shadow_rays = c[3];
            s->stats.volume_steps = c[6];
        }
    }
}

tr_status finish_frame(tr_scene* s, std::chrono::steady_clock::time_point t0, uint64_t paths)
{
    TR_HIP(hipStreamSynchronize(s->stream));
    auto t1 = std::chrono::steady_clock::now();
    collect_stats(s, paths);
    s->stats.frame_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    s->frame_pending = false;
    return TR_OK;
}

tr_status start_frame(tr_scene* s)
{
    if (!s->committed) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    uint64_t npx = (uint64_t)s->width * s->height;
    TR_HIP(s->accum.reserve(npx * 3));
    if (s->rgba.count < npx) {
        /* a pixel that TransparentBackground never writes stays 0 */
        TR_HIP(s->rgba.reserve(npx));
        TR_HIP(hipMemsetAsync(s->rgba.ptr, 0, npx * sizeof(uint32_t), s->stream));
    }
    if (s->transparent) TR_HIP(s->hitmask.reserve(npx));
    s->expo_used = s->post == TR_POST_AUTO_EXPOSURE_SRGB || s->post == TR_POST_AUTO_EXPOSURE_HDR;
    if (s->expo_used) {
        TR_HIP(s->expo_partial.reserve((npx + 255) / 256 * 3));
        TR_HIP(s->exposure.reserve(1));
    }
    if (s->accum_samples == 0) TR_HIP(hipMemsetAsync(s->accum.ptr, 0, npx * 3 * sizeof(double), s->stream));
    BatchOut out;
    memset(&out, 0, sizeof(out));
    out.accum = s->accum.ptr;
    out.hitmask = s->transparent ? s->hitmask.ptr : nullptr;
    tr_status st = enqueue_frame(s, npx, out);
    if (st != TR_OK) return st;
    s->accum_samples += s->spp;
    s->frame_index += 1;
    launch_resolve(s->stream, s->accum.ptr, 1.0 / (double)s->accum_samples, s->rgba.ptr, (uint32_t)npx, s->post,
                   s->expo_partial.ptr, s->exposure.ptr, out.hitmask);
    TR_HIP(hipGetLastError());
    s->frame_pending = true;
    return TR_OK;
}

bool valid(const tr_scene* s) { return s != nullptr; }

/* Bounding box of all primitives: the cache's position input is the
 * position within it. */
void scene_box(tr_scene* s)
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const PrimRec& p : s->prims_in) {
        float pl[3], ph[3];
        prim_box(p, pl, ph);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pl[a]);
            hi[a] = std::max(hi[a], ph[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (!(lo[a] <= hi[a])) { lo[a] = 0.0f; hi[a] = 1.0f; }
        s->box_lo[a] = lo[a];
        s->box_hi[a] = hi[a];
    }
}

}  // namespace

extern "C" {

const char* tr_last_error_message(void) { return g_error.c_str(); }

tr_status tr_material_init(tr_material* m, uint32_t model)
{
    if (!m) return fail(TR_ERROR_INVALID_ARG, "null material");
    memset(m, 0, sizeof(*m));
    m->model = model;
    m->fresnel[0] = 1.0f;
    m->fresnel[1] = 0.0f;
    m->sggx_depth = 2;
    m->sggx_sample_count = 256;
    m->sggx_basis[3] = 1.0f;
    return TR_OK;
}

/* RTMaterial::setup -> DefaultMicrofacetMaterialSetup (illumination-models.hh:599-614)
 * for the models whose setup entry is that function (:373-409). */
tr_status tr_material_setup(tr_material* m)
{
    if (!m) return fail(TR_ERROR_INVALID_ARG, "null material");
    uint32_t model = m->model;
    bool default_setup = model == TR_MODEL_BLINN_PHONG || model == TR_MODEL_ASHIKHMIN_SHIRLEY ||
                         (model >= TR_MODEL_GGX_MICROFACET && model <= TR_MODEL_GGX_MICROFACET_DIELECTRIC) ||
                         (model >= TR_MODEL_GGX_MICROFACET_ANISOTROPIC &&
                          model <= TR_MODEL_GGX_MICROFACET_DIELECTRIC_ANISOTROPIC) ||
                         model == TR_MODEL_SGGX_SURFACE || model == TR_MODEL_SGGX_PSEUDO_VOLUME ||
                         (model >= TR_MODEL_BECKMANN_MICROFACET &&
                          model <= TR_MODEL_BECKMANN_MICROFACET_DIELECTRIC_ANISOTROPIC);
    if (!default_setup) return TR_OK;
    /* SampleIncidentLightLookup is non-null for all of these models */
    float sum[3] = {m->diffuse[0] + m->specular[0], m->diffuse[1] + m->specular[1], m->diffuse[2] + m->specular[2]};
    float max_value = std::max(std::max(sum[0], sum[1]), sum[2]);
    if (max_value > 1.0f) {
        float r = 1.0f / max_value;
        for (int i = 0; i < 3; ++i) { m->diffuse[i] *= r; m->specular[i] *= r; }
    }
    m->normalization = 1.0f;
    return TR_OK;
}

tr_status tr_scene_create(uint32_t width, uint32_t height, const float vpi[16], const tr_settings* settings,
                          tr_scene** out)
{
    if (!out || !vpi || width == 0 || height == 0) return fail(TR_ERROR_INVALID_ARG, "invalid scene arguments");
    tr_scene* s = new (std::nothrow) tr_scene();
    if (!s) return fail(TR_ERROR_OUT_OF_MEMORY, "scene allocation failed");
    s->width = width;
    s->height = height;
    memcpy(s->vpi, vpi, sizeof(s->vpi));
    if (settings && settings->trace_waves_per_cu > 32) {
        delete s;
        return fail(TR_ERROR_INVALID_ARG, "trace_waves_per_cu above 32");
    }
    if (settings) s->settings = *settings;
    else { s->settings.data_pool_size = 32u << 20; s->settings.path_pool_size = 0; s->settings.trace_waves_per_cu = 0;
           s->settings.shade_order = 0; }
    if (s->settings.shade_order > 1) {
        delete s;
        return fail(TR_ERROR_INVALID_ARG, "shade_order above 1");
    }
    memset(&s->stats, 0, sizeof(s->stats));
    *out = s;
    return TR_OK;
}

void tr_scene_destroy(tr_scene* s) { delete s; }

#define SETTER(name, field, type)                                               \
    tr_status name(tr_scene* s, type v)                                         \
    {                                                                           \
        if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");         \
        s->field = v;                                                           \
        return TR_OK;                                                           \
    }
SETTER(tr_scene_set_chunk_size, chunk, uint32_t)
SETTER(tr_scene_set_samples_camera, spp, uint32_t)
SETTER(tr_scene_set_max_ray_depth, max_depth, uint32_t)
SETTER(tr_scene_set_russian_roulette, rr, float)
tr_status tr_scene_set_picture_post_process(tr_scene* s, uint32_t mode)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (mode > TR_POST_ACES) return fail(TR_ERROR_INVALID_ARG, "unknown picture post-process");
    s->post = mode;
    return TR_OK;
}
SETTER(tr_scene_set_transparent_background, transparent, int32_t)
tr_status tr_scene_set_render_mode(tr_scene* s, uint32_t mode)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (mode > 4) return fail(TR_ERROR_INVALID_ARG, "RenderMode is 0 (Normal) to 4 (DebugLighting)");
    s->render_mode = mode;
    return TR_OK;
}
SETTER(tr_scene_set_seed, seed, uint64_t)
SETTER(tr_scene_set_frame_index, frame_index, uint32_t)
#undef SETTER

tr_status tr_scene_set_samples_global_illumination(tr_scene* s, uint32_t v)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    s->n_gi = v;
    return TR_OK;
}

tr_status tr_scene_set_samples_local_area_light(tr_scene* s, uint32_t v)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    s->n_light = v;
    return TR_OK;
}

tr_status tr_scene_set_estimator(tr_scene* s, uint32_t mode)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (mode != TR_ESTIMATOR_REFERENCE && mode != TR_ESTIMATOR_UNBIASED)
        return fail(TR_ERROR_INVALID_ARG, "unknown estimator");
    s->estimator = mode;
    return TR_OK;
}

tr_status tr_scene_set_radiance_cache(tr_scene* s, const tr_cache_settings* c)
{
    if (!valid(s) || !c) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (c->enabled) {
        if (!(c->rough_alpha >= 0.0f) || c->train_paths == 0 || c->batch == 0)
            return fail(TR_ERROR_INVALID_ARG, "cache settings need train_paths > 0, batch > 0, rough_alpha >= 0");
        if (!s->net) {
            tr_status st = tr_mlp_create(c->seed, &s->net);
            if (st != TR_OK) return st;
        }
    }
    s->cache = *c;
    return TR_OK;
}

tr_status tr_scene_get_radiance_cache(tr_scene* s, tr_mlp** out)
{
    if (!valid(s) || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->net) return fail(TR_ERROR_NOT_COMMITTED, "the radiance cache has not been enabled");
    *out = s->net;
    return TR_OK;
}

tr_status tr_scene_read_cache_records(tr_scene* s, size_t count, float* inputs, float* targets, float* valid_out)
{
    if (!valid(s) || !inputs || !targets || !valid_out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->net || s->net->st.train_cap < count) return fail(TR_ERROR_INVALID_ARG, "fewer records than requested");
    TR_HIP(hipStreamSynchronize(s->stream));
    MlpState& ms = s->net->st;
    size_t cap = ms.train_cap;
    std::vector<float> in((size_t)kMlpIn * cap), tg((size_t)3 * cap), vd(cap);
    TR_HIP(hipMemcpy(in.data(), ms.s_in.ptr, in.size() * sizeof(float), hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(tg.data(), ms.s_target.ptr, tg.size() * sizeof(float), hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(vd.data(), ms.s_valid.ptr, vd.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < count; ++r) {
        for (int k = 0; k < kMlpIn; ++k) inputs[r * kMlpIn + k] = in[(size_t)k * cap + r];
        for (int c = 0; c < 3; ++c) targets[r * 3 + c] = tg[(size_t)c * cap + r];
        valid_out[r] = vd[r];
    }
    return TR_OK;
}

tr_status tr_scene_set_counting(tr_scene* s, int32_t enabled)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    s->counting = enabled != 0;
    return TR_OK;
}

tr_status tr_scene_set_timing(tr_scene* s, int32_t enabled)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    s->timing = enabled != 0;
    return TR_OK;
}

tr_status tr_scene_set_background_spectrum(tr_scene* s, const float rgb[3])
{
    if (!valid(s) || !rgb) return fail(TR_ERROR_INVALID_ARG, "null argument");
    memcpy(s->background, rgb, sizeof(s->background));
    return TR_OK;
}

tr_status tr_scene_repaint(tr_scene* s)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    s->accum_samples = 0;
    return TR_OK;
}

tr_status tr_scene_add_light_source(tr_scene* s, const tr_light* light, uint32_t* out_index)
{
    if (!valid(s) || !light) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (light->type != TR_LIGHT_DIRECTIONAL && light->type != TR_LIGHT_POINT)
        return fail(TR_ERROR_UNSUPPORTED, "addLightSource takes directional and point lights; use addSphereLightSource");
    LightRec L;
    memset(&L, 0, sizeof(L));
    L.src = *light;
    L.geom = TR_INVALID_GEOMETRY;
    L.area = 0;
    if (out_index) *out_index = (uint32_t)s->lights.size();
    s->lights.push_back(L);
    s->lights_dirty = true;
    s->accum_samples = 0;
    return TR_OK;
}

tr_status tr_scene_update_light_source(tr_scene* s, uint32_t index, const tr_light* light)
{
    if (!valid(s) || !light) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (index >= s->lights.size()) return fail(TR_ERROR_INVALID_ARG, "light index out of range");
    LightRec& L = s->lights[index];
    if (L.src.type != light->type) return fail(TR_ERROR_INVALID_ARG, "updateLightSource cannot change the light type");
    uint32_t gid = L.src.geometry_id;
    L.src = *light;
    L.src.geometry_id = gid;
    s->lights_dirty = true;
    s->accum_samples = 0; /* repaint(), ray-tracing.hh:327 */
    return TR_OK;
}

tr_status tr_scene_clear_light_sources(tr_scene* s)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    s->lights.clear();
    s->lights_dirty = true;
    s->accum_samples = 0;
    return TR_OK;
}

static uint32_t add_sphere_prim(tr_scene* s, const float c[3], float r, uint32_t type, uint32_t material)
{
    GeomRec g;
    memset(&g, 0, sizeof(g));
    g.type = type;
    g.material = material;
    g.sphere[0] = c[0]; g.sphere[1] = c[1]; g.sphere[2] = c[2]; g.sphere[3] = r;
    uint32_t slot = (uint32_t)s->geoms.size();
    s->geoms.push_back(g);
    PrimRec p;
    memset(&p, 0, sizeof(p));
    p.v0[0] = c[0]; p.v0[1] = c[1]; p.v0[2] = c[2];
    p.v1[0] = r;
    p.prim = 0;
    p.geom = slot;
    p.flags = PRIM_SPHERE;
    s->prims_in.push_back(p);
    return slot;
}

tr_status tr_scene_add_sphere_light_source(tr_scene* s, const tr_light* light, uint32_t* out_geom_id)
{
    if (!valid(s) || !light) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "scene already committed");
    if (!(light->sphere_radius > 0.0f)) return fail(TR_ERROR_INVALID_ARG, "sphere light radius must be positive");
    tr_material m;
    tr_material_init(&m, TR_MODEL_EMISSIVE);
    memcpy(m.diffuse, light->radiance, sizeof(m.diffuse));
    s->mat_keys.push_back(nullptr);
    s->mats.push_back(m);
    uint32_t slot = add_sphere_prim(s, light->sphere_center, light->sphere_radius, GEOM_SPHERE_LIGHT,
                                    (uint32_t)s->mats.size() - 1);
    LightRec L;
    memset(&L, 0, sizeof(L));
    L.src = *light;
    L.src.type = TR_LIGHT_SPHERE_AREA;
    L.src.geometry_id = slot;
    L.geom = slot;
    L.area = 1;
    s->lights.push_back(L);
    s->lights_dirty = true;
    if (out_geom_id) *out_geom_id = slot;
    return TR_OK;
}

tr_status tr_scene_add_triangle_mesh(tr_scene* s, const float world[16], size_t submesh_count, const tr_submesh* subs,
                                     size_t index_count, const int32_t* tris, size_t vert_size, const void* verts,
                                     const tr_mesh_options* opts, uint64_t* geom_ids)
{
    if (!valid(s) || !world || (submesh_count && (!subs || !tris || !verts)))
        return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "scene already committed");
    for (size_t si = 0; si < submesh_count; ++si) {
        const tr_submesh& sm = subs[si];
        if (!sm.material) return fail(TR_ERROR_INVALID_ARG, "submesh without material");
        if (!model_supported(sm.material->model)) return fail(TR_ERROR_UNSUPPORTED, "illumination model not supported by the GPU tracer");
        if (sm.stride < 24 || sm.stride % 4) return fail(TR_ERROR_INVALID_ARG, "vertex stride must be a multiple of 4, at least 24");
        if (sm.vertex_count % 3) return fail(TR_ERROR_INVALID_ARG, "submesh index count must be a multiple of 3");
        if ((size_t)sm.base_index + sm.vertex_count > index_count) return fail(TR_ERROR_INVALID_ARG, "submesh indices out of range");
        if (sm.vertex_offset > vert_size) return fail(TR_ERROR_INVALID_ARG, "vertex offset out of range");
        size_t nverts = (vert_size - sm.vertex_offset) / sm.stride;
        for (uint32_t k = 0; k < sm.vertex_count; ++k) {
            int32_t idx = tris[sm.base_index + k];
            if (idx < 0 || (size_t)idx >= nverts) return fail(TR_ERROR_INVALID_ARG, "vertex index out of range");
        }
    }
    uint64_t vbase = s->vpool.size();
    s->vpool.insert(s->vpool.end(), (const uint8_t*)verts, (const uint8_t*)verts + vert_size);
    while (s->vpool.size() % 16) s->vpool.push_back(0);
    uint32_t inst = s->instances++;
    for (size_t si = 0; si < submesh_count; ++si) {
        const tr_submesh& sm = subs[si];
        GeomRec g;
        memset(&g, 0, sizeof(g));
        g.type = GEOM_MESH;
        g.material = intern_material(s, sm.material);
        g.two_sided = opts && opts->two_sided ? 1u : 0u;
        g.stride = sm.stride;
        g.vbase = vbase + sm.vertex_offset;
        g.index_base = (uint32_t)s->ipool.size();
        g.tri_count = sm.vertex_count / 3;
        s->ipool.insert(s->ipool.end(), tris + sm.base_index, tris + sm.base_index + sm.vertex_count);
        uint32_t slot = (uint32_t)s->geoms.size();
        const uint8_t* base = (const uint8_t*)verts + sm.vertex_offset;
        for (uint32_t t = 0; t < g.tri_count; ++t) {
            PrimRec p;
            memset(&p, 0, sizeof(p));
            float* dst[3] = {p.v0, p.v1, p.v2};
            for (int k = 0; k < 3; ++k) {
                const float* f = reinterpret_cast<const float*>(base + (size_t)tris[sm.base_index + t * 3 + k] * sm.stride);
                for (int r = 0; r < 3; ++r)
                    dst[k][r] = ((world[r] * f[0] + world[4 + r] * f[1]) + world[8 + r] * f[2]) + world[12 + r];
            }
            p.prim = t;
            p.geom = slot;
            p.flags = g.two_sided ? PRIM_TWO_SIDED : 0u;
            s->prims_in.push_back(p);
        }
        s->geoms.push_back(g);
        if (s->mats[g.material].model == TR_MODEL_EMISSIVE) {
            LightRec L;
            memset(&L, 0, sizeof(L));
            L.src.type = TR_LIGHT_MESH_AREA;
            L.src.geometry_id = slot;
            L.geom = slot;
            L.area = 0;
            s->lights.push_back(L);
            s->lights_dirty = true;
        }
        if (s->ext_ids.size() < s->geoms.size()) s->ext_ids.resize(s->geoms.size(), UINT64_MAX);
        s->ext_ids[slot] = ((uint64_t)inst << 32) | (uint64_t)si;
        if (geom_ids) geom_ids[si] = ((uint64_t)inst << 32) | (uint64_t)si;
    }
    return TR_OK;
}

tr_status tr_scene_add_sphere(tr_scene* s, const float center[3], float radius, const tr_material* m, uint64_t* out)
{
    if (!valid(s) || !center || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "scene already committed");
    if (!(radius > 0.0f)) return fail(TR_ERROR_INVALID_ARG, "sphere radius must be positive");
    if (!model_supported(m->model)) return fail(TR_ERROR_UNSUPPORTED, "illumination model not supported by the GPU tracer");
    uint32_t slot = add_sphere_prim(s, center, radius, GEOM_SPHERE, intern_material(s, m));
    if (out) *out = slot;
    return TR_OK;
}

/* ---------------- user shapes (ray-tracing.cc:1315-1446) ---------------- */

namespace {

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

/* MatrixToQuaternion (quaternion.hh:59-70) of the matrix with columns c0,
 * c1, c2; matrix(i, j) is component j of column i. */
void matrix_to_quat(V3 c0, V3 c1, V3 c2, float q[4])
{
    q[3] = sqrtf(std::max(0.0f, 1.0f + c0.x + c1.y + c2.z)) / 2.0f;
    q[0] = sqrtf(std::max(0.0f, 1.0f + c0.x - c1.y - c2.z)) / 2.0f;
    q[1] = sqrtf(std::max(0.0f, 1.0f - c0.x + c1.y - c2.z)) / 2.0f;
    q[2] = sqrtf(std::max(0.0f, 1.0f - c0.x - c1.y + c2.z)) / 2.0f;
    q[0] = std::copysign(q[0], c1.z - c2.y);
    q[1] = std::copysign(q[1], c2.x - c0.z);
    q[2] = std::copysign(q[2], c0.y - c1.x);
}

/* A shape's geometry slot, record and primitive; its box [lo, hi] is
 * widened by one float ulp-scale step outward. */
uint32_t add_shape_prim(tr_scene* s, ShapeRec rec, uint32_t material, const double lo[3], const double hi[3])
{
    GeomRec g;
    memset(&g, 0, sizeof(g));
    g.type = GEOM_SHAPE;
    g.material = material;
    g.index_base = (uint32_t)s->shapes.size();
    const uint32_t slot = (uint32_t)s->geoms.size();
    s->geoms.push_back(g);
    rec.geom = slot;
    s->shapes.push_back(rec);
    PrimRec p;
    memset(&p, 0, sizeof(p));
    p.v0[0] = tr_u2f(g.index_base);
    for (int a = 0; a < 3; ++a) {
        const double pad = 1e-6 * (std::fabs(lo[a]) + std::fabs(hi[a]) + (hi[a] - lo[a])) + 1e-30;
        p.v1[a] = std::nextafter((float)(lo[a] - pad), -INFINITY);
        p.v2[a] = std::nextafter((float)(hi[a] + pad), INFINITY);
    }
    p.prim = 0;
    p.geom = slot;
    p.flags = PRIM_SHAPE | (rec.kind << 8);
    s->prims_in.push_back(p);
    return slot;
}

/* Rect and blocker: Orientation = ToQuaternion(Matrix3(tan, norm x tan,
 * norm)) and its ToMatrix3 columns; Rect3Bounds (shapes.hh:260-266). */
tr_status add_rect_shape(tr_scene* s, const float pos[3], const float tan[3], const float norm[3], const float size[2],
                         uint32_t material, const float tc[4], uint32_t tangent_map, uint64_t* out)
{
    if (!finite3(pos) || !finite3(tan) || !finite3(norm) || !std::isfinite(size[0]) || !std::isfinite(size[1]) ||
        !(size[0] > 0.0f) || !(size[1] > 0.0f))
        return fail(TR_ERROR_INVALID_ARG, "rect: finite position, tangent and normal and positive half sizes needed");
    const V3 T = mk(tan[0], tan[1], tan[2]), N = mk(norm[0], norm[1], norm[2]);
    float q[4];
    matrix_to_quat(T, cross(N, T), N, q);
    const M3 o = quat_to_m3(q[0], q[1], q[2], q[3]);
    ShapeRec r;
    memset(&r, 0, sizeof(r));
    r.kind = SHAPE_RECT;
    r.tangent_map = tangent_map;
    const V3 cols[3] = {o.c0, o.c1, o.c2};
    for (int a = 0; a < 3; ++a) r.f[a] = pos[a];
    for (int k = 0; k < 3; ++k) {
        r.f[3 + 3 * k] = cols[k].x;
        r.f[4 + 3 * k] = cols[k].y;
        r.f[5 + 3 * k] = cols[k].z;
    }
    r.f[12] = size[0];
    r.f[13] = size[1];
    if (tc) {
        r.f[14] = tc[0];
        r.f[15] = tc[1];
        r.f[16] = tc[2] - tc[0];
        r.f[17] = tc[3] - tc[1];
    } else {
        r.f[16] = 1.0f;
        r.f[17] = 1.0f;
    }
    double lo[3], hi[3];
    const double ext[3] = {std::fabs((double)size[0] * o.c0.x) + std::fabs((double)size[1] * o.c1.x),
                           std::fabs((double)size[0] * o.c0.y) + std::fabs((double)size[1] * o.c1.y),
                           std::fabs((double)size[0] * o.c0.z) + std::fabs((double)size[1] * o.c1.z)};
    for (int a = 0; a < 3; ++a) { lo[a] = pos[a] - ext[a]; hi[a] = pos[a] + ext[a]; }
    const uint32_t slot = add_shape_prim(s, r, material, lo, hi);
    if (out) *out = slot;
    return TR_OK;
}

tr_status shape_common(tr_scene* s, const tr_material* m)
{
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "scene already committed");
    if (m && !model_supported(m->model)) return fail(TR_ERROR_UNSUPPORTED, "illumination model not supported by the GPU tracer");
    return TR_OK;
}

/* The cylinder's box: the axis moves by bending * dz over the height
 * (ObliqueCylinderBoundsFunction, ray-tracing.cc:981-990, widened by
 * |Tilt * TiltDirection| per axis instead of Tilt). */
void cylinder_box(const ShapeRec& r, double tx, double ty, double lo[3], double hi[3])
{
    const double ext[3] = {(double)r.f[3] + std::fabs(tx), (double)r.f[3] + std::fabs(ty), (double)r.f[4]};
    for (int a = 0; a < 3; ++a) { lo[a] = r.f[a] - ext[a]; hi[a] = r.f[a] + ext[a]; }
}

}  // namespace

tr_status tr_scene_add_rect(tr_scene* s, const float pos[3], const float tan[3], const float norm[3],
                            const float size[2], const tr_material* m, const float tc_box[4], uint32_t tangent_map,
                            uint64_t* out_geom_id)
{
    if (!valid(s) || !pos || !tan || !norm || !size || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    tr_status st = shape_common(s, m);
    if (st != TR_OK) return st;
    if (tangent_map > s->tex_recs.size()) return fail(TR_ERROR_INVALID_ARG, "rect: unknown tangent-map texture");
    return add_rect_shape(s, pos, tan, norm, size, intern_material(s, m), tc_box, tangent_map, out_geom_id);
}

tr_status tr_scene_add_blocker(tr_scene* s, const float pos[3], const float tan[3], const float norm[3],
                               const float size[2], uint64_t* out_geom_id)
{
    if (!valid(s) || !pos || !tan || !norm || !size) return fail(TR_ERROR_INVALID_ARG, "null argument");
    tr_status st = shape_common(s, nullptr);
    if (st != TR_OK) return st;
    return add_rect_shape(s, pos, tan, norm, size, kNoMaterial, nullptr, 0, out_geom_id);
}

tr_status tr_scene_add_disk(tr_scene* s, const float pos[3], const float norm[3], float inner_radius,
                            float outer_radius, const tr_material* m, uint64_t* out_geom_id)
{
    if (!valid(s) || !pos || !norm || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    tr_status st = shape_common(s, m);
    if (st != TR_OK) return st;
    const double nl = std::sqrt((double)norm[0] * norm[0] + (double)norm[1] * norm[1] + (double)norm[2] * norm[2]);
    if (!finite3(pos) || !finite3(norm) || !(nl > 0.0) || !(inner_radius >= 0.0f) || !(outer_radius > inner_radius) ||
        !std::isfinite(outer_radius))
        return fail(TR_ERROR_INVALID_ARG, "disk: finite centre, non-zero normal and 0 <= inner < outer radius needed");
    ShapeRec r;
    memset(&r, 0, sizeof(r));
    r.kind = SHAPE_DISK;
    for (int a = 0; a < 3; ++a) { r.f[a] = pos[a]; r.f[3 + a] = norm[a]; }
    r.f[6] = inner_radius;
    r.f[7] = outer_radius;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const double na = norm[a] / nl;
        const double e = outer_radius * std::sqrt(std::max(0.0, 1.0 - na * na));
        lo[a] = pos[a] - e;
        hi[a] = pos[a] + e;
    }
    const uint32_t slot = add_shape_prim(s, r, intern_material(s, m), lo, hi);
    if (out_geom_id) *out_geom_id = slot;
    return TR_OK;
}

tr_status tr_scene_add_cylinder(tr_scene* s, const float center[3], float radius, float half_height,
                                const tr_material* m, uint64_t* out_geom_id)
{
    if (!valid(s) || !center || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    tr_status st = shape_common(s, m);
    if (st != TR_OK) return st;
    if (!finite3(center) || !(radius > 0.0f) || !(half_height > 0.0f) || !std::isfinite(radius) ||
        !std::isfinite(half_height))
        return fail(TR_ERROR_INVALID_ARG, "cylinder: finite centre, positive radius and half height needed");
    ShapeRec r;
    memset(&r, 0, sizeof(r));
    r.kind = SHAPE_CYLINDER;
    for (int a = 0; a < 3; ++a) r.f[a] = center[a];
    r.f[3] = radius;
    r.f[4] = half_height;
    double lo[3], hi[3];
    cylinder_box(r, 0.0, 0.0, lo, hi);
    const uint32_t slot = add_shape_prim(s, r, intern_material(s, m), lo, hi);
    if (out_geom_id) *out_geom_id = slot;
    return TR_OK;
}

tr_status tr_scene_add_oblique_cylinder(tr_scene* s, const float center[3], float radius, float half_height,
                                        const float tilt_direction[2], float tilt, const tr_material* m,
                                        uint64_t* out_geom_id)
{
    if (!valid(s) || !center || !tilt_direction || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    tr_status st = shape_common(s, m);
    if (st != TR_OK) return st;
    if (!finite3(center) || !(radius > 0.0f) || !(half_height > 0.0f) || !std::isfinite(radius) ||
        !std::isfinite(half_height) || !std::isfinite(tilt) || !std::isfinite(tilt_direction[0]) ||
        !std::isfinite(tilt_direction[1]))
        return fail(TR_ERROR_INVALID_ARG, "oblique cylinder: finite parameters, positive radius and half height needed");
    ShapeRec r;
    memset(&r, 0, sizeof(r));
    r.kind = SHAPE_OBLIQUE_CYLINDER;
    for (int a = 0; a < 3; ++a) r.f[a] = center[a];
    r.f[3] = radius;
    r.f[4] = half_height;
    /* bending = Tilt * TiltDirection / HalfHeight (ray-tracing.cc:999) */
    r.f[5] = (tilt * tilt_direction[0]) / half_height;
    r.f[6] = (tilt * tilt_direction[1]) / half_height;
    double lo[3], hi[3];
    cylinder_box(r, (double)tilt * tilt_direction[0], (double)tilt * tilt_direction[1], lo, hi);
    const uint32_t slot = add_shape_prim(s, r, intern_material(s, m), lo, hi);
    if (out_geom_id) *out_geom_id = slot;
    return TR_OK;
}

tr_status tr_scene_add_ellipsoid(tr_scene* s, const float center[3], const float scale[3],
                                 const float orientation[4], const tr_material* m, uint64_t* out_geom_id)
{
    if (!valid(s) || !center || !scale || !orientation || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    tr_status st = shape_common(s, m);
    if (st != TR_OK) return st;
    const double ql = std::sqrt((double)orientation[0] * orientation[0] + (double)orientation[1] * orientation[1] +
                                (double)orientation[2] * orientation[2] + (double)orientation[3] * orientation[3]);
    if (!finite3(center) || !finite3(scale) || !(scale[0] > 0.0f) || !(scale[1] > 0.0f) || !(scale[2] > 0.0f) ||
        !(ql > 0.0) || !std::isfinite(ql))
        return fail(TR_ERROR_INVALID_ARG, "ellipsoid: finite centre, positive scale and a non-zero quaternion needed");
    ShapeRec r;
    memset(&r, 0, sizeof(r));
    r.kind = SHAPE_ELLIPSOID;
    for (int a = 0; a < 3; ++a) { r.f[a] = center[a]; r.f[3 + a] = scale[a]; }
    for (int k = 0; k < 4; ++k) r.f[6 + k] = orientation[k];
    /* Inverse = Conjugate (quaternion.hh:215-241) */
    r.f[10] = -orientation[0];
    r.f[11] = -orientation[1];
    r.f[12] = -orientation[2];
    r.f[13] = orientation[3];
    /* ToNormal (quaternion.hh:170-181) */
    {
        const float x = orientation[0], y = orientation[1], z = orientation[2], w = orientation[3];
        const float xx = x * x, xz = x * z, xw = x * w, yy = y * y, yz = y * z, yw = y * w;
        r.f[14] = 2.0f * (xz + yw);
        r.f[15] = 2.0f * (yz - xw);
        r.f[16] = 1.0f - 2.0f * (xx + yy);
    }
    /* box: the surface is c + L^-1 S u, |u| = 1, with L the linear map
     * v -> Transform(conj q, v); half extent per axis = row norm of L^-1 S */
    double L[3][3];
    for (int j = 0; j < 3; ++j) {
        const double e[3] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0};
        const double qv[3] = {-(double)orientation[0], -(double)orientation[1], -(double)orientation[2]};
        const double w = orientation[3];
        const double c1[3] = {qv[1] * e[2] - qv[2] * e[1] + w * e[0], qv[2] * e[0] - qv[0] * e[2] + w * e[1],
                              qv[0] * e[1] - qv[1] * e[0] + w * e[2]};
        const double c2[3] = {qv[1] * c1[2] - qv[2] * c1[1], qv[2] * c1[0] - qv[0] * c1[2], qv[0] * c1[1] - qv[1] * c1[0]};
        for (int i = 0; i < 3; ++i) L[i][j] = e[i] + 2.0 * c2[i];
    }
    const double det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) - L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                       L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
    if (!(std::fabs(det) > 1e-12)) return fail(TR_ERROR_INVALID_ARG, "ellipsoid: degenerate orientation");
    double inv[3][3];
    inv[0][0] = (L[1][1] * L[2][2] - L[1][2] * L[2][1]) / det;
    inv[0][1] = (L[0][2] * L[2][1] - L[0][1] * L[2][2]) / det;
    inv[0][2] = (L[0][1] * L[1][2] - L[0][2] * L[1][1]) / det;
    inv[1][0] = (L[1][2] * L[2][0] - L[1][0] * L[2][2]) / det;
    inv[1][1] = (L[0][0] * L[2][2] - L[0][2] * L[2][0]) / det;
    inv[1][2] = (L[0][2] * L[1][0] - L[0][0] * L[1][2]) / det;
    inv[2][0] = (L[1][0] * L[2][1] - L[1][1] * L[2][0]) / det;
    inv[2][1] = (L[0][1] * L[2][0] - L[0][0] * L[2][1]) / det;
    inv[2][2] = (L[0][0] * L[1][1] - L[0][1] * L[1][0]) / det;
    double lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        double e2 = 0.0;
        for (int j = 0; j < 3; ++j) e2 += (inv[i][j] * scale[j]) * (inv[i][j] * scale[j]);
        const double e = std::sqrt(e2) * (1.0 + 1e-5);
        lo[i] = center[i] - e;
        hi[i] = center[i] + e;
    }
    const uint32_t slot = add_shape_prim(s, r, intern_material(s, m), lo, hi);
    if (out_geom_id) *out_geom_id = slot;
    return TR_OK;
}

tr_status tr_scene_add_grid_volume(tr_scene* s, const tr_grid_volume* v, const tr_material* m, uint64_t* out)
{
    if (!valid(s) || !v || !m || !v->density || !v->direction) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "scene already committed");
    if (m->model != TR_MODEL_SGGX_MICRO_FLAKE)
        return fail(TR_ERROR_UNSUPPORTED, "grid volumes take the SGGXMicroFlake volume model");
    for (int a = 0; a < 3; ++a) {
        if (v->dims[a] < 2) return fail(TR_ERROR_INVALID_ARG, "volume dims must be at least 2 per axis");
        if (!(v->max_corner[a] > v->min_corner[a]) || !std::isfinite(v->min_corner[a]) ||
            !std::isfinite(v->max_corner[a]))
            return fail(TR_ERROR_INVALID_ARG, "volume bounds must be finite with min < max");
        if (!(m->sggx_stddev[a] > 0.0f) || !std::isfinite(m->sggx_stddev[a]))
            return fail(TR_ERROR_INVALID_ARG, "sggx_stddev must be positive and finite");
    }
    uint64_t n = (uint64_t)v->dims[0] * v->dims[1] * v->dims[2];
    if (n > (1ull << 31)) return fail(TR_ERROR_INVALID_ARG, "volume larger than 2^31 samples");
    float rho_max = 0.0f;
    for (uint64_t i = 0; i < n; ++i) {
        float r = h2f(v->density[i]);
        if (!(r >= 0.0f) || !std::isfinite(r))
            return fail(TR_ERROR_INVALID_ARG, "volume density must be finite and non-negative");
        rho_max = std::max(rho_max, r);
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(h2f(v->direction[i * 3 + a])))
                return fail(TR_ERROR_INVALID_ARG, "volume direction must be finite");
    }
    if (v->max_density != 0.0f) {
        if (!(v->max_density >= rho_max) || !std::isfinite(v->max_density))
            return fail(TR_ERROR_INVALID_ARG, "max_density below the grid's largest density");
        rho_max = v->max_density;
    }
    if (!(rho_max > 0.0f)) return fail(TR_ERROR_INVALID_ARG, "volume density is zero everywhere");
    VolumeRec r;
    memset(&r, 0, sizeof(r));
    float smax = std::max(m->sggx_stddev[0], std::max(m->sggx_stddev[1], m->sggx_stddev[2]));
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = v->min_corner[a];
        r.hi[a] = v->max_corner[a];
        r.scale[a] = (float)(v->dims[a] - 1u) / (v->max_corner[a] - v->min_corner[a]);
        r.stddev[a] = m->sggx_stddev[a];
        r.dims[a] = v->dims[a];
    }
    r.rcp_rho_max = 1.0f / rho_max;
    r.rcp_majorant = 1.0f / (rho_max * smax);
    r.rcp_smax = 1.0f / smax;
    /* tiles of 4 x 2 x 2 samples from a 16-word boundary (voxel_tile_index) */
    s->voxels.resize((s->voxels.size() + 15u) & ~(size_t)15u, 0u);
    r.voxel_base = s->voxels.size();
    s->voxels.resize(r.voxel_base + voxel_tile_words(r.dims), 0u);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t x = (uint32_t)(i % r.dims[0]), y = (uint32_t)((i / r.dims[0]) % r.dims[1]);
        const uint32_t z = (uint32_t)(i / ((uint64_t)r.dims[0] * r.dims[1]));
        s->voxels[r.voxel_base + voxel_tile_index(x, y, z, r.dims)] =
            (uint64_t)v->direction[i * 3] | ((uint64_t)v->direction[i * 3 + 1] << 16) |
            ((uint64_t)v->direction[i * 3 + 2] << 32) | ((uint64_t)v->density[i] << 48);
    }
    GeomRec g;
    memset(&g, 0, sizeof(g));
    g.type = GEOM_VOLUME;
    g.material = intern_material(s, m);
    g.tri_count = (uint32_t)s->vols.size();
    uint32_t slot = (uint32_t)s->geoms.size();
    s->geoms.push_back(g);
    s->vols.push_back(r);
    PrimRec p;
    memset(&p, 0, sizeof(p));
    for (int a = 0; a < 3; ++a) { p.v0[a] = r.lo[a]; p.v1[a] = r.hi[a]; }
    p.prim = 0;
    p.geom = slot;
    p.flags = PRIM_VOLUME;
    s->prims_in.push_back(p);
    if (out) *out = slot;
    return TR_OK;
}

namespace {

/* An integer 0..255 as binary16 bits (exact). */
uint64_t u8_half_bits(uint32_t v)
{
    if (v == 0) return 0;
    uint32_t e = 31u - (uint32_t)__builtin_clz(v);
    return (uint64_t)(((e + 15u) << 10) | ((v << (10u - e)) & 0x3ffu));
}

/* s_cosPhi, s_sinPhi, s_cosTheta, s_sinTheta (ray-tracing.cc:777-788):
 * angle i * pi / 255 (theta) and twice that (phi) for i < 255, 0 at 255,
 * with the reference's truncated pi. */
void append_direction_tables(std::vector<uint64_t>* voxels)
{
    float tab[1024];
    for (int i = 0; i < 255; ++i) {
        float angle = (float)i * (kPi / 255.0f);
        tab[i] = std::cos(2.0f * angle);
        tab[256 + i] = std::sin(2.0f * angle);
        tab[512 + i] = std::cos(angle);
        tab[768 + i] = std::sin(angle);
    }
    tab[255] = tab[511] = tab[767] = tab[1023] = 0.0f;
    const size_t at = voxels->size();
    voxels->resize(at + 512);
    memcpy(voxels->data() + at, tab, sizeof(tab));
}

}  // namespace

tr_status tr_scene_add_hierarchical_volume(tr_scene* s, const tr_volume_root* root, const tr_material* m,
                                           uint64_t* out)
{
    if (!valid(s) || !root || !m) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "scene already committed");
    if (m->model != TR_MODEL_SGGX_MICRO_FLAKE)
        return fail(TR_ERROR_UNSUPPORTED, "hierarchical volumes take the SGGXMicroFlake volume model");
    for (int a = 0; a < 3; ++a) {
        if (root->dims[a] < 1) return fail(TR_ERROR_INVALID_ARG, "volume root dims must be positive");
        if (!(root->max_corner[a] > root->min_corner[a]) || !std::isfinite(root->min_corner[a]) ||
            !std::isfinite(root->max_corner[a]))
            return fail(TR_ERROR_INVALID_ARG, "volume bounds must be finite with min < max");
        if (!(m->sggx_stddev[a] > 0.0f) || !std::isfinite(m->sggx_stddev[a]))
            return fail(TR_ERROR_INVALID_ARG, "sggx_stddev must be positive and finite");
    }
    const uint64_t nb = (uint64_t)root->dims[0] * root->dims[1] * root->dims[2];
    if (nb > (1u << 20)) return fail(TR_ERROR_INVALID_ARG, "more than 2^20 bricks");
    if (nb && !root->bricks) return fail(TR_ERROR_INVALID_ARG, "null brick array");
    uint64_t total = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        const tr_volume_brick& br = root->bricks[b];
        if (!br.data) continue;
        if (br.dims[0] < 1 || br.dims[1] < 1 || br.dims[2] < 1) return fail(TR_ERROR_INVALID_ARG, "empty brick dims");
        total += (uint64_t)br.dims[0] * br.dims[1] * br.dims[2];
    }
    if (total > (1ull << 31)) return fail(TR_ERROR_INVALID_ARG, "volume larger than 2^31 samples");
    if (s->dir_table == UINT64_MAX) {
        s->dir_table = s->voxels.size();
        append_direction_tables(&s->voxels);
    }
    /* vol_size = (MaxCorner - MinCorner) / Dimensions (ray-tracing.cc:1455);
     * brick corners Box{x, y, z} * vol_size + MinCorner (:1486-1487) */
    float vs[3];
    for (int a = 0; a < 3; ++a) vs[a] = (root->max_corner[a] - root->min_corner[a]) / (float)root->dims[a];
    const uint32_t mat = intern_material(s, m);
    uint64_t first = UINT64_MAX;
    for (uint64_t b = 0; b < nb; ++b) {
        const tr_volume_brick& br = root->bricks[b];
        if (!br.data) continue;
        const uint32_t idx[3] = {(uint32_t)(b % (uint64_t)root->dims[0]),
                                 (uint32_t)((b / (uint64_t)root->dims[0]) % (uint64_t)root->dims[1]),
                                 (uint32_t)(b / ((uint64_t)root->dims[0] * root->dims[1]))};
        VolumeRec r;
        memset(&r, 0, sizeof(r));
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = (float)idx[a] * vs[a] + root->min_corner[a];
            r.hi[a] = r.lo[a] + vs[a];
            r.dims[a] = (uint32_t)br.dims[a];
            r.scale[a] = (float)(br.dims[a] - 1) / (r.hi[a] - r.lo[a]);
            r.stddev[a] = m->sggx_stddev[a];
        }
        r.format = kVoxelQuantized;
        r.rcp_rho_max = 1.0f / 255.0f;
        r.rcp_majorant = 1.0f / kDensityScale;  /* RcpMaxExtinction (:1463) */
        r.rcp_smax = 1.0f;
        r.table_base = s->dir_table;
        /* tiles of 4 x 2 x 2 samples from a 16-word boundary (voxel_tile_index) */
        s->voxels.resize((s->voxels.size() + 15u) & ~(size_t)15u, 0u);
        r.voxel_base = s->voxels.size();
        const uint64_t n = (uint64_t)br.dims[0] * br.dims[1] * br.dims[2];
        s->voxels.resize(r.voxel_base + voxel_tile_words(r.dims), 0u);
        for (uint64_t i = 0; i < n; ++i) {
            const uint8_t* q = br.data + n + 2 * i;
            const uint32_t x = (uint32_t)(i % r.dims[0]), y = (uint32_t)((i / r.dims[0]) % r.dims[1]);
            const uint32_t z = (uint32_t)(i / ((uint64_t)r.dims[0] * r.dims[1]));
            s->voxels[r.voxel_base + voxel_tile_index(x, y, z, r.dims)] =
                (uint64_t)q[0] | ((uint64_t)q[1] << 8) | (u8_half_bits(br.data[i]) << 48);
        }
        GeomRec g;
        memset(&g, 0, sizeof(g));
        g.type = GEOM_VOLUME;
        g.material = mat;
        g.tri_count = (uint32_t)s->vols.size();
        uint32_t slot = (uint32_t)s->geoms.size();
        s->geoms.push_back(g);
        s->vols.push_back(r);
        PrimRec p;
        memset(&p, 0, sizeof(p));
        for (int a = 0; a < 3; ++a) { p.v0[a] = r.lo[a]; p.v1[a] = r.hi[a]; }
        p.prim = 0;
        p.geom = slot;
        p.flags = PRIM_VOLUME;
        s->prims_in.push_back(p);
        if (first == UINT64_MAX) first = slot;
    }
    if (out) *out = first;
    return TR_OK;
}

tr_status tr_scene_bind_texture(tr_scene* s, const tr_texture_desc* d, const void* texels, uint32_t* out)
{
    if (!valid(s) || !d || !texels || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "textures are bound before commitScene");
    if (d->width < 1 || d->height < 1 || d->width > 65535 || d->height > 65535)
        return fail(TR_ERROR_INVALID_ARG, "texture sides are 1..65535");
    const uint32_t tb = texel_bytes(d->format);
    if (tb == 0 || d->format > TR_TEX_RGBA32F) return fail(TR_ERROR_INVALID_ARG, "unknown texture format");
    if (d->sampling != TR_TEX_BILINEAR && d->sampling != TR_TEX_NEAREST)
        return fail(TR_ERROR_INVALID_ARG, "unknown texture sampling");
    const size_t bytes = (size_t)d->width * d->height * tb;
    TexRec r;
    memset(&r, 0, sizeof(r));
    r.width = d->width;
    r.height = d->height;
    r.format = d->format;
    r.sampling = d->sampling;
    r.offset = (s->texels.size() + 15) & ~(size_t)15;
    s->texels.resize(r.offset + bytes);
    memcpy(s->texels.data() + r.offset, texels, bytes);
    s->tex_recs.push_back(r);
    *out = (uint32_t)s->tex_recs.size();
    return TR_OK;
}

tr_status tr_scene_bind_mesh_blob_textures(tr_scene* s, tr_mesh_blob* blob)
{
    if (!valid(s) || !blob) return fail(TR_ERROR_INVALID_ARG, "null argument");
    std::vector<uint32_t> handle(blob->texture_count + 1, 0);
    for (uint32_t i = 0; i < blob->texture_count; ++i) {
        const tr_image& im = blob->textures[i];
        tr_texture_desc d = {im.width, im.height, im.format, TR_TEX_BILINEAR};
        tr_status st = tr_scene_bind_texture(s, &d, im.data, &handle[i + 1]);
        if (st != TR_OK) return st;
    }
    for (uint32_t i = 0; i < blob->material_count; ++i) {
        uint32_t& m = blob->materials[i].diffuse_map;
        m = m <= blob->texture_count ? handle[m] : 0;
    }
    return TR_OK;
}

tr_status tr_scene_set_global_cube_map(tr_scene* s, uint32_t side, const uint32_t* const faces[6])
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (s->committed) return fail(TR_ERROR_INVALID_ARG, "the cube map is set before commitScene");
    if (!faces) {
        s->cube.clear();
        s->cube_side = 0;
        return TR_OK;
    }
    if (side <= 16 || side > 16384) return fail(TR_ERROR_INVALID_ARG, "cube map sides are 17..16384");
    for (int f = 0; f < 6; ++f)
        if (!faces[f]) return fail(TR_ERROR_INVALID_ARG, "null cube map face");
    const size_t pitch = (size_t)side * 4;
    s->cube.resize(6 * pitch * side);
    for (int f = 0; f < 6; ++f) {
        uint8_t* dst = s->cube.data() + (size_t)f * pitch * side;
        const uint8_t* src = reinterpret_cast<const uint8_t*>(faces[f]);
        for (size_t y = 0; y < side; ++y) memcpy(dst + y * pitch, src + (side - 1 - y) * pitch, pitch);
    }
    s->cube_side = side;
    return TR_OK;
}

namespace {

/* A material's texture handles against the bound textures and the
 * extractor each model applies to them. */
tr_status check_maps(const tr_scene* s, const tr_material& m)
{
    const uint32_t n = (uint32_t)s->tex_recs.size();
    auto ok = [&](uint32_t h, int kind) {
        return h == 0 || (h <= n && texel_kind_ok(s->tex_recs[h - 1].format, kind));
    };
    if (m.model == TR_MODEL_BLINN_PHONG && !ok(m.diffuse_map, TEX_SPECTRUM))
        return fail(TR_ERROR_INVALID_ARG, "DiffuseMap: unknown texture or a format SampleSpectrum does not read");
    if (m.model == TR_MODEL_SGGX_PSEUDO_VOLUME) {
        /* the BRDF runs SampleCount walks of up to Depth scatterings per
         * evaluation: 0 walks divide by zero, and the bounds keep one
         * evaluation within a kernel's time */
        if (m.sggx_sample_count < 1 || m.sggx_sample_count > 4096 || m.sggx_depth > 64)
            return fail(TR_ERROR_INVALID_ARG, "pseudo-volume SampleCount must be 1..4096 and Depth at most 64");
    }
    if (m.model == TR_MODEL_SGGX_SURFACE || m.model == TR_MODEL_SGGX_PSEUDO_VOLUME) {
        if (m.model == TR_MODEL_SGGX_SURFACE && (!ok(m.diffuse_map, TEX_SPECTRUM) || !ok(m.specular_map, TEX_SPECTRUM)))
            return fail(TR_ERROR_INVALID_ARG, "Diffuse/SpecularMap: unknown texture or unreadable format");
        if (!ok(m.stddev_map, TEX_RG))
            return fail(TR_ERROR_INVALID_ARG, "StandardDeviationMap: unknown texture or a format SampleRG does not read");
        if (!ok(m.basis_map, TEX_RGBA))
            return fail(TR_ERROR_INVALID_ARG, "BasisMap: unknown texture or a format SampleRGBA does not read");
    }
    return TR_OK;
}

/* SRGBToSpectrum of 8-bit channels: ToVector3's c * (1 / 255)
 * (spectrum.hh:815-819), then ConvertSLuminanceToLinear (:156-162). */
void srgb_table(std::vector<float>* t)
{
    t->resize(256);
    for (int i = 0; i < 256; ++i) {
        float c = (float)i * (1.0f / 255.0f);
        (*t)[i] = c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
    }
}

}  // namespace

tr_status tr_scene_commit(tr_scene* s)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (!s->committed) {
        for (const tr_material& m : s->mats) {
            tr_status ms = check_maps(s, m);
            if (ms != TR_OK) return ms;
        }
        for (const ShapeRec& r : s->shapes)
            if (r.tangent_map && !texel_kind_ok(s->tex_recs[r.tangent_map - 1].format, TEX_RGB))
                return fail(TR_ERROR_INVALID_ARG, "rect tangent map: a format sampleRGB does not read");
        if (s->cube_side) {
            s->cube_offset = (s->texels.size() + 15) & ~(size_t)15;
            s->texels.resize(s->cube_offset + s->cube.size());
            memcpy(s->texels.data() + s->cube_offset, s->cube.data(), s->cube.size());
        }
        auto t0 = std::chrono::steady_clock::now();
        std::string err;
        if (!build_bvh8(s->prims_in, s->nodes, s->prims, &s->bvh_stats, &err)) return fail(TR_ERROR_INVALID_ARG, err);
        scene_box(s);
        auto t1 = std::chrono::steady_clock::now();
        s->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        s->committed = true;
        s->lights_dirty = true;
    }
    tr_status st = ensure_device(s);
    if (st != TR_OK) return st;
    if ((st = upload(s, s->d_nodes, s->nodes)) != TR_OK) return st;
    if ((st = upload(s, s->d_prims, s->prims)) != TR_OK) return st;
    for (GeomRec& g : s->geoms) g.shade_class = shade_class_of(g, g.material == kNoMaterial ? nullptr : &s->mats[g.material]);
    if ((st = upload(s, s->d_geoms, s->geoms)) != TR_OK) return st;
    if ((st = upload(s, s->d_mats, s->mats)) != TR_OK) return st;
    if ((st = upload(s, s->d_vpool, s->vpool)) != TR_OK) return st;
    if ((st = upload(s, s->d_ipool, s->ipool)) != TR_OK) return st;
    if ((st = upload(s, s->d_vols, s->vols)) != TR_OK) return st;
    if ((st = upload(s, s->d_voxels, s->voxels)) != TR_OK) return st;
    if ((st = upload(s, s->d_shapes, s->shapes)) != TR_OK) return st;
    if ((st = upload(s, s->d_tex, s->tex_recs)) != TR_OK) return st;
    if ((st = upload(s, s->d_texels, s->texels)) != TR_OK) return st;
    srgb_table(&s->srgb);
    if ((st = upload(s, s->d_srgb, s->srgb)) != TR_OK) return st;
    s->lights_dirty = true;
    if ((st = sync_lights(s)) != TR_OK) return st;
    TR_HIP(hipStreamSynchronize(s->stream));
    s->accum_samples = 0;
    return TR_OK;
}

tr_status tr_scene_get_info(const tr_scene* s, tr_scene_info* info)
{
    if (!valid(s) || !info) return fail(TR_ERROR_INVALID_ARG, "null argument");
    memset(info, 0, sizeof(*info));
    info->primitives = s->prims_in.size();
    info->bvh_nodes = s->nodes.size();
    info->bvh_bytes = s->nodes.size() * sizeof(Node8);
    info->prim_bytes = s->prims.size() * sizeof(PrimRec);
    info->bvh_depth = s->bvh_stats.depth;
    info->geometries = (uint32_t)s->geoms.size();
    info->lights = (uint32_t)s->lights.size();
    info->materials = (uint32_t)s->mats.size();
    info->build_ms = s->build_ms;
    return TR_OK;
}

tr_status tr_scene_draw_once(tr_scene* s)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    auto t0 = std::chrono::steady_clock::now();
    tr_status st = start_frame(s);
    if (st != TR_OK) return st;
    return finish_frame(s, t0, (uint64_t)s->width * s->height * s->spp);
}

static std::chrono::steady_clock::time_point g_draw_start;

tr_status tr_scene_draw(tr_scene* s, uint32_t width, uint32_t height, const float vpi[16])
{
    if (!valid(s) || !vpi) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    if (width != s->width || height != s->height || memcmp(vpi, s->vpi, sizeof(s->vpi)) != 0) {
        s->width = width;
        s->height = height;
        memcpy(s->vpi, vpi, sizeof(s->vpi));
        s->accum_samples = 0;
    }
    g_draw_start = std::chrono::steady_clock::now();
    return start_frame(s);
}

tr_status tr_scene_frame_wait(tr_scene* s)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (!s->frame_pending) return TR_OK;
    return finish_frame(s, g_draw_start, (uint64_t)s->width * s->height * s->spp);
}

tr_status tr_scene_read_backbuffer(tr_scene* s, uint32_t* rgba, size_t count)
{
    if (!valid(s) || !rgba) return fail(TR_ERROR_INVALID_ARG, "null argument");
    size_t npx = (size_t)s->width * s->height;
    if (count < npx) return fail(TR_ERROR_INVALID_ARG, "buffer too small");
    if (s->accum_samples == 0 || !s->rgba.ptr) return fail(TR_ERROR_NOT_COMMITTED, "no frame rendered");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    TR_HIP(hipMemcpy(rgba, s->rgba.ptr, npx * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return TR_OK;
}

tr_status tr_scene_save_image(tr_scene* s, const char* path)
{
    if (!valid(s) || !path) return fail(TR_ERROR_INVALID_ARG, "null argument");
    std::vector<uint32_t> px((size_t)s->width * s->height);
    tr_status st = tr_scene_read_backbuffer(s, px.data(), px.size());
    if (st != TR_OK) return st;
    return tr_image_save(path, s->width, s->height, px.data());
}

tr_status tr_scene_read_accum(tr_scene* s, double* rgb_sum, size_t count, uint64_t* spp_out)
{
    if (!valid(s) || !rgb_sum) return fail(TR_ERROR_INVALID_ARG, "null argument");
    size_t npx = (size_t)s->width * s->height;
    if (count < npx) return fail(TR_ERROR_INVALID_ARG, "buffer too small");
    if (s->accum_samples == 0 || !s->accum.ptr) return fail(TR_ERROR_NOT_COMMITTED, "no frame rendered");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    TR_HIP(hipMemcpy(rgb_sum, s->accum.ptr, npx * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (spp_out) *spp_out = s->accum_samples;
    return TR_OK;
}

tr_status tr_scene_trace_rays(tr_scene* s, size_t n, const tr_ray* rays, tr_hit* hits)
{
    if (!valid(s) || (n && (!rays || !hits))) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed || !s->device_ready) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    if (n == 0) return TR_OK;
    if (n > 0xffffffffull - 64) return fail(TR_ERROR_INVALID_ARG, "too many rays");
    DevBuf<tr_ray> d_rays;
    DevBuf<tr_hit> d_hits;
    if (d_rays.reserve(n) != hipSuccess || d_hits.reserve(n) != hipSuccess) {
        d_rays.release(); d_hits.release();
        return fail(TR_ERROR_OUT_OF_MEMORY, "device allocation failed");
    }
    tr_status st = TR_OK;
    if (s->spill.reserve((size_t)persistent_grid(s) * kTraceBlock * kStackSpill) != hipSuccess)
        st = fail(TR_ERROR_OUT_OF_MEMORY, "device allocation failed");
    if (st == TR_OK && hipMemcpyAsync(d_rays.ptr, rays, n * sizeof(tr_ray), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        st = fail(TR_ERROR_DEVICE, "ray upload failed");
    if (st == TR_OK) {
        launch_trace_rays(s->stream, scene_view(s), d_rays.ptr, d_hits.ptr, (uint32_t)n, s->spill.ptr,
                          persistent_grid(s));
        if (hipGetLastError() != hipSuccess) st = fail(TR_ERROR_DEVICE, "trace launch failed");
    }
    if (st == TR_OK && hipMemcpyAsync(hits, d_hits.ptr, n * sizeof(tr_hit), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        st = fail(TR_ERROR_DEVICE, "hit download failed");
    if (hipStreamSynchronize(s->stream) != hipSuccess && st == TR_OK) st = fail(TR_ERROR_DEVICE, "trace failed");
    d_rays.release(); d_hits.release();
    return st;
}

tr_status tr_scene_trace_rays_noalloc(tr_scene* s, size_t n, const tr_ray* device_rays, tr_hit* device_hits,
                                      void* stream)
{
    if (!valid(s) || (n && (!device_rays || !device_hits))) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed || !s->device_ready) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    if (n == 0) return TR_OK;
    if (n > 0xffffffffull - 64) return fail(TR_ERROR_INVALID_ARG, "too many rays");
    if (s->bvh_stats.depth + 2 > (uint32_t)kStackLds + kNoAllocSpill)
        return fail(TR_ERROR_UNSUPPORTED, "the allocation-free trace holds BVHs of at most " +
                                              std::to_string(kStackLds + kNoAllocSpill - 2) + " levels; this one has " +
                                              std::to_string(s->bvh_stats.depth));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    if (launch_trace_rays_noalloc(st, scene_view(s), device_rays, device_hits, (uint32_t)n, persistent_grid(s)) !=
        hipSuccess)
        return fail(TR_ERROR_DEVICE, "trace launch failed");
    if (!stream && hipStreamSynchronize(s->stream) != hipSuccess) return fail(TR_ERROR_DEVICE, "trace failed");
    return TR_OK;
}

tr_status tr_scene_ray_query(tr_scene* s, const float tc[2], tr_sample_data* sd, uint64_t* out_geom_id)
{
    if (!valid(s) || !tc || !sd || !out_geom_id) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed || !s->device_ready) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    if (!std::isfinite(tc[0]) || !std::isfinite(tc[1])) return fail(TR_ERROR_INVALID_ARG, "non-finite screen point");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    TR_HIP(s->query_out.reserve(1));
    launch_ray_query(s->stream, scene_view(s), s->vpi, tc[0], tc[1], s->query_out.ptr);
    TR_HIP(hipGetLastError());
    RayQueryOut o;
    TR_HIP(hipMemcpyAsync(&o, s->query_out.ptr, sizeof(o), hipMemcpyDeviceToHost, s->stream));
    TR_HIP(hipStreamSynchronize(s->stream));
    if (o.geom == TR_INVALID_GEOMETRY) {
        *out_geom_id = TR_INVALID_GEOMETRY;
        return TR_OK;
    }
    memset(sd, 0, sizeof(*sd));
    const GeomRec& g = s->geoms[o.geom];
    sd->material = g.material == kNoMaterial ? nullptr : s->mat_keys[g.material];
    memcpy(sd->normal, o.normal, sizeof(sd->normal));
    memcpy(sd->tangent, o.tangent, sizeof(sd->tangent));
    memcpy(sd->binormal, o.binormal, sizeof(sd->binormal));
    memcpy(sd->texcoord, o.tc, sizeof(sd->texcoord));
    sd->t = o.t;
    *out_geom_id = o.geom < s->ext_ids.size() && s->ext_ids[o.geom] != UINT64_MAX ? s->ext_ids[o.geom] : o.geom;
    return TR_OK;
}

tr_status tr_scene_render_paths(tr_scene* s, size_t pixel_count, const uint32_t* pixels, float* out_rgb,
                                uint32_t* out_hash)
{
    if (!valid(s) || (pixel_count && (!pixels || !out_rgb))) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed || !s->device_ready) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    if (pixel_count == 0) return TR_OK;
    uint64_t npx = (uint64_t)s->width * s->height;
    for (size_t i = 0; i < pixel_count; ++i)
        if (pixels[i] >= npx) return fail(TR_ERROR_INVALID_ARG, "pixel index out of range");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    size_t nsamples = pixel_count * s->spp;
    DevBuf<uint32_t> d_pix, d_hash;
    DevBuf<float> d_rgb;
    if (d_pix.reserve(pixel_count) != hipSuccess || d_rgb.reserve(nsamples * 3) != hipSuccess ||
        d_hash.reserve(nsamples) != hipSuccess) {
        d_pix.release(); d_rgb.release(); d_hash.release();
        return fail(TR_ERROR_OUT_OF_MEMORY, "device allocation failed");
    }
    auto t0 = std::chrono::steady_clock::now();
    tr_status st = TR_OK;
    if (hipMemcpyAsync(d_pix.ptr, pixels, pixel_count * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        st = fail(TR_ERROR_DEVICE, "pixel upload failed");
    BatchOut out;
    memset(&out, 0, sizeof(out));
    out.path_rgb = d_rgb.ptr;
    out.path_hash = d_hash.ptr;
    out.pixel_list = d_pix.ptr;
    if (st == TR_OK) st = enqueue_frame(s, pixel_count, out);
    if (st == TR_OK) {
        if (hipMemcpyAsync(out_rgb, d_rgb.ptr, nsamples * 3 * sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
            st = fail(TR_ERROR_DEVICE, "download failed");
        if (out_hash && st == TR_OK &&
            hipMemcpyAsync(out_hash, d_hash.ptr, nsamples * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
            st = fail(TR_ERROR_DEVICE, "download failed");
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess && st == TR_OK) st = fail(TR_ERROR_DEVICE, "render failed");
    if (st == TR_OK) {
        collect_stats(s, nsamples);
        s->stats.frame_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    d_pix.release(); d_rgb.release(); d_hash.release();
    return st;
}

namespace {

/* Cuts the next batches of at most `pool` paths, ending at pixel
 * boundaries, from position (request r, pixel lp of it), appending up to
 * `max_batches` to `out`.  Returns false if a single pixel's paths exceed
 * the pool. */
bool cut_batches(const std::vector<ReqRec>& reqs, const std::vector<uint64_t>& path_off,
                 const std::vector<uint64_t>& pix_off, uint32_t n_gi, uint32_t pool, uint32_t max_batches,
                 uint32_t* r, uint64_t* lp, std::vector<ReqBatch>* out)
{
    const uint32_t n = (uint32_t)reqs.size();
    for (uint32_t k = 0; k < max_batches && *r < n; ++k) {
        ReqBatch b;
        const uint64_t per0 = (uint64_t)reqs[*r].spp * n_gi;
        b.path_base = path_off[*r] + *lp * per0;
        b.pix_base = pix_off[*r] + *lp;
        b.n_slots = 0;
        b.n_pix = 0;
        uint64_t room = pool;
        while (*r < n) {
            const uint64_t per = (uint64_t)reqs[*r].spp * n_gi;
            const uint64_t pixels = (uint64_t)reqs[*r].width * reqs[*r].height;
            const uint64_t take = std::min(room / per, pixels - *lp);
            if (take == 0) break;
            b.n_pix += (uint32_t)take;
            b.n_slots += (uint32_t)(take * per);
            room -= take * per;
            *lp += take;
            if (*lp < pixels) break;
            ++*r;
            *lp = 0;
        }
        if (b.n_slots == 0) return false;
        out->push_back(b);
    }
    return true;
}

}  // namespace

tr_status tr_scene_render_requests(tr_scene* s, size_t count, const tr_request* req, const tr_request_options* opt,
                                   float* out_mean, float* out_image, tr_request_stats* out_stats)
{
    if (!valid(s) || (count && !req)) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed || !s->device_ready) return fail(TR_ERROR_NOT_COMMITTED, "commitScene has not been called");
    if (count == 0) return TR_OK;
    if (count >= (1u << 24)) return fail(TR_ERROR_INVALID_ARG, "at most 2^24 - 1 requests");
    if (s->lights.empty() || s->lights[0].src.type != TR_LIGHT_DIRECTIONAL)
        return fail(TR_ERROR_INVALID_ARG, "render requests turn light 0, which must be a directional light");
    if (s->render_mode != 0) return fail(TR_ERROR_UNSUPPORTED, "render requests take the Normal render mode only");
    if (!s->vols.empty()) return fail(TR_ERROR_UNSUPPORTED, "render requests do not take grid volumes");
    if (s->cache.enabled) return fail(TR_ERROR_UNSUPPORTED, "render requests do not take the radiance cache");
    tr_request_options o;
    memset(&o, 0, sizeof(o));
    o.use_graph = 1;
    if (opt) o = *opt;
    if (o.pool_slots == 0) o.pool_slots = 1u << 24;
    if (o.trace_waves_per_cu == 0) o.trace_waves_per_cu = 16;
    if (o.trace_waves_per_cu > 32) return fail(TR_ERROR_INVALID_ARG, "trace_waves_per_cu above 32");
    if (s->frame_pending) {
        tr_status st = tr_scene_frame_wait(s);
        if (st != TR_OK) return st;
    }
    auto t0 = std::chrono::steady_clock::now();
    /* request table and the path / pixel index spaces */
    std::vector<ReqRec> reqs(count);
    std::vector<uint64_t> path_off(count + 1), pix_off(count + 1);
    uint32_t max_per_pixel = 0;
    for (size_t i = 0; i < count; ++i) {
        const tr_request& q = req[i];
        if (q.width < 2 || q.height < 2 || q.samples_camera == 0 || q.samples_camera > 65536)
            return fail(TR_ERROR_INVALID_ARG, "requests need width, height >= 2 and 1..65536 camera samples");
        for (int k = 0; k < 16; ++k)
            if (!std::isfinite(q.view_proj_inv[k])) return fail(TR_ERROR_INVALID_ARG, "non-finite camera matrix");
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(q.light_direction[k])) return fail(TR_ERROR_INVALID_ARG, "non-finite light direction");
        ReqRec& r = reqs[i];
        memcpy(r.vpi, q.view_proj_inv, sizeof(r.vpi));
        r.width = q.width;
        r.height = q.height;
        r.spp = q.samples_camera;
        memcpy(r.light_dir, q.light_direction, sizeof(r.light_dir));
        const uint64_t px = (uint64_t)q.width * q.height;
        pix_off[i + 1] = pix_off[i] + px;
        path_off[i + 1] = path_off[i] + px * q.samples_camera * s->n_gi;
        max_per_pixel = std::max(max_per_pixel, q.samples_camera * s->n_gi);
    }
    const uint64_t total_pix = pix_off[count], total_paths = path_off[count];
    if (total_pix >= (1ull << 32)) return fail(TR_ERROR_INVALID_ARG, "more than 2^32 - 1 pixels over all requests");
    /* candidate configurations: (paths per batch, traversal waves per CU) */
    std::vector<std::pair<uint32_t, uint32_t>> cands;
    if (o.autotune) {
        for (uint32_t pool : {1u << 22, 1u << 23, 1u << 24, 1u << 25})
            for (uint32_t w : {16u, 24u}) cands.push_back({pool, w});
    } else {
        cands.push_back({o.pool_slots, o.trace_waves_per_cu});
    }
    for (auto& c : cands) c.first = std::max(c.first, max_per_pixel);
    uint32_t pool_max = 0;
    for (auto& c : cands) pool_max = std::max(pool_max, c.first);
    tr_status st = sync_lights(s);
    if (st != TR_OK) return st;
    {
        /* the pool sized for the largest candidate */
        uint32_t keep = s->settings.path_pool_size;
        s->settings.path_pool_size = pool_max;
        st = ensure_pool(s, pool_max);
        s->settings.path_pool_size = keep;
        if (st != TR_OK) return st;
    }
    if (s->capacity < max_per_pixel) return fail(TR_ERROR_OUT_OF_MEMORY, "path pool below one pixel's samples");
    for (auto& c : cands) c.first = std::min(c.first, s->capacity);
    TR_HIP(s->slot_req.reserve(s->capacity));
    /* traversal stack spill for the widest persistent grid */
    uint32_t max_waves = 16;
    for (auto& c : cands) max_waves = std::max(max_waves, c.second);
    TR_HIP(s->spill.reserve((size_t)std::max(1, s->cu_count) * max_waves * kTraceBlock * kStackSpill));
    /* batch table: the smallest candidate cuts the most batches; each
     * trial phase may end one batch early */
    uint32_t pool_min = s->capacity;
    for (auto& c : cands) pool_min = std::min(pool_min, c.first);
    uint64_t batch_cap = cands.size() + 2;
    {
        std::vector<ReqBatch> tmp;
        uint32_t r = 0;
        uint64_t lp = 0;
        while (r < count) {
            tmp.clear();
            if (!cut_batches(reqs, path_off, pix_off, s->n_gi, pool_min, 1u << 16, &r, &lp, &tmp))
                return fail(TR_ERROR_INVALID_ARG, "a pixel's samples exceed the batch size");
            batch_cap += tmp.size();
        }
    }
    DevBuf<ReqRec> d_reqs;
    DevBuf<uint64_t> d_path_off, d_pix_off;
    DevBuf<ReqBatch> d_batches;
    DevBuf<uint32_t> d_cursor;
    DevBuf<double> d_accum;
    DevBuf<float> d_mean, d_image;
    TR_HIP(d_reqs.reserve(count));
    TR_HIP(d_path_off.reserve(count + 1));
    TR_HIP(d_pix_off.reserve(count + 1));
    TR_HIP(d_batches.reserve(batch_cap));
    TR_HIP(d_cursor.reserve(1));
    TR_HIP(d_accum.reserve(total_pix * 3));
    TR_HIP(d_mean.reserve(count * 3));
    if (out_image) TR_HIP(d_image.reserve(total_pix * 3));
    hipStream_t stream = s->stream;
    TR_HIP(hipMemcpyAsync(d_reqs.ptr, reqs.data(), count * sizeof(ReqRec), hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemcpyAsync(d_path_off.ptr, path_off.data(), (count + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemcpyAsync(d_pix_off.ptr, pix_off.data(), (count + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemsetAsync(d_cursor.ptr, 0, sizeof(uint32_t), stream));
    TR_HIP(hipMemsetAsync(d_accum.ptr, 0, total_pix * 3 * sizeof(double), stream));

    FrameParams p = frame_params(s);
    p.req.reqs = d_reqs.ptr;
    p.req.path_off = d_path_off.ptr;
    p.req.pix_off = d_pix_off.ptr;
    p.req.batches = d_batches.ptr;
    p.req.cursor = d_cursor.ptr;
    p.req.n = (uint32_t)count;
    PathPool pool = path_pool(s);
    BatchOut out;
    memset(&out, 0, sizeof(out));
    CacheRecords rec;
    memset(&rec, 0, sizeof(rec));

    /* the launches of one batch: every per-batch quantity is read on the
     * device through the cursor, so the sequence can be captured once */
    auto enqueue_batch = [&](uint32_t slots, uint32_t waves, uint32_t max_pix) {
        const uint32_t grid = (uint32_t)std::max(1, s->cu_count) * waves;
        launch_raygen(stream, p, pool, out, 0, slots, rec);
        int cur = 0, nxt = 1;
        for (uint32_t v = 0; v <= s->max_depth; ++v) {
            launch_trace(stream, p, pool, out, 0, v, cur, nxt, grid, grid, nullptr);
            launch_shade(stream, p, pool, out, 0, v, cur, nxt, slots, rec);
            launch_compact(stream, pool, slots, cur, nxt, false);
            if (p.sc.n_lights > 0) launch_shadow(stream, p, pool, out, 0, v, grid, nullptr, waves > 16);
            std::swap(cur, nxt);
        }
        launch_accumulate_requests(stream, p, pool, d_accum.ptr, max_pix, d_cursor.ptr);
    };
    std::vector<ReqBatch> batches;
    batches.reserve(batch_cap);   /* stable storage while uploads are in flight */
    uint32_t pos_r = 0;
    uint64_t pos_lp = 0;
    uint32_t launched = 0;
    /* runs the batches [launched, batches.size()) with one configuration */
    auto run = [&](uint32_t slots, uint32_t waves) -> tr_status {
        const uint32_t b0 = launched, b1 = (uint32_t)batches.size();
        if (b1 <= b0) return TR_OK;
        if (b1 > batch_cap) return fail(TR_ERROR_INVALID_ARG, "batch table overflow");
        TR_HIP(hipMemcpyAsync(d_batches.ptr + b0, batches.data() + b0, (b1 - b0) * sizeof(ReqBatch),
                              hipMemcpyHostToDevice, stream));
        uint32_t max_pix = 1;
        for (uint32_t k = b0; k < b1; ++k) max_pix = std::max(max_pix, batches[k].n_pix);
        if (o.use_graph) {
            /* one graph holding the bounce loops of all these batches, one
             * after another, launched once (each batch's parameters come
             * from the device batch table through the cursor) */
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            TR_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            for (uint32_t k = b0; k < b1; ++k) enqueue_batch(slots, waves, max_pix);
            hipError_t e = hipStreamEndCapture(stream, &graph);
            if (e != hipSuccess) return fail(TR_ERROR_DEVICE, "graph capture failed");
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return fail(TR_ERROR_DEVICE, "graph instantiation failed");
            }
            e = hipGraphLaunch(exec, stream);
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(TR_ERROR_DEVICE, "graph launch failed");
        } else {
            for (uint32_t k = b0; k < b1; ++k) enqueue_batch(slots, waves, max_pix);
            TR_HIP(hipGetLastError());
        }
        launched = b1;
        return TR_OK;
    };
    /* autotuning: two trial batches per candidate, timed with events */
    size_t best = 0;
    double tune_ms = 0.0;
    uint32_t timed = 0;
    if (cands.size() > 1) {
        double best_rate = -1.0;
        for (size_t c = 0; c < cands.size() && pos_r < count; ++c) {
            const uint32_t b0 = (uint32_t)batches.size();
            if (!cut_batches(reqs, path_off, pix_off, s->n_gi, cands[c].first, 2, &pos_r, &pos_lp, &batches))
                return fail(TR_ERROR_INVALID_ARG, "a pixel's samples exceed the batch size");
            uint64_t paths = 0;
            for (uint32_t k = b0; k < batches.size(); ++k) paths += batches[k].n_slots;
            hipEvent_t a = get_event(s), b = get_event(s);
            if (!a || !b) return fail(TR_ERROR_DEVICE, "event creation failed");
            TR_HIP(hipEventRecord(a, stream));
            if ((st = run(cands[c].first, cands[c].second)) != TR_OK) return st;
            TR_HIP(hipEventRecord(b, stream));
            TR_HIP(hipEventSynchronize(b));
            float ms = 0.0f;
            TR_HIP(hipEventElapsedTime(&ms, a, b));
            s->ev_pool.push_back(a);
            s->ev_pool.push_back(b);
            tune_ms += ms;
            ++timed;
            const double rate = (double)paths / std::max(1e-6, (double)ms);
            if (rate > best_rate) {
                best_rate = rate;
                best = c;
            }
        }
    }
    /* the rest with the chosen configuration */
    if (!cut_batches(reqs, path_off, pix_off, s->n_gi, cands[best].first, 0xffffffffu, &pos_r, &pos_lp, &batches))
        return fail(TR_ERROR_INVALID_ARG, "a pixel's samples exceed the batch size");
    if ((st = run(cands[best].first, cands[best].second)) != TR_OK) return st;
    launch_request_outputs(stream, p.req, s->n_gi, d_accum.ptr, d_mean.ptr, out_image ? d_image.ptr : nullptr,
                           total_pix);
    TR_HIP(hipGetLastError());
    if (out_mean)
        TR_HIP(hipMemcpyAsync(out_mean, d_mean.ptr, count * 3 * sizeof(float), hipMemcpyDeviceToHost, stream));
    if (out_image)
        TR_HIP(hipMemcpyAsync(out_image, d_image.ptr, total_pix * 3 * sizeof(float), hipMemcpyDeviceToHost, stream));
    TR_HIP(hipStreamSynchronize(stream));
    if (out_stats) {
        memset(out_stats, 0, sizeof(*out_stats));
        out_stats->paths = total_paths;
        out_stats->pixels = total_pix;
        out_stats->batches = (uint32_t)batches.size();
        out_stats->pool_slots = cands[best].first;
        out_stats->trace_waves_per_cu = cands[best].second;
        out_stats->candidates = timed;
        out_stats->tune_ms = tune_ms;
        out_stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return TR_OK;
}

tr_status tr_scene_get_stats(const tr_scene* s, tr_stats* out)
{
    if (!valid(s) || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    *out = s->stats;
    return TR_OK;
}

/* ---------------- debug exports (tempest_rt_debug.h) ---------------- */

tr_status tr_debug_build_scene_host(tr_scene* s)
{
    if (!valid(s)) return fail(TR_ERROR_INVALID_ARG, "null scene");
    if (s->committed) return TR_OK;
    std::string err;
    auto t0 = std::chrono::steady_clock::now();
    if (!build_bvh8(s->prims_in, s->nodes, s->prims, &s->bvh_stats, &err)) return fail(TR_ERROR_INVALID_ARG, err);
    s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s->committed = true;
    return TR_OK;
}

tr_status tr_debug_trace_rays_host(const tr_scene* s, size_t n, const tr_ray* rays, tr_hit* hits, uint64_t* node_visits,
                                   uint64_t* prim_tests)
{
    if (!valid(s) || (n && (!rays || !hits))) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed) return fail(TR_ERROR_NOT_COMMITTED, "scene not built");
    uint64_t nv = 0, np = 0;
    HostStack* stack = new HostStack();
    for (size_t i = 0; i < n; ++i) {
        RayPre r = ray_setup(rays[i].org, rays[i].dir);
        stack->sp = 0;
        uint32_t v = 0, p = 0;
        Hit h = traverse_closest(s->nodes.data(), s->prims.data(), s->shapes.data(), r, rays[i].tnear, rays[i].tfar, *stack, &v, &p);
        nv += v;
        np += p;
        hits[i].t = h.t; hits[i].u = h.u; hits[i].v = h.v; hits[i].geom = h.geom; hits[i].prim = h.prim;
    }
    delete stack;
    if (node_visits) *node_visits = nv;
    if (prim_tests) *prim_tests = np;
    return TR_OK;
}

tr_status tr_debug_bvh_check(const tr_scene* s, uint64_t* violations)
{
    if (!valid(s) || !violations) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (!s->committed) return fail(TR_ERROR_NOT_COMMITTED, "scene not built");
    /* every primitive's box lies inside its leaf slot's decoded box, every
     * internal slot points at an existing node, every primitive is
     * referenced exactly once */
    uint64_t bad = 0;
    std::vector<char> seen(s->prims.size(), 0);
    for (const Node8& n : s->nodes) {
        float sc[3] = {exp2i(n.exp[0]), exp2i(n.exp[1]), exp2i(n.exp[2])};
        uint32_t rank = 0;
        for (int slot = 0; slot < 8; ++slot) {
            float lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = n.origin[a] + half_value(n.qlo[a][slot]) * sc[a];
                hi[a] = n.origin[a] + half_value(n.qhi[a][slot]) * sc[a];
            }
            if ((n.imask >> slot) & 1u) {
                if (n.child_base + rank >= s->nodes.size()) ++bad;
                ++rank;
            } else if (n.meta[slot]) {
                uint32_t first = n.prim_base + (n.meta[slot] & 31u), cnt = n.meta[slot] >> 5u;
                for (uint32_t k = 0; k < cnt; ++k) {
                    if (first + k >= s->prims.size()) { ++bad; continue; }
                    seen[first + k]++;
                    const PrimRec& p = s->prims[first + k];
                    float pl[3], ph[3];
                    prim_box(p, pl, ph);
                    for (int a = 0; a < 3; ++a)
                        if (pl[a] < lo[a] || ph[a] > hi[a]) ++bad;
                }
            } else if (!(lo[0] > hi[0] && lo[1] > hi[1] && lo[2] > hi[2])) {
                ++bad; /* an empty slot's box must be inverted */
            }
        }
    }
    for (char c : seen)
        if (c != 1) ++bad;
    *violations = bad;
    return TR_OK;
}

/* The pseudo-volume BRDF microbenchmark (ComputeImage,
 * tests/benchmark/pseudo-volume-performance-test.cu:16-41): `repeats` timed
 * width x height launches after one warm-up. */
tr_status tr_bench_pseudo_volume_image(const tr_material* m, uint32_t width, uint32_t height, uint32_t repeats,
                                       uint32_t* rgba, float probe[4], float* ms_per_image)
{
    if (!m || !rgba || !probe || !ms_per_image) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (m->model != TR_MODEL_SGGX_PSEUDO_VOLUME) return fail(TR_ERROR_INVALID_ARG, "the material is not a pseudo-volume");
    if (m->stddev_map || m->basis_map) return fail(TR_ERROR_INVALID_ARG, "the benchmark takes constant parameters");
    if (m->sggx_sample_count < 1 || m->sggx_sample_count > 4096 || m->sggx_depth > 64)
        return fail(TR_ERROR_INVALID_ARG, "pseudo-volume SampleCount must be 1..4096 and Depth at most 64");
    if (width < 1 || height < 1 || width > 16384 || height > 16384 || repeats < 1 || repeats > 1000)
        return fail(TR_ERROR_INVALID_ARG, "image size or repeat count out of range");
    const size_t n = (size_t)width * height;
    DevBuf<tr_material> dm;
    DevBuf<uint32_t> dimg;
    DevBuf<float> dprobe;
    TR_HIP(dm.reserve(1));
    TR_HIP(dimg.reserve(n));
    TR_HIP(dprobe.reserve(4));
    TR_HIP(hipMemcpy(dm.ptr, m, sizeof(tr_material), hipMemcpyHostToDevice));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TR_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(TR_ERROR_DEVICE, "hipEventCreate failed");
    }
    launch_pseudo_volume_image(nullptr, dm.ptr, width, height, dimg.ptr, dprobe.ptr);
    (void)hipEventRecord(e0, nullptr);
    for (uint32_t r = 0; r < repeats; ++r) launch_pseudo_volume_image(nullptr, dm.ptr, width, height, dimg.ptr, dprobe.ptr);
    (void)hipEventRecord(e1, nullptr);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TR_HIP(e);
    TR_HIP(hipGetLastError());
    TR_HIP(hipMemcpy(rgba, dimg.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(probe, dprobe.ptr, 4 * sizeof(float), hipMemcpyDeviceToHost));
    *ms_per_image = ms / (float)repeats;
    return TR_OK;
}

}  // extern "C"
