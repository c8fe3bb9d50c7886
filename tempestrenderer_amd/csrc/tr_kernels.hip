/*
 * tr_kernels.hip - wavefront path-tracing kernels for gfx950.
 *
 * The path loop restated here is the reference CPU integrator
 * RayTracerScene::subdraw (src/graphics/ray-tracing/ray-tracing.cc:
 * 1981-2379) with next-event estimation EvaluateLightSources (:1809-1930)
 * and the balance heuristic EvaluateTotalBalanceHeuristic (:1791-1806),
 * split into stages over a pool of path slots instead of a per-pixel loop:
 *   raygen -> { trace -> shade -> shadow } per vertex -> accumulate.
 * Traversal kernels are persistent: a fixed grid of one-wave blocks walks
 * the ray queue with a static interleaved stride (no shared work counter,
 * whose single-address atomics would serialise thousands of waves); each
 * lane walks the 8-wide BVH with a node-group stack whose first kStackLds
 * entries live in LDS.
 *
 * Philox counter of every random number: (pixel, camera sample,
 * split << 16 | vertex << 8 | stream, sub).  The CPU oracle uses the same
 * layout, so a path is identical on both sides.
 *
 * Scenes with grid volumes use the VOL variants of the traversal kernels:
 * a closest hit on a volume box runs delta tracking in the same thread and
 * either turns into a scattering vertex or re-traces beyond the volume;
 * shadow rays through volumes take the reference's two-repetition Woodcock
 * estimate.
 */
#include "tr_kernels.h"
#include "tr_shading.h"
#include "tr_traverse.h"
#include "tr_volume.h"

#include <hip/hip_ext.h>

#include <mutex>

namespace trk {

namespace {

constexpr uint32_t kFnvBasis = 2166136261u;
constexpr uint32_t kFnvPrime = 16777619u;

__device__ inline uint32_t fnv_hit(uint32_t h, uint32_t geom, uint32_t prim)
{
    h = (h ^ geom) * kFnvPrime;
    return (h ^ prim) * kFnvPrime;
}

/* The LDS side moves an entry as two 32-bit words and the spill side as
 * one 64-bit word (in DevStack and LaneStack), so that the compiler keeps
 * an LDS access (ds_*) and a global one instead of merging both into one
 * flat access through a selected address, which waits on both memory
 * counters. */
struct DevStack {
    uint2* lds;     /* &stk[0][lane], stride kTraceBlock */
    uint2* spill;   /* kStackSpill entries of this lane */
    int sp;
    __device__ void push(uint32_t b, uint32_t m)
    {
        if (sp < kStackLds) lds[sp * kTraceBlock] = make_uint2(b, m);
        else *reinterpret_cast<uint64_t*>(spill + (sp - kStackLds)) = (uint64_t)b | ((uint64_t)m << 32);
        ++sp;
    }
    __device__ void pop(uint32_t* b, uint32_t* m)
    {
        --sp;
        if (sp < kStackLds) {
            const uint2 e = lds[sp * kTraceBlock];
            *b = e.x;
            *m = e.y;
        } else {
            const uint64_t w = *reinterpret_cast<const uint64_t*>(spill + (sp - kStackLds));
            *b = (uint32_t)w;
            *m = (uint32_t)(w >> 32);
        }
    }
    __device__ bool empty() const { return sp == 0; }
};

/* The refill kernel's stack: the same entries as DevStack, addressed from
 * the lane and the wave's spill base on each access, so that no per-lane
 * pointer is held in registers. */
struct LaneStack {
    uint2 (*lds)[kTraceBlock];   /* the block's LDS stack array */
    uint2* spill;                /* the wave's spill area: lane l at l * kStackSpill */
    int sp;
    __device__ void push(uint32_t b, uint32_t m)
    {
        if (sp < kStackLds) {
            lds[sp][threadIdx.x] = make_uint2(b, m);
        } else {
            uint64_t* w = reinterpret_cast<uint64_t*>(spill + threadIdx.x * kStackSpill + (sp - kStackLds));
            *w = (uint64_t)b | ((uint64_t)m << 32);
        }
        ++sp;
    }
    __device__ void pop(uint32_t* b, uint32_t* m)
    {
        --sp;
        if (sp < kStackLds) {
            const uint2 e = lds[sp][threadIdx.x];
            *b = e.x;
            *m = e.y;
        } else {
            const uint64_t w = *reinterpret_cast<const uint64_t*>(spill + threadIdx.x * kStackSpill + (sp - kStackLds));
            *b = (uint32_t)w;
            *m = (uint32_t)(w >> 32);
        }
    }
    __device__ bool empty() const { return sp == 0; }
};

/* The number of set bits of a wave mask below this lane. */
__device__ inline uint32_t bits_below_lane(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* Segment-local append of one entry to each of the three shading queues
 * (next path, shadow, cache query) for the threads whose predicate holds:
 * the block's entries go to its own segment in thread order and the
 * block's first threads write the three counts.  Queues 0 and 1 receive
 * `value` (the path slot); queue 2 receives the entry's own position,
 * which is returned: the query record is stored there, so the queries of
 * a launch lie densely in segment order.  Every thread of the block must
 * call it.  No atomics: one counter bumped by every wave of a launch
 * serialises (tens of thousands of same-address atomics per launch),
 * while k_compact concatenates the segments in a few microseconds. */
__device__ inline uint32_t segment_push(const PathPool& pool, uint32_t seg, bool e0, bool e1, bool e2, uint32_t value)
{
    __shared__ uint32_t wc[3][kSegment / 64];
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    const uint64_t m[3] = {__ballot(e0), __ballot(e1), __ballot(e2)};
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) wc[k][w] = (uint32_t)__popcll(m[k]);
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    const bool e[3] = {e0, e1, e2};
    const size_t seg_stride = (size_t)pool.seg_cap * kSegment;
    uint32_t pos2 = 0xffffffffu;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t o = 0;
        for (uint32_t j = 0; j < w; ++j) o += wc[k][j];
        const uint32_t pos = seg * kSegment + o + (uint32_t)__popcll(m[k] & lt);
        if (k == 2) pos2 = pos;
        if (e[k]) pool.seg[k * seg_stride + pos] = k == 2 ? pos : value;
    }
    if (threadIdx.x < 3) {
        uint32_t t = 0;
        for (uint32_t j = 0; j < kSegment / 64; ++j) t += wc[threadIdx.x][j];
        pool.seg_n[threadIdx.x * pool.seg_cap + seg] = t;
    }
    return pos2;
}

/* Request mode: the request whose range of `off` (path or pixel offsets,
 * increasing, every request non-empty) holds q. */
__device__ inline uint32_t find_request(const uint64_t* off, uint32_t n, uint64_t q)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* Request mode coordinates of slot `slot` of the current batch, a path of
 * request r: pixel over all requests (the random-number key), pixel within
 * the request, camera sample and GI split. */
__device__ inline void request_coords(const FrameParams& p, uint32_t r, uint32_t slot, uint32_t* gpix,
                                      uint32_t* lpix, uint32_t* cs, uint32_t* split)
{
    const ReqBatch& b = p.req.batches[*p.req.cursor];
    const uint64_t local = b.path_base + slot - p.req.path_off[r];
    const uint64_t per_pixel = (uint64_t)p.req.reqs[r].spp * p.n_gi;
    /* a 32-bit quotient where it fits (a 64-bit division is a long
     * instruction sequence); the same value */
    const uint64_t lp = ((local | per_pixel) >> 32) == 0 ? (uint64_t)((uint32_t)local / (uint32_t)per_pixel)
                                                          : local / per_pixel;
    const uint32_t s = (uint32_t)(local - lp * per_pixel);
    *lpix = (uint32_t)lp;
    *gpix = (uint32_t)(p.req.pix_off[r] + lp);
    *cs = s / p.n_gi;
    *split = s % p.n_gi;
}

__device__ inline void slot_coords(const FrameParams& p, const BatchOut& out, uint64_t slot, uint32_t* pixel,
                                   uint32_t* cs, uint32_t* split, uint64_t* pseq)
{
    if (p.req.n) {
        const ReqBatch& b = p.req.batches[*p.req.cursor];
        const uint32_t r = find_request(p.req.path_off, p.req.n, b.path_base + slot);
        uint32_t lp;
        request_coords(p, r, (uint32_t)slot, pixel, &lp, cs, split);
        *pseq = lp;
        return;
    }
    uint64_t per_pixel = (uint64_t)p.spp * p.n_gi;
    uint64_t ps;
    uint32_t sample, split_;
    if ((slot >> 32) == 0 && (per_pixel >> 32) == 0) {
        /* 32-bit quotients (a 64-bit division is a long instruction
         * sequence); the same values */
        const uint32_t s32 = (uint32_t)slot;
        const uint32_t q = s32 / p.n_gi;
        ps = q / p.spp;
        sample = q - (uint32_t)ps * p.spp;
        split_ = s32 - q * p.n_gi;
    } else {
        ps = slot / per_pixel;
        sample = (uint32_t)((slot / p.n_gi) % p.spp);
        split_ = (uint32_t)(slot % p.n_gi);
    }
    *pseq = ps;
    *pixel = out.pixel_list ? out.pixel_list[ps] : (uint32_t)ps;
    *cs = p.frame * p.spp + sample;
    *split = split_;
}

__device__ inline V3 ld4(float4 v) { return mk(v.x, v.y, v.z); }
__device__ inline float4 st4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

/* ---------------- radiance cache helpers ---------------- */

/* Roughness input of the cache: the microfacet alpha (sqrt of
 * SpecularPowerToGeometryFactorSq, microfacet-common.hh:109-112; the same
 * map of the Blinn-Phong exponent; the geometric mean of the two alphas of
 * the anisotropic models), 1 for lobes without a specular part
 * and for SGGX surfaces. */
__device__ inline float cache_roughness(const tr_material* m)
{
    if (!is_microfacet(m->model) && !is_aniso(m->model) && m->model != TR_MODEL_BLINN_PHONG) return 1.0f;
    if (m->specular[0] == 0.0f && m->specular[1] == 0.0f && m->specular[2] == 0.0f) return 1.0f;
    float a = sqrtf(geom_factor_sq(m->specular_power[0]));
    if (is_aniso(m->model)) a = sqrtf(a * sqrtf(geom_factor_sq(m->specular_power[1])));
    return a < 1.0f ? a : 1.0f;
}

__device__ inline float unit_clamp(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

__device__ inline void cache_inputs(const FrameParams& p, V3 pos, V3 dir, const Sample& sd, float in[kMlpIn])
{
    in[0] = unit_clamp((pos.x - p.box_lo[0]) * p.box_rcp[0]);
    in[1] = unit_clamp((pos.y - p.box_lo[1]) * p.box_rcp[1]);
    in[2] = unit_clamp((pos.z - p.box_lo[2]) * p.box_rcp[2]);
    in[3] = (1.0f - dir.x) * 0.5f;
    in[4] = (1.0f - dir.y) * 0.5f;
    in[5] = (1.0f - dir.z) * 0.5f;
    in[6] = (sd.normal.x + 1.0f) * 0.5f;
    in[7] = (sd.normal.y + 1.0f) * 0.5f;
    in[8] = (sd.normal.z + 1.0f) * 0.5f;
    in[9] = cache_roughness(sd.mat);
}

/* Training record of a path, or ~0 if it is not a training path: path
 * p = pixel sequence * spp + camera sample of the frame trains iff
 * (p + frame) % train_stride == 0, as record (p + frame) / train_stride. */
__device__ inline uint32_t train_record(const FrameParams& p, uint64_t pseq, uint32_t cs)
{
    uint64_t path = pseq * p.spp + (cs - p.frame * p.spp) + p.frame;
    if (path % p.train_stride != 0) return 0xffffffffu;
    return (uint32_t)(path / p.train_stride);
}

/* ---------------- ray generation ---------------- */

/* The camera ray of camera sample cs of image pixel lpix (pixel: the
 * random-number key): through the jittered pixel (the pixel's corner grid
 * with one sample), from the near to the far plane of the inverse
 * view-projection m. */
__device__ inline void camera_ray(const FrameParams& p, uint32_t pixel, uint32_t lpix, uint32_t cs, uint32_t width,
                                  uint32_t height, uint32_t spp, const float* m, V3* start_out, V3* dir_out)
{
    uint32_t x = lpix % width, y = lpix / width;
    float sx, sy;
    if (spp > 1) {
        /* camera jitter draws with split 0: all GI splits share the camera ray */
        U4 r = philox(pixel, cs, (uint32_t)STREAM_CAMERA, 0u, p.key0, p.key1);
        sx = 2.0f * ((float)x + u01(r.a)) / (float)width - 1.0f;
        sy = 2.0f * ((float)y + u01(r.b)) / (float)height - 1.0f;
    } else {
        sx = 2.0f * (float)x / (float)(width - 1) - 1.0f;
        sy = 2.0f * (float)y / (float)(height - 1) - 1.0f;
    }
    /* Matrix4 * Vector4 (matrix4.hh:858-880), ToVector3 (vector4.hh:125-130) */
    float p0[4], p1[4];
    for (int r = 0; r < 4; ++r) {
        p0[r] = m[r] * sx;
        p1[r] = m[r] * sx;
        p0[r] = p0[r] + m[4 + r] * sy;
        p1[r] = p1[r] + m[4 + r] * sy;
        p0[r] = p0[r] + m[8 + r] * -1.0f;
        p1[r] = p1[r] + m[8 + r] * 1.0f;
        p0[r] = p0[r] + m[12 + r] * 1.0f;
        p1[r] = p1[r] + m[12 + r] * 1.0f;
    }
    float w0 = 1.0f / p0[3], w1 = 1.0f / p1[3];
    V3 start = mk(p0[0] * w0, p0[1] * w0, p0[2] * w0);
    V3 end = mk(p1[0] * w1, p1[1] * w1, p1[2] * w1);
    *start_out = start;
    *dir_out = normalize(sub(end, start));
}

__global__ __launch_bounds__(256) void k_raygen(FrameParams p, PathPool pool, BatchOut out, uint64_t slot_base,
                                                uint32_t n_slots, CacheRecords rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (p.req.n) n_slots = p.req.batches[*p.req.cursor].n_slots;
    if (i == 0) pool.counters[0] = n_slots;
    if (p.camera_fused || i >= n_slots) return;
    uint32_t pixel, cs, split, lpix;
    uint64_t pseq;
    uint32_t width = p.width, height = p.height, spp = p.spp;
    const float* m = p.vpi;
    if (p.req.n) {
        /* request mode: the slot's request, its camera and image size */
        const ReqBatch& b = p.req.batches[*p.req.cursor];
        const uint32_t r = find_request(p.req.path_off, p.req.n, b.path_base + i);
        pool.slot_req[i] = r;
        const ReqRec& R = p.req.reqs[r];
        request_coords(p, r, i, &pixel, &lpix, &cs, &split);
        pseq = lpix;
        width = R.width;
        height = R.height;
        spp = R.spp;
        m = R.vpi;
    } else {
        slot_coords(p, out, slot_base + i, &pixel, &cs, &split, &pseq);
        lpix = pixel;
    }
    V3 start, dir;
    camera_ray(p, pixel, lpix, cs, width, height, spp, m, &start, &dir);
    pool.ray_o[i] = st4(start, 0.0f);
    pool.ray_d[i] = st4(dir, 0.0f);
    /* throughput 1, balance 1, zero radiance and the hash basis are the
     * state of every path at its camera vertex: k_shade starts vertex 0
     * from them instead of reading them back, and writes all four */
    pool.queue[0][i] = i;
    if (p.cache) {
        uint32_t r = train_record(p, pseq, cs);
        pool.pflags[i] = r != 0xffffffffu ? PF_TRAIN : 0u;
        if (r != 0xffffffffu) {
            rec.slot[r] = i;
            rec.smp.valid[r] = 0.0f;
        }
    }
}

/* ---------------- closest-hit trace (persistent) ---------------- */

template <bool COUNT, bool VOL>
__global__ __launch_bounds__(kTraceBlock) void k_trace(FrameParams p, PathPool pool, BatchOut out, uint64_t slot_base,
                                                       uint32_t vertex, int cur, int nxt, unsigned long long* counts)
{
    __shared__ uint2 stk[kStackLds][kTraceBlock];
    const SceneView& sc = p.sc;
    uint32_t lane = threadIdx.x;
    uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    const uint32_t n = pool.counters[cur];
    const uint32_t* queue = pool.queue[cur];
    DevStack st;
    st.lds = &stk[0][lane];
    st.spill = pool.spill + (size_t)gtid * kStackSpill;
    uint32_t nv = 0, np = 0, nr = 0, ns = 0;
    /* static interleaved assignment: wave w takes rays w*64 + k*grid*64 */
    const uint32_t stride = gridDim.x * kTraceBlock;
    for (uint32_t base = blockIdx.x * kTraceBlock; base < n; base += stride) {
        uint32_t qi = base + lane;
        if (qi < n) {
            uint32_t slot = queue[qi];
            float4 o = pool.ray_o[slot], d = pool.ray_d[slot];
            float org[3] = {o.x, o.y, o.z}, dir[3] = {d.x, d.y, d.z};
            RayPre r = ray_setup(org, dir);
            st.sp = 0;
            Hit h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, 0.0f, INFINITY, st, COUNT ? &nv : nullptr,
                                     COUNT ? &np : nullptr);
            if (VOL) {
                /* the tracking loops of subdraw (ray-tracing.cc:2040-2128, 2212-2272) */
                uint32_t step = 0, crossings = 0;
                Rng rng;
                bool have_rng = false;
                while (h.geom != TR_INVALID_GEOMETRY && sc.geoms[h.geom].type == GEOM_VOLUME) {
                    if (!have_rng) {
                        uint32_t pixel, cs, split;
                        uint64_t pseq;
                        slot_coords(p, out, slot_base + slot, &pixel, &cs, &split, &pseq);
                        rng.k0 = p.key0; rng.k1 = p.key1; rng.pixel = pixel; rng.sample = cs; rng.split = split;
                        have_rng = true;
                    }
                    const VolumeRec& V = sc.volumes[sc.geoms[h.geom].tri_count];
                    float tc;
                    if (delta_track_any(sc.voxels, V, ld4(o), ld4(d), h.t, h.u, rng, vertex, &step, &tc,
                                    COUNT ? &ns : nullptr)) {
                        h.t = tc;
                        h.u = h.v = 0.0f;
                        break;
                    }
                    if (++crossings >= kMaxVolumeCrossings) {
                        h.geom = h.prim = TR_INVALID_GEOMETRY;
                        h.t = INFINITY;
                        break;
                    }
                    st.sp = 0;
                    h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, tc, INFINITY, st, COUNT ? &nv : nullptr,
                                         COUNT ? &np : nullptr);
                }
            }
            pool.hit[slot] = make_float4(h.t, h.u, h.v, __uint_as_float(h.geom));
            pool.hit_prim[slot] = h.prim;
            ++nr;
        }
    }
    if (COUNT) {
        atomicAdd(&counts[0], (unsigned long long)nv);
        atomicAdd(&counts[1], (unsigned long long)np);
        atomicAdd(&counts[2], (unsigned long long)nr);
        if (VOL) atomicAdd(&counts[6], (unsigned long long)ns);
    }
}

/* Closest hit for scenes without volumes, with ray refill: a lane whose
 * ray has finished takes the wave's next ray once at least kRefill lanes
 * are idle, so the traversal loop runs with mostly full waves instead of
 * waiting for the slowest of a fixed group of 64 rays.  The queue is dealt
 * to the waves in chunks of kRayChunk consecutive rays (wave w takes
 * chunks w, w + grid, ...).  Each iteration makes one memory round trip
 * per lane: a lane either tests the next primitive of the hit leaves of
 * its last node (a leaf cursor) or pops and tests the next node, and the
 * fetches of both kinds are in flight together.  The visiting order is
 * that of traverse_closest: a node's hit leaves, in visiting order, before
 * the next node is popped.  LEAVES instead tests a node and then all of
 * its hit leaves' primitives in one iteration: for scenes of a few hundred
 * primitives (at most kLeafLoopPrims), which stay in L1, the primitive
 * fetches are short and the per-iteration work of the loop dominates
 * (box scene, 12 primitives: 28.5 us per launch against 38.4 us; the
 * 7,938-triangle gonioreflectometer patch is 5 % faster with one round
 * trip per iteration).  The queue is dealt in chunks of 16 rays, or 64
 * with LEAVES (box: 34.8 us with chunks of 16). */
constexpr uint32_t kRefill = 16;
constexpr uint32_t kLeafLoopPrims = 1024;

template <bool COUNT, bool SHAPES, bool LEAVES>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kRefillWavesPerSimd))) void k_trace_refill(FrameParams p, PathPool pool, int cur,
                                                              unsigned long long* counts)
{
    __shared__ uint2 stk[kStackLds][kTraceBlock];
    const SceneView& sc = p.sc;
    const uint32_t n = pool.counters[cur];
    const uint32_t* queue = pool.queue[cur];
    const uint32_t grid = gridDim.x;
    LaneStack st;
    st.lds = stk;
    st.spill = pool.spill + (size_t)blockIdx.x * kTraceBlock * kStackSpill;
    st.sp = 0;
    uint32_t nv = 0, np = 0, nr = 0;
    constexpr uint32_t kRayChunk = LEAVES ? 64u : 16u;
    auto ray_index = [&](uint32_t k) -> uint32_t {
        return (blockIdx.x + (k / kRayChunk) * grid) * kRayChunk + (k % kRayChunk);
    };
    uint32_t next = 0;                /* the wave's next unassigned ray (uniform) */
    bool active = false;
    uint32_t slot = 0;
    RayPre r;
    Hit best;
    /* leaf cursor: hit leaves still to test (visiting order), the meta words
     * and primitive base of their node, the position in the current leaf */
    /* lv: bits 0-7 the leaf mask, bits 8-10 the position in the leaf */
    uint32_t lv = 0, lm0 = 0, lm1 = 0, lpb = 0;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const bool more = ray_index(next) < n;
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (more && (n_idle >= kRefill || idle == ~0ull)) {
            if (!active) {
                const uint32_t qi = ray_index(next + bits_below_lane(idle));
                if (qi < n) {
                    slot = queue[qi];
                    float4 o = pool.ray_o[slot], d = pool.ray_d[slot];
                    float org[3] = {o.x, o.y, o.z}, dir[3] = {d.x, d.y, d.z};
                    r = ray_setup(org, dir);
                    best.t = INFINITY;
                    best.u = best.v = 0.0f;
                    best.geom = best.prim = TR_INVALID_GEOMETRY;
                    st.sp = 0;
                    st.push(0u, (1u << r.octant()) | (1u << 8));
                    lv = 0;
                    active = true;
                    ++nr;
                }
            }
            next += n_idle;
        }
        if (__ballot(active) == 0ull) {
            if (ray_index(next) < n) continue;
            break;
        }
        if (LEAVES && active) {
            /* one node of the ray's traversal and its hit leaves */
            const uint32_t oct = r.octant();
            uint32_t base, masks;
            st.pop(&base, &masks);
            uint32_t vis = masks & 0xffu;
            const uint32_t imask = masks >> 8;
            const uint32_t i = ctz32(vis);
            vis &= vis - 1u;
            if (vis) st.push(base, vis | (imask << 8));
            const uint32_t cslot = i ^ oct;
            const uint32_t node_index = base + popc8(imask & ((1u << cslot) - 1u));
            NodeRegs nd = load_node(sc.nodes, node_index);
            if (COUNT) ++nv;
            const float tmax = best.geom == TR_INVALID_GEOMETRY ? INFINITY : best.t;
            const uint32_t hit = intersect_children(nd, r, 0.0f, tmax);
            const uint32_t nimask = node_imask(nd);
            const uint32_t inner = hit & nimask;
            uint32_t leaves = to_visit_order(hit & ~nimask, oct);
            while (leaves) {
                const uint32_t li = ctz32(leaves);
                leaves &= leaves - 1u;
                const uint32_t meta = node_meta(nd, li ^ oct);
                const uint32_t first = nd.d[5] + (meta & 31u);
                const uint32_t count = meta >> 5;
                for (uint32_t k = 0; k < count; ++k) {
                    if (COUNT) ++np;
                    test_prim<SHAPES, false>(load_prim(sc.prims, first + k), sc.shapes, r, 0.0f, INFINITY, &best);
                }
            }
            if (inner) st.push(nd.d[4], to_visit_order(inner, oct) | (nimask << 8));
            if (st.empty()) {
                if (best.geom == TR_INVALID_GEOMETRY) best.t = INFINITY;
                pool.hit[slot] = make_float4(best.t, best.u, best.v, __uint_as_float(best.geom));
                pool.hit_prim[slot] = best.prim;
                active = false;
            }
        } else if (!LEAVES && active) {
            const uint32_t oct = r.octant();
            const bool prim_step = (lv & 0xffu) != 0;
            uint32_t count = 0, index;
            if (prim_step) {
                const uint64_t m = (uint64_t)lm0 | ((uint64_t)lm1 << 32);
                const uint32_t meta = (uint32_t)(m >> (8u * (ctz32(lv & 0xffu) ^ oct))) & 0xffu;
                count = meta >> 5;
                index = lpb + (meta & 31u) + (lv >> 8);
            } else {
                uint32_t base, masks;
                st.pop(&base, &masks);
                uint32_t vis = masks & 0xffu;
                const uint32_t imask = masks >> 8;
                const uint32_t i = ctz32(vis);
                vis &= vis - 1u;
                if (vis) st.push(base, vis | (imask << 8));
                index = base + popc8(imask & ((1u << (i ^ oct)) - 1u));
            }
            /* the fetch: the first three 16-byte loads are common to a
             * primitive and a node, a node takes five more (a skipped empty
             * leaf reads node 0, which always exists) */
            const uint4* src = prim_step && count ? reinterpret_cast<const uint4*>(sc.prims + index)
                                                  : reinterpret_cast<const uint4*>(sc.nodes + (prim_step ? 0u : index));
            uint4 q[8];
            q[0] = src[0];
            q[1] = src[1];
            q[2] = src[2];
            if (!prim_step) {
#pragma unroll
                for (int k = 3; k < 8; ++k) q[k] = src[k];
            }
            if (prim_step) {
                if (count) {
                    if (COUNT) ++np;
                    PrimRec pr;
                    pr.v0[0] = __uint_as_float(q[0].x); pr.v0[1] = __uint_as_float(q[0].y);
                    pr.v0[2] = __uint_as_float(q[0].z); pr.prim = q[0].w;
                    pr.v1[0] = __uint_as_float(q[1].x); pr.v1[1] = __uint_as_float(q[1].y);
                    pr.v1[2] = __uint_as_float(q[1].z); pr.geom = q[1].w;
                    pr.v2[0] = __uint_as_float(q[2].x); pr.v2[1] = __uint_as_float(q[2].y);
                    pr.v2[2] = __uint_as_float(q[2].z); pr.flags = q[2].w;
                    test_prim<SHAPES, false>(pr, sc.shapes, r, 0.0f, INFINITY, &best);
                }
                if ((lv >> 8) + 1u >= count) {
                    lv = (lv & 0xffu) & ((lv & 0xffu) - 1u);
                } else {
                    lv += 1u << 8;
                }
            } else {
                NodeRegs nd;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    nd.d[4 * k + 0] = q[k].x; nd.d[4 * k + 1] = q[k].y;
                    nd.d[4 * k + 2] = q[k].z; nd.d[4 * k + 3] = q[k].w;
                }
                if (COUNT) ++nv;
                const float tmax = best.geom == TR_INVALID_GEOMETRY ? INFINITY : best.t;
                const uint32_t hit = intersect_children(nd, r, 0.0f, tmax);
                const uint32_t nimask = node_imask(nd);
                const uint32_t inner = hit & nimask;
                lv = to_visit_order(hit & ~nimask, oct);
                lm0 = nd.d[6];
                lm1 = nd.d[7];
                lpb = nd.d[5];
                if (inner) st.push(nd.d[4], to_visit_order(inner, oct) | (nimask << 8));
            }
            if ((lv & 0xffu) == 0 && st.empty()) {
                if (best.geom == TR_INVALID_GEOMETRY) best.t = INFINITY;
                pool.hit[slot] = make_float4(best.t, best.u, best.v, __uint_as_float(best.geom));
                pool.hit_prim[slot] = best.prim;
                active = false;
            }
        }
    }
    if (COUNT) {
        atomicAdd(&counts[0], (unsigned long long)nv);
        atomicAdd(&counts[1], (unsigned long long)np);
        atomicAdd(&counts[2], (unsigned long long)nr);
    }
}

/* Closest hit of camera rays: a wave traces 64 consecutive queue entries
 * together, without refill.  The camera vertex's queue is the batch in
 * slot order, pixel-major, so the 64 rays of a wave are samples of one
 * or two pixels and visit nearly the same nodes: each node load serves
 * the whole wave from one or two cache lines.  Refill would mix the rays
 * of several pixels in a wave as they finish at different times. */
template <bool SHAPES>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kRefillWavesPerSimd))) void k_trace_packet(
    FrameParams p, PathPool pool, BatchOut out, uint64_t slot_base, int cur)
{
    __shared__ uint2 stk[kStackLds][kTraceBlock];
    const SceneView& sc = p.sc;
    const uint32_t n = pool.counters[cur];
    const uint32_t* queue = pool.queue[cur];
    /* entries addressed from the lane and a wave-uniform spill base (fewer
     * registers than per-lane pointers) */
    LaneStack st;
    st.lds = stk;
    st.spill = pool.spill + (size_t)blockIdx.x * kTraceBlock * kStackSpill;
    st.sp = 0;
    for (uint32_t base = blockIdx.x * kTraceBlock; base < n; base += gridDim.x * kTraceBlock) {
        const uint32_t qi = base + threadIdx.x;
        if (qi < n) {
            uint32_t slot;
            float org[3], dir[3];
            if (p.camera_fused) {
                /* the camera vertex's queue is the batch in slot order */
                slot = qi;
                uint32_t pixel, cs, split;
                uint64_t pseq;
                slot_coords(p, out, slot_base + slot, &pixel, &cs, &split, &pseq);
                V3 o, d;
                camera_ray(p, pixel, pixel, cs, p.width, p.height, p.spp, p.vpi, &o, &d);
                org[0] = o.x; org[1] = o.y; org[2] = o.z;
                dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
            } else {
                slot = queue[qi];
                const float4 o = pool.ray_o[slot], d = pool.ray_d[slot];
                org[0] = o.x; org[1] = o.y; org[2] = o.z;
                dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
            }
            const RayPre r = ray_setup(org, dir);
            st.sp = 0;
            const Hit h = traverse_closest(sc.nodes, sc.prims, SHAPES ? sc.shapes : nullptr, r, 0.0f, INFINITY, st,
                                           (uint32_t*)nullptr, (uint32_t*)nullptr);
            pool.hit[slot] = make_float4(h.t, h.u, h.v, __uint_as_float(h.geom));
            pool.hit_prim[slot] = h.prim;
        }
    }
}

/* Closest hit with delta tracking for scenes with volumes, with ray refill
 * (k_trace_refill's scheme: a lane whose ray has finished takes the wave's
 * next ray once kRefill lanes are idle; the queue is dealt in chunks of
 * kVolChunk rays).  A lane either searches its ray's closest hit from its
 * current near bound (a whole traversal; volume scenes hold few
 * primitives) or, inside a volume box, takes one pair of tentative
 * collisions per iteration, so lanes whose walk ends early take new rays
 * instead of waiting for the longest walk of a fixed group of 64.  Each
 * ray follows k_trace's sequence exactly: a walk that leaves the box
 * re-traces from beyond its exit, at most kMaxVolumeCrossings times. */
constexpr uint32_t kVolChunk = 32;   /* 16: 0.9 % slower on the volume frame, 64: 2.7 % */

template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_vol(FrameParams p, PathPool pool, BatchOut out,
                                                           uint64_t slot_base, uint32_t vertex, int cur,
                                                           unsigned long long* counts)
{
    __shared__ uint2 stk[kStackLds][kTraceBlock];
    const SceneView& sc = p.sc;
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    const uint32_t n = pool.counters[cur];
    const uint32_t* queue = pool.queue[cur];
    const uint32_t grid = gridDim.x;
    DevStack st;
    st.lds = &stk[0][lane];
    st.spill = pool.spill + (size_t)gtid * kStackSpill;
    st.sp = 0;
    uint32_t nv = 0, np = 0, nr = 0, ns = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    auto ray_index = [&](uint32_t k) -> uint32_t {
        return (blockIdx.x + (k / kVolChunk) * grid) * kVolChunk + (k % kVolChunk);
    };
    uint32_t next = 0;
    bool active = false, tracking = false;
    uint32_t slot = 0, step = 0, crossings = 0, vol = 0;
    float tnear = 0.0f, tc = 0.0f, t_exit = 0.0f;
    V3 org = mk(0.0f, 0.0f, 0.0f), dir = mk(0.0f, 0.0f, 0.0f);
    RayPre r;
    Rng rng;
    Hit h;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const bool more = ray_index(next) < n;
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (more && (n_idle >= kRefill || idle == ~0ull)) {
            if (!active) {
                const uint32_t qi = ray_index(next + (uint32_t)__popcll(idle & lt));
                if (qi < n) {
                    slot = queue[qi];
                    const float4 o = pool.ray_o[slot], d = pool.ray_d[slot];
                    org = ld4(o);
                    dir = ld4(d);
                    const float ao[3] = {o.x, o.y, o.z}, ad[3] = {d.x, d.y, d.z};
                    r = ray_setup(ao, ad);
                    uint32_t pixel, cs, split;
                    uint64_t pseq;
                    slot_coords(p, out, slot_base + slot, &pixel, &cs, &split, &pseq);
                    rng.k0 = p.key0; rng.k1 = p.key1; rng.pixel = pixel; rng.sample = cs; rng.split = split;
                    tnear = 0.0f;
                    step = 0;
                    crossings = 0;
                    tracking = false;
                    active = true;
                    ++nr;
                }
            }
            next += n_idle;
        }
        if (__ballot(active) == 0ull) {
            if (ray_index(next) < n) continue;
            break;
        }
        bool done = false;
        if (active && !tracking) {
            st.sp = 0;
            h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, tnear, INFINITY, st, COUNT ? &nv : nullptr,
                                 COUNT ? &np : nullptr);
            if (h.geom != TR_INVALID_GEOMETRY && sc.geoms[h.geom].type == GEOM_VOLUME) {
                /* the tracking loops of subdraw (ray-tracing.cc:2040-2128, 2212-2272) */
                tracking = true;
                vol = sc.geoms[h.geom].tri_count;
                tc = h.t;
                t_exit = h.u;
            } else {
                done = true;
            }
        }
        if (active && tracking) {
            /* an accepted collision hands its cell's fibre orientation to the
             * shading kernel in the hit's u, v and primitive words (a volume
             * box is primitive 0 of its geometry); NaN when the step budget
             * ended the walk without an accepted collision */
            V3 orient = mk(__uint_as_float(0x7fc00000u), 0.0f, 0.0f);
            const int res = delta_track_pair_any(sc.voxels, sc.volumes[vol], org, dir, &tc, t_exit, rng, vertex,
                                                 &step, COUNT ? &ns : nullptr, &orient);
            if (res > 0) {
                h.t = tc;
                h.u = orient.x;
                h.v = orient.y;
                h.prim = __float_as_uint(orient.z);
                done = true;
            } else if (res < 0) {
                tracking = false;
                if (++crossings >= kMaxVolumeCrossings) {
                    h.geom = h.prim = TR_INVALID_GEOMETRY;
                    h.t = INFINITY;
                    done = true;
                } else {
                    tnear = tc;
                }
            }
        }
        if (done) {
            pool.hit[slot] = make_float4(h.t, h.u, h.v, __uint_as_float(h.geom));
            pool.hit_prim[slot] = h.prim;
            active = false;
        }
    }
    if (COUNT) {
        atomicAdd(&counts[0], (unsigned long long)nv);
        atomicAdd(&counts[1], (unsigned long long)np);
        atomicAdd(&counts[2], (unsigned long long)nr);
        atomicAdd(&counts[6], (unsigned long long)ns);
    }
}

/* ---------------- shading ---------------- */

/* The texture view of k_shade: the scene's textures, or none. */
template <bool TEX> struct TexSel;
template <> struct TexSel<true> {
    typedef TexView type;
    __device__ static const TexView& get(const SceneView& sc) { return sc.tex; }
};
template <> struct TexSel<false> {
    typedef NoTex type;
    __device__ static NoTex get(const SceneView&) { return NoTex(); }
};

/* The debug outputs of RenderMode at the camera vertex (ray-tracing.cc:
 * 2078-2089): tangent, binormal or normal * 0.5 + 0.5, or max(0, N . -dir). */
__device__ inline V3 debug_color(uint32_t mode, V3 tangent, V3 binormal, V3 normal, V3 dir)
{
    if (mode == 4u) {
        const float c = fmax_(0.0f, dot(normal, neg(dir)));
        return mk(c, c, c);
    }
    const V3 v = mode == 1u ? tangent : (mode == 2u ? binormal : normal);
    return add(muls(v, 0.5f), mk(0.5f, 0.5f, 0.5f));
}

/* Material-coherent shading: a stable counting sort of a 256-thread
 * block's queue entries by the shading class of their hit (misses first,
 * then paths without a BSDF, then one BSDF family after another), in LDS,
 * so that the lanes of a wave run one BSDF's code.  Entries past the end of
 * the queue take the last class, so thread i of the block still holds an
 * entry iff i is below the block's entry count.  The order changes neither
 * a path's result nor the queue contents, only the queue order.  It costs
 * coalescing: a wave's slots then come from the block's whole range of
 * 256 instead of 64 consecutive queue entries. */
__device__ inline uint32_t sort_by_class(const SceneView& sc, const PathPool& pool, uint32_t slot, bool active,
                                         float4* hit)
{
    constexpr uint32_t kWaves = 256 / 64;
    __shared__ uint32_t s_off[kShadeClasses * kWaves];   /* [class][wave]: counts, then offsets */
    __shared__ uint32_t s_slot[256];
    __shared__ float4 s_hit[256];                        /* the hit record moves with its slot */
    uint32_t key = SC_NONE;
    float4 h = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (active) {
        h = pool.hit[slot];
        const uint32_t geom = __float_as_uint(h.w);
        key = geom == TR_INVALID_GEOMETRY ? (uint32_t)SC_MISS : sc.geoms[geom].shade_class;
    }
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t rank = 0;
#pragma unroll
    for (uint32_t k = 0; k < kShadeClasses; ++k) {
        const uint64_t m = __ballot(key == k);
        if (key == k) rank = (uint32_t)__popcll(m & lt);
        if (lane == 0) s_off[k * kWaves + w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (w == 0) {
        /* exclusive scan of the 36 counts in class-major, wave-minor order */
        const uint32_t c = lane < kShadeClasses * kWaves ? s_off[lane] : 0u;
        uint32_t v = c;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t u = __shfl_up(v, d, 64);
            if (lane >= d) v += u;
        }
        if (lane < kShadeClasses * kWaves) s_off[lane] = v - c;
    }
    __syncthreads();
    const uint32_t pos = s_off[key * kWaves + w] + rank;
    s_slot[pos] = slot;
    s_hit[pos] = h;
    __syncthreads();
    *hit = s_hit[threadIdx.x];
    return s_slot[threadIdx.x];
}

/* Shades queue entries seg * kSegment .. + kSegment - 1 (those below n),
 * one per thread of the block, and writes segment seg of the next queues. */
template <bool VOL, bool CACHE, bool TEX, bool LIGHTS>
__device__ __forceinline__ void shade_segment(const FrameParams& p, const PathPool& pool, const BatchOut& out,
                                              uint64_t slot_base, uint32_t vertex, int cur, const CacheRecords& rec,
                                              uint32_t seg, uint32_t n)
{
    uint32_t qi = seg * kSegment + threadIdx.x;
    bool active = qi < n;
    bool extend = false;
    bool shadow = false;
    bool cache_query = false;
    float qin[kMlpIn];
    V3 qthr = mk(0.0f, 0.0f, 0.0f);
    /* the camera vertex's queue is the batch in slot order */
    const bool fused = vertex == 0 && p.camera_fused;
    uint32_t slot = active ? (fused ? qi : pool.queue[cur][qi]) : 0u;
    float4 sorted_hit;
    if (p.shade_sort) slot = sort_by_class(p.sc, pool, slot, active, &sorted_hit);
    if (active) {
        const SceneView& sc = p.sc;
        const typename TexSel<TEX>::type tv = TexSel<TEX>::get(sc);
        uint32_t pixel, cs, split;
        uint64_t pseq;
        uint32_t req = 0;
        if (p.req.n) {
            uint32_t lp;
            req = pool.slot_req[slot];
            request_coords(p, req, slot, &pixel, &lp, &cs, &split);
            pseq = lp;
        } else {
            slot_coords(p, out, slot_base + slot, &pixel, &cs, &split, &pseq);
        }
        Rng rng;
        rng.k0 = p.key0; rng.k1 = p.key1; rng.pixel = pixel; rng.sample = cs; rng.split = split;
        const bool unbiased = p.estimator == TR_ESTIMATOR_UNBIASED;
        /* Scenes without light sources and without the cache write a path's
         * radiance of later vertices only where a vertex adds to it, and
         * zero when the path ends with none; the ray direction's .w carries
         * "written" from vertex to vertex (the camera vertex writes none). */
        constexpr bool kLazyGlobal = !CACHE && !LIGHTS;
        bool has_global = false;
        V3 org, dir;
        if (fused) {
            camera_ray(p, pixel, pixel, cs, p.width, p.height, p.spp, p.vpi, &org, &dir);
        } else {
            org = ld4(pool.ray_o[slot]);
            const float4 d4 = pool.ray_d[slot];
            dir = ld4(d4);
            has_global = kLazyGlobal && d4.w != 0.0f;
        }
        float4 h4 = p.shade_sort ? sorted_hit : pool.hit[slot];
        const float4 t4 = vertex == 0 ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : pool.thr[slot];
        V3 T = ld4(t4);
        float bal = t4.w;
        uint32_t geom = __float_as_uint(h4.w);
        /* The hit-sequence hash is kept only for a batch that outputs it. */
        const bool want_hash = out.path_hash != nullptr;
        uint32_t hash = vertex == 0 || !want_hash ? kFnvBasis : pool.hash[slot];
        V3 local = mk(0.0f, 0.0f, 0.0f);
        /* The radiance of later vertices is read and written only by the
         * vertices that add to it (a miss, an emitter) and by the camera
         * vertex, which clears it; the cache mode reads it at every vertex. */
        V3 global = CACHE && vertex != 0 ? ld4(pool.global[slot]) : mk(0.0f, 0.0f, 0.0f);
        bool write_global = CACHE || (vertex == 0 && !kLazyGlobal);
        /* the camera vertex writes the slot's camera-vertex radiance, zero or not */
        bool write_local = vertex == 0;
        uint32_t pf = 0;
        if (CACHE) {
            pf = pool.pflags[slot];
            if (pf & PF_BEFORE) {
                /* the vertex after a training path's query vertex: the
                 * radiance so far includes the query vertex's next-event
                 * estimate, the part the network does not predict */
                uint32_t r = train_record(p, pseq, cs);
                uint32_t st = rec.smp.stride;
                rec.before[r] = global.x;
                rec.before[st + r] = global.y;
                rec.before[2 * st + r] = global.z;
                pf &= ~PF_BEFORE;
                pool.pflags[slot] = pf;
            }
        }
        if (geom == TR_INVALID_GEOMETRY) {
            hash = (hash ^ 0xffffffffu) * kFnvPrime;
            /* m_GlobalCubeMap->sampleSpectrum(ray direction) or the
             * background spectrum (ray-tracing.cc:2060, 2224) */
            V3 env = (TEX && sc.tex.cube_side) ? cube_sample(sc.tex, dir) : ld3(p.background);
            if (vertex == 0) {
                local = env;
                write_local = true;
            } else {
                if (!CACHE && (!kLazyGlobal || has_global)) global = ld4(pool.global[slot]);
                global = add(global, mulv(env, T));
                write_global = true;
            }
        } else if (sc.geoms[geom].material == kNoMaterial) {
            /* a blocker: the path ends with nothing (ray-tracing.cc:2072-2077,
             * 2236-2240); its sample() frame is zero (:1239-1244) */
            /* a volume hit's primitive word carries its fibre orientation;
             * a volume box is primitive 0 */
            if (want_hash)
                hash = fnv_hit(hash, geom, VOL && sc.geoms[geom].type == GEOM_VOLUME ? 0u : pool.hit_prim[slot]);
            if (vertex == 0 && p.render_mode) {
                const V3 z = mk(0.0f, 0.0f, 0.0f);
                local = debug_color(p.render_mode, z, z, z, dir);
            }
        } else {
            uint32_t prim = pool.hit_prim[slot];
            Sample sd;
            sd.dir_density = 0.0f;
            V3 pos;
            const GeomRec& G = sc.geoms[geom];
            bool volume = VOL && G.type == GEOM_VOLUME;
            hash = fnv_hit(hash, geom, volume ? 0u : prim);
            if (volume) {
                /* RTVolumeSet::sampleScattering's frame and DirectionalDensity
                 * at the accepted collision (ray-tracing.cc:327-337): the
                 * closest-hit kernel's acceptance test computed the cell's
                 * fibre orientation and passed it in (u, v, primitive); NaN
                 * when the walk's step budget ended it */
                const VolumeRec& V = sc.volumes[G.tri_count];
                pos = ray_point(org, dir, h4.x);
                M3 b;
                if (h4.y == h4.y) b = make_basis_tangent(mk(h4.y, h4.z, __uint_as_float(prim)));
                else b = volume_frame_any(fetch_cell(sc.voxels, V, pos), sc.voxels, V);
                sd.mat = &sc.mats[G.material];
                sd.tangent = b.c0;
                sd.binormal = b.c1;
                sd.normal = b.c2;
                sd.wi = dir;
                sd.tcx = sd.tcy = 0.0f;
                sd.dir_density = sggx_projected_area(ld3(V.stddev), b, dir);
                /* throughput *= albedo; a path below TransmittanceThreshold
                 * ends (ray-tracing.cc:2136-2142, 2277-2284) */
                T = mulv(T, ld3(sd.mat->albedo));
            } else {
                sample_geometry<TEX>(sc, geom, prim, h4.x, h4.y, h4.z, org, dir, &sd);
                pos = add(org, smul(h4.x, dir));
                sd.wi = dir;
                material_cache(&sd, tv);
            }
            if (!VOL && vertex == 0 && p.render_mode) {
                /* RenderMode debug output instead of shading (the camera ray ends) */
                local = debug_color(p.render_mode, sd.tangent, sd.binormal, sd.normal, dir);
            } else if (volume && T.x < 1e-3f && T.y < 1e-3f && T.z < 1e-3f) {
                /* absorbed: no contribution from this vertex on */
            } else if (is_emissive(sd.mat->model)) {
                V3 le = ld3(sd.mat->diffuse);
                if (vertex == 0) {
                    local = mulv(T, le);
                    write_local = true;
                } else {
                    float w = bal;
                    if (unbiased) {
                        /* balance heuristic of the BSDF strategy against the
                         * light that owns the hit point; bal holds ss * pdf */
                        float lp = hit_light_pdf(sc, geom, prim, org, dir, h4.x);
                        float ls_prev = vertex == 1 ? (float)p.n_light : 1.0f;
                        w = bal / (bal + ls_prev * lp);
                    }
                    if (!CACHE && (!kLazyGlobal || has_global)) global = ld4(pool.global[slot]);
                    global = add(global, mulv(muls(T, w), le));
                    write_global = true;
                }
            } else {
                sd.wo = neg(dir);
                /* radiance cache: the vertex after a path's first rough one
                 * (below the last bounce) is its query vertex, where the
                 * network predicts the radiance leaving it (with cache_nee:
                 * the radiance beyond the vertex's own next-event estimate)
                 * and the path ends; a training path records it instead
                 * and goes on */
                if (CACHE) {
                    if ((pf & PF_ROUGH) && vertex > 0 && vertex < p.max_depth) {
                        float in[kMlpIn];
                        cache_inputs(p, pos, dir, sd, in);
                        if (!(pf & PF_TRAIN)) {
                            cache_query = true;
#pragma unroll
                            for (int k = 0; k < kMlpIn; ++k) qin[k] = in[k];
                            qthr = T;
                        } else {
                            uint32_t r = train_record(p, pseq, cs);
                            uint32_t st = rec.smp.stride;
#pragma unroll
                            for (int k = 0; k < kMlpIn; ++k) rec.smp.in[(size_t)k * st + r] = in[k];
                            rec.thr[r] = T.x;
                            rec.thr[st + r] = T.y;
                            rec.thr[2 * st + r] = T.z;
                            rec.smp.valid[r] = 1.0f;
                            pf = (pf & ~PF_ROUGH) | PF_RECORDED;
                            if (p.cache_nee) {
                                pf |= PF_BEFORE;
                            } else {
                                rec.before[r] = global.x;
                                rec.before[st + r] = global.y;
                                rec.before[2 * st + r] = global.z;
                            }
                            pool.pflags[slot] = pf;
                        }
                    }
                }
                uint32_t ls = vertex == 0 ? p.n_light : 1u;
                uint32_t ss = vertex == 0 ? p.n_gi : 1u;
                bool apply_balance = vertex == 0 ? (p.max_depth != 0 && !volume) : (vertex != p.max_depth);
                /* next-event estimation: the camera vertex's is evaluated once, by split 0;
                 * a query vertex has one only with cache_nee */
                if ((vertex != 0 || split == 0) && (!cache_query || p.cache_nee)) {
                    uint32_t ne = 0;
                    for (uint32_t li = 0; li < (LIGHTS ? sc.n_lights : 0u); ++li) {
                        const LightRec& L = sc.lights[li];
                        /* request mode: the request's direction of light 0 */
                        const float* dir_override = p.req.n && li == 0 ? p.req.reqs[req].light_dir : nullptr;
                        bool multi = multi_sample_light(L, unbiased);
                        uint32_t samples = multi ? ls : 1u;
                        for (uint32_t k = 0; k < samples; ++k) {
                            U4 xr = rng.draw(vertex, STREAM_LIGHT, (li << 16) | k);
                            LightSample smp = sample_light(sc, L, pos, xr.a, xr.b, xr.c, unbiased, dir_override);
                            if (smp.pdf == 0.0f) continue;
                            Sample lsd = sd;
                            lsd.wi = smp.dir;
                            V3 curT = T;
                            if (apply_balance && multi) {
                                float balance;
                                if (unbiased) {
                                    float a = (float)samples * smp.pdf;
                                    balance = a / (a + (float)ss * bsdf_pdf(lsd, tv));
                                } else {
                                    float total = total_balance(sc, (float)ls, (float)ss, lsd, pos, tv, LIGHTS);
                                    balance = (light_pdf(sc, L, pos, lsd.wi) * (float)samples) / total;
                                }
                                curT = muls(curT, balance);
                            }
                            V3 contrib = divs(mulv(mulv(smp.radiance, curT), bsdf_eval(lsd, tv)), smp.pdf);
                            size_t e = (size_t)ne * pool.capacity + slot;
                            pool.sh_o[e] = st4(pos, smp.dist);
                            pool.sh_d[e] = st4(smp.dir, __uint_as_float(smp.geom));
                            pool.sh_c[e] = st4(contrib, __uint_as_float((li << 16) | k));
                            ++ne;
                        }
                    }
                    pool.sh_n[slot] = ne;
                    if (VOL) pool.sh_tmax[slot] = max_value(T);
                    shadow = ne > 0;
                }
                bool alive = !cache_query;
                if (vertex > 0) {
                    T = muls(T, (unbiased ? 1.0f : bal) / p.rr);
                    if (p.rr < 1.0f) {
                        U4 xq = rng.draw(vertex, STREAM_RR, 0u);
                        if (u01(xq.a) > p.rr) alive = false;
                    }
                }
                if (alive && vertex < p.max_depth) {
                    Strata sta;
                    if (vertex == 0) {
                        sta.x = (float)(split % p.rays_sq);
                        sta.y = (float)(split / p.rays_sq);
                        sta.tx = sta.ty = (float)p.rays_sq;
                    } else {
                        sta.x = sta.y = 0.0f;
                        sta.tx = sta.ty = 1.0f;
                    }
                    U4 xb = rng.draw(vertex, STREAM_BSDF, 0u);
                    bsdf_sample(sta, &sd, u01(xb.a), u01(xb.b), u01(xb.c), tv, rng, vertex);
                    if (sd.pdf != 0.0f) {
                        /* volume vertices keep balance 1 and the throughput
                         * (phase / pdf = 1, ray-tracing.cc:2182-2201) */
                        float nb = 1.0f;
                        if (!volume) {
                            if (unbiased) {
                                nb = (float)ss * sd.pdf;
                            } else {
                                float total = total_balance(sc, (float)ls, (float)ss, sd, pos, tv, LIGHTS);
                                nb = total > 0.0f ? ((float)ss * sd.pdf) / total : 0.0f;
                            }
                            T = mulv(T, divs(bsdf_eval(sd, tv), sd.pdf));
                        }
                        pool.ray_o[slot] = st4(pos, 0.0f);
                        pool.ray_d[slot] = st4(sd.wi, kLazyGlobal && (has_global || write_global) ? 1.0f : 0.0f);
                        pool.thr[slot] = st4(T, nb);
                        extend = true;
                        if (CACHE && !(pf & PF_RECORDED) && cache_roughness(sd.mat) >= p.rough_alpha)
                            pool.pflags[slot] = pf | PF_ROUGH;
                    }
                }
            }
        }
        if (want_hash) pool.hash[slot] = hash;
        /* a path that ends with nothing written beyond its camera vertex: zero */
        if (kLazyGlobal && !extend && !has_global) write_global = true;
        if (write_global) pool.global[slot] = st4(global, 0.0f);
        /* .w: the camera ray hit geometry (TransparentBackground) */
        if (write_local) pool.local[slot] = st4(local, geom != TR_INVALID_GEOMETRY ? 1.0f : 0.0f);
    }
    const uint32_t qpos = segment_push(pool, seg, extend, shadow, cache_query, slot);
    if (CACHE && cache_query) {
        const MlpQueries& q = rec.queries;
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) q.in[(size_t)k * q.stride + qpos] = qin[k];
        q.thr[qpos] = qthr.x;
        q.thr[q.stride + qpos] = qthr.y;
        q.thr[2 * (size_t)q.stride + qpos] = qthr.z;
        q.slot[qpos] = slot;
    }
}

/* One block per segment of the largest possible queue (the queue size is
 * known only on the device); blocks past the queue's end return at once
 * and write no segment.
 * LIGHTS false is the instance for scenes without light sources (no
 * next-event estimation, no light terms in the balance heuristic): without
 * that code the kernel holds under 80 VGPRs, 6 waves per SIMD, which the
 * latency-bound shading of whole-frame launches needs.  The instance with
 * light sources (no cache, no textures) is held to 5 waves per SIMD (96
 * VGPRs, a 28-byte spill; 2.5 % faster than 4 waves on the
 * gonioreflectometer requests and the box frame); the cache instance to 4
 * waves (128 VGPRs, no spill; cached box frames 3 % faster than at 3
 * waves; it spills 104 bytes at 5); the texture instances keep their
 * registers (192 bytes of spills at 5 waves). */
template <bool VOL, bool CACHE, bool TEX, bool LIGHTS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LIGHTS ? (TEX ? 1 : (CACHE ? 4 : 5)) : 6))) void k_shade(
    FrameParams p, PathPool pool, BatchOut out, uint64_t slot_base, uint32_t vertex, int cur, int nxt,
    CacheRecords rec)
{
    (void)nxt;
    const uint32_t n = pool.counters[cur];
    if (blockIdx.x * kSegment >= n) return;
    shade_segment<VOL, CACHE, TEX, LIGHTS>(p, pool, out, slot_base, vertex, cur, rec, blockIdx.x, n);
}

/* ---------------- request mode: accumulation and outputs ---------------- */

/* threads of an accumulation block: one wave */
constexpr uint32_t kAccBlock = 64;

/* Per-pixel sums of the current batch (pixels b.pix_base ..), in sample
 * order, each added once to the pixel's FP64 sum: batches end at pixel
 * boundaries, so the result does not depend on the batch size. */
__global__ __launch_bounds__(kAccBlock) void k_accumulate_req(FrameParams p, PathPool pool, double* accum,
                                                              const uint32_t* cursor)
{
    __shared__ float4 s_l[8][kAccBlock + 1], s_g[8][kAccBlock + 1];
    __shared__ uint32_t s_first_of[kAccBlock], s_spp[kAccBlock];
    const ReqBatch& b = p.req.batches[*cursor];
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kAccBlock + lane;
    if (blockIdx.x * kAccBlock >= b.n_pix) return;
    const bool valid = i < b.n_pix;
    const uint64_t gp = b.pix_base + i;
    uint32_t spp = 0, s_first = 0;
    if (valid) {
        const uint32_t r = find_request(p.req.pix_off, p.req.n, gp);
        spp = p.req.reqs[r].spp;
        const uint64_t per_pixel = (uint64_t)spp * p.n_gi;
        s_first = (uint32_t)(p.req.path_off[r] + (gp - p.req.pix_off[r]) * per_pixel - b.path_base);
    }
    const float rcp_gi = 1.0f / (float)p.n_gi;
    double sum[3] = {0.0, 0.0, 0.0};
    if (p.n_gi == 1) {
        /* k_accumulate's staging through LDS, with each pixel's own first
         * slot and sample count (requests differ in both) */
        s_first_of[lane] = s_first;
        s_spp[lane] = spp;
        uint32_t max_spp = spp;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) max_spp = max(max_spp, (uint32_t)__shfl_xor((int)max_spp, d));
        __syncthreads();
        const uint32_t sub = lane & 7u, row = lane >> 3;
        for (uint32_t c = 0; c < max_spp; c += 8) {
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t q = row + 8 * k;
                if (c + sub < s_spp[q]) {
                    const uint32_t sl = s_first_of[q] + c + sub;
                    s_g[sub][q] = pool.global[sl];
                    s_l[sub][q] = pool.local[sl];
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
                if (c + u < spp) {
                    V3 v = add(ld4(s_l[u][lane]), muls(ld4(s_g[u][lane]), rcp_gi));
                    sum[0] += (double)v.x;
                    sum[1] += (double)v.y;
                    sum[2] += (double)v.z;
                }
            }
            __syncthreads();
        }
    } else {
        for (uint32_t c = 0; c < spp; ++c) {
            const uint32_t s0 = s_first + c * p.n_gi;
            V3 g = ld4(pool.global[s0]);
            for (uint32_t j = 1; j < p.n_gi; ++j) g = add(g, ld4(pool.global[s0 + j]));
            V3 v = add(ld4(pool.local[s0]), muls(g, rcp_gi));
            sum[0] += (double)v.x;
            sum[1] += (double)v.y;
            sum[2] += (double)v.z;
        }
    }
    if (!valid) return;
    accum[gp * 3 + 0] += sum[0];
    accum[gp * 3 + 1] += sum[1];
    accum[gp * 3 + 2] += sum[2];
}

/* Moves the request cursor to the next batch (after the batch's last
 * kernel, in stream order). */
__global__ void k_req_next(uint32_t* cursor)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) cursor[0] = cursor[0] + 1u;
}

/* Mean RGB of request blockIdx.x over its pixels and camera samples: the
 * block sums its pixels in a fixed order (strided per thread, then a tree). */
__global__ __launch_bounds__(256) void k_req_mean(ReqView rv, const double* accum, float* mean)
{
    __shared__ double red[3][256];
    const uint32_t r = blockIdx.x;
    const uint64_t p0 = rv.pix_off[r], p1 = rv.pix_off[r + 1];
    double s[3] = {0.0, 0.0, 0.0};
    for (uint64_t q = p0 + threadIdx.x; q < p1; q += blockDim.x)
        for (int k = 0; k < 3; ++k) s[k] += accum[q * 3 + k];
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = (double)(p1 - p0) * (double)rv.reqs[r].spp;
        for (int k = 0; k < 3; ++k) mean[r * 3 + k] = (float)(red[k][0] / n);
    }
}

/* Per-pixel mean radiance of every request, in request order. */
__global__ __launch_bounds__(256) void k_req_image(ReqView rv, const double* accum, float* image, uint64_t n_pix)
{
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_pix) return;
    const uint32_t r = find_request(rv.pix_off, rv.n, q);
    const double rcp = 1.0 / (double)rv.reqs[r].spp;
    for (int k = 0; k < 3; ++k) image[q * 3 + k] = (float)(accum[q * 3 + k] * rcp);
}

/* ---------------- queue compaction ---------------- */

/* Exclusive prefix sum over the block (blockDim.x == kSegment) of one value
 * per thread; *total receives the block's sum. */
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t ws[kSegment / 64];
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSegment / 64; ++j) {
        if (j < w) before += ws[j];
        sum += ws[j];
    }
    __syncthreads();
    *total = sum;
    return before + x - v;
}

/* The shading launch wrote segments 0 .. n_seg - 1, one per kSegment
 * entries of path queue cur; they fall into groups of `group` segments,
 * one per block of the compaction grid (whose x size is the same in the
 * three kernels), at least 16 segments to a group. */
__device__ inline void compact_groups(const PathPool& pool, int cur, uint32_t* n_seg, uint32_t* group)
{
    const uint32_t n = pool.counters[cur];
    *n_seg = (n + kSegment - 1) / kSegment;
    *group = max(16u, (*n_seg + gridDim.x - 1) / gridDim.x);
}

/* Block (x, k) sums queue k's segment counts of group x (segments
 * [x * group, (x + 1) * group)), in a fixed order. */
__global__ __launch_bounds__(kSegment) void k_compact_sums(PathPool pool, int cur, uint32_t* group_sum)
{
    const uint32_t k = blockIdx.y;
    const uint32_t* cnt = pool.seg_n + (size_t)k * pool.seg_cap;
    uint32_t n_seg, group;
    compact_groups(pool, cur, &n_seg, &group);
    const uint32_t s0 = blockIdx.x * group;
    const uint32_t s1 = min(s0 + group, n_seg);
    uint32_t part = 0;
    for (uint32_t j = s0 + threadIdx.x; j < s1; j += kSegment) part += cnt[j];
    uint32_t total;
    (void)block_exclusive_scan(part, &total);
    if (threadIdx.x == 0) group_sum[k * gridDim.x + blockIdx.x] = total;
}

/* Exclusive scan of queue k's group sums (block k, up to kCompactGroups
 * of them), in place: group x's sum becomes the number of entries in the
 * groups before it; the queue size goes to its counter. */
__global__ __launch_bounds__(kSegment) void k_compact_scan(PathPool pool, uint32_t n_groups, int nxt,
                                                           uint32_t* group_sum)
{
    const uint32_t k = blockIdx.x;
    uint32_t* gs = group_sum + k * n_groups;
    constexpr uint32_t kPer = kCompactGroups / kSegment;
    uint32_t v[kPer];
    uint32_t part = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t g = threadIdx.x * kPer + j;
        v[j] = g < n_groups ? gs[g] : 0u;
        part += v[j];
    }
    uint32_t total;
    uint32_t before = block_exclusive_scan(part, &total);
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t g = threadIdx.x * kPer + j;
        if (g < n_groups) gs[g] = before;
        before += v[j];
    }
    if (threadIdx.x == 0) {
        const uint32_t slot = k == 0 ? (uint32_t)nxt : k + 1;
        pool.counters[slot] = total;
    }
}

/* Block (x, k) copies queue k's segments of group x to the dense queue at
 * the offset k_compact_scan left in the group's sum. */
__global__ __launch_bounds__(kSegment) void k_compact(PathPool pool, int cur, int nxt, const uint32_t* group_off)
{
    const uint32_t k = blockIdx.y;
    uint32_t n_seg, group;
    compact_groups(pool, cur, &n_seg, &group);
    const uint32_t* cnt = pool.seg_n + (size_t)k * pool.seg_cap;
    const uint32_t* src = pool.seg + (size_t)k * pool.seg_cap * kSegment;
    uint32_t* dst = k == 0 ? pool.queue[nxt] : (k == 1 ? pool.shadow_queue : pool.query_queue);
    const uint32_t s0 = blockIdx.x * group;
    const uint32_t s1 = min(s0 + group, n_seg);
    uint32_t base = group_off[k * gridDim.x + blockIdx.x];
    for (uint32_t c0 = s0; c0 < s1; c0 += kSegment) {
        __shared__ uint32_t sc[kSegment], so[kSegment];
        uint32_t c = c0 + threadIdx.x < s1 ? cnt[c0 + threadIdx.x] : 0u;
        uint32_t sum;
        uint32_t o = block_exclusive_scan(c, &sum);
        sc[threadIdx.x] = c;
        so[threadIdx.x] = o;
        __syncthreads();
        /* thread t moves entry t of each segment; the loads of 16
         * segments are issued before their stores */
        const uint32_t nb = min(s1 - c0, kSegment);
        for (uint32_t j0 = 0; j0 < nb; j0 += 16) {
            uint32_t v[16];
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const uint32_t j = j0 + u;
                if (j < nb && threadIdx.x < sc[j]) v[u] = src[(size_t)(c0 + j) * kSegment + threadIdx.x];
            }
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const uint32_t j = j0 + u;
                if (j < nb && threadIdx.x < sc[j]) dst[base + so[j] + threadIdx.x] = v[u];
            }
        }
        __syncthreads();
        base += sum;
    }
}

/* ---------------- radiance cache: training targets ---------------- */

/* Target of a training record: the radiance its path gathered beyond the
 * query vertex's next-event estimate, divided by the throughput at the
 * query vertex (the quantity a query supplies); 0 when the path ended
 * there.  Records whose path never reached a query vertex, or whose
 * throughput is zero in a channel, get weight 0. */
__global__ __launch_bounds__(256) void k_cache_targets(PathPool pool, CacheRecords rec, uint32_t r0, uint32_t r1)
{
    uint32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    uint32_t st = rec.smp.stride;
    if (rec.smp.valid[r] == 0.0f) {
        rec.smp.target[r] = rec.smp.target[st + r] = rec.smp.target[2 * st + r] = 0.0f;
        return;
    }
    uint32_t slot = rec.slot[r];
    const bool ended = (pool.pflags[slot] & PF_BEFORE) != 0;
    float4 g = pool.global[slot];
    float gc[3] = {g.x, g.y, g.z};
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float t = rec.thr[(size_t)c * st + r];
        ok = ok && t > 1e-8f;
        float v = 0.0f;
        if (ok && !ended) v = (gc[c] - rec.before[(size_t)c * st + r]) / t;
        rec.smp.target[(size_t)c * st + r] = v;
    }
    if (!ok) rec.smp.valid[r] = 0.0f;
}

/* ---------------- shadow rays (persistent) ---------------- */

/* WIDE: the instance for persistent grids above 16 waves per CU, held to 6
 * waves per SIMD (80 VGPRs, a 48-byte spill) so that all of them stay
 * resident; at 16 waves per CU the instance without the spill (97 VGPRs,
 * 4 waves per SIMD) is 4 % faster, at 24 the WIDE one 3.5 %. */
template <bool COUNT, bool VOL, bool WIDE = false>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(WIDE ? 6 : 1))) void k_shadow(FrameParams p, PathPool pool, BatchOut out, uint64_t slot_base,
                                                        uint32_t vertex, unsigned long long* counts)
{
    __shared__ uint2 stk[kStackLds][kTraceBlock];
    uint32_t lane = threadIdx.x;
    uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    const uint32_t n = pool.counters[2];
    const SceneView& sc = p.sc;
    DevStack st;
    st.lds = &stk[0][lane];
    st.spill = pool.spill + (size_t)gtid * kStackSpill;
    uint32_t nv = 0, np = 0, nr = 0, ns = 0;
    const uint32_t stride = gridDim.x * kTraceBlock;
    for (uint32_t base = blockIdx.x * kTraceBlock; base < n; base += stride) {
        uint32_t qi = base + lane;
        if (qi < n) {
            uint32_t slot = pool.shadow_queue[qi];
            uint32_t ne = pool.sh_n[slot];
            Rng rng;
            float tmax_T = 0.0f;
            if (VOL) {
                uint32_t pixel, cs, split;
                uint64_t pseq;
                slot_coords(p, out, slot_base + slot, &pixel, &cs, &split, &pseq);
                rng.k0 = p.key0; rng.k1 = p.key1; rng.pixel = pixel; rng.sample = cs; rng.split = split;
                tmax_T = pool.sh_tmax[slot];
            }
            uint32_t ls = vertex == 0 ? p.n_light : 1u;
            V3 radiance = mk(0.0f, 0.0f, 0.0f);
            uint32_t e = 0;
            for (uint32_t li = 0; li < sc.n_lights; ++li) {
                uint32_t samples = multi_sample_light(sc.lights[li], p.estimator == TR_ESTIMATOR_UNBIASED) ? ls : 1u;
                V3 color = mk(0.0f, 0.0f, 0.0f);
                while (e < ne) {
                    size_t idx = (size_t)e * pool.capacity + slot;
                    float4 c4 = pool.sh_c[idx];
                    uint32_t tag = __float_as_uint(c4.w);
                    if ((tag >> 16) != li) break;
                    float4 o4 = pool.sh_o[idx], d4 = pool.sh_d[idx];
                    float org[3] = {o4.x, o4.y, o4.z}, dir[3] = {d4.x, d4.y, d4.z};
                    RayPre r = ray_setup(org, dir);
                    uint32_t light_geom = __float_as_uint(d4.w);
                    st.sp = 0;
                    Hit h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, 0.0f, o4.w, st, COUNT ? &nv : nullptr,
                                             COUNT ? &np : nullptr);
                    if (!VOL) {
                        if (h.geom == TR_INVALID_GEOMETRY || h.geom == light_geom) color = add(color, ld4(c4));
                    } else {
                        /* EvaluateLightSources' occlusion loop with Woodcock
                         * transmittance (ray-tracing.cc:1845-1926) */
                        float dist = o4.w, tf = 1.0f, step_through = 0.0f;
                        uint32_t step = 0, crossings = 0;
                        uint32_t sub_hi = ((tag >> 16) << 24) | ((tag & 0xffu) << 16);
                        bool reached = false;
                        for (;;) {
                            if (h.geom == TR_INVALID_GEOMETRY || h.geom == light_geom) {
                                reached = true;
                                break;
                            }
                            const GeomRec& G = sc.geoms[h.geom];
                            if (G.type != GEOM_VOLUME || ++crossings > kMaxVolumeCrossings) break;
                            const VolumeRec& V = sc.volumes[G.tri_count];
                            uint32_t succ = woodcock_any(sc.voxels, V, ld4(o4), ld4(d4), h.t, h.u, rng, vertex, sub_hi,
                                                     &step, &step_through, COUNT ? &ns : nullptr);
                            tf = tf * ((float)succ / 2.0f);
                            if (tmax_T * tf < 1e-3f) break;
                            if (step_through >= dist - 1e-3f) {
                                reached = true;
                                break;
                            }
                            float tn = dist - 1e-3f < step_through ? dist - 1e-3f : step_through;
                            st.sp = 0;
                            h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, tn, dist, st, COUNT ? &nv : nullptr,
                                                 COUNT ? &np : nullptr);
                        }
                        if (reached) color = add(color, muls(ld4(c4), tf));
                    }
                    ++e;
                    ++nr;
                }
                radiance = add(radiance, divs(color, (float)samples));
            }
            if (vertex == 0) {
                /* .w = 1: the camera ray hit geometry (it has shadow rays) */
                pool.local[slot] = st4(radiance, 1.0f);
            } else {
                float4 g = pool.global[slot];
                pool.global[slot] = st4(add(ld4(g), radiance), 0.0f);
            }
        }
    }
    if (COUNT) {
        atomicAdd(&counts[4], (unsigned long long)nv);
        atomicAdd(&counts[5], (unsigned long long)np);
        atomicAdd(&counts[3], (unsigned long long)nr);
        if (VOL) atomicAdd(&counts[6], (unsigned long long)ns);
    }
}

/* ---------------- accumulation ---------------- */

/* One wave per 64 pixels (pixel i = blockIdx.x * 64 + lane).  With one GI
 * split and no per-path output, the wave stages 8 camera samples of its 64
 * pixels at a time through LDS: 8 lanes read one pixel's 8 consecutive
 * slots (one 128-byte line of each array), so every load instruction
 * covers whole lines, instead of each lane reading 16 bytes of 64 lines
 * per instruction.  Each lane then adds its pixel's samples in sample
 * order, the order of the per-lane loop below. */
__global__ __launch_bounds__(kAccBlock) void k_accumulate(FrameParams p, PathPool pool, BatchOut out, uint64_t slot_base,
                                                          uint32_t n_pixels)
{
    __shared__ float4 s_l[8][kAccBlock + 1], s_g[8][kAccBlock + 1];
    const uint32_t lane = threadIdx.x;
    const uint32_t i0 = blockIdx.x * kAccBlock;
    const uint32_t i = i0 + lane;
    const bool valid = i < n_pixels;
    uint64_t per_pixel = (uint64_t)p.spp * p.n_gi;
    uint64_t first_slot = slot_base + (uint64_t)i * per_pixel;
    uint64_t pseq = first_slot / per_pixel;
    double sum[3] = {0.0, 0.0, 0.0};
    float rcp_gi = 1.0f / (float)p.n_gi;
    uint32_t c = 0;
    bool any_hit = false;
    if (p.n_gi == 1 && !out.path_rgb) {
        const uint32_t sub = lane & 7u, row = lane >> 3;
        for (; c + 8 <= p.spp; c += 8) {
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t q = row + 8 * k;
                if (i0 + q < n_pixels) {
                    const uint32_t sl = (i0 + q) * p.spp + c + sub;
                    s_g[sub][q] = pool.global[sl];
                    s_l[sub][q] = pool.local[sl];
                }
            }
            __syncthreads();
            if (valid) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float4 l = s_l[u][lane], g = s_g[u][lane];
                    V3 v = add(ld4(l), muls(ld4(g), rcp_gi));
                    sum[0] += (double)v.x;
                    sum[1] += (double)v.y;
                    sum[2] += (double)v.z;
                    any_hit |= l.w != 0.0f;
                }
            }
            __syncthreads();
        }
    }
    if (!valid) return;
    uint32_t pixel = out.pixel_list ? out.pixel_list[pseq] : (uint32_t)pseq;
    for (; c < p.spp; ++c) {
        uint32_t s0 = (uint32_t)(first_slot - slot_base) + c * p.n_gi;
        V3 g = ld4(pool.global[s0]);
        uint32_t h = out.path_hash ? pool.hash[s0] : 0u;
        for (uint32_t j = 1; j < p.n_gi; ++j) {
            g = add(g, ld4(pool.global[s0 + j]));
            if (out.path_hash) h = (h * 0x9E3779B1u) ^ pool.hash[s0 + j];
        }
        const float4 l0 = pool.local[s0];
        any_hit |= l0.w != 0.0f;
        V3 v = add(ld4(l0), muls(g, rcp_gi));
        sum[0] += (double)v.x;
        sum[1] += (double)v.y;
        sum[2] += (double)v.z;
        if (out.path_rgb) {
            size_t k = (size_t)pseq * p.spp + c;
            out.path_rgb[k * 3 + 0] = v.x;
            out.path_rgb[k * 3 + 1] = v.y;
            out.path_rgb[k * 3 + 2] = v.z;
            if (out.path_hash) out.path_hash[k] = h;
        }
    }
    if (out.hitmask) out.hitmask[pixel] = any_hit ? 1u : 0u;
    if (out.accum) {
        out.accum[(size_t)pixel * 3 + 0] += sum[0];
        out.accum[(size_t)pixel * 3 + 1] += sum[1];
        out.accum[(size_t)pixel * 3 + 2] += sum[2];
    }
}

/* ---------------- fixed ray sets ---------------- */

/* spill_stride: spill entries per lane (the host checks that the BVH depth
 * fits kStackLds + spill_stride) */
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(SceneView sc, const tr_ray* rays, tr_hit* hits, uint32_t n,
                                                            uint2* spill, uint32_t spill_stride)
{
    __shared__ uint2 stk[kStackLds][kTraceBlock];
    uint32_t lane = threadIdx.x;
    uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    DevStack st;
    st.lds = &stk[0][lane];
    st.spill = spill + (size_t)gtid * spill_stride;
    const uint32_t stride = gridDim.x * kTraceBlock;
    for (uint32_t base = blockIdx.x * kTraceBlock; base < n; base += stride) {
        uint32_t i = base + lane;
        if (i < n) {
            const float4* rv = reinterpret_cast<const float4*>(rays + i);
            float4 a = rv[0], b = rv[1];
            float org[3] = {a.x, a.y, a.z}, dir[3] = {b.x, b.y, b.z};
            RayPre r = ray_setup(org, dir);
            st.sp = 0;
            Hit h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, a.w, b.w, st, (uint32_t*)nullptr, (uint32_t*)nullptr);
            tr_hit o;
            o.t = h.t; o.u = h.u; o.v = h.v; o.geom = h.geom; o.prim = h.prim;
            hits[i] = o;
        }
    }
}

/* ---------------- rayQuery ---------------- */

/* RayTracerScene::rayQuery (ray-tracing.cc:1932-1979): the ray through
 * screen point (2 tc.x - 1, 1 - 2 tc.y) from the near to the far plane,
 * its closest hit and the hit geometry's sample() frame.  One thread. */
__global__ void k_ray_query(SceneView sc, Mat16 vpi, float tx, float ty, RayQueryOut* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const float sx = 2.0f * tx - 1.0f, sy = 1.0f - 2.0f * ty;
    const float* m = vpi.m;
    float p0[4], p1[4];
    for (int r = 0; r < 4; ++r) {
        p0[r] = m[r] * sx;
        p1[r] = m[r] * sx;
        p0[r] = p0[r] + m[4 + r] * sy;
        p1[r] = p1[r] + m[4 + r] * sy;
        p0[r] = p0[r] + m[8 + r] * -1.0f;
        p1[r] = p1[r] + m[8 + r] * 1.0f;
        p0[r] = p0[r] + m[12 + r] * 1.0f;
        p1[r] = p1[r] + m[12 + r] * 1.0f;
    }
    const float w0 = 1.0f / p0[3], w1 = 1.0f / p1[3];
    const V3 start = mk(p0[0] * w0, p0[1] * w0, p0[2] * w0);
    const V3 end = mk(p1[0] * w1, p1[1] * w1, p1[2] * w1);
    const V3 dir = normalize(sub(end, start));
    const float org[3] = {start.x, start.y, start.z}, d[3] = {dir.x, dir.y, dir.z};
    RayPre r = ray_setup(org, d);
    HostStack st;
    Hit h = traverse_closest(sc.nodes, sc.prims, sc.shapes, r, 0.0f, INFINITY, st, (uint32_t*)nullptr,
                             (uint32_t*)nullptr);
    RayQueryOut o;
    o.t = h.t; o.u = h.u; o.v = h.v; o.geom = h.geom; o.prim = h.prim;
    for (int k = 0; k < 3; ++k) o.normal[k] = o.tangent[k] = o.binormal[k] = 0.0f;
    o.tc[0] = o.tc[1] = 0.0f;
    if (h.geom != TR_INVALID_GEOMETRY && sc.geoms[h.geom].material != kNoMaterial &&
        sc.geoms[h.geom].type != GEOM_VOLUME) {
        Sample sd;
        sample_geometry<true>(sc, h.geom, h.prim, h.t, h.u, h.v, start, dir, &sd);
        o.normal[0] = sd.normal.x; o.normal[1] = sd.normal.y; o.normal[2] = sd.normal.z;
        o.tangent[0] = sd.tangent.x; o.tangent[1] = sd.tangent.y; o.tangent[2] = sd.tangent.z;
        o.binormal[0] = sd.binormal.x; o.binormal[1] = sd.binormal.y; o.binormal[2] = sd.binormal.z;
        o.tc[0] = sd.tcx; o.tc[1] = sd.tcy;
    }
    *out = o;
}

/* ---------------- backbuffer ---------------- */

/* ConvertLinearToSLuminance (spectrum.hh:146-152; pow via tr_fmath) and
 * ToColor (spectrum.hh:792-798). */
__device__ inline float to_srgb(float v)
{
    if (v <= 0.0031308f) return v * 12.92f;
    return 1.055f * tr_expf(tr_logf(v) * (1.0f / 2.4f)) - 0.055f;
}
__device__ inline uint32_t to_color(float x, float y, float z)
{
    x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    y = y < 0.0f ? 0.0f : (y > 1.0f ? 1.0f : y);
    z = z < 0.0f ? 0.0f : (z > 1.0f ? 1.0f : z);
    return (uint32_t)(255.0f * x) | ((uint32_t)(255.0f * y) << 8u) | ((uint32_t)(255.0f * z) << 16u) | (255u << 24u);
}

/* ComputeImage (tests/benchmark/pseudo-volume-performance-test.cu:16-41):
 * every pixel evaluates the pseudo-volume BRDF at wi = (0, sin 45, cos 45),
 * wo = (0, -sin 45, cos 45) in the unit frame, after SGGXSurfaceCache, and
 * stores ToColor of it.  TexCoord is (0, 0), so every pixel runs the same
 * walks (the benchmark's "most optimistic case"). */
__global__ void k_pseudo_volume_image(const tr_material* m, uint32_t width, uint32_t height, uint32_t* rgba,
                                      float* probe)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    float sn, cs;
    tr_sincos_turns(0.125f, &sn, &cs);
    Sample sd{};
    sd.mat = m;
    sd.wi = mk(0.0f, sn, cs);
    sd.wo = mk(0.0f, -sn, cs);
    sd.tangent = mk(1.0f, 0.0f, 0.0f);
    sd.binormal = mk(0.0f, 1.0f, 0.0f);
    sd.normal = mk(0.0f, 0.0f, 1.0f);
    sd.tcx = sd.tcy = 0.0f;
    const NoTex tv{};
    sd.dir_density = sggx_surface_density(sd, tv);
    V3 f = sggx_pseudo_volume_brdf(sd, tv);
    rgba[(size_t)y * width + x] = to_color(f.x, f.y, f.z);
    if (x == 0 && y == 0) {
        probe[0] = f.x;
        probe[1] = f.y;
        probe[2] = f.z;
        probe[3] = sd.dir_density;
    }
}

/* ACESFilm (math/hdr.hh:76-84): x (a x + b) / (x (c x + d) + e), clamped
 * to [0, 1]. */
__device__ inline float aces_film(float x)
{
    float v = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

/* Auto exposure, first pass: per block the FP64 sums of the per-channel
 * log(Epsilon + pixel), Epsilon = 1e-3 (ray-tracing.cc:61, 2375), over
 * finite pixels; second pass (one block): the blocks' sums in order, then
 * the exposure 0.18 / RGBToLuminance(exp(sum / pixels)) (ray-tracing.cc:
 * 1649-1659, spectrum.hh:328-331). */
__global__ __launch_bounds__(256) void k_log_sum(const double* accum, double rcp_count, uint32_t n, double* partial,
                                                 const uint8_t* hitmask)
{
    __shared__ double red[3][256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < n && (!hitmask || hitmask[i])) {
        const float v[3] = {(float)(accum[(size_t)i * 3 + 0] * rcp_count), (float)(accum[(size_t)i * 3 + 1] * rcp_count),
                            (float)(accum[(size_t)i * 3 + 2] * rcp_count)};
        if (isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))
            for (int k = 0; k < 3; ++k) s[k] = (double)tr_logf(1e-3f + v[k]);
    }
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partial[(size_t)blockIdx.x * 3 + k] = red[k][0];
}

__global__ __launch_bounds__(256) void k_exposure(const double* partial, uint32_t blocks, uint32_t n, float* exposure)
{
    __shared__ double red[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
        for (int k = 0; k < 3; ++k) s[k] += partial[(size_t)b * 3 + k];
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float m[3];
        for (int k = 0; k < 3; ++k) m[k] = tr_expf((float)(red[k][0] / (double)n));
        const float lum = 0.2126f * m[0] + 0.7152f * m[1] + 0.0722f * m[2];
        exposure[0] = 0.18f / lum;
    }
}

/* The backbuffer: per-pixel mean radiance through the PicturePostProcess
 * (ray-tracing.cc:2367-2375 and postprocess :1638-1696): Linear, sRGB,
 * ACESFilm then sRGB, and the auto-exposure modes (exposure * value, with
 * the Reinhard operator x / (1 + x) for AutoExposureHDR, math/hdr.hh:
 * 32-35) then sRGB; ToColor truncates 255 * clamp(v, 0, 1)
 * (spectrum.hh:792-798). */
__global__ __launch_bounds__(256) void k_resolve(const double* accum, double rcp_count, uint32_t* rgba, uint32_t n,
                                                 uint32_t post, const float* exposure, const uint8_t* hitmask)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (hitmask && !hitmask[i]) return;
    float c[3] = {(float)(accum[(size_t)i * 3 + 0] * rcp_count), (float)(accum[(size_t)i * 3 + 1] * rcp_count),
                  (float)(accum[(size_t)i * 3 + 2] * rcp_count)};
    if (post == TR_POST_LINEAR) {
        rgba[i] = to_color(c[0], c[1], c[2]);
        return;
    }
    if (post == TR_POST_ACES) {
        for (int k = 0; k < 3; ++k) c[k] = aces_film(c[k]);
    } else if (post == TR_POST_AUTO_EXPOSURE_SRGB || post == TR_POST_AUTO_EXPOSURE_HDR) {
        const float e = exposure[0];
        for (int k = 0; k < 3; ++k) {
            c[k] = e * c[k];
            if (post == TR_POST_AUTO_EXPOSURE_HDR) c[k] = c[k] / (1.0f + c[k]);
        }
    }
    rgba[i] = to_color(to_srgb(c[0]), to_srgb(c[1]), to_srgb(c[2]));
}

}  // namespace

/* Plain launch, or one whose dispatch records the timing events. */
template <typename K, typename... A>
void launch_k(K kernel, uint32_t grid, uint32_t block, hipStream_t s, const LaunchEvents* ev, A... args)
{
    if (ev) hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, ev->start, ev->stop, 0u, args...);
    else kernel<<<dim3(grid), dim3(block), 0, s>>>(args...);
}

void launch_pseudo_volume_image(hipStream_t s, const tr_material* m, uint32_t width, uint32_t height, uint32_t* rgba,
                                float* probe)
{
    dim3 block(8, 8, 1);
    dim3 grid((width + 7) / 8, (height + 7) / 8, 1);
    hipLaunchKernelGGL(k_pseudo_volume_image, grid, block, 0, s, m, width, height, rgba, probe);
}

void launch_raygen(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out, uint64_t slot_base,
                   uint32_t n_slots, const CacheRecords& rec, const LaunchEvents* ev)
{
    /* with camera_fused only the queue size is set */
    launch_k(k_raygen, p.camera_fused ? 1u : (n_slots + 255) / 256, 256, s, ev, p, pool, out, slot_base, n_slots, rec);
}

bool camera_rays_in_kernels(const FrameParams& p, bool counting)
{
    /* launch_trace's choice of k_trace_packet at the camera vertex */
    return !counting && !p.cache && p.req.n == 0 && p.sc.n_volumes == 0 && p.sc.n_prims > kLeafLoopPrims;
}

void launch_trace(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out, uint64_t slot_base,
                  uint32_t vertex, int cur, int nxt, uint32_t grid, uint32_t grid_refill, uint64_t* counts,
                  const LaunchEvents* ev)
{
    unsigned long long* c = (unsigned long long*)counts;
    bool vol = p.sc.n_volumes > 0;
    bool shapes = p.sc.shapes != nullptr;
    bool small = p.sc.n_prims <= kLeafLoopPrims;
    if (counts) {
        if (vol) launch_k(k_trace_vol<true>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, cur, c);
        else if (small) {
            if (shapes) launch_k(k_trace_refill<true, true, true>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
            else launch_k(k_trace_refill<true, false, true>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
        } else if (shapes) launch_k(k_trace_refill<true, true, false>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
        else launch_k(k_trace_refill<true, false, false>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
    } else {
        if (vol) launch_k(k_trace_vol<false>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, cur, c);
        else if (vertex == 0 && !small) {
            if (shapes) launch_k(k_trace_packet<true>, grid_refill, kTraceBlock, s, ev, p, pool, out, slot_base, cur);
            else launch_k(k_trace_packet<false>, grid_refill, kTraceBlock, s, ev, p, pool, out, slot_base, cur);
        } else if (small) {
            if (shapes) launch_k(k_trace_refill<false, true, true>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
            else launch_k(k_trace_refill<false, false, true>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
        } else if (shapes) launch_k(k_trace_refill<false, true, false>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
        else launch_k(k_trace_refill<false, false, false>, grid_refill, kTraceBlock, s, ev, p, pool, cur, c);
    }
}

void launch_shade(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out, uint64_t slot_base,
                  uint32_t vertex, int cur, int nxt, uint32_t n_slots, const CacheRecords& rec, const LaunchEvents* ev)
{
    const uint32_t grid = (n_slots + kSegment - 1) / kSegment;
    /* material maps are bound only to scenes that have textures */
    /* the texture instances also carry the pseudo-volume BRDF's walks */
    const bool tex = p.sc.tex.count > 0 || p.sc.tex.cube_side > 0 || p.sc.tex.pseudo_volume;
    /* scenes without light sources take the instance without next-event
     * estimation (texture-free, cache-free) */
    const bool lights = p.sc.n_lights > 0;
    if (p.sc.n_volumes > 0) {
        if (tex) {
            if (p.cache) launch_k(k_shade<true, true, true, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
            else launch_k(k_shade<true, false, true, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
        } else {
            if (p.cache) launch_k(k_shade<true, true, false, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
            else if (lights) launch_k(k_shade<true, false, false, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
            else launch_k(k_shade<true, false, false, false>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
        }
    } else {
        if (tex) {
            if (p.cache) launch_k(k_shade<false, true, true, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
            else launch_k(k_shade<false, false, true, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
        } else {
            if (p.cache) launch_k(k_shade<false, true, false, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
            else if (lights) launch_k(k_shade<false, false, false, true>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
            else launch_k(k_shade<false, false, false, false>, grid, 256, s, ev, p, pool, out, slot_base, vertex, cur, nxt, rec);
        }
    }
}

void launch_compact(hipStream_t s, const PathPool& pool, uint32_t n_slots, int cur, int nxt, bool queries,
                    const LaunchEvents* ev)
{
    /* one block per group of at least 16 segments of the largest possible
     * queue, at most kCompactGroups per queue: sums, one scan of the sums,
     * then the copies; the kernels size the groups to the shaded queue */
    const uint32_t n_seg = (n_slots + kSegment - 1) / kSegment;
    const uint32_t blocks = std::max(1u, std::min(kCompactGroups, (n_seg + 15) / 16));
    const uint32_t nq = queries ? 3 : 2;
    dim3 grid(blocks, nq);
    if (ev) {
        hipExtLaunchKernelGGL(k_compact_sums, grid, dim3(kSegment), 0, s, ev->start, nullptr, 0u, pool, cur,
                              pool.group_sum);
        hipExtLaunchKernelGGL(k_compact_scan, dim3(nq), dim3(kSegment), 0, s, nullptr, nullptr, 0u, pool, blocks, nxt,
                              pool.group_sum);
        hipExtLaunchKernelGGL(k_compact, grid, dim3(kSegment), 0, s, nullptr, ev->stop, 0u, pool, cur, nxt,
                              (const uint32_t*)pool.group_sum);
    } else {
        k_compact_sums<<<grid, dim3(kSegment), 0, s>>>(pool, cur, pool.group_sum);
        k_compact_scan<<<dim3(nq), dim3(kSegment), 0, s>>>(pool, blocks, nxt, pool.group_sum);
        k_compact<<<grid, dim3(kSegment), 0, s>>>(pool, cur, nxt, (const uint32_t*)pool.group_sum);
    }
}

void launch_cache_infer(hipStream_t s, const MlpWeights& w, const PathPool& pool, const CacheRecords& rec,
                        uint32_t grid, const LaunchEvents* ev)
{
    (void)ev;
    launch_mlp_infer_counted(s, w, rec.queries, pool.query_queue, &pool.counters[3], pool.global, grid);
}

void launch_cache_targets(hipStream_t s, const PathPool& pool, const CacheRecords& rec, uint32_t r0, uint32_t r1,
                          const LaunchEvents* ev)
{
    if (r1 <= r0) return;
    launch_k(k_cache_targets, (r1 - r0 + 255) / 256, 256, s, ev, pool, rec, r0, r1);
}

void launch_shadow(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out,
                   uint64_t slot_base, uint32_t vertex, uint32_t grid, uint64_t* counts, bool wide,
                   const LaunchEvents* ev)
{
    unsigned long long* c = (unsigned long long*)counts;
    bool vol = p.sc.n_volumes > 0;
    if (counts) {
        if (vol) launch_k(k_shadow<true, true>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, c);
        else launch_k(k_shadow<true, false>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, c);
    } else {
        if (vol) launch_k(k_shadow<false, true>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, c);
        else if (wide) launch_k(k_shadow<false, false, true>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, c);
        else launch_k(k_shadow<false, false>, grid, kTraceBlock, s, ev, p, pool, out, slot_base, vertex, c);
    }
}

void launch_accumulate(hipStream_t s, const FrameParams& p, const PathPool& pool, const BatchOut& out,
                       uint64_t slot_base, uint32_t n_pixels, const LaunchEvents* ev)
{
    launch_k(k_accumulate, (n_pixels + kAccBlock - 1) / kAccBlock, kAccBlock, s, ev, p, pool, out, slot_base, n_pixels);
}

void launch_accumulate_requests(hipStream_t s, const FrameParams& p, const PathPool& pool, double* accum,
                                uint32_t max_pix, uint32_t* cursor, const LaunchEvents* ev)
{
    launch_k(k_accumulate_req, (max_pix + kAccBlock - 1) / kAccBlock, kAccBlock, s, ev, p, pool, accum,
             (const uint32_t*)cursor);
    k_req_next<<<1, 64, 0, s>>>(cursor);
}

void launch_request_outputs(hipStream_t s, const ReqView& rv, uint32_t n_gi, const double* accum, float* mean,
                            float* image, uint64_t n_pix_total)
{
    (void)n_gi;
    if (mean) k_req_mean<<<rv.n, 256, 0, s>>>(rv, accum, mean);
    if (image) k_req_image<<<(uint32_t)((n_pix_total + 255) / 256), 256, 0, s>>>(rv, accum, image, n_pix_total);
}

void launch_trace_rays(hipStream_t s, const SceneView& sc, const tr_ray* rays, tr_hit* hits, uint32_t n, uint2* spill,
                       uint32_t grid)
{
    hipLaunchKernelGGL(k_trace_rays, dim3(grid), dim3(kTraceBlock), 0, s, sc, rays, hits, n, spill,
                       (uint32_t)kStackSpill);
}

/* Scratch of the allocation-free ray-set trace: the traversal spill stack
 * of a persistent grid of at most kNoAllocWaves one-wave blocks,
 * kNoAllocSpill entries per lane (16 MiB).  Each device holds its own copy
 * until the process ends; host code uses the per-device address from
 * hipGetSymbolAddress, never the array's name. */
__device__ __align__(256) uint2 scratch[(size_t)kNoAllocWaves * kTraceBlock * kNoAllocSpill];

namespace {
constexpr int kMaxDevices = 64;
uint2* g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;
}  // namespace

hipError_t noalloc_scratch(uint2** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<uint2*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t launch_trace_rays_noalloc(hipStream_t s, const SceneView& sc, const tr_ray* rays, tr_hit* hits, uint32_t n,
                                     uint32_t grid)
{
    uint2* spill = nullptr;
    hipError_t e = noalloc_scratch(&spill);
    if (e != hipSuccess) return e;
    if (grid > kNoAllocWaves) grid = kNoAllocWaves;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_trace_rays, dim3(grid), dim3(kTraceBlock), 0, s, sc, rays, hits, n, spill,
                       (uint32_t)kNoAllocSpill);
    return hipGetLastError();
}

void launch_ray_query(hipStream_t s, const SceneView& sc, const float vpi[16], float tx, float ty, RayQueryOut* out)
{
    Mat16 m;
    for (int k = 0; k < 16; ++k) m.m[k] = vpi[k];
    hipLaunchKernelGGL(k_ray_query, dim3(1), dim3(64), 0, s, sc, m, tx, ty, out);
}

void launch_resolve(hipStream_t s, const double* accum, double rcp_count, uint32_t* rgba, uint32_t n, uint32_t post,
                    double* partial, float* exposure, const uint8_t* hitmask)
{
    uint32_t grid = (n + 255) / 256;
    if (post == TR_POST_AUTO_EXPOSURE_SRGB || post == TR_POST_AUTO_EXPOSURE_HDR) {
        hipLaunchKernelGGL(k_log_sum, dim3(grid), dim3(256), 0, s, accum, rcp_count, n, partial, hitmask);
        hipLaunchKernelGGL(k_exposure, dim3(1), dim3(256), 0, s, (const double*)partial, grid, n, exposure);
    }
    hipLaunchKernelGGL(k_resolve, dim3(grid), dim3(256), 0, s, accum, rcp_count, rgba, n, post, (const float*)exposure,
                       hitmask);
}

}  // namespace trk
