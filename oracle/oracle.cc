/*
 * oracle.cc - CPU oracle of the path tracer.
 *
 * ORACLE: test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, as the checker or as the timed
 * CPU baseline; the product never calls it.
 *
 * It restates the reference CPU integrator RayTracerScene::subdraw
 * (src/graphics/ray-tracing/ray-tracing.cc:1981-2379), its next-event
 * estimation EvaluateLightSources (:1809-1930) and balance heuristic
 * EvaluateTotalBalanceHeuristic (:1791-1806), the light samplers of
 * include/tempest/graphics/ray-tracing/ray-tracing-common.hh:567-687 and
 * the geometry samplers DefaultMeshSample (:375-435), SphereGeometry::sample
 * (:1166-1180) and SphereAreaLightGeometry::sample (:1775-1783).  Embree's
 * traversal is replaced by a binary BVH over the same primitives with the
 * closest-hit rule of orc_geom.h; orc_trace_rays can also brute-force.
 *
 * Philox counter layout (shared with the GPU kernels):
 *   ctr = (pixel, camera sample index, split<<16 | vertex<<8 | stream, sub)
 *   key = (seed low, seed high)
 * Streams: 0 camera jitter, 1 BSDF sample (x, y, lobe), 2 light sample
 * (sub = light<<16 | sample; triangle pick, ra, rb), 3 Russian roulette,
 * 4 delta tracking of a path vertex (sub = step pair), 5 Woodcock tracking
 * of a shadow ray (sub = light<<24 | sample<<16 | step pair).
 *
 * Grid volumes (orc_volume.h) are one box primitive each; a closest hit on
 * a box runs delta tracking along the ray and either scatters there or
 * traces on beyond the box (the free-path retrace of subdraw).
 */
#include "oracle.h"
#include "orc_philox.h"
#include "orc_math.h"
#include "orc_bsdf.h"
#include "orc_geom.h"
#include "orc_volume.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

using namespace orc;

namespace {

/* When set, render_camera_sample prints every vertex of the path. */
thread_local FILE* g_trace = nullptr;
#define ORC_TRACE(...) do { if (g_trace) fprintf(g_trace, __VA_ARGS__); } while (0)

enum { GEOM_MESH = 0, GEOM_SPHERE = 1, GEOM_SPHERE_LIGHT = 2, GEOM_VOLUME = 3, GEOM_SHAPE = 4 };
/* the material of a blocker (addBlocker, ray-tracing.cc:1386-1400 and
 * BlockerGeometry::sample :1239-1244): none */
constexpr uint32_t ORC_NO_MATERIAL = 0xffffffffu;
enum { STREAM_CAMERA = 0, STREAM_BSDF = 1, STREAM_LIGHT = 2, STREAM_RR = 3 };

struct Geom {
    uint32_t type;
    uint32_t material;
    uint32_t two_sided;
    uint32_t vbuf;          /* index into Scene::vbufs */
    uint32_t vertex_offset; /* bytes */
    uint32_t stride;
    uint32_t index_base;    /* into Scene::indices (triplets of this submesh) */
    uint32_t tri_count;
    float sphere[4];
    uint32_t light;         /* sphere light: light index */
    uint32_t volume;        /* grid volume: index into orc_scene::volumes */
    uint32_t shape;         /* user shape: index into orc_scene::shapes */
};

struct Prim {
    float v0[3], v1[3], v2[3]; /* triangle; spheres: v0 = centre, v1[0] = radius; volumes: box v0..v1;
                                  shapes: box v1..v2 */
    uint32_t geom, prim;
    uint32_t is_sphere, two_sided, is_volume, is_shape;
};

struct Light {
    tr_light src;
    uint32_t geom;          /* mesh lights and sphere lights: geometry slot */
};

struct BNode {
    float lo[3], hi[3];
    uint32_t left_or_first; /* internal: left child; leaf: first prim index */
    uint32_t count;         /* 0 = internal (right = left + 1 is not assumed: see right) */
    uint32_t right;
};

}  // namespace

struct orc_scene {
    uint32_t width, height;
    float vpi[16];
    uint32_t spp = 1, n_gi = 64, n_light = 4, max_depth = 1;
    float rr = 1.0f;
    float background[3] = {0, 0, 0};
    uint32_t key[2] = {1, 0};
    uint32_t estimator = TR_ESTIMATOR_REFERENCE;
    uint32_t render_mode = 0;  /* RenderMode (ray-tracing-common.hh:132-139) */
    /* added to the pixel index in the random-number key: a request of a
     * batched render draws the numbers of its pixels' position over all
     * requests */
    uint32_t pixel_key_base = 0;
    /* radiance-cache mode with a constant network output (the checker of
     * the path logic of cache mode) and the training records it writes */
    struct {
        bool enabled = false;
        float rough_alpha = 0.2f;
        uint32_t train_stride = 1;
        uint32_t frame = 0;
        float value[3] = {0, 0, 0};
        bool query_nee = false;
        float box_lo[3] = {0, 0, 0}, box_rcp[3] = {1, 1, 1};
        std::vector<float> in, target, valid;   /* [r][10], [r][3], [r] */
    } cache;
    std::vector<tr_material> materials;
    std::vector<const tr_material*> material_keys;
    std::vector<std::vector<uint8_t>> vbufs;
    std::vector<int32_t> indices;
    std::vector<Geom> geoms;
    std::vector<Prim> prims;
    std::vector<uint32_t> prim_order;
    std::vector<BNode> nodes;
    std::vector<Light> lights;
    std::vector<GridVolume> volumes;
    std::vector<OrcShape> shapes;
    OrcTexSet texset;          /* bound textures and the global cube map */
    uint32_t instances = 0;
    bool committed = false;
};

namespace {

uint32_t intern_material(orc_scene* s, const tr_material* m)
{
    for (size_t i = 0; i < s->material_keys.size(); ++i)
        if (s->material_keys[i] == m && memcmp(&s->materials[i], m, sizeof(*m)) == 0) return (uint32_t)i;
    s->material_keys.push_back(m);
    s->materials.push_back(*m);
    return (uint32_t)(s->materials.size() - 1);
}

inline const float* vertex_ptr(const orc_scene* s, const Geom& g, int32_t idx)
{
    return reinterpret_cast<const float*>(s->vbufs[g.vbuf].data() + g.vertex_offset + (size_t)idx * g.stride);
}

/* ---------------- BVH (binary, median split) ---------------- */

void prim_bounds(const Prim& p, float lo[3], float hi[3])
{
    if (p.is_sphere) {
        for (int a = 0; a < 3; ++a) { lo[a] = p.v0[a] - p.v1[0]; hi[a] = p.v0[a] + p.v1[0]; }
    } else if (p.is_shape) {
        for (int a = 0; a < 3; ++a) { lo[a] = p.v1[a]; hi[a] = p.v2[a]; }
    } else if (p.is_volume) {
        for (int a = 0; a < 3; ++a) { lo[a] = p.v0[a]; hi[a] = p.v1[a]; }
    } else {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(p.v0[a], std::min(p.v1[a], p.v2[a]));
            hi[a] = std::max(p.v0[a], std::max(p.v1[a], p.v2[a]));
        }
    }
}

uint32_t build_node(orc_scene* s, uint32_t begin, uint32_t end, std::vector<float>& cen,
                    std::vector<float>& blo, std::vector<float>& bhi)
{
    BNode n;
    for (int a = 0; a < 3; ++a) { n.lo[a] = INFINITY; n.hi[a] = -INFINITY; }
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = begin; i < end; ++i) {
        uint32_t p = s->prim_order[i];
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = std::min(n.lo[a], blo[p * 3 + a]);
            n.hi[a] = std::max(n.hi[a], bhi[p * 3 + a]);
            clo[a] = std::min(clo[a], cen[p * 3 + a]);
            chi[a] = std::max(chi[a], cen[p * 3 + a]);
        }
    }
    /* pad the box outward so rounding in the slab test cannot cull a hit */
    for (int a = 0; a < 3; ++a) {
        float pad = (n.hi[a] - n.lo[a]) * 1e-5f + 1e-5f * (fabsf(n.lo[a]) + fabsf(n.hi[a])) + 1e-30f;
        n.lo[a] -= pad;
        n.hi[a] += pad;
    }
    uint32_t idx = (uint32_t)s->nodes.size();
    s->nodes.push_back(n);
    if (end - begin <= 4) {
        s->nodes[idx].left_or_first = begin;
        s->nodes[idx].count = end - begin;
        s->nodes[idx].right = 0;
        return idx;
    }
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > ext) { ext = chi[a] - clo[a]; axis = a; }
    uint32_t mid = (begin + end) / 2;
    std::nth_element(s->prim_order.begin() + begin, s->prim_order.begin() + mid, s->prim_order.begin() + end,
                     [&](uint32_t x, uint32_t y) {
                         float cx = cen[x * 3 + axis], cy = cen[y * 3 + axis];
                         return cx < cy || (cx == cy && x < y);
                     });
    uint32_t l = build_node(s, begin, mid, cen, blo, bhi);
    uint32_t r = build_node(s, mid, end, cen, blo, bhi);
    s->nodes[idx].left_or_first = l;
    s->nodes[idx].count = 0;
    s->nodes[idx].right = r;
    return idx;
}

/* Slab test in float64 on the padded box. */
inline bool box_hit(const BNode& n, const double o[3], const double inv[3], double tnear, double tfar, double* tmin_out)
{
    double t0 = tnear, t1 = tfar;
    for (int a = 0; a < 3; ++a) {
        double ta = (n.lo[a] - o[a]) * inv[a];
        double tb = (n.hi[a] - o[a]) * inv[a];
        if (ta > tb) std::swap(ta, tb);
        if (ta != ta) ta = -INFINITY;
        if (tb != tb) tb = INFINITY;
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
    }
    *tmin_out = t0;
    return t0 <= t1;
}

/* The box of a grid volume: FP32 slab test with the reciprocal direction
 * (components below 1e-20 in magnitude taken as +-1e-20).  Entry =
 * max(slab entry, tnear) if the slab exit lies beyond tnear and the entry
 * before tfar; the exit is returned as the hit's u. */
inline int volume_box_hit(const float org[3], const float dir[3], const float lo[3], const float hi[3], float tnear,
                          float tfar, float* t_out, float* t_exit)
{
    float t0 = -INFINITY, t1 = INFINITY;
    for (int a = 0; a < 3; ++a) {
        float d = dir[a];
        if (fabsf(d) < 1e-20f) d = d < 0.0f ? -1e-20f : 1e-20f;
        float inv = 1.0f / d;
        float ta = (lo[a] - org[a]) * inv, tb = (hi[a] - org[a]) * inv;
        t0 = std::max(t0, std::min(ta, tb));
        t1 = std::min(t1, std::max(ta, tb));
    }
    if (!(t0 <= t1) || !(t1 > tnear)) return 0;
    float t = std::max(t0, tnear);
    if (!(t < tfar)) return 0;
    *t_out = t;
    *t_exit = t1;
    return 1;
}

inline void test_prim(const orc_scene* s, uint32_t pi, const WatertightRay& wr, const float org[3],
                      const float dir[3], float tnear, float tfar, tr_hit* best)
{
    const Prim& p = s->prims[pi];
    float t, u = 0.0f, v = 0.0f;
    int hit;
    if (p.is_sphere) hit = sphere_hit(org, dir, p.v0, p.v1[0], tnear, tfar, &t);
    else if (p.is_shape) {
        float n[3];
        hit = orc_shape_hit(s->shapes[s->geoms[p.geom].shape], org, dir, &t, &u, &v, n) && t > tnear && t < tfar;
    } else if (p.is_volume) hit = volume_box_hit(org, dir, p.v0, p.v1, tnear, tfar, &t, &u);
    else hit = watertight_triangle(wr, org, p.v0, p.v1, p.v2, (int)p.two_sided, tnear, tfar, &t, &u, &v);
    if (!hit) return;
    if (best->geom != TR_INVALID_GEOMETRY && !hit_better(t, p.geom, p.prim, best->t, best->geom, best->prim)) return;
    best->t = t; best->u = u; best->v = v; best->geom = p.geom; best->prim = p.prim;
}

tr_hit trace_closest(const orc_scene* s, const float org[3], const float dir[3], float tnear, float tfar, bool brute)
{
    tr_hit best;
    best.t = tfar; best.u = best.v = 0.0f; best.geom = TR_INVALID_GEOMETRY; best.prim = TR_INVALID_GEOMETRY;
    if (s->prims.empty()) return best;
    WatertightRay wr = watertight_setup(dir);
    if (brute) {
        for (uint32_t i = 0; i < s->prims.size(); ++i) test_prim(s, i, wr, org, dir, tnear, tfar, &best);
    } else {
        double o[3] = {org[0], org[1], org[2]};
        double inv[3];
        for (int a = 0; a < 3; ++a) inv[a] = dir[a] != 0.0f ? 1.0 / (double)dir[a] : copysign(1e300, (double)dir[a]);
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const BNode& n = s->nodes[stack[--sp]];
            double tmin;
            double tmax = best.geom == TR_INVALID_GEOMETRY ? (double)tfar : (double)best.t;
            if (!box_hit(n, o, inv, (double)tnear, tmax * (1.0 + 1e-6) + 1e-30, &tmin)) continue;
            if (n.count) {
                for (uint32_t i = 0; i < n.count; ++i)
                    test_prim(s, s->prim_order[n.left_or_first + i], wr, org, dir, tnear, tfar, &best);
            } else {
                double tl, tr;
                const BNode& L = s->nodes[n.left_or_first];
                const BNode& R = s->nodes[n.right];
                bool hl = box_hit(L, o, inv, (double)tnear, tmax * (1.0 + 1e-6) + 1e-30, &tl);
                bool hr = box_hit(R, o, inv, (double)tnear, tmax * (1.0 + 1e-6) + 1e-30, &tr);
                if (hl && hr) {
                    if (tl <= tr) { stack[sp++] = n.right; stack[sp++] = n.left_or_first; }
                    else { stack[sp++] = n.left_or_first; stack[sp++] = n.right; }
                } else if (hl) {
                    stack[sp++] = n.left_or_first;
                } else if (hr) {
                    stack[sp++] = n.right;
                }
            }
        }
    }
    if (best.geom == TR_INVALID_GEOMETRY) best.t = tfar;
    return best;
}

/* ---------------- RNG ---------------- */

struct Rng {
    uint32_t key[2];
    uint32_t pixel, sample, split;
    void draw(uint32_t vertex, uint32_t stream, uint32_t sub, uint32_t out[4]) const
    {
        uint32_t ctr[4] = {pixel, sample, (split << 16) | (vertex << 8) | stream, sub};
        orc_philox4x32_10(ctr, key, out);
    }
    StepDraws steps(uint32_t vertex, uint32_t stream) const
    {
        StepDraws d;
        d.ctr[0] = pixel; d.ctr[1] = sample; d.ctr[2] = (split << 16) | (vertex << 8) | stream; d.ctr[3] = 0;
        d.key[0] = key[0]; d.key[1] = key[1];
        return d;
    }
};

bool is_volume_geom(const orc_scene* s, uint32_t geom)
{
    return geom != TR_INVALID_GEOMETRY && s->geoms[geom].type == GEOM_VOLUME;
}

/* Closest hit of a path segment with delta tracking through grid volumes:
 * the primary and secondary tracking loops of subdraw (ray-tracing.cc:
 * 2040-2128, 2212-2272).  A volume hit becomes a collision at distance t
 * (u = v = 0), or the trace continues beyond the point where the walk left
 * the volume. */
tr_hit trace_segment(const orc_scene* s, const float org[3], const float dir[3], const Rng& rng, uint32_t vertex)
{
    tr_hit h = trace_closest(s, org, dir, 0.0f, INFINITY, false);
    if (s->volumes.empty()) return h;
    StepDraws dr = rng.steps(vertex, ORC_STREAM_VOLUME);
    uint32_t step = 0, crossings = 0;
    V3<float> o = v3(org[0], org[1], org[2]), d = v3(dir[0], dir[1], dir[2]);
    while (is_volume_geom(s, h.geom)) {
        const GridVolume& g = s->volumes[s->geoms[h.geom].volume];
        float tc;
        if (delta_track(g, o, d, h.t, h.u, dr, &step, &tc)) {
            h.t = tc;
            h.u = h.v = 0.0f;
            break;
        }
        if (++crossings >= kOrcMaxVolumeCrossings) {
            h.geom = h.prim = TR_INVALID_GEOMETRY;
            h.t = INFINITY;
            break;
        }
        h = trace_closest(s, org, dir, tc, INFINITY, false);
    }
    return h;
}

/* ---------------- geometry sampling ---------------- */

/* The sample() of the user shapes (computed in FP32, as the hit):
 * RectGeometry::sample (ray-tracing.cc:1246-1272), DiskGeometry::sample
 * (:1274-1283), CylinderGeometry::sample (:1142-1164, also for the oblique
 * cylinder, whose own sample() is commented out), EllipsoidGeometry::sample
 * (:1182-1208).  The normal is the intersection's, recomputed from the same
 * ray. */
template <class T>
void shape_sample(const orc_scene* s, const OrcShape& sh, const tr_hit& h, V3<T> org, V3<T> dir, Sample<T>* sd)
{
    float of[3] = {(float)org.x, (float)org.y, (float)org.z};
    float df[3] = {(float)dir.x, (float)dir.y, (float)dir.z};
    float t, u, v, n[3] = {0.0f, 0.0f, 1.0f};
    orc_shape_hit(sh, of, df, &t, &u, &v, n);
    const float* f = sh.f;
    float N[3] = {n[0], n[1], n[2]}, Tn[3], B[3];
    float tcx = 0.0f, tcy = 0.0f;
    if (sh.kind == ORC_RECT) {
        tcx = f[14] + h.u * f[16];
        tcy = f[15] + h.v * f[17];
        M3<float> o;
        o.c0 = v3(f[3], f[4], f[5]);
        o.c1 = v3(f[6], f[7], f[8]);
        o.c2 = v3(f[9], f[10], f[11]);
        if (sh.tangent_map) {
            Rgba rgb = orc_tex_sample(s->texset, sh.tangent_map, tcx, tcy, ORC_TEX_RGB);
            M3<float> rel = make_basis_tangent(normalize(v3(rgb.v[0], rgb.v[1], rgb.v[2])));
            o = mul(o, rel);
        }
        Tn[0] = o.c0.x; Tn[1] = o.c0.y; Tn[2] = o.c0.z;
        B[0] = o.c1.x; B[1] = o.c1.y; B[2] = o.c1.z;
        N[0] = o.c2.x; N[1] = o.c2.y; N[2] = o.c2.z;
    } else if (sh.kind == ORC_DISK) {
        V3<float> hp = v3(h.t * df[0] + of[0], h.t * df[1] + of[1], h.t * df[2] + of[2]);
        V3<float> bn = normalize(v3(f[0], f[1], f[2]) - hp);
        V3<float> tn = cross(bn, v3(n[0], n[1], n[2]));
        Tn[0] = tn.x; Tn[1] = tn.y; Tn[2] = tn.z;
        B[0] = bn.x; B[1] = bn.y; B[2] = bn.z;
    } else if (sh.kind == ORC_ELLIPSOID) {
        V3<float> qn = v3(f[14], f[15], f[16]), nn = v3(n[0], n[1], n[2]);
        V3<float> bn, tn;
        if (fabsf(dot(qn, nn)) < 1e-6f) {
            M3<float> bs = make_basis(nn);
            bn = bs.c1;
            tn = bs.c0;
        } else {
            bn = normalize(cross(qn, nn));
            tn = cross(bn, nn);
        }
        Tn[0] = tn.x; Tn[1] = tn.y; Tn[2] = tn.z;
        B[0] = bn.x; B[1] = bn.y; B[2] = bn.z;
    } else {
        if (fabsf(n[2]) == 1.0f) {
            float p[2] = {of[0] + h.t * df[0], of[1] + h.t * df[1]};
            float tv[2] = {p[0] - f[0], p[1] - f[1]};
            float l = sqrtf(tv[0] * tv[0] + tv[1] * tv[1]);
            if (l != 0.0f) { tv[0] /= l; tv[1] /= l; }
            Tn[0] = tv[0]; Tn[1] = tv[1]; Tn[2] = 0.0f;
        } else {
            Tn[0] = 0.0f; Tn[1] = 0.0f; Tn[2] = 1.0f;
        }
        cross3f(n, Tn, B);
    }
    sd->normal = v3((T)N[0], (T)N[1], (T)N[2]);
    sd->tangent = v3((T)Tn[0], (T)Tn[1], (T)Tn[2]);
    sd->binormal = v3((T)B[0], (T)B[1], (T)B[2]);
    sd->tcx = (T)tcx;
    sd->tcy = (T)tcy;
}

template <class T>
void sample_geometry(const orc_scene* s, const tr_hit& h, V3<T> org, V3<T> dir, Sample<T>* sd)
{
    const Geom& g = s->geoms[h.geom];
    sd->mat = &s->materials[g.material];
    sd->tex = &s->texset;
    sd->tcx = sd->tcy = (T)0;
    if (g.type == GEOM_MESH) {
        /* RTMesh::sample + DefaultMeshSample (ray-tracing.cc:375-461) */
        const int32_t* ib = &s->indices[g.index_base + (size_t)h.prim * 3];
        const uint8_t* base = s->vbufs[g.vbuf].data() + g.vertex_offset;
        const uint8_t* p0 = base + (size_t)ib[0] * g.stride;
        const uint8_t* p1 = base + (size_t)ib[1] * g.stride;
        const uint8_t* p2 = base + (size_t)ib[2] * g.stride;
        uint32_t no = g.stride - 12;
        const float* n0 = reinterpret_cast<const float*>(p0 + no);
        const float* n1 = reinterpret_cast<const float*>(p1 + no);
        const float* n2 = reinterpret_cast<const float*>(p2 + no);
        T u = (T)h.u, v = (T)h.v;
        T w = ((T)1 - u) - v;
        V3<T> N0 = v3((T)n0[0], (T)n0[1], (T)n0[2]);
        V3<T> N1 = v3((T)n1[0], (T)n1[1], (T)n1[2]);
        V3<T> N2 = v3((T)n2[0], (T)n2[1], (T)n2[2]);
        V3<T> norm = normalize(N0 * w + N1 * u + N2 * v);
        sd->tangent = v3((T)0, (T)0, (T)0);
        sd->binormal = v3((T)0, (T)0, (T)0);
        sd->tcx = sd->tcy = (T)0;
        if (g.stride >= 32) {
            /* TexCoord = tc0 * w + tc1 * u + tc2 * v (ray-tracing.cc:388-394) */
            const float* c0 = reinterpret_cast<const float*>(p0 + 12);
            const float* c1 = reinterpret_cast<const float*>(p1 + 12);
            const float* c2 = reinterpret_cast<const float*>(p2 + 12);
            sd->tcx = (T)c0[0] * w + (T)c1[0] * u + (T)c2[0] * v;
            sd->tcy = (T)c0[1] * w + (T)c1[1] * u + (T)c2[1] * v;
        }
        if (g.stride >= 44) {
            const float* t0 = reinterpret_cast<const float*>(p0 + 20);
            const float* t1 = reinterpret_cast<const float*>(p1 + 20);
            const float* t2 = reinterpret_cast<const float*>(p2 + 20);
            V3<T> tan = v3((T)t0[0], (T)t0[1], (T)t0[2]) * w + v3((T)t1[0], (T)t1[1], (T)t1[2]) * u +
                        v3((T)t2[0], (T)t2[1], (T)t2[2]) * v;
            M3<T> b = make_basis_orthogonalize(tan, norm);
            sd->tangent = b.c0;
            sd->binormal = b.c1;
        }
        sd->normal = norm;
    } else if (g.type == GEOM_SHAPE) {
        shape_sample(s, s->shapes[g.shape], h, org, dir, sd);
    } else {
        float of[3] = {(float)org.x, (float)org.y, (float)org.z};
        float df[3] = {(float)dir.x, (float)dir.y, (float)dir.z};
        if (g.type == GEOM_SPHERE) {
            /* SphereGeometry::sample (ray-tracing.cc:1166-1180); the normal
             * is the one IntersectSphere wrote into ray.Ng */
            float n[3];
            sphere_normal(of, df, g.sphere, h.t, n);
            V3<T> N = v3((T)n[0], (T)n[1], (T)n[2]);
            sd->normal = N;
            sd->binormal = normalize(cross(v3((T)0, (T)0, (T)1), N));
            sd->tangent = cross(sd->binormal, sd->normal);
        } else {
            /* SphereAreaLightGeometry::sample (ray-tracing.cc:1775-1783) */
            V3<T> c = v3((T)g.sphere[0], (T)g.sphere[1], (T)g.sphere[2]);
            sd->normal = normalize(org + dir * (T)h.t - c);
            sd->tangent = v3((T)0, (T)0, (T)0);
            sd->binormal = v3((T)0, (T)0, (T)0);
        }
    }
}

/* ---------------- lights ---------------- */

bool is_area_light(uint32_t type) { return type == TR_LIGHT_SPHERE_AREA; }

/* Lights sampled SamplesLocalAreaLight times and balance-weighted in NEE:
 * sphere lights in the reference estimator (IsAreaLight), sphere and mesh
 * lights in the unbiased one. */
bool multi_sample(uint32_t type, bool unbiased)
{
    return type == TR_LIGHT_SPHERE_AREA || (unbiased && type == TR_LIGHT_MESH_AREA);
}

/* Unbiased estimator: solid-angle density dist^2 / (cos * area) / count of
 * area-uniform sampling of a triangle with n = e0 x e1, zero from behind. */
template <class T> T mesh_light_pdf_u(V3<T> n, V3<T> dir, T dist, uint32_t tri_count)
{
    T c = dot(n, dir);
    if (!(c < (T)-1e-6f)) return (T)0;
    return (dist * dist) / ((T)-0.5 * c) / (T)tri_count;
}

/* ComputeSphereAreaPDF (sampling3.hh:229-262) */
template <class T> T sphere_light_pdf(V3<T> c, T radius, V3<T> pos, V3<T> dir)
{
    V3<T> dc = c - pos;
    T dist = xsqrt(dot(dc, dc));
    if (dist <= radius) return (T)1 / ((T)4 * C<T>::pi);
    V3<T> to_c = dc / dist;
    T sin_rc = radius / dist;
    T cos_rc = xsqrt((T)1 - sin_rc * sin_rc);
    T cos_theta = dot(dir, to_c);
    if (cos_theta < cos_rc) return (T)0;
    return (T)1 / ((T)2 * C<T>::pi * ((T)1 - cos_rc));
}

template <class T> struct LightSample {
    uint32_t geom;
    V3<T> dir;
    T dist;
    V3<T> radiance;
    T pdf;
};

/* TriangleArea of the projected vertices, sampling3.hh:206-212 */
template <class T> T projected_triangle_area(V3<T> pos, V3<T> p0, V3<T> p1, V3<T> p2)
{
    V3<T> a = normalize(p0 - pos), b = normalize(p1 - pos), c = normalize(p2 - pos);
    return (T)0.5 * length(cross(b - a, c - a));
}

template <class T> V3<T> vpos(const orc_scene* s, const Geom& g, uint32_t tri, int k)
{
    const int32_t* ib = &s->indices[g.index_base + (size_t)tri * 3];
    const float* f = vertex_ptr(s, g, ib[k]);
    return v3((T)f[0], (T)f[1], (T)f[2]);
}

/* Light sampling (ray-tracing-common.hh:567-647).  xi: triangle pick, ra, rb. */
/* UniformSampleTriangleArea (sampling3.hh:214-227): direction, distance and
 * (facing ? 1/projected area : 0). */
template <class T>
void triangle_area_sample(V3<T> pos, V3<T> p0, V3<T> p1, V3<T> p2, T ra, T rb, V3<T>* dir, T* dist, T* pdf)
{
    T sqrt_r = xsqrt(ra);
    V3<T> e0 = p1 - p0, e1 = p2 - p0;
    V3<T> point = p0 + (sqrt_r * rb) * e0 + ((T)1 - sqrt_r) * e1;
    V3<T> du = point - pos;
    T len = length(du);
    *dir = du / len;
    *dist = len;
    *pdf = dot(cross(e0, e1), *dir) < -(T)1e-6f ? (T)1 / projected_triangle_area(pos, p0, p1, p2) : (T)0;
}

/* Light sampling (ray-tracing-common.hh:567-647).  pick: the triangle
 * selector word of mesh lights; ra, rb: the two sample coordinates. */
template <class T>
LightSample<T> sample_light(const orc_scene* s, const Light& L, V3<T> pos, uint32_t pick, T ra, T rb,
                            bool unbiased = false)
{
    LightSample<T> r;
    const tr_light& l = L.src;
    if (l.type == TR_LIGHT_DIRECTIONAL) {
        r.dist = (T)INFINITY;
        r.dir = spec3<T>(l.direction);
        r.radiance = spec3<T>(l.radiance);
        r.pdf = (T)1;
        r.geom = TR_INVALID_GEOMETRY;
    } else if (l.type == TR_LIGHT_POINT) {
        V3<T> dv = spec3<T>(l.position) - pos;
        T dist = length(dv);
        r.dist = (T)INFINITY;
        r.dir = dv / dist;
        r.radiance = spec3<T>(l.radiance) / (C<T>::tau * dist * dist);
        r.pdf = (T)1;
        r.geom = TR_INVALID_GEOMETRY;
    } else if (l.type == TR_LIGHT_SPHERE_AREA) {
        /* SampleSphereArea, sampling3.hh:157-203 */
        V3<T> c = spec3<T>(l.sphere_center);
        T radius = (T)l.sphere_radius;
        V3<T> dc = c - pos;
        T dist2 = dot(dc, dc);
        T dist = xsqrt(dist2);
        r.radiance = spec3<T>(l.radiance);
        r.geom = L.geom;
        if (dist <= radius) {
            /* UniformSampleSphereArea, sampling3.hh:142-154 */
            T sin_phi, cos_phi;
            sincos_turns(rb, &sin_phi, &cos_phi);
            T r2 = (T)2 * radius;
            T q = xsqrt(ra * ((T)1 - ra));
            V3<T> res = c - pos + v3(r2 * cos_phi * q, r2 * sin_phi * q, radius * ((T)1 - (T)2 * ra));
            T t = length(res);
            r.dir = res / t;
            r.dist = t;
            r.pdf = (T)1 / ((T)4 * C<T>::pi);
            return r;
        }
        V3<T> to_c = dc / dist;
        T sin_rc = radius / dist;
        T cos_rc = xsqrt((T)1 - sin_rc * sin_rc);
        T sin_phi, cos_phi;
        sincos_turns(ra, &sin_phi, &cos_phi);
        T cos_theta = (T)1 - rb + rb * cos_rc;
        T sin_theta = xsqrt((T)1 - cos_theta * cos_theta);
        M3<T> basis = make_basis(to_c);
        V3<T> d = normalize(transform(basis, v3(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta)));
        T b = dist / cos_theta;
        T cc = dist2 - radius * radius;
        T D = b * b - cc;
        T t = b;
        if (D >= (T)0) t -= xsqrt(D);
        r.dir = d;
        r.dist = t;
        r.pdf = (T)1 / ((T)2 * C<T>::pi * ((T)1 - cos_rc));
    } else {
        /* MeshAreaSampleLightSource + UniformSampleTriangleArea
         * (ray-tracing-common.hh:622-647, sampling3.hh:214-227) */
        const Geom& g = s->geoms[L.geom];
        uint32_t tri = (uint32_t)(((uint64_t)pick * g.tri_count) >> 32);
        V3<T> p0 = vpos<T>(s, g, tri, 0), p1 = vpos<T>(s, g, tri, 1), p2 = vpos<T>(s, g, tri, 2);
        T pdf;
        triangle_area_sample(pos, p0, p1, p2, ra, rb, &r.dir, &r.dist, &pdf);
        r.pdf = unbiased ? mesh_light_pdf_u(cross(p1 - p0, p2 - p0), r.dir, r.dist, g.tri_count)
                         : pdf / (T)g.tri_count;
        r.radiance = spec3<T>(s->materials[g.material].diffuse);
        r.geom = L.geom;
    }
    return r;
}

/* LightSourceComputePDFLookup (ray-tracing-common.hh:582-687) */
template <class T> T light_pdf(const orc_scene* s, const Light& L, V3<T> pos, V3<T> dir)
{
    const tr_light& l = L.src;
    if (l.type == TR_LIGHT_DIRECTIONAL || l.type == TR_LIGHT_POINT) return (T)1;
    if (l.type == TR_LIGHT_SPHERE_AREA) return sphere_light_pdf(spec3<T>(l.sphere_center), (T)l.sphere_radius, pos, dir);
    /* MeshAreaLightComputePDF, ray-tracing-common.hh:650-687 */
    const Geom& g = s->geoms[L.geom];
    T total = (T)0;
    for (uint32_t tri = 0; tri < g.tri_count; ++tri) {
        V3<T> p0 = vpos<T>(s, g, tri, 0), p1 = vpos<T>(s, g, tri, 1), p2 = vpos<T>(s, g, tri, 2);
        V3<T> d0 = p0 - p1, d1 = p2 - p1;
        V3<T> un = cross(d1, d0);
        if (dot(dir, un) >= (T)0) continue;
        total += projected_triangle_area(pos, p0, p1, p2);
    }
    return total != (T)0 ? (T)1 / total : (T)0;
}

/* Unbiased estimator: density with which the light owning an emitter hit
 * produces the hit point from `pos` along dir at distance t (0 for
 * emitters that are not lights). */
template <class T> T hit_light_pdf(const orc_scene* s, const tr_hit& h, V3<T> pos, V3<T> dir)
{
    const Geom& g = s->geoms[h.geom];
    if (g.type == GEOM_SPHERE_LIGHT)
        return sphere_light_pdf(v3((T)g.sphere[0], (T)g.sphere[1], (T)g.sphere[2]), (T)g.sphere[3], pos, dir);
    if (g.type != GEOM_MESH) return (T)0;
    V3<T> p0 = vpos<T>(s, g, h.prim, 0), p1 = vpos<T>(s, g, h.prim, 1), p2 = vpos<T>(s, g, h.prim, 2);
    return mesh_light_pdf_u(cross(p1 - p0, p2 - p0), dir, (T)h.t, g.tri_count);
}

/* EvaluateTotalBalanceHeuristic (ray-tracing.cc:1791-1806) */
template <class T> T total_balance(const orc_scene* s, T light_samples, T surface_samples, const Sample<T>& sd, V3<T> pos)
{
    T total = (T)0;
    for (const Light& L : s->lights) {
        T n = is_area_light(L.src.type) ? light_samples : (T)1;
        total += light_pdf(s, L, pos, sd.wi) * n;
    }
    total += bsdf_pdf(sd) * surface_samples;
    return total;
}

/* Hash of a camera sample whose GI splits all share one hit sequence. */
inline uint32_t repeat_hash(uint32_t h, uint32_t n)
{
    uint32_t c = h;
    for (uint32_t j = 1; j < n; ++j) c = (c * 0x9E3779B1u) ^ h;
    return c;
}

inline void hash_hit(uint32_t* h, const tr_hit& hit)
{
    *h = (*h ^ hit.geom) * 16777619u;
    *h = (*h ^ hit.prim) * 16777619u;
}

/* EvaluateLightSources (ray-tracing.cc:1809-1930). */
template <class T>
V3<T> evaluate_lights(const orc_scene* s, const Rng& rng, uint32_t vertex, const Sample<T>& sd,
                      uint32_t max_light_samples, uint32_t max_surface_samples, V3<T> pos, V3<T> throughput,
                      bool apply_balance)
{
    V3<T> radiance = v3((T)0, (T)0, (T)0);
    for (uint32_t li = 0; li < s->lights.size(); ++li) {
        const Light& L = s->lights[li];
        bool unbiased = s->estimator == TR_ESTIMATOR_UNBIASED;
        bool area = multi_sample(L.src.type, unbiased);
        uint32_t samples = area ? max_light_samples : 1;
        V3<T> color = v3((T)0, (T)0, (T)0);
        Sample<T> lsd = sd;
        for (uint32_t k = 0; k < samples; ++k) {
            uint32_t xi[4];
            rng.draw(vertex, STREAM_LIGHT, (li << 16) | k, xi);
            LightSample<T> smp = sample_light(s, L, pos, xi[0], (T)orc_u01(xi[1]), (T)orc_u01(xi[2]), unbiased);
            if (smp.pdf == (T)0) continue;
            lsd.wi = smp.dir;
            V3<T> cur = throughput;
            float of[3] = {(float)pos.x, (float)pos.y, (float)pos.z};
            float df[3] = {(float)smp.dir.x, (float)smp.dir.y, (float)smp.dir.z};
            float dist = (float)smp.dist;
            tr_hit h = trace_closest(s, of, df, 0.0f, dist, false);
            /* occlusion loop: volumes in the way scale the throughput by the
             * Woodcock success ratio (ray-tracing.cc:1845-1926) */
            float tf = 1.0f, step_through = 0.0f;
            uint32_t step = 0, crossings = 0;
            bool reached = false;
            StepDraws dr = rng.steps(vertex, ORC_STREAM_SHADOW_VOLUME);
            T tmax = max_value(throughput);
            for (;;) {
                if (h.geom == TR_INVALID_GEOMETRY || h.geom == smp.geom) { reached = true; break; }
                if (!is_volume_geom(s, h.geom) || ++crossings > kOrcMaxVolumeCrossings) break;
                const GridVolume& g = s->volumes[s->geoms[h.geom].volume];
                uint32_t succ = woodcock_crossing(g, v3(of[0], of[1], of[2]), v3(df[0], df[1], df[2]), h.t, h.u, dr,
                                                  (li << 24) | ((k & 0xffu) << 16), &step, &step_through);
                tf = tf * ((float)succ / 2.0f);
                if ((float)tmax * tf < 1e-3f) break;
                if (step_through >= dist - 1e-3f) { reached = true; break; }
                h = trace_closest(s, of, df, std::min(dist - 1e-3f, step_through), dist, false);
            }
            if (reached) {
                if (apply_balance && area) {
                    T balance;
                    if (unbiased) {
                        /* balance heuristic of this light against the BSDF */
                        T a = (T)samples * smp.pdf;
                        balance = a / (a + (T)max_surface_samples * bsdf_pdf(lsd));
                    } else {
                        T total = total_balance(s, (T)max_light_samples, (T)max_surface_samples, lsd, pos);
                        balance = (light_pdf(s, L, pos, lsd.wi) * (T)samples) / total;
                    }
                    cur *= balance;
                }
                color += (smp.radiance * cur * bsdf_eval(lsd) / smp.pdf) * (T)tf;
            }
        }
        radiance += color / (T)samples;
    }
    return radiance;
}

template <class T> inline V3<T> to_v3(const float* f) { return v3((T)f[0], (T)f[1], (T)f[2]); }

/* What a ray that leaves the scene sees: m_GlobalCubeMap->sampleSpectrum
 * of its direction, or the background spectrum (ray-tracing.cc:2060,
 * 2224). */
template <class T> inline V3<T> env_radiance(const orc_scene* s, const float dir[3])
{
    if (!s->texset.cube_side) return to_v3<T>(s->background);
    float c[3];
    orc_cube_sample(s->texset, dir, c);
    return to_v3<T>(c);
}

/* Spectrum < TransmittanceThreshold: every component below 1e-3
 * (vector3.hh:205, ray-tracing.cc:58). */
template <class T> inline bool below_threshold(V3<T> t)
{
    return t.x < (T)1e-3f && t.y < (T)1e-3f && t.z < (T)1e-3f;
}

/* The shading point of a hit: surfaces through geom_desc->sample and the
 * material cache; a volume collision through the fibre frame at the
 * collision and DirectionalDensity = sigma(ray direction)
 * (ray-tracing.cc:327-337).  Moves *pos to the hit; returns true for a
 * volume. */
template <class T> bool surface_or_volume(const orc_scene* s, const tr_hit& h, V3<T>* pos, Sample<T>* sd)
{
    const Geom& g = s->geoms[h.geom];
    if (g.type == GEOM_VOLUME) {
        const GridVolume& gv = s->volumes[g.volume];
        sd->mat = &s->materials[g.material];
        *pos = *pos + (T)h.t * sd->wi;
        M3<T> b = grid_frame(gv, *pos);
        sd->tangent = b.c0;
        sd->binormal = b.c1;
        sd->normal = b.c2;
        sd->dir_density = sggx_projected_area(spec3<T>(sd->mat->sggx_stddev), b, sd->wi);
        sd->position = *pos;
        return true;
    }
    sample_geometry(s, h, *pos, sd->wi, sd);
    *pos = *pos + (T)h.t * sd->wi;
    sd->position = *pos;
    material_cache(sd);
    return false;
}

/* Cache mode: roughness input (microfacet or Blinn-Phong alpha =
 * sqrt(2 / (power + 2)), the geometric mean of the two alphas of an
 * anisotropic microfacet model, 1 without a specular part or for SGGX surfaces)
 * and the network inputs of a shading point. */
inline float cache_roughness(const tr_material* m)
{
    if (!is_microfacet(m->model) && !is_aniso(m->model) && m->model != TR_MODEL_BLINN_PHONG) return 1.0f;
    if (m->specular[0] == 0.0f && m->specular[1] == 0.0f && m->specular[2] == 0.0f) return 1.0f;
    float a = sqrtf(2.0f / (m->specular_power[0] + 2.0f));
    if (is_aniso(m->model)) a = sqrtf(a * sqrtf(2.0f / (m->specular_power[1] + 2.0f)));
    return a < 1.0f ? a : 1.0f;
}

inline float unit_clamp(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

template <class T> void cache_inputs(const orc_scene* s, V3<T> pos, V3<T> dir, const Sample<T>& sd, float in[10])
{
    float p[3] = {(float)pos.x, (float)pos.y, (float)pos.z};
    for (int a = 0; a < 3; ++a) in[a] = unit_clamp((p[a] - s->cache.box_lo[a]) * s->cache.box_rcp[a]);
    in[3] = (1.0f - (float)dir.x) * 0.5f;
    in[4] = (1.0f - (float)dir.y) * 0.5f;
    in[5] = (1.0f - (float)dir.z) * 0.5f;
    in[6] = ((float)sd.normal.x + 1.0f) * 0.5f;
    in[7] = ((float)sd.normal.y + 1.0f) * 0.5f;
    in[8] = ((float)sd.normal.z + 1.0f) * 0.5f;
    in[9] = cache_roughness(sd.mat);
}

/* The ray from the near to the far plane through screen point (sx, sy):
 * Matrix4 * Vector4 (matrix4.hh:858-880), ToVector3 (vector4.hh:125-130)
 * and the normalized difference. */
void screen_ray(const orc_scene* s, float sx, float sy, float start_f[3], float dir_f[3])
{
    const float* m = s->vpi;
    float in0[4] = {sx, sy, -1.0f, 1.0f}, in1[4] = {sx, sy, 1.0f, 1.0f};
    float p0[4], p1[4];
    for (int r = 0; r < 4; ++r) {
        p0[r] = m[r] * in0[0]; p1[r] = m[r] * in1[0];
        p0[r] += m[4 + r] * in0[1]; p1[r] += m[4 + r] * in1[1];
        p0[r] += m[8 + r] * in0[2]; p1[r] += m[8 + r] * in1[2];
        p0[r] += m[12 + r] * in0[3]; p1[r] += m[12 + r] * in1[3];
    }
    float w0 = 1.0f / p0[3], w1 = 1.0f / p1[3];
    float end_f[3] = {p1[0] * w1, p1[1] * w1, p1[2] * w1};
    start_f[0] = p0[0] * w0; start_f[1] = p0[1] * w0; start_f[2] = p0[2] * w0;
    for (int a = 0; a < 3; ++a) dir_f[a] = end_f[a] - start_f[a];
    float l = sqrtf(dir_f[0] * dir_f[0] + dir_f[1] * dir_f[1] + dir_f[2] * dir_f[2]);
    if (l != 0.0f) { float r = 1.0f / l; dir_f[0] *= r; dir_f[1] *= r; dir_f[2] *= r; }
}

/* The camera ray of pixel (x, y) for the sample whose Rng is `rng`
 * (subdraw, ray-tracing.cc:2011-2037): jittered when SamplesCamera > 1,
 * else through the pixel's corner grid point. */
void camera_ray(const orc_scene* s, const Rng& rng, uint32_t x, uint32_t y, float start_f[3], float dir_f[3])
{
    float sx, sy;
    if (s->spp > 1) {
        uint32_t xi[4];
        rng.draw(0, STREAM_CAMERA, 0, xi);
        sx = 2.0f * ((float)x + orc_u01(xi[0])) / (float)s->width - 1.0f;
        sy = 2.0f * ((float)y + orc_u01(xi[1])) / (float)s->height - 1.0f;
    } else {
        sx = 2.0f * (float)x / (float)(s->width - 1) - 1.0f;
        sy = 2.0f * (float)y / (float)(s->height - 1) - 1.0f;
    }
    screen_ray(s, sx, sy, start_f, dir_f);
}

/* RenderMode: the tangent, binormal or normal * 0.5 + 0.5, or
 * max(0, N . -dir) (ray-tracing.cc:2078-2085). */
template <class T> V3<T> debug_color(uint32_t mode, V3<T> t, V3<T> b, V3<T> n, V3<T> dir)
{
    if (mode == 4) {
        T c = dot(n, -dir);
        c = (T)0 > c ? (T)0 : c;
        return v3(c, c, c);
    }
    V3<T> v = mode == 1 ? t : (mode == 2 ? b : n);
    return v * (T)0.5f + v3((T)0.5f, (T)0.5f, (T)0.5f);
}

/* One camera sample of one pixel: the body of the camera loop of subdraw
 * (ray-tracing.cc:2011-2345) for surfaces. */
template <class T>
V3<T> render_camera_sample(const orc_scene* s, uint32_t pixel, uint32_t cs, uint32_t* hash_out, uint64_t pseq = 0)
{
    /* cache mode: training path iff (path index + frame) % stride == 0 */
    const bool cache = s->cache.enabled;
    uint32_t record = 0xffffffffu;
    if (cache) {
        uint64_t path = pseq * s->spp + (cs - s->cache.frame * s->spp) + s->cache.frame;
        if (path % s->cache.train_stride == 0) record = (uint32_t)(path / s->cache.train_stride);
    }
    bool rough = false, recorded = false;
    float rec_in[10];
    V3<T> rec_thr = v3((T)0, (T)0, (T)0), rec_before = rec_thr;
    uint32_t x = pixel % s->width, y = pixel / s->width;
    Rng rng;
    rng.key[0] = s->key[0]; rng.key[1] = s->key[1];
    rng.pixel = pixel + s->pixel_key_base; rng.sample = cs; rng.split = 0;
    uint32_t hash = 2166136261u;
    float start_f[3], dir_f[3];
    camera_ray(s, rng, x, y, start_f, dir_f);
    V3<T> pos = to_v3<T>(start_f);
    Sample<T> sd;
    sd.wi = to_v3<T>(dir_f);
    sd.dir_density = (T)0;
    sd.total_density = (T)0;
    float of[3] = {(float)pos.x, (float)pos.y, (float)pos.z};
    float df[3] = {(float)sd.wi.x, (float)sd.wi.y, (float)sd.wi.z};
    tr_hit h = trace_segment(s, of, df, rng, 0);
    if (h.geom == TR_INVALID_GEOMETRY) {
        hash = (hash ^ 0xffffffffu) * 16777619u;
        *hash_out = repeat_hash(hash, s->n_gi);
        return env_radiance<T>(s, df);
    }
    hash_hit(&hash, h);
    if (s->geoms[h.geom].material == ORC_NO_MATERIAL) {
        /* a blocker ends the camera sample (ray-tracing.cc:2072-2077); its
         * sample() frame is zero (:1239-1244) */
        *hash_out = repeat_hash(hash, s->n_gi);
        V3<T> z = v3((T)0, (T)0, (T)0);
        return s->render_mode ? debug_color<T>(s->render_mode, z, z, z, sd.wi) : z;
    }
    bool volume = surface_or_volume(s, h, &pos, &sd);
    if (s->render_mode && !volume) {
        /* RenderMode debug output at the camera vertex (ray-tracing.cc:2078-2089) */
        *hash_out = repeat_hash(hash, s->n_gi);
        return debug_color<T>(s->render_mode, sd.tangent, sd.binormal, sd.normal, sd.wi);
    }
    V3<T> start_T = v3((T)1, (T)1, (T)1);
    if (volume) {
        /* albedo * density / density (ray-tracing.cc:2134-2142) */
        start_T *= spec3<T>(sd.mat->albedo);
        if (below_threshold(start_T)) {
            *hash_out = repeat_hash(hash, s->n_gi);
            return v3((T)0, (T)0, (T)0);
        }
    }
    sd.wo = -sd.wi;
    if (is_emissive(sd.mat->model)) {
        *hash_out = repeat_hash(hash, s->n_gi);
        return start_T * spec3<T>(sd.mat->diffuse);
    }
    V3<T> local = evaluate_lights(s, rng, 0, sd, s->n_light, s->n_gi, pos, start_T, s->max_depth != 0 && !volume);
    V3<T> global = v3((T)0, (T)0, (T)0);
    const bool unbiased = s->estimator == TR_ESTIMATOR_UNBIASED;
    uint32_t camera_hash = hash;
    uint32_t combined = 0;
    uint32_t rays_sq = (uint32_t)sqrtf((float)s->n_gi);
    for (uint32_t split = 0; split < s->n_gi; ++split) {
        /* each split keeps its own radiance sum and hit hash; the GPU kernels
         * combine them in the same order */
        V3<T> gsplit = v3((T)0, (T)0, (T)0);
        hash = camera_hash;
        Rng srng = rng;
        srng.split = split;
        Sample<T> cur = sd;
        bool cur_volume = volume;
        uint32_t ss = s->n_gi, ls = s->n_light;
        V3<T> cpos = pos;
        V3<T> ct = start_T;
        Strata<T> st;
        st.x = (T)(float)(split % rays_sq);
        st.y = (T)(float)(split / rays_sq);
        st.tx = st.ty = (T)(float)rays_sq;
        for (uint32_t d = 0; d < s->max_depth; ++d) {
            uint32_t xr[4];
            srng.draw(d, STREAM_BSDF, 0, xr);
            T xi[5] = {(T)orc_u01(xr[0]), (T)orc_u01(xr[1]), (T)orc_u01(xr[2]), (T)0, (T)0};
            if (is_aniso(cur.mat->model)) {
                /* the anisotropic specular lobe's second pair: sub-counter 1 */
                srng.draw(d, STREAM_BSDF, 1, xr);
                xi[3] = (T)orc_u01(xr[0]);
                xi[4] = (T)orc_u01(xr[1]);
            }
            bsdf_sample(st, &cur, xi);
            ORC_TRACE("d=%u model=%u n=(%g %g %g) wo=(%g %g %g) xi=(%g %g %g) wi=(%g %g %g) pdf=%g\n", d,
                      cur.mat->model, (double)cur.normal.x, (double)cur.normal.y, (double)cur.normal.z,
                      (double)cur.wo.x, (double)cur.wo.y, (double)cur.wo.z, (double)xi[0], (double)xi[1],
                      (double)xi[2], (double)cur.wi.x, (double)cur.wi.y, (double)cur.wi.z, (double)cur.pdf);
            if (cur.pdf == (T)0) break;
            T balance = (T)1;
            if (!cur_volume) {
                if (unbiased) {
                    balance = (T)ss * cur.pdf;   /* the BSDF side of the emitter-hit weight */
                } else {
                    T total = total_balance(s, (T)ls, (T)ss, cur, cpos);
                    balance = total > (T)0 ? ((T)ss * cur.pdf) / total : (T)0;
                }
                ct *= bsdf_eval(cur) / cur.pdf;
                ORC_TRACE("   balance=%g T=(%g %g %g) f=(%g %g %g)\n", (double)balance, (double)ct.x, (double)ct.y,
                          (double)ct.z, (double)bsdf_eval(cur).x, (double)bsdf_eval(cur).y, (double)bsdf_eval(cur).z);
            }
            if (cache && !recorded && cache_roughness(cur.mat) >= s->cache.rough_alpha) rough = true;
            float o2[3] = {(float)cpos.x, (float)cpos.y, (float)cpos.z};
            float d2[3] = {(float)cur.wi.x, (float)cur.wi.y, (float)cur.wi.z};
            tr_hit hh = trace_segment(s, o2, d2, srng, d + 1);
            if (hh.geom == TR_INVALID_GEOMETRY) {
                hash = (hash ^ 0xffffffffu) * 16777619u;
                gsplit += env_radiance<T>(s, d2) * ct;
                break;
            }
            hash_hit(&hash, hh);
            ORC_TRACE("   hit geom=%u prim=%u t=%g\n", hh.geom, hh.prim, (double)hh.t);
            if (s->geoms[hh.geom].material == ORC_NO_MATERIAL) break;  /* blocker (ray-tracing.cc:2236-2240) */
            V3<T> prev_pos = cpos;
            cur_volume = surface_or_volume(s, hh, &cpos, &cur);
            if (cur_volume) {
                ct *= spec3<T>(cur.mat->albedo);
                if (below_threshold(ct)) break;
            }
            if (is_emissive(cur.mat->model)) {
                T w = balance;
                if (unbiased) {
                    T lp = hit_light_pdf(s, hh, prev_pos, cur.wi);
                    T ls_prev = d == 0 ? (T)s->n_light : (T)1;
                    w = balance / (balance + ls_prev * lp);
                }
                gsplit += ct * w * spec3<T>(cur.mat->diffuse);
                break;
            }
            cur.wo = -cur.wi;
            ss = 1;
            ls = 1;
            bool capture = false;
            if (cache && rough && d + 1 < s->max_depth) {
                /* the query vertex: the constant prediction (of the radiance
                 * leaving the vertex, or with query_nee of the radiance
                 * beyond its next-event estimate, which then follows; the
                 * GPU adds them in this order) ends the path; a training
                 * path records the vertex and goes on */
                if (record == 0xffffffffu) {
                    gsplit += ct * v3((T)xmax(s->cache.value[0], 0.0f), (T)xmax(s->cache.value[1], 0.0f),
                                      (T)xmax(s->cache.value[2], 0.0f));
                    if (s->cache.query_nee)
                        gsplit += evaluate_lights(s, srng, d + 1, cur, ls, ss, cpos, ct, d + 1 != s->max_depth);
                    break;
                }
                cache_inputs(s, cpos, cur.wi, cur, rec_in);
                rec_thr = ct;
                rough = false;
                recorded = true;
                if (s->cache.query_nee) capture = true;
                else rec_before = gsplit;
            }
            gsplit += evaluate_lights(s, srng, d + 1, cur, ls, ss, cpos, ct, d + 1 != s->max_depth);
            if (capture) rec_before = gsplit;
            ct *= (unbiased ? (T)1 : balance) / (T)s->rr;
            st.x = st.y = (T)0;
            st.tx = st.ty = (T)1;
            if (s->rr < 1.0f) {
                uint32_t xq[4];
                srng.draw(d + 1, STREAM_RR, 0, xq);
                if ((T)orc_u01(xq[0]) > (T)s->rr) break;
            }
        }
        if (cache && record != 0xffffffffu && record < s->cache.valid.size()) {
            orc_scene* ms = const_cast<orc_scene*>(s);
            float* tg = &ms->cache.target[(size_t)record * 3];
            float vt = 0.0f;
            if (recorded) {
                V3<T> after = gsplit;
                T th[3] = {rec_thr.x, rec_thr.y, rec_thr.z}, g1[3] = {after.x, after.y, after.z},
                  g0[3] = {rec_before.x, rec_before.y, rec_before.z};
                bool ok = true;
                for (int c = 0; c < 3; ++c) {
                    ok = ok && (float)th[c] > 1e-8f;
                    tg[c] = ok ? ((float)g1[c] - (float)g0[c]) / (float)th[c] : 0.0f;
                }
                vt = ok ? 1.0f : 0.0f;
                memcpy(&ms->cache.in[(size_t)record * 10], rec_in, sizeof(rec_in));
            } else {
                tg[0] = tg[1] = tg[2] = 0.0f;
            }
            ms->cache.valid[record] = vt;
        }
        if (split == 0) { global = gsplit; combined = hash; }
        else { global += gsplit; combined = (combined * 0x9E3779B1u) ^ hash; }
    }
    hash = combined;
    global = global / (T)s->n_gi;
    *hash_out = hash;
    return local + global;
}

template <class F> void parallel_for(uint32_t n, int nthreads, uint32_t chunk, F fn)
{
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next(0);
    auto worker = [&]() {
        for (;;) {
            uint32_t b = next.fetch_add(chunk);
            if (b >= n) break;
            uint32_t e = std::min(n, b + chunk);
            for (uint32_t i = b; i < e; ++i) fn(i);
        }
    };
    if (nthreads == 1) { worker(); return; }
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
}

}  // namespace

extern "C" {

orc_scene* orc_scene_create(uint32_t width, uint32_t height, const float vpi[16])
{
    orc_scene* s = new orc_scene();
    s->width = width;
    s->height = height;
    memcpy(s->vpi, vpi, sizeof(s->vpi));
    s->materials.reserve(1024);
    return s;
}

void orc_scene_destroy(orc_scene* s) { delete s; }
void orc_set_samples_camera(orc_scene* s, uint32_t n) { s->spp = n; }
void orc_set_samples_global_illumination(orc_scene* s, uint32_t n) { s->n_gi = n; }
void orc_set_samples_local_area_light(orc_scene* s, uint32_t n) { s->n_light = n; }
void orc_set_max_ray_depth(orc_scene* s, uint32_t d) { s->max_depth = d; }
void orc_set_russian_roulette(orc_scene* s, float rr) { s->rr = rr; }
void orc_set_background_spectrum(orc_scene* s, const float rgb[3]) { memcpy(s->background, rgb, 12); }
void orc_set_seed(orc_scene* s, uint64_t seed) { s->key[0] = (uint32_t)seed; s->key[1] = (uint32_t)(seed >> 32); }
void orc_set_pixel_key_base(orc_scene* s, uint32_t base) { s->pixel_key_base = base; }
void orc_set_estimator(orc_scene* s, uint32_t mode) { s->estimator = mode; }
void orc_set_render_mode(orc_scene* s, uint32_t mode) { s->render_mode = mode; }

/* Per pixel of frame `frame`: 1 if one of its camera rays hit geometry
 * (after tracking through volumes), the reference's `intersections`
 * behind TransparentBackground (ray-tracing.cc:2062-2066, 2364). */
void orc_camera_hits(const orc_scene* s, uint32_t frame, uint8_t* out)
{
    const uint32_t n = s->width * s->height;
    parallel_for(n, 0, 64, [&](uint32_t p) {
        uint8_t any = 0;
        for (uint32_t k = 0; k < s->spp && !any; ++k) {
            Rng rng;
            rng.key[0] = s->key[0]; rng.key[1] = s->key[1];
            rng.pixel = p + s->pixel_key_base; rng.sample = frame * s->spp + k; rng.split = 0;
            float o[3], d[3];
            camera_ray(s, rng, p % s->width, p / s->width, o, d);
            any = trace_segment(s, o, d, rng, 0).geom != TR_INVALID_GEOMETRY;
        }
        out[p] = any;
    });
}

/* rayQuery (ray-tracing.cc:1932-1979): out = t, normal, tangent, binormal,
 * TexCoord; returns the geometry slot or TR_INVALID_GEOMETRY. */
uint32_t orc_ray_query(const orc_scene* s, const float tc[2], float out[12])
{
    float o[3], d[3];
    screen_ray(s, 2.0f * tc[0] - 1.0f, 1.0f - 2.0f * tc[1], o, d);
    tr_hit h = trace_closest(s, o, d, 0.0f, INFINITY, false);
    for (int k = 0; k < 12; ++k) out[k] = 0.0f;
    if (h.geom == TR_INVALID_GEOMETRY) return h.geom;
    out[0] = h.t;
    const Geom& g = s->geoms[h.geom];
    if (g.material != ORC_NO_MATERIAL && g.type != GEOM_VOLUME) {
        Sample<float> sd{};
        sample_geometry<float>(s, h, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), &sd);
        V3<float> f[3] = {sd.normal, sd.tangent, sd.binormal};
        for (int k = 0; k < 3; ++k) { out[1 + 3 * k] = f[k].x; out[2 + 3 * k] = f[k].y; out[3 + 3 * k] = f[k].z; }
        out[10] = sd.tcx;
        out[11] = sd.tcy;
    }
    return h.geom;
}

/* Cache mode with a constant network output `value`; the scene box of the
 * position input is the box of all primitives (set at commit).  Training
 * records for `records` training paths are kept for orc_cache_records. */
void orc_set_cache(orc_scene* s, int enabled, float rough_alpha, uint32_t train_stride, uint32_t frame,
                   const float value[3], uint32_t records, int query_nee)
{
    s->cache.enabled = enabled != 0;
    s->cache.query_nee = query_nee != 0;
    s->cache.rough_alpha = rough_alpha;
    s->cache.train_stride = train_stride ? train_stride : 1;
    s->cache.frame = frame;
    memcpy(s->cache.value, value, sizeof(s->cache.value));
    s->cache.in.assign((size_t)records * 10, 0.0f);
    s->cache.target.assign((size_t)records * 3, 0.0f);
    s->cache.valid.assign(records, 0.0f);
}

void orc_cache_records(const orc_scene* s, size_t count, float* in, float* target, float* valid)
{
    for (size_t r = 0; r < count && r < s->cache.valid.size(); ++r) {
        memcpy(in + r * 10, &s->cache.in[r * 10], 10 * sizeof(float));
        memcpy(target + r * 3, &s->cache.target[r * 3], 3 * sizeof(float));
        valid[r] = s->cache.valid[r];
    }
}

int orc_add_triangle_mesh(orc_scene* s, const float world[16], size_t submesh_count, const tr_submesh* subs,
                          size_t index_count, const int32_t* tris, size_t vert_size, const void* verts,
                          const tr_mesh_options* opts, uint64_t* geom_ids)
{
    if (s->committed) return -1;
    if (s->materials.size() + submesh_count > s->materials.capacity()) return -1; /* keep pointers stable */
    uint32_t vb = (uint32_t)s->vbufs.size();
    s->vbufs.emplace_back((const uint8_t*)verts, (const uint8_t*)verts + vert_size);
    uint32_t inst = s->instances++;
    for (size_t si = 0; si < submesh_count; ++si) {
        const tr_submesh& sm = subs[si];
        Geom g;
        memset(&g, 0, sizeof(g));
        g.type = GEOM_MESH;
        g.material = intern_material(s, sm.material);
        g.two_sided = opts && opts->two_sided ? 1 : 0;
        g.vbuf = vb;
        g.vertex_offset = sm.vertex_offset;
        g.stride = sm.stride;
        g.index_base = (uint32_t)s->indices.size();
        g.tri_count = sm.vertex_count / 3;
        s->indices.insert(s->indices.end(), tris + sm.base_index, tris + sm.base_index + (size_t)g.tri_count * 3);
        /* world transform applied to positions (column-major 4x4) */
        const uint8_t* base = s->vbufs[vb].data() + sm.vertex_offset;
        uint32_t slot = (uint32_t)s->geoms.size();
        for (uint32_t t = 0; t < g.tri_count; ++t) {
            Prim p;
            memset(&p, 0, sizeof(p));
            float* dst[3] = {p.v0, p.v1, p.v2};
            for (int k = 0; k < 3; ++k) {
                const float* f = reinterpret_cast<const float*>(base + (size_t)s->indices[g.index_base + t * 3 + k] * sm.stride);
                for (int r = 0; r < 3; ++r)
                    dst[k][r] = ((world[r] * f[0] + world[4 + r] * f[1]) + world[8 + r] * f[2]) + world[12 + r];
            }
            p.geom = slot;
            p.prim = t;
            p.two_sided = g.two_sided;
            s->prims.push_back(p);
        }
        s->geoms.push_back(g);
        if (s->materials[g.material].model == TR_MODEL_EMISSIVE) {
            Light L;
            memset(&L, 0, sizeof(L));
            L.src.type = TR_LIGHT_MESH_AREA;
            L.geom = slot;
            s->lights.push_back(L);
        }
        if (geom_ids) geom_ids[si] = ((uint64_t)inst << 32) | (uint64_t)si;
    }
    (void)index_count;
    return 0;
}

static uint32_t add_sphere_geom(orc_scene* s, const float c[3], float r, uint32_t type, uint32_t material, uint32_t light)
{
    Geom g;
    memset(&g, 0, sizeof(g));
    g.type = type;
    g.material = material;
    g.sphere[0] = c[0]; g.sphere[1] = c[1]; g.sphere[2] = c[2]; g.sphere[3] = r;
    g.light = light;
    uint32_t slot = (uint32_t)s->geoms.size();
    s->geoms.push_back(g);
    Prim p;
    memset(&p, 0, sizeof(p));
    p.v0[0] = c[0]; p.v0[1] = c[1]; p.v0[2] = c[2];
    p.v1[0] = r;
    p.geom = slot;
    p.prim = 0;
    p.is_sphere = 1;
    s->prims.push_back(p);
    return slot;
}

int orc_add_sphere(orc_scene* s, const float center[3], float radius, const tr_material* m, uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    uint32_t slot = add_sphere_geom(s, center, radius, GEOM_SPHERE, intern_material(s, m), 0);
    if (out) *out = slot;
    return 0;
}

/* ---- user shapes (ray-tracing.cc:1315-1446), parameters in FP32 ---- */

static uint32_t add_shape_geom(orc_scene* s, const OrcShape& sh, uint32_t material, const double lo[3],
                               const double hi[3])
{
    Geom g;
    memset(&g, 0, sizeof(g));
    g.type = GEOM_SHAPE;
    g.material = material;
    g.shape = (uint32_t)s->shapes.size();
    s->shapes.push_back(sh);
    uint32_t slot = (uint32_t)s->geoms.size();
    s->geoms.push_back(g);
    Prim p;
    memset(&p, 0, sizeof(p));
    for (int a = 0; a < 3; ++a) {
        double pad = 1e-6 * (fabs(lo[a]) + fabs(hi[a]) + (hi[a] - lo[a])) + 1e-30;
        p.v1[a] = std::nextafter((float)(lo[a] - pad), -INFINITY);
        p.v2[a] = std::nextafter((float)(hi[a] + pad), INFINITY);
    }
    p.geom = slot;
    p.prim = 0;
    p.is_shape = 1;
    s->prims.push_back(p);
    return slot;
}

/* Orientation = ToQuaternion(Matrix3(tan, norm x tan, norm))
 * (MatrixToQuaternion, quaternion.hh:59-70), then ToMatrix3 of it. */
static int add_rect_geom(orc_scene* s, const float pos[3], const float tan[3], const float norm[3], const float size[2],
                         uint32_t material, const float tc[4], uint32_t tangent_map, uint64_t* out)
{
    float b[3];
    cross3f(norm, tan, b);
    const float* c0 = tan;
    const float* c1 = b;
    const float* c2 = norm;
    float q[4];
    q[3] = sqrtf(std::max(0.0f, 1.0f + c0[0] + c1[1] + c2[2])) / 2.0f;
    q[0] = sqrtf(std::max(0.0f, 1.0f + c0[0] - c1[1] - c2[2])) / 2.0f;
    q[1] = sqrtf(std::max(0.0f, 1.0f - c0[0] + c1[1] - c2[2])) / 2.0f;
    q[2] = sqrtf(std::max(0.0f, 1.0f - c0[0] - c1[1] + c2[2])) / 2.0f;
    q[0] = std::copysign(q[0], c1[2] - c2[1]);
    q[1] = std::copysign(q[1], c2[0] - c0[2]);
    q[2] = std::copysign(q[2], c0[1] - c1[0]);
    M3<float> o = quat_to_m3(q);
    OrcShape sh;
    memset(&sh, 0, sizeof(sh));
    sh.kind = ORC_RECT;
    sh.tangent_map = tangent_map;
    for (int a = 0; a < 3; ++a) sh.f[a] = pos[a];
    V3<float> cols[3] = {o.c0, o.c1, o.c2};
    for (int k = 0; k < 3; ++k) {
        sh.f[3 + 3 * k] = cols[k].x;
        sh.f[4 + 3 * k] = cols[k].y;
        sh.f[5 + 3 * k] = cols[k].z;
    }
    sh.f[12] = size[0];
    sh.f[13] = size[1];
    if (tc) {
        sh.f[14] = tc[0]; sh.f[15] = tc[1];
        sh.f[16] = tc[2] - tc[0]; sh.f[17] = tc[3] - tc[1];
    } else {
        sh.f[16] = 1.0f; sh.f[17] = 1.0f;
    }
    /* Rect3Bounds (shapes.hh:260-266) */
    double lo[3], hi[3];
    double t[3] = {o.c0.x, o.c0.y, o.c0.z}, bn[3] = {o.c1.x, o.c1.y, o.c1.z};
    for (int a = 0; a < 3; ++a) {
        double e = fabs(size[0] * t[a]) + fabs(size[1] * bn[a]);
        lo[a] = pos[a] - e;
        hi[a] = pos[a] + e;
    }
    uint32_t slot = add_shape_geom(s, sh, material, lo, hi);
    if (out) *out = slot;
    return 0;
}

int orc_add_rect(orc_scene* s, const float pos[3], const float tan[3], const float norm[3], const float size[2],
                 const tr_material* m, const float tc[4], uint32_t tangent_map, uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    return add_rect_geom(s, pos, tan, norm, size, intern_material(s, m), tc, tangent_map, out);
}

int orc_add_blocker(orc_scene* s, const float pos[3], const float tan[3], const float norm[3], const float size[2],
                    uint64_t* out)
{
    if (s->committed) return -1;
    return add_rect_geom(s, pos, tan, norm, size, ORC_NO_MATERIAL, nullptr, 0, out);
}

int orc_add_disk(orc_scene* s, const float pos[3], const float norm[3], float inner, float outer, const tr_material* m,
                 uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    OrcShape sh;
    memset(&sh, 0, sizeof(sh));
    sh.kind = ORC_DISK;
    for (int a = 0; a < 3; ++a) { sh.f[a] = pos[a]; sh.f[3 + a] = norm[a]; }
    sh.f[6] = inner;
    sh.f[7] = outer;
    double nl = sqrt((double)norm[0] * norm[0] + (double)norm[1] * norm[1] + (double)norm[2] * norm[2]);
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        double na = norm[a] / nl, e = outer * sqrt(std::max(0.0, 1.0 - na * na));
        lo[a] = pos[a] - e;
        hi[a] = pos[a] + e;
    }
    uint32_t slot = add_shape_geom(s, sh, intern_material(s, m), lo, hi);
    if (out) *out = slot;
    return 0;
}

static int add_cylinder_geom(orc_scene* s, const float c[3], float r, float hh, const float* tilt_dir, float tilt,
                             const tr_material* m, uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    OrcShape sh;
    memset(&sh, 0, sizeof(sh));
    sh.kind = tilt_dir ? ORC_OBLIQUE : ORC_CYLINDER;
    for (int a = 0; a < 3; ++a) sh.f[a] = c[a];
    sh.f[3] = r;
    sh.f[4] = hh;
    double ex = 0.0, ey = 0.0;
    if (tilt_dir) {
        /* bending = Tilt * TiltDirection / HalfHeight (ray-tracing.cc:999) */
        sh.f[5] = (tilt * tilt_dir[0]) / hh;
        sh.f[6] = (tilt * tilt_dir[1]) / hh;
        ex = fabs((double)tilt * tilt_dir[0]);
        ey = fabs((double)tilt * tilt_dir[1]);
    }
    double lo[3] = {c[0] - r - ex, c[1] - r - ey, (double)c[2] - hh};
    double hi[3] = {c[0] + r + ex, c[1] + r + ey, (double)c[2] + hh};
    uint32_t slot = add_shape_geom(s, sh, intern_material(s, m), lo, hi);
    if (out) *out = slot;
    return 0;
}

int orc_add_cylinder(orc_scene* s, const float c[3], float r, float hh, const tr_material* m, uint64_t* out)
{
    return add_cylinder_geom(s, c, r, hh, nullptr, 0.0f, m, out);
}

int orc_add_oblique_cylinder(orc_scene* s, const float c[3], float r, float hh, const float tilt_dir[2], float tilt,
                             const tr_material* m, uint64_t* out)
{
    return add_cylinder_geom(s, c, r, hh, tilt_dir, tilt, m, out);
}

int orc_add_ellipsoid(orc_scene* s, const float c[3], const float scale[3], const float q[4], const tr_material* m,
                      uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    OrcShape sh;
    memset(&sh, 0, sizeof(sh));
    sh.kind = ORC_ELLIPSOID;
    for (int a = 0; a < 3; ++a) { sh.f[a] = c[a]; sh.f[3 + a] = scale[a]; }
    for (int k = 0; k < 4; ++k) sh.f[6 + k] = q[k];
    sh.f[10] = -q[0]; sh.f[11] = -q[1]; sh.f[12] = -q[2]; sh.f[13] = q[3];   /* Inverse = Conjugate */
    {   /* ToNormal (quaternion.hh:170-181) */
        float x = q[0], y = q[1], z = q[2], w = q[3];
        float xx = x * x, xz = x * z, xw = x * w, yy = y * y, yz = y * z, yw = y * w;
        sh.f[14] = 2.0f * (xz + yw);
        sh.f[15] = 2.0f * (yz - xw);
        sh.f[16] = 1.0f - 2.0f * (xx + yy);
    }
    /* box: the surface is c + M^-1 S u, |u| = 1, with M v = Transform(conj q,
     * v); the half extent along axis i is the norm of row i of M^-1 S */
    double Mm[3][3];
    for (int j = 0; j < 3; ++j) {
        double e[3] = {0, 0, 0};
        e[j] = 1.0;
        double qv[3] = {-(double)q[0], -(double)q[1], -(double)q[2]}, w = q[3];
        double c1[3] = {qv[1] * e[2] - qv[2] * e[1] + w * e[0], qv[2] * e[0] - qv[0] * e[2] + w * e[1],
                        qv[0] * e[1] - qv[1] * e[0] + w * e[2]};
        double c2[3] = {qv[1] * c1[2] - qv[2] * c1[1], qv[2] * c1[0] - qv[0] * c1[2], qv[0] * c1[1] - qv[1] * c1[0]};
        for (int i = 0; i < 3; ++i) Mm[i][j] = e[i] + 2.0 * c2[i];
    }
    double det = Mm[0][0] * (Mm[1][1] * Mm[2][2] - Mm[1][2] * Mm[2][1]) -
                 Mm[0][1] * (Mm[1][0] * Mm[2][2] - Mm[1][2] * Mm[2][0]) +
                 Mm[0][2] * (Mm[1][0] * Mm[2][1] - Mm[1][1] * Mm[2][0]);
    if (!(fabs(det) > 1e-12)) return -1;
    double inv[3][3] = {{(Mm[1][1] * Mm[2][2] - Mm[1][2] * Mm[2][1]) / det, (Mm[0][2] * Mm[2][1] - Mm[0][1] * Mm[2][2]) / det,
                         (Mm[0][1] * Mm[1][2] - Mm[0][2] * Mm[1][1]) / det},
                        {(Mm[1][2] * Mm[2][0] - Mm[1][0] * Mm[2][2]) / det, (Mm[0][0] * Mm[2][2] - Mm[0][2] * Mm[2][0]) / det,
                         (Mm[0][2] * Mm[1][0] - Mm[0][0] * Mm[1][2]) / det},
                        {(Mm[1][0] * Mm[2][1] - Mm[1][1] * Mm[2][0]) / det, (Mm[0][1] * Mm[2][0] - Mm[0][0] * Mm[2][1]) / det,
                         (Mm[0][0] * Mm[1][1] - Mm[0][1] * Mm[1][0]) / det}};
    double lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        double e2 = 0.0;
        for (int j = 0; j < 3; ++j) e2 += (inv[i][j] * scale[j]) * (inv[i][j] * scale[j]);
        double e = sqrt(e2) * (1.0 + 1e-5);
        lo[i] = c[i] - e;
        hi[i] = c[i] + e;
    }
    uint32_t slot = add_shape_geom(s, sh, intern_material(s, m), lo, hi);
    if (out) *out = slot;
    return 0;
}

/* A grid volume (the build's addGridVolume): samples decoded from
 * binary16, majorant = largest density (or max_density) * max(stddev). */
int orc_add_grid_volume(orc_scene* s, const tr_grid_volume* v, const tr_material* m, uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    GridVolume g;
    size_t n = (size_t)v->dims[0] * v->dims[1] * v->dims[2];
    g.rho.resize(n);
    g.dir.resize(n * 3);
    float rho_max = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        g.rho[i] = half_bits_to_float(v->density[i]);
        rho_max = std::max(rho_max, g.rho[i]);
        for (int a = 0; a < 3; ++a) g.dir[i * 3 + a] = half_bits_to_float(v->direction[i * 3 + a]);
    }
    if (v->max_density != 0.0f) rho_max = v->max_density;
    float smax = std::max(m->sggx_stddev[0], std::max(m->sggx_stddev[1], m->sggx_stddev[2]));
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = v->min_corner[a];
        g.hi[a] = v->max_corner[a];
        g.scale[a] = (float)(v->dims[a] - 1) / (v->max_corner[a] - v->min_corner[a]);
        g.dims[a] = v->dims[a];
        g.stddev[a] = m->sggx_stddev[a];
    }
    g.rcp_rho_max = 1.0f / rho_max;
    g.rcp_majorant = 1.0f / (rho_max * smax);
    g.rcp_smax = 1.0f / smax;
    Geom gm;
    memset(&gm, 0, sizeof(gm));
    gm.type = GEOM_VOLUME;
    gm.material = intern_material(s, m);
    gm.volume = (uint32_t)s->volumes.size();
    uint32_t slot = (uint32_t)s->geoms.size();
    s->geoms.push_back(gm);
    s->volumes.push_back(std::move(g));
    Prim p;
    memset(&p, 0, sizeof(p));
    for (int a = 0; a < 3; ++a) { p.v0[a] = v->min_corner[a]; p.v1[a] = v->max_corner[a]; }
    p.geom = slot;
    p.prim = 0;
    p.is_volume = 1;
    s->prims.push_back(p);
    if (out) *out = slot;
    return 0;
}

/* addHierarchicalVolume (ray-tracing.cc:1448-1504): one grid per non-empty
 * brick at Box{x, y, z} * vol_size + MinCorner, u8 densities, directions
 * (cos phi sin theta, sin phi sin theta, cos theta) from the s_cosPhi /
 * s_sinPhi / s_cosTheta / s_sinTheta tables (:777-788), RcpMaxExtinction =
 * 1 / DensityScale (:60, 1463). */
int orc_add_hierarchical_volume(orc_scene* s, const tr_volume_root* root, const tr_material* m, uint64_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    float cos_phi[256], sin_phi[256], cos_theta[256], sin_theta[256];
    for (int i = 0; i < 255; ++i) {
        float angle = (float)i * (3.141592f / 255.0f);
        cos_phi[i] = std::cos(2.0f * angle);
        sin_phi[i] = std::sin(2.0f * angle);
        cos_theta[i] = std::cos(angle);
        sin_theta[i] = std::sin(angle);
    }
    cos_phi[255] = sin_phi[255] = cos_theta[255] = sin_theta[255] = 0.0f;
    float vs[3];
    for (int a = 0; a < 3; ++a) vs[a] = (root->max_corner[a] - root->min_corner[a]) / (float)root->dims[a];
    uint64_t first = UINT64_MAX;
    const size_t nb = (size_t)root->dims[0] * root->dims[1] * root->dims[2];
    const uint32_t mat = intern_material(s, m);
    for (size_t b = 0; b < nb; ++b) {
        const tr_volume_brick& br = root->bricks[b];
        if (!br.data) continue;
        size_t idx[3] = {b % (size_t)root->dims[0], (b / (size_t)root->dims[0]) % (size_t)root->dims[1],
                         b / ((size_t)root->dims[0] * root->dims[1])};
        GridVolume g;
        g.brick = 1;
        size_t n = (size_t)br.dims[0] * br.dims[1] * br.dims[2];
        g.rho.resize(n);
        g.dir.resize(3 * n);
        for (size_t i = 0; i < n; ++i) {
            g.rho[i] = (float)br.data[i];
            uint8_t th = br.data[n + 2 * i], ph = br.data[n + 2 * i + 1];
            g.dir[3 * i + 0] = cos_phi[ph] * sin_theta[th];
            g.dir[3 * i + 1] = sin_phi[ph] * sin_theta[th];
            g.dir[3 * i + 2] = cos_theta[th];
        }
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = (float)idx[a] * vs[a] + root->min_corner[a];
            g.hi[a] = g.lo[a] + vs[a];
            g.dims[a] = (uint32_t)br.dims[a];
            g.scale[a] = (float)(br.dims[a] - 1) / (g.hi[a] - g.lo[a]);
            g.stddev[a] = m->sggx_stddev[a];
        }
        g.rcp_rho_max = 1.0f / 255.0f;
        g.rcp_majorant = 1.0f / 100.0f;
        g.rcp_smax = 1.0f;
        Geom gm;
        memset(&gm, 0, sizeof(gm));
        gm.type = GEOM_VOLUME;
        gm.material = mat;
        gm.volume = (uint32_t)s->volumes.size();
        uint32_t slot = (uint32_t)s->geoms.size();
        s->geoms.push_back(gm);
        Prim p;
        memset(&p, 0, sizeof(p));
        for (int a = 0; a < 3; ++a) { p.v0[a] = g.lo[a]; p.v1[a] = g.hi[a]; }
        s->volumes.push_back(std::move(g));
        p.geom = slot;
        p.prim = 0;
        p.is_volume = 1;
        s->prims.push_back(p);
        if (first == UINT64_MAX) first = slot;
    }
    if (out) *out = first;
    return 0;
}

/* bindTexture: a copy of the texels; handles count from 1. */
int orc_bind_texture(orc_scene* s, const tr_texture_desc* d, const void* texels, uint32_t* out)
{
    if (s->committed || orc_texel_size(d->format) == 0) return -1;
    OrcTexture t;
    t.w = d->width;
    t.h = d->height;
    t.format = d->format;
    t.sampling = d->sampling;
    size_t bytes = (size_t)t.w * t.h * orc_texel_size(t.format);
    t.data.assign((const uint8_t*)texels, (const uint8_t*)texels + bytes);
    s->texset.tex.push_back(std::move(t));
    *out = (uint32_t)s->texset.tex.size();
    return 0;
}

/* CubeMap(Texture**) (src/graphics/cube-map.cc:32-71): six faces, each
 * copied with its rows in reverse order. */
int orc_set_global_cube_map(orc_scene* s, uint32_t side, const uint32_t* const faces[6])
{
    if (s->committed) return -1;
    if (!faces) {
        s->texset.cube_side = 0;
        s->texset.cube.clear();
        return 0;
    }
    size_t pitch = (size_t)side * 4;
    s->texset.cube.assign(6 * pitch * side, 0);
    for (int f = 0; f < 6; ++f)
        for (size_t y = 0; y < side; ++y)
            memcpy(s->texset.cube.data() + (size_t)f * pitch * side + y * pitch,
                   (const uint8_t*)faces[f] + (side - 1 - y) * pitch, pitch);
    s->texset.cube_side = side;
    return 0;
}

/* One texture lookup (tests): kind 0 spectrum, 1 RG, 2 RGBA. */
void orc_texture_sample(const orc_scene* s, uint32_t handle, float u, float v, int kind, float out[4])
{
    Rgba r = orc_tex_sample(s->texset, handle, u, v, kind);
    memcpy(out, r.v, 16);
}

/* One cube-map lookup (tests). */
void orc_cube_map_sample(const orc_scene* s, const float dir[3], float out[3]) { orc_cube_sample(s->texset, dir, out); }

int orc_add_light_source(orc_scene* s, const tr_light* light, uint32_t* out)
{
    Light L;
    memset(&L, 0, sizeof(L));
    L.src = *light;
    L.geom = TR_INVALID_GEOMETRY;
    if (out) *out = (uint32_t)s->lights.size();
    s->lights.push_back(L);
    return 0;
}

int orc_add_sphere_light_source(orc_scene* s, const tr_light* light, uint32_t* out)
{
    if (s->committed || s->materials.size() + 1 > s->materials.capacity()) return -1;
    tr_material m;
    memset(&m, 0, sizeof(m));
    m.model = TR_MODEL_EMISSIVE;
    memcpy(m.diffuse, light->radiance, 12);
    m.fresnel[0] = 1.0f;
    s->material_keys.push_back(nullptr);
    s->materials.push_back(m);
    uint32_t mat = (uint32_t)s->materials.size() - 1;
    Light L;
    memset(&L, 0, sizeof(L));
    L.src = *light;
    L.src.type = TR_LIGHT_SPHERE_AREA;
    uint32_t slot = add_sphere_geom(s, light->sphere_center, light->sphere_radius, GEOM_SPHERE_LIGHT, mat,
                                    (uint32_t)s->lights.size());
    L.geom = slot;
    L.src.geometry_id = slot;
    s->lights.push_back(L);
    if (out) *out = slot;
    return 0;
}

int orc_commit(orc_scene* s)
{
    uint32_t n = (uint32_t)s->prims.size();
    {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (const Prim& p : s->prims) {
            float pl[3], ph[3];
            prim_bounds(p, pl, ph);
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], pl[a]); hi[a] = std::max(hi[a], ph[a]); }
        }
        for (int a = 0; a < 3; ++a) {
            if (!(lo[a] <= hi[a])) { lo[a] = 0.0f; hi[a] = 1.0f; }
            float ext = hi[a] - lo[a];
            s->cache.box_lo[a] = lo[a];
            s->cache.box_rcp[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
        }
    }
    s->prim_order.resize(n);
    s->nodes.clear();
    if (n == 0) { s->committed = true; return 0; }
    std::vector<float> cen(n * 3), blo(n * 3), bhi(n * 3);
    for (uint32_t i = 0; i < n; ++i) {
        s->prim_order[i] = i;
        prim_bounds(s->prims[i], &blo[i * 3], &bhi[i * 3]);
        for (int a = 0; a < 3; ++a) cen[i * 3 + a] = 0.5f * (blo[i * 3 + a] + bhi[i * 3 + a]);
    }
    s->nodes.reserve(2 * n);
    build_node(s, 0, n, cen, blo, bhi);
    s->committed = true;
    return 0;
}

void orc_trace_rays(const orc_scene* s, size_t n, const tr_ray* rays, tr_hit* hits, int brute, int nthreads)
{
    parallel_for((uint32_t)n, nthreads, 256, [&](uint32_t i) {
        hits[i] = trace_closest(s, rays[i].org, rays[i].dir, rays[i].tnear, rays[i].tfar, brute != 0);
    });
}

void orc_trace_path(const orc_scene* s, uint32_t pixel, uint32_t cs)
{
    g_trace = stderr;
    uint32_t hash;
    V3<float> c = render_camera_sample<float>(s, pixel, cs, &hash);
    fprintf(stderr, "result (%g %g %g) hash %08x\n", (double)c.x, (double)c.y, (double)c.z, hash);
    g_trace = nullptr;
}

void orc_render_paths(const orc_scene* s, size_t pixel_count, const uint32_t* pixels, uint32_t frame,
                      float* out_rgb, uint32_t* out_hash, int nthreads)
{
    uint32_t total = (uint32_t)(pixel_count * s->spp);
    parallel_for(total, nthreads, 16, [&](uint32_t i) {
        uint32_t pi = i / s->spp, k = i % s->spp;
        uint32_t hash;
        V3<float> c = render_camera_sample<float>(s, pixels[pi], frame * s->spp + k, &hash, pi);
        out_rgb[i * 3 + 0] = c.x;
        out_rgb[i * 3 + 1] = c.y;
        out_rgb[i * 3 + 2] = c.z;
        if (out_hash) out_hash[i] = hash;
    });
}

/* Per-pixel sums and sums of squares over the frame's camera samples of the
 * listed pixels (the pixel-list twin of orc_render_rows). */
double orc_render_pixels(const orc_scene* s, size_t pixel_count, const uint32_t* pixels, uint32_t frame,
                         int use_fp64, double* out_sum, double* out_sumsq, int nthreads)
{
    auto t0 = std::chrono::steady_clock::now();
    parallel_for((uint32_t)pixel_count, nthreads, 4, [&](uint32_t i) {
        double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
        for (uint32_t k = 0; k < s->spp; ++k) {
            uint32_t hash;
            double c[3];
            if (use_fp64) {
                V3<double> v = render_camera_sample<double>(s, pixels[i], frame * s->spp + k, &hash, i);
                c[0] = v.x; c[1] = v.y; c[2] = v.z;
            } else {
                V3<float> v = render_camera_sample<float>(s, pixels[i], frame * s->spp + k, &hash, i);
                c[0] = v.x; c[1] = v.y; c[2] = v.z;
            }
            for (int a = 0; a < 3; ++a) { sum[a] += c[a]; sq[a] += c[a] * c[a]; }
        }
        for (int a = 0; a < 3; ++a) {
            out_sum[(size_t)i * 3 + a] = sum[a];
            if (out_sumsq) out_sumsq[(size_t)i * 3 + a] = sq[a];
        }
    });
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

double orc_render_rows(const orc_scene* s, uint32_t row_begin, uint32_t row_end, uint32_t frame, int use_fp64,
                       double* out_sum, double* out_sumsq, int nthreads)
{
    auto t0 = std::chrono::steady_clock::now();
    uint32_t first = row_begin * s->width;
    uint32_t count = (row_end - row_begin) * s->width;
    parallel_for(count, nthreads, 16, [&](uint32_t i) {
        uint32_t p = first + i;
        double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
        for (uint32_t k = 0; k < s->spp; ++k) {
            uint32_t hash;
            double c[3];
            if (use_fp64) {
                V3<double> v = render_camera_sample<double>(s, p, frame * s->spp + k, &hash);
                c[0] = v.x; c[1] = v.y; c[2] = v.z;
            } else {
                V3<float> v = render_camera_sample<float>(s, p, frame * s->spp + k, &hash);
                c[0] = v.x; c[1] = v.y; c[2] = v.z;
            }
            for (int a = 0; a < 3; ++a) { sum[a] += c[a]; sq[a] += c[a] * c[a]; }
        }
        for (int a = 0; a < 3; ++a) {
            out_sum[(size_t)p * 3 + a] = sum[a];
            if (out_sumsq) out_sumsq[(size_t)p * 3 + a] = sq[a];
        }
    });
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"

/* ---------------- unit exports for the test suite ---------------- */

extern "C" {

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { orc_philox4x32_10(ctr, key, out); }

float orc_u01f(uint32_t x) { return orc_u01(x); }

/* op: 0 sin(2 pi x), 1 cos(2 pi x), 2 log, 3 exp */
void orc_fmath(int op, size_t n, const float* in, float* out)
{
    for (size_t i = 0; i < n; ++i) {
        float s, c;
        switch (op) {
        case 0: tr_sincos_turns(in[i], &s, &c); out[i] = s; break;
        case 1: tr_sincos_turns(in[i], &s, &c); out[i] = c; break;
        case 2: out[i] = tr_logf(in[i]); break;
        default: out[i] = tr_expf(in[i]); break;
        }
    }
}

/* The backbuffer resolve, restated from the reference's per-pixel colour
 * conversion (src/graphics/ray-tracing/ray-tracing.cc:2366-2375) and its
 * auto-exposure post pass (:1638-1696): mean = accum * rcp_count in FP64,
 * rounded to float; Linear -> ToColor; sRGB -> ToColor(sRGB); ACES ->
 * ToColor(sRGB(ACESFilm)) (include/tempest/math/hdr.hh:76-84); the
 * auto-exposure modes scale by 0.18 / RGBToLuminance(exp(mean of
 * log(1e-3 + v) over the pixels)), with log taken over finite pixels only
 * and the mean over all of them (spectrum.hh:328-331), then Reinhard
 * x / (1 + x) for AutoExposureHDR (hdr.hh:32-35) and sRGB.  ToColor
 * truncates 255 * clamp(v, 0, 1) (spectrum.hh:792-798); sRGB is
 * ConvertLinearToSLuminance (spectrum.hh:147-153) with pow written as
 * exp(log(v) / 2.4) in the shared tr_logf/tr_expf.  `exposure_in` > 0
 * replaces the computed exposure (to compare pixels with a device that
 * summed the logs in another order); the computed one is returned. */
static float orc_to_srgb(float v)
{
    if (v <= 0.0031308f) return v * 12.92f;
    return 1.055f * tr_expf(tr_logf(v) * (1.0f / 2.4f)) - 0.055f;
}

static uint32_t orc_to_color(const float c[3])
{
    uint32_t out = 255u << 24;
    for (int k = 0; k < 3; ++k) {
        float v = c[k] < 0.0f ? 0.0f : (c[k] > 1.0f ? 1.0f : c[k]);
        out |= (uint32_t)(255.0f * v) << (8 * k);
    }
    return out;
}

static float orc_aces(float x)
{
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    float v = (x * (a * x + b)) / (x * (c * x + d) + e);
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

float orc_resolve_masked(size_t n, const double* accum, double rcp_count, uint32_t post, float exposure_in,
                         const uint8_t* mask, uint32_t* rgba);

float orc_resolve(size_t n, const double* accum, double rcp_count, uint32_t post, float exposure_in, uint32_t* rgba)
{
    return orc_resolve_masked(n, accum, rcp_count, post, exposure_in, nullptr, rgba);
}

/* With TransparentBackground (ray-tracing.cc:2364-2365): pixels whose mask
 * is 0 are neither written nor part of the log sum (whose mean still
 * divides by all n pixels). */
float orc_resolve_masked(size_t n, const double* accum, double rcp_count, uint32_t post, float exposure_in,
                         const uint8_t* mask, uint32_t* rgba)
{
    auto mean = [&](size_t i, float c[3]) {
        for (int k = 0; k < 3; ++k) c[k] = (float)(accum[i * 3 + k] * rcp_count);
    };
    const bool autoexp = post == TR_POST_AUTO_EXPOSURE_SRGB || post == TR_POST_AUTO_EXPOSURE_HDR;
    float exposure = 0.0f;
    if (autoexp) {
        double log_sum[3] = {0.0, 0.0, 0.0};
        for (size_t i = 0; i < n; ++i) {
            if (mask && !mask[i]) continue;
            float c[3];
            mean(i, c);
            if (!(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2]))) continue;
            for (int k = 0; k < 3; ++k) log_sum[k] += (double)tr_logf(1e-3f + c[k]);
        }
        float m[3];
        for (int k = 0; k < 3; ++k) m[k] = tr_expf((float)(log_sum[k] / (double)n));
        exposure = 0.18f / (0.2126f * m[0] + 0.7152f * m[1] + 0.0722f * m[2]);
    }
    const float e = exposure_in > 0.0f ? exposure_in : exposure;
    for (size_t i = 0; i < n; ++i) {
        if (mask && !mask[i]) continue;
        float c[3];
        mean(i, c);
        if (post == TR_POST_LINEAR) {
            rgba[i] = orc_to_color(c);
            continue;
        }
        for (int k = 0; k < 3; ++k) {
            if (post == TR_POST_ACES) c[k] = orc_aces(c[k]);
            else if (autoexp) {
                c[k] = e * c[k];
                if (post == TR_POST_AUTO_EXPOSURE_HDR) c[k] = c[k] / (1.0f + c[k]);
            }
            c[k] = orc_to_srgb(c[k]);
        }
        rgba[i] = orc_to_color(c);
    }
    return exposure;
}

static Sample<float> unit_sample(const tr_material* m, const float* n, const float* t, const float* b, const float* wo,
                                 const float* wi, float dir_density)
{
    Sample<float> s{};
    s.mat = m;
    s.normal = v3(n[0], n[1], n[2]);
    s.tangent = t ? v3(t[0], t[1], t[2]) : v3(0.0f, 0.0f, 0.0f);
    s.binormal = b ? v3(b[0], b[1], b[2]) : v3(0.0f, 0.0f, 0.0f);
    s.wo = v3(wo[0], wo[1], wo[2]);
    s.wi = wi ? v3(wi[0], wi[1], wi[2]) : v3(0.0f, 0.0f, 0.0f);
    s.dir_density = dir_density;
    return s;
}

/* BRDF*cos (out[0..2]) and PDF (out[3]) of a material; when dir_density < 0
 * the material cache (SGGX surface density) is evaluated with wi = -wo. */
void orc_bsdf_eval(const tr_material* m, const float* n, const float* t, const float* b, const float* wo,
                   const float* wi, float dir_density, float* out)
{
    Sample<float> s = unit_sample(m, n, t, b, wo, wi, dir_density);
    if (dir_density < 0.0f) {
        Sample<float> c = s;
        c.wi = -s.wo;
        material_cache(&c);
        s.dir_density = c.dir_density;
    }
    V3<float> f = bsdf_eval(s);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    out[3] = bsdf_pdf(s);
}

/* Sampled direction (out[0..2]) and PDF (out[3]); xi holds five uniforms
 * (the last two read by the anisotropic microfacet models only). */
void orc_bsdf_sample(const tr_material* m, const float* n, const float* t, const float* b, const float* wo,
                     const float* strata, const float* xi, float dir_density, float* out)
{
    Sample<float> s = unit_sample(m, n, t, b, wo, nullptr, dir_density);
    Strata<float> st;
    st.x = strata[0]; st.y = strata[1]; st.tx = strata[2]; st.ty = strata[3];
    bsdf_sample(st, &s, xi);
    out[0] = s.wi.x; out[1] = s.wi.y; out[2] = s.wi.z; out[3] = s.pdf;
}

/* n values of the pseudo-volume's LCG (FloatIntelFastRand) from `seed`. */
void orc_lcg15(uint32_t seed, uint32_t n, float* out)
{
    OrcLcg g{seed};
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next();
}

float orc_fresnel_conductor(float eta, float k, float c) { return fresnel_conductor(eta, k, c); }
float orc_fresnel_schlick(float r, float c) { return fresnel_schlick(r, c); }
float orc_fresnel_dielectric(float eta, float ci, float ct) { return fresnel_dielectric(eta, ci, ct); }

void orc_make_basis(const float* n, float* out9)
{
    M3<float> m = make_basis(v3(n[0], n[1], n[2]));
    out9[0] = m.c0.x; out9[1] = m.c0.y; out9[2] = m.c0.z;
    out9[3] = m.c1.x; out9[4] = m.c1.y; out9[5] = m.c1.z;
    out9[6] = m.c2.x; out9[7] = m.c2.y; out9[8] = m.c2.z;
}

/* Sphere-light sample (SampleSphereArea) at pos: out = dir[3], dist, pdf. */
void orc_sample_sphere_light(const float* pos, const float* c, float r, float ra, float rb, float* out)
{
    Light L;
    memset(&L, 0, sizeof(L));
    L.src.type = TR_LIGHT_SPHERE_AREA;
    L.src.sphere_center[0] = c[0]; L.src.sphere_center[1] = c[1]; L.src.sphere_center[2] = c[2];
    L.src.sphere_radius = r;
    LightSample<float> smp = sample_light<float>(nullptr, L, v3(pos[0], pos[1], pos[2]), 0u, ra, rb);
    out[0] = smp.dir.x; out[1] = smp.dir.y; out[2] = smp.dir.z; out[3] = smp.dist; out[4] = smp.pdf;
}

float orc_sphere_light_pdf(const float* pos, const float* dir, const float* c, float r)
{
    Light L;
    memset(&L, 0, sizeof(L));
    L.src.type = TR_LIGHT_SPHERE_AREA;
    L.src.sphere_center[0] = c[0]; L.src.sphere_center[1] = c[1]; L.src.sphere_center[2] = c[2];
    L.src.sphere_radius = r;
    return light_pdf<float>(nullptr, L, v3(pos[0], pos[1], pos[2]), v3(dir[0], dir[1], dir[2]));
}

void orc_sample_triangle_area(const float* pos, const float* p0, const float* p1, const float* p2, float ra, float rb,
                              float* out)
{
    V3<float> d;
    float dist, pdf;
    triangle_area_sample(v3(pos[0], pos[1], pos[2]), v3(p0[0], p0[1], p0[2]), v3(p1[0], p1[1], p1[2]),
                         v3(p2[0], p2[1], p2[2]), ra, rb, &d, &dist, &pdf);
    out[0] = d.x; out[1] = d.y; out[2] = d.z; out[3] = dist; out[4] = pdf;
}

float orc_projected_triangle_area(const float* pos, const float* p0, const float* p1, const float* p2)
{
    return projected_triangle_area(v3(pos[0], pos[1], pos[2]), v3(p0[0], p0[1], p0[2]), v3(p1[0], p1[1], p1[2]),
                                   v3(p2[0], p2[1], p2[2]));
}

/* Grid-volume lookups of the first volume of a committed scene at pos:
 * out[0] density, out[1..9] fibre frame (tangent, binormal, normal),
 * out[10] sigma(dir), out[11] sampleScattering's decision for `threshold`
 * (0 rejected, 1 accepted, 2 zero density). */
void orc_grid_volume_lookup(const orc_scene* s, const float* pos, const float* dir, float threshold, float* out)
{
    const GridVolume& g = s->volumes[0];
    V3<float> p = v3(pos[0], pos[1], pos[2]), d = v3(dir[0], dir[1], dir[2]);
    GridCell<float> c = grid_cell(g, p);
    out[0] = grid_density(g, c);
    M3<float> b = make_basis_tangent(grid_orientation(g, c));
    float* o = out + 1;
    o[0] = b.c0.x; o[1] = b.c0.y; o[2] = b.c0.z; o[3] = b.c1.x; o[4] = b.c1.y; o[5] = b.c1.z;
    o[6] = b.c2.x; o[7] = b.c2.y; o[8] = b.c2.z;
    out[10] = sggx_projected_area(v3(g.stddev[0], g.stddev[1], g.stddev[2]), b, d);
    out[11] = (float)grid_accept(g, p, d, threshold);
}

/* Delta tracking of n rays through the scene with the Philox stream of
 * (pixel = ray index, sample, split 0, vertex 0): the segment's closest hit
 * after tracking (collision: geom of the volume, t = collision distance). */
void orc_trace_segments(const orc_scene* s, size_t n, const tr_ray* rays, uint32_t sample, tr_hit* hits, int nthreads)
{
    parallel_for((uint32_t)n, nthreads, 64, [&](uint32_t i) {
        Rng rng;
        rng.key[0] = s->key[0]; rng.key[1] = s->key[1];
        rng.pixel = i; rng.sample = sample; rng.split = 0;
        hits[i] = trace_segment(s, rays[i].org, rays[i].dir, rng, 0);
    });
}

/* Woodcock transmittance estimate (0, 0.5 or 1 per crossing) of shadow
 * rays org -> org + dist * dir through the scene's volumes, with throughput
 * 1 and light geometry `light_geom` (TR_INVALID_GEOMETRY for none): the
 * occlusion loop of EvaluateLightSources with the streams of (pixel = ray
 * index, sample, split 0, vertex 0, light 0, sample 0). */
void orc_shadow_transmittance(const orc_scene* s, size_t n, const tr_ray* rays, uint32_t sample, uint32_t light_geom,
                              float* out, int nthreads)
{
    parallel_for((uint32_t)n, nthreads, 64, [&](uint32_t i) {
        Rng rng;
        rng.key[0] = s->key[0]; rng.key[1] = s->key[1];
        rng.pixel = i; rng.sample = sample; rng.split = 0;
        const float* of = rays[i].org;
        const float* df = rays[i].dir;
        float dist = rays[i].tfar;
        tr_hit h = trace_closest(s, of, df, 0.0f, dist, false);
        float tf = 1.0f, step_through = 0.0f;
        uint32_t step = 0, crossings = 0;
        bool reached = false;
        StepDraws dr = rng.steps(0, ORC_STREAM_SHADOW_VOLUME);
        for (;;) {
            if (h.geom == TR_INVALID_GEOMETRY || h.geom == light_geom) { reached = true; break; }
            if (!is_volume_geom(s, h.geom) || ++crossings > kOrcMaxVolumeCrossings) break;
            const GridVolume& g = s->volumes[s->geoms[h.geom].volume];
            uint32_t succ = woodcock_crossing(g, v3(of[0], of[1], of[2]), v3(df[0], df[1], df[2]), h.t, h.u, dr, 0u,
                                              &step, &step_through);
            tf = tf * ((float)succ / 2.0f);
            if (tf < 1e-3f) break;
            if (step_through >= dist - 1e-3f) { reached = true; break; }
            h = trace_closest(s, of, df, std::min(dist - 1e-3f, step_through), dist, false);
        }
        out[i] = reached ? tf : 0.0f;
    });
}

/* Matrix3 operations of the phase sampling; matrices are 9 floats,
 * column by column. */
static M3<float> m9(const float* f) { return m3(v3(f[0], f[1], f[2]), v3(f[3], f[4], f[5]), v3(f[6], f[7], f[8])); }
static void w9(const M3<float>& m, float* f)
{
    const V3<float>* c[3] = {&m.c0, &m.c1, &m.c2};
    for (int i = 0; i < 3; ++i) { f[3 * i] = c[i]->x; f[3 * i + 1] = c[i]->y; f[3 * i + 2] = c[i]->z; }
}
void orc_make_basis_tangent(const float* t, float* out9) { w9(make_basis_tangent(v3(t[0], t[1], t[2])), out9); }
void orc_cholesky(const float* a9, float* out9) { w9(cholesky(m9(a9)), out9); }
void orc_transpose(const float* a9, float* out9) { w9(transpose(m9(a9)), out9); }
void orc_mat_mul(const float* a9, const float* b9, float* out9) { w9(mul(m9(a9), m9(b9)), out9); }
void orc_transform_covariance(const float* a9, const float* var, float* out9)
{
    w9(transform_covariance(m9(a9), v3(var[0], var[1], var[2])), out9);
}
void orc_transform_rotation_inverse(const float* a9, const float* v, float* out)
{
    V3<float> r = transform_rotation_inverse(m9(a9), v3(v[0], v[1], v[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* bsdf_eval / bsdf_pdf of one shading frame for `count` incident
 * directions: out[4 i .. 4 i + 3] = BRDF*cos (RGB), PDF. */
void orc_bsdf_eval_batch(const tr_material* m, const float* n, const float* t, const float* b, const float* wo,
                         size_t count, const float* wi, float dir_density, float* out)
{
    Sample<float> s = unit_sample(m, n, t, b, wo, nullptr, dir_density);
    for (size_t i = 0; i < count; ++i) {
        s.wi = v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        V3<float> f = bsdf_eval(s);
        out[4 * i] = f.x; out[4 * i + 1] = f.y; out[4 * i + 2] = f.z;
        out[4 * i + 3] = bsdf_pdf(s);
    }
}

/* A user shape's intersection without range: kind 1-5 (rect, disk,
 * cylinder, oblique cylinder, ellipsoid) and its 20 parameters; out = t, u,
 * v, normal.  Returns 1 on a hit. */
int orc_shape_intersect(uint32_t kind, const float* params, const float* org, const float* dir, float* out)
{
    OrcShape sh;
    memset(&sh, 0, sizeof(sh));
    sh.kind = kind;
    memcpy(sh.f, params, sizeof(sh.f));
    float n[3] = {0.0f, 0.0f, 0.0f};
    int hit = orc_shape_hit(sh, org, dir, &out[0], &out[1], &out[2], n);
    out[3] = n[0]; out[4] = n[1]; out[5] = n[2];
    return hit;
}

/* SGGX helpers: out[0] = SGGXMicroFlakeNDF(stddev, basis, dir),
 * out[1] = SGGXProjectedArea(stddev, basis, dir); basis from a quaternion. */
void orc_sggx(const float* stddev, const float* quat, const float* dir, float* out)
{
    M3<float> basis = quat_to_m3(quat);
    V3<float> sd = v3(stddev[0], stddev[1], stddev[2]);
    V3<float> d = v3(dir[0], dir[1], dir[2]);
    out[0] = sggx_ndf(sd, basis, d);
    out[1] = sggx_projected_area(sd, basis, d);
}

/* Microfacet terms at FP32: kind 0 G1_GGX, 1 G1_Beckmann (the rational
 * fit), 2 the GGX distribution, 3 the Beckmann distribution
 * (ggx-models-impl.hh:126-151, beckmann-models-impl.hh:34-70).  For G1,
 * a = cos(v, m) and b = cos(v, n); for the distributions a = cos(n, h). */
float orc_microfacet_term(int kind, float g2, float a, float b)
{
    switch (kind) {
    case 0: return ggx_g1(g2, a, b);
    case 1: return beckmann_g1(g2, a, b);
    case 2: return ggx_d(g2, a);
    default: return beckmann_d(g2, a);
    }
}

}  // extern "C"

