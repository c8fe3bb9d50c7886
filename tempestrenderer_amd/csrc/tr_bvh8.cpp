/*
 * tr_bvh8.cpp - host builder of the 8-wide BVH (Node8) used by the kernels.
 *
 * The reference hands BVH construction to Embree (rtcCommit,
 * ray-tracing.cc:1633-1636) on the CPU and builds a Morton-ordered binary
 * LBVH for its CUDA path (src/mesh/lbvh2.cc:37-122, sslbvh2.cc:37-124).
 * This builder instead makes a binned-SAH binary tree (16 bins on each
 * axis, leaves of at most 4 primitives) and collapses it into 8-wide nodes
 * by repeatedly opening the child with the largest surface area.  Children
 * are placed in octant slots so that traversal can order them by the
 * ray's direction signs, their boxes are stored as binary16 offsets from
 * the node's centre at a power-of-two scale, rounded outward, and the primitives of a
 * node's leaf children are stored contiguously in leaf order.
 */
#include "tr_bvh8.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

namespace trk {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() { for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; } }
    void grow(const Box& b)
    {
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
    }
    float area() const
    {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Box box;
    int32_t left = -1, right = -1;
    uint32_t first = 0, count = 0; /* leaf when count > 0 */
};

struct Builder {
    std::vector<Box> boxes;
    std::vector<float> cent;
    std::vector<uint32_t> order;
    std::vector<BNode> nodes;

};

constexpr int kBins = 16;
constexpr uint32_t kMaxLeaf = 4;
constexpr float kNodeCost = 0.25f;

/* Builds the subtree over order[begin, end) into `out` (a local node list),
 * returning the index of its root in `out`. */
uint32_t build_rec(Builder& b, std::vector<BNode>& out, uint32_t begin, uint32_t end)
{
    BNode node;
    node.box.reset();
    Box cb;
    cb.reset();
    for (uint32_t i = begin; i < end; ++i) {
        uint32_t p = b.order[i];
        node.box.grow(b.boxes[p]);
        for (int a = 0; a < 3; ++a) {
            cb.lo[a] = std::min(cb.lo[a], b.cent[p * 3 + a]);
            cb.hi[a] = std::max(cb.hi[a], b.cent[p * 3 + a]);
        }
    }
    uint32_t n = end - begin;
    uint32_t idx = (uint32_t)out.size();
    out.push_back(node);
    if (n <= 1) {
        out[idx].first = begin;
        out[idx].count = n;
        return idx;
    }
    /* binned SAH over all three axes */
    float best_cost = INFINITY;
    int best_axis = -1, best_split = -1;
    for (int a = 0; a < 3; ++a) {
        float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0.0f)) continue;
        Box bin_box[kBins];
        uint32_t bin_cnt[kBins] = {0};
        for (int k = 0; k < kBins; ++k) bin_box[k].reset();
        float scale = (float)kBins / ext;
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t p = b.order[i];
            int k = (int)((b.cent[p * 3 + a] - cb.lo[a]) * scale);
            k = std::min(std::max(k, 0), kBins - 1);
            bin_cnt[k]++;
            bin_box[k].grow(b.boxes[p]);
        }
        float right_area[kBins];
        uint32_t right_cnt[kBins];
        Box acc;
        acc.reset();
        uint32_t c = 0;
        for (int k = kBins - 1; k > 0; --k) {
            acc.grow(bin_box[k]);
            c += bin_cnt[k];
            right_area[k] = acc.area();
            right_cnt[k] = c;
        }
        acc.reset();
        c = 0;
        for (int k = 0; k < kBins - 1; ++k) {
            acc.grow(bin_box[k]);
            c += bin_cnt[k];
            if (c == 0 || right_cnt[k + 1] == 0) continue;
            float cost = acc.area() * (float)c + right_area[k + 1] * (float)right_cnt[k + 1];
            if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = k; }
        }
    }
    /* SAH with a node-test cost of a quarter of a primitive test: after the
     * collapse into 8-wide nodes most binary nodes disappear, and a
     * primitive in a leaf costs a memory round trip of its own in the
     * traversal (heightfield: 1.81 -> 1.67 primitives per segment at 5.57
     * nodes, 0.9 % faster; a cost of 1 / 2 / 3 primitive tests gave
     * 1.81 / 2.71 / 3.19 primitives per segment and 0 / -11 / -16 %) */
    float leaf_cost = node.box.area() * (float)n;
    if (n <= kMaxLeaf && (best_axis < 0 || leaf_cost <= best_cost + kNodeCost * node.box.area())) {
        out[idx].first = begin;
        out[idx].count = n;
        return idx;
    }
    uint32_t mid;
    if (best_axis < 0) {
        mid = begin + n / 2; /* all centroids coincide: split by index */
    } else {
        float ext = cb.hi[best_axis] - cb.lo[best_axis];
        float scale = (float)kBins / ext;
        auto it = std::partition(b.order.begin() + begin, b.order.begin() + end, [&](uint32_t p) {
            int k = (int)((b.cent[p * 3 + best_axis] - cb.lo[best_axis]) * scale);
            k = std::min(std::max(k, 0), kBins - 1);
            return k <= best_split;
        });
        mid = (uint32_t)(it - b.order.begin());
        if (mid == begin || mid == end) mid = begin + n / 2;
    }
    uint32_t l = build_rec(b, out, begin, mid);
    uint32_t r = build_rec(b, out, mid, end);
    out[idx].left = (int32_t)l;
    out[idx].right = (int32_t)r;
    return idx;
}

/* Builds in parallel: the top `depth` levels split at the centroid median of
 * the longest axis and hand the halves to threads; each subtree is built
 * into its own list and appended to `out` with its indices relocated. */
void append_relocated(std::vector<BNode>& out, std::vector<BNode>& sub)
{
    int32_t off = (int32_t)out.size();
    for (auto& nd : sub) {
        if (nd.count == 0) { nd.left += off; nd.right += off; }
        out.push_back(nd);
    }
}

uint32_t build_par(Builder& b, std::vector<BNode>& out, uint32_t begin, uint32_t end, int depth)
{
    uint32_t n = end - begin;
    if (depth <= 0 || n < 65536) return build_rec(b, out, begin, end);
    Box cb, nb;
    cb.reset();
    nb.reset();
    for (uint32_t i = begin; i < end; ++i) {
        uint32_t p = b.order[i];
        nb.grow(b.boxes[p]);
        for (int a = 0; a < 3; ++a) {
            cb.lo[a] = std::min(cb.lo[a], b.cent[p * 3 + a]);
            cb.hi[a] = std::max(cb.hi[a], b.cent[p * 3 + a]);
        }
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
    uint32_t mid = begin + n / 2;
    std::nth_element(b.order.begin() + begin, b.order.begin() + mid, b.order.begin() + end, [&](uint32_t x, uint32_t y) {
        float cx = b.cent[x * 3 + axis], cy = b.cent[y * 3 + axis];
        return cx < cy || (cx == cy && x < y);
    });
    std::vector<BNode> lnodes, rnodes;
    uint32_t lroot = 0, rroot = 0;
    /* the halves touch disjoint ranges of b.order */
    std::thread tl([&]() { lroot = build_par(b, lnodes, begin, mid, depth - 1); });
    rroot = build_par(b, rnodes, mid, end, depth - 1);
    tl.join();
    uint32_t pidx = (uint32_t)out.size();
    BNode parent;
    parent.box = nb;
    out.push_back(parent);
    uint32_t loff = (uint32_t)out.size();
    append_relocated(out, lnodes);
    uint32_t roff = (uint32_t)out.size();
    append_relocated(out, rnodes);
    out[pidx].left = (int32_t)(loff + lroot);
    out[pidx].right = (int32_t)(roff + rroot);
    return pidx;
}

/* Binary16 codes are handled through an ordered key: key k >= 0 is the
 * pattern k (value >= 0), k < 0 the pattern 0x8000 | -k, so that keys
 * -0x7BFF .. 0x7BFF enumerate the finite values in increasing order. */
constexpr int kKeyMax = 0x7BFF;

inline uint16_t key_bits(int k) { return k >= 0 ? (uint16_t)k : (uint16_t)(0x8000 | -k); }

/* origin + value(k) * 2^e in FP32, as the node is decoded. */
inline float decode(float origin, int k, int e)
{
    float s;
    uint32_t bits = (uint32_t)(e + 127) << 23;
    memcpy(&s, &bits, 4);
    return origin + half_value(key_bits(k)) * s;
}

/* Key of a binary16 value near x (|x| <= 65504 after clamping); the caller
 * steps it to the exact outward bound. */
int near_key(double x)
{
    const double a = std::fabs(x);
    int k;
    if (a >= 65504.0) k = kKeyMax;
    else if (a < 6.103515625e-05) k = (int)std::lround(a * 16777216.0); /* subnormal: m * 2^-24 */
    else {
        int E;
        const double f = std::frexp(a, &E); /* a = f 2^E, f in [0.5, 1) */
        long m = std::lround((f * 2.0 - 1.0) * 1024.0);
        int ef = E + 14;
        if (m == 1024) { m = 0; ++ef; }
        k = std::min((ef << 10) | (int)m, kKeyMax);
    }
    return x < 0.0 ? -k : k;
}

/* Codes for child bounds [lo, hi] (already padded) on one axis, as
 * offsets from `origin` (the node's centre) at a power-of-two scale. */
bool quantise_axis(float origin, const float* clo, const float* chi, int nchild, int8_t* exp_out, uint16_t* qlo,
                   uint16_t* qhi)
{
    double maxoff = 0.0;
    for (int c = 0; c < nchild; ++c)
        maxoff = std::max(maxoff, std::max(std::fabs((double)clo[c] - origin), std::fabs((double)chi[c] - origin)));
    int e = -126;
    if (maxoff > 0.0) {
        int ee;
        std::frexp(maxoff / 32768.0, &ee); /* maxoff / 2^ee <= 32768 */
        e = std::max(ee, -126);
    }
    for (; e <= 127; ++e) {
        bool ok = true;
        const double sc = std::ldexp(1.0, e);
        for (int c = 0; c < nchild && ok; ++c) {
            int l = near_key(((double)clo[c] - origin) / sc);
            while (l > -kKeyMax && decode(origin, l, e) > clo[c]) --l;
            while (l < kKeyMax && decode(origin, l + 1, e) <= clo[c]) ++l;
            int h = near_key(((double)chi[c] - origin) / sc);
            while (h < kKeyMax && decode(origin, h, e) < chi[c]) ++h;
            while (h > -kKeyMax && decode(origin, h - 1, e) >= chi[c]) --h;
            if (decode(origin, l, e) > clo[c] || decode(origin, h, e) < chi[c]) { ok = false; break; }
            qlo[c] = key_bits(l);
            qhi[c] = key_bits(h);
        }
        if (ok) { *exp_out = (int8_t)e; return true; }
    }
    return false;
}

}  // namespace

bool build_bvh8(const std::vector<PrimRec>& in, std::vector<Node8>& nodes_out, std::vector<PrimRec>& prims_out,
                Bvh8Stats* stats, std::string* err)
{
    nodes_out.clear();
    prims_out.clear();
    uint32_t n = (uint32_t)in.size();
    if (n == 0) {
        Node8 root;
        memset(&root, 0, sizeof(root));
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 8; ++k) { root.qlo[a][k] = kHalfMax; root.qhi[a][k] = kHalfMinusMax; }
        nodes_out.push_back(root);
        if (stats) { stats->depth = 1; stats->nodes = 1; }
        return true;
    }
    Builder b;
    b.boxes.resize(n);
    b.cent.resize((size_t)n * 3);
    b.order.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const PrimRec& p = in[i];
        Box bx;
        prim_box(p, bx.lo, bx.hi);
        /* relative outward padding: slack for rounding in the slab test */
        for (int a = 0; a < 3; ++a) {
            float pad = (std::fabs(bx.lo[a]) + std::fabs(bx.hi[a]) + (bx.hi[a] - bx.lo[a])) * 7.62939453e-06f + 1e-30f;
            bx.lo[a] -= pad;
            bx.hi[a] += pad;
        }
        b.boxes[i] = bx;
        for (int a = 0; a < 3; ++a) b.cent[i * 3 + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
        b.order[i] = i;
    }
    b.nodes.reserve(2 * (size_t)n / 3 + 16);
    uint32_t broot = build_par(b, b.nodes, 0, n, 3);

    /* collapse into 8-wide nodes, breadth first so siblings are contiguous */
    struct Pending {
        uint32_t wide;
        uint32_t bnode;
        uint32_t depth;
    };
    std::vector<Pending> queue;
    nodes_out.emplace_back();
    queue.push_back({0, broot, 1});
    uint32_t max_depth = 1;
    for (size_t qi = 0; qi < queue.size(); ++qi) {
        Pending pd = queue[qi];
        max_depth = std::max(max_depth, pd.depth);
        std::vector<uint32_t> kids;
        const BNode& bn = b.nodes[pd.bnode];
        if (bn.count > 0) {
            kids.push_back(pd.bnode);
        } else {
            kids.push_back((uint32_t)bn.left);
            kids.push_back((uint32_t)bn.right);
        }
        while (kids.size() < 8) {
            int best = -1;
            float best_area = -1.0f;
            for (size_t k = 0; k < kids.size(); ++k) {
                const BNode& c = b.nodes[kids[k]];
                if (c.count > 0) continue;
                float a = c.box.area();
                if (a > best_area) { best_area = a; best = (int)k; }
            }
            if (best < 0) break;
            uint32_t open = kids[best];
            kids[best] = (uint32_t)b.nodes[open].left;
            kids.push_back((uint32_t)b.nodes[open].right);
        }
        /* octant slot assignment (greedy on alignment of centroid offsets) */
        Box nb;
        nb.reset();
        for (uint32_t k : kids) nb.grow(b.nodes[k].box);
        float nc[3] = {0.5f * (nb.lo[0] + nb.hi[0]), 0.5f * (nb.lo[1] + nb.hi[1]), 0.5f * (nb.lo[2] + nb.hi[2])};
        int slot_of[8];
        bool slot_used[8] = {false};
        bool kid_done[8] = {false};
        for (size_t round = 0; round < kids.size(); ++round) {
            float best = -INFINITY;
            int bk = -1, bs = -1;
            for (size_t k = 0; k < kids.size(); ++k) {
                if (kid_done[k]) continue;
                const Box& cbx = b.nodes[kids[k]].box;
                float d[3] = {0.5f * (cbx.lo[0] + cbx.hi[0]) - nc[0], 0.5f * (cbx.lo[1] + cbx.hi[1]) - nc[1],
                              0.5f * (cbx.lo[2] + cbx.hi[2]) - nc[2]};
                for (int s = 0; s < 8; ++s) {
                    if (slot_used[s]) continue;
                    float sc = ((s & 1) ? d[0] : -d[0]) + ((s & 2) ? d[1] : -d[1]) + ((s & 4) ? d[2] : -d[2]);
                    if (sc > best) { best = sc; bk = (int)k; bs = s; }
                }
            }
            kid_done[bk] = true;
            slot_used[bs] = true;
            slot_of[bk] = bs;
        }
        int kid_in_slot[8];
        for (int s = 0; s < 8; ++s) kid_in_slot[s] = -1;
        for (size_t k = 0; k < kids.size(); ++k) kid_in_slot[slot_of[k]] = (int)k;

        Node8 node;
        memset(&node, 0, sizeof(node));
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 8; ++k) { node.qlo[a][k] = kHalfMax; node.qhi[a][k] = kHalfMinusMax; }
        float clo[3][8], chi[3][8];
        int nslot = 0;
        int slot_list[8];
        for (int s = 0; s < 8; ++s) {
            if (kid_in_slot[s] < 0) continue;
            const Box& cbx = b.nodes[kids[kid_in_slot[s]]].box;
            for (int a = 0; a < 3; ++a) { clo[a][nslot] = cbx.lo[a]; chi[a][nslot] = cbx.hi[a]; }
            slot_list[nslot++] = s;
        }
        for (int a = 0; a < 3; ++a) {
            node.origin[a] = 0.5f * (nb.lo[a] + nb.hi[a]);
            uint16_t ql[8], qh[8];
            if (!quantise_axis(node.origin[a], clo[a], chi[a], nslot, &node.exp[a], ql, qh)) {
                if (err) *err = "BVH8 quantisation failed";
                return false;
            }
            for (int i = 0; i < nslot; ++i) {
                node.qlo[a][slot_list[i]] = ql[i];
                node.qhi[a][slot_list[i]] = qh[i];
            }
        }
        /* internal children: contiguous node indices in slot order */
        uint32_t child_base = (uint32_t)nodes_out.size();
        uint32_t ninner = 0;
        for (int s = 0; s < 8; ++s) {
            int k = kid_in_slot[s];
            if (k < 0) continue;
            const BNode& c = b.nodes[kids[k]];
            if (c.count == 0) {
                node.imask |= (uint8_t)(1u << s);
                ++ninner;
            }
        }
        node.child_base = child_base;
        uint32_t pos = 0;
        for (int s = 0; s < 8; ++s) {
            int k = kid_in_slot[s];
            if (k < 0) continue;
            const BNode& c = b.nodes[kids[k]];
            if (c.count == 0) {
                nodes_out.emplace_back();
                queue.push_back({child_base + pos, (uint32_t)kids[k], pd.depth + 1});
                ++pos;
            }
        }
        /* leaf children: primitives appended contiguously */
        node.prim_base = (uint32_t)prims_out.size();
        uint32_t off = 0;
        for (int s = 0; s < 8; ++s) {
            int k = kid_in_slot[s];
            if (k < 0) continue;
            const BNode& c = b.nodes[kids[k]];
            if (c.count == 0) continue;
            if (c.count > 7 || off > 31) {
                if (err) *err = "BVH8 leaf layout overflow";
                return false;
            }
            node.meta[s] = (uint8_t)((c.count << 5) | off);
            for (uint32_t i = 0; i < c.count; ++i) prims_out.push_back(in[b.order[c.first + i]]);
            off += c.count;
        }
        nodes_out[pd.wide] = node;
        (void)ninner;
    }
    if (stats) {
        stats->depth = max_depth;
        stats->nodes = (uint32_t)nodes_out.size();
    }
    return true;
}

}  // namespace trk
