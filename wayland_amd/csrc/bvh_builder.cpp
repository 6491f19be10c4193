// bvh_builder.cpp — host SAH builder: binned-SAH BVH2, greedy collapse to
// BVH8, octant slot ordering and 16-bit outward-rounded box quantisation.
// Replaces the closed GPU builder behind optixAccelBuild
// (src/Backend/Optix/Core/AccelStructure.cpp:133-162).
#include "bvh_builder.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <limits>

namespace wl
{
namespace
{

constexpr int kBins = 32;
// SAH cost of a traversal step relative to one triangle test.  Measured on
// the 2M-triangle soup (closest hit steps one node or one triangle): 0.7
// keeps incoherent closest hit level with 1.0 and makes the path tracer's
// extension and shadow rays 2 % faster (fewer triangle tests per ray).
constexpr double kSahTraversal = 0.7;
constexpr uint32_t kParallelMin = 1u << 16;

struct Aabb
{
    float lo[3] = { INFINITY, INFINITY, INFINITY };
    float hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    void grow(const Aabb &o)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], o.lo[k]);
            hi[k] = std::max(hi[k], o.hi[k]);
        }
    }
    void grow_point(const float *p)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double area() const
    {
        double d[3];
        for (int k = 0; k < 3; ++k)
        {
            d[k] = (double)hi[k] - (double)lo[k];
            if (!(d[k] >= 0))
                return 0.0;
        }
        return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

struct Node2
{
    Aabb box;
    uint32_t left = 0, right = 0; // inner
    uint32_t first = 0, count = 0; // leaf (count > 0)
};

struct Bvh2Builder
{
    const float *tris;
    uint32_t n;
    std::vector<Aabb> pbox;
    std::vector<float> cen;
    std::vector<uint32_t> prims;
    std::vector<Node2> nodes;
    std::atomic<uint32_t> next_node{ 0 };
    std::atomic<int> threads_left{ 0 };

    uint32_t alloc_nodes(uint32_t k) { return next_node.fetch_add(k); }

    void build_range(uint32_t node, uint32_t begin, uint32_t end)
    {
        // iterative on the right child, recursive (maybe async) on the left
        for (;;)
        {
            Node2 &nd = nodes[node];
            Aabb box, cbox;
            for (uint32_t i = begin; i < end; ++i)
            {
                uint32_t p = prims[i];
                box.grow(pbox[p]);
                cbox.grow_point(&cen[3 * (size_t)p]);
            }
            nd.box = box;
            uint32_t cnt = end - begin;
            int best_axis = -1, best_bin = 0;
            double best_cost = std::numeric_limits<double>::infinity();
            double leaf_cost = (double)cnt;
            double parea = box.area();
            if (cnt > 1)
            {
                for (int axis = 0; axis < 3; ++axis)
                {
                    float lo = cbox.lo[axis], hi = cbox.hi[axis];
                    if (!(hi > lo))
                        continue;
                    Aabb bins[kBins];
                    uint32_t counts[kBins] = {};
                    double scale = kBins / ((double)hi - lo);
                    for (uint32_t i = begin; i < end; ++i)
                    {
                        uint32_t p = prims[i];
                        int b = (int)(((double)cen[3 * (size_t)p + axis] - lo) * scale);
                        b = std::min(std::max(b, 0), kBins - 1);
                        counts[b]++;
                        bins[b].grow(pbox[p]);
                    }
                    double rarea[kBins];
                    uint32_t rcnt[kBins];
                    Aabb acc;
                    uint32_t c = 0;
                    for (int b = kBins - 1; b > 0; --b)
                    {
                        acc.grow(bins[b]);
                        c += counts[b];
                        rarea[b] = acc.area();
                        rcnt[b] = c;
                    }
                    acc = Aabb();
                    c = 0;
                    for (int b = 0; b < kBins - 1; ++b)
                    {
                        acc.grow(bins[b]);
                        c += counts[b];
                        if (c == 0 || rcnt[b + 1] == 0)
                            continue;
                        double cost =
                            kSahTraversal + (acc.area() * c + rarea[b + 1] * rcnt[b + 1]) /
                                      (parea > 0 ? parea : 1.0);
                        if (cost < best_cost)
                        {
                            best_cost = cost;
                            best_axis = axis;
                            best_bin = b;
                        }
                    }
                }
            }
            bool make_leaf = cnt <= WL_BVH8_MAX_LEAF_TRIS &&
                             (best_axis < 0 || leaf_cost <= best_cost);
            if (cnt == 0 || make_leaf)
            {
                nd.first = begin;
                nd.count = cnt;
                return;
            }
            uint32_t mid;
            if (best_axis >= 0)
            {
                float lo = cbox.lo[best_axis], hi = cbox.hi[best_axis];
                double scale = kBins / ((double)hi - lo);
                auto it = std::partition(
                    prims.begin() + begin, prims.begin() + end, [&](uint32_t p) {
                        int b = (int)(((double)cen[3 * (size_t)p + best_axis] - lo) *
                                      scale);
                        b = std::min(std::max(b, 0), kBins - 1);
                        return b <= best_bin;
                    });
                mid = (uint32_t)(it - prims.begin());
            }
            else
            {
                // all centroids coincide: split the range in half
                mid = begin + cnt / 2;
            }
            if (mid == begin || mid == end)
                mid = begin + cnt / 2;
            uint32_t l = alloc_nodes(2);
            nd.left = l;
            nd.right = l + 1;
            nd.count = 0;
            if (cnt >= kParallelMin && threads_left.fetch_sub(1) > 0)
            {
                auto fut = std::async(std::launch::async,
                                      [this, l, begin, mid] { build_range(l, begin, mid); });
                build_range(l + 1, mid, end);
                fut.get();
                return;
            }
            build_range(l, begin, mid);
            node = l + 1;
            begin = mid;
        }
    }
};

inline float exp2i(int e) { return std::ldexp(1.0f, e); }

// Quantised planes are integers 0..kQMax, stored as binary16 (exact for
// integers up to 2048): the traversal's fma then takes them as a
// mixed-precision operand with no separate conversion.
constexpr int kQMax = 2047;

// Per-axis exponent so that (hi - p) / 2^e <= kQMax - 2 (the outward
// rounding and one quantum of padding on either side then stay in range).
int choose_exponent(float lo, float hi)
{
    const double room = kQMax - 2;
    double ext = (double)hi - (double)lo;
    if (!(ext > 0))
        return -120;
    int x;
    std::frexp(ext / room, &x); // ext/room = m * 2^x, m in [0.5, 1)
    int e = x;
    if (std::ldexp(1.0, e - 1) >= ext / room)
        e -= 1;
    while (std::ldexp(1.0, e) * room < ext)
        ++e;
    return std::min(std::max(e, -120), 127);
}

// binary16 bits of an integer 0 <= n <= 2047 (exact: 11 significant bits).
uint16_t half_of_int(int n)
{
    if (n <= 0)
        return 0;
    int e = 0;
    while ((n >> (e + 1)) != 0)
        ++e; // 2^e <= n < 2^(e+1), e <= 10
    const uint32_t mant = ((uint32_t)n << (10 - e)) & 0x3FFu;
    return (uint16_t)(((uint32_t)(e + 15) << 10) | mant);
}
constexpr uint16_t kHalfMaxFinite = 0x7BFF; // 65504: the empty slot's lo plane

} // namespace

bool build_bvh8(const float *tris, uint32_t ntris, int nthreads, Bvh8Host *out,
                std::string *err)
{
    auto t0 = std::chrono::steady_clock::now();
    Bvh2Builder b;
    b.tris = tris;
    b.n = ntris;
    b.pbox.resize(ntris);
    b.cen.resize(3 * (size_t)ntris);
    b.prims.resize(ntris);
    for (uint32_t i = 0; i < ntris; ++i)
    {
        Aabb a;
        for (int v = 0; v < 3; ++v)
            a.grow_point(tris + 9 * (size_t)i + 3 * v);
        b.pbox[i] = a;
        for (int k = 0; k < 3; ++k)
            b.cen[3 * (size_t)i + k] = 0.5f * (a.lo[k] + a.hi[k]);
        b.prims[i] = i;
    }
    b.nodes.resize(std::max<size_t>(1, 2 * (size_t)ntris));
    b.next_node = 1;
    b.threads_left = std::max(0, nthreads - 1);
    b.build_range(0, 0, ntris);

    // ---- collapse to BVH8 (breadth-first, so the top levels sit together)
    out->nodes.clear();
    out->tris.clear();
    // the root owns a group of 8 slots like every other node, so that every
    // child group starts at a multiple of 8 (the stack stores base / 8)
    out->nodes.resize(8);
    out->tris.reserve(ntris);
    out->depth = 1;
    out->inner_nodes = 0;
    out->leaves = 0;
    double root_area = b.nodes[0].box.area();
    double sah = 0.0;
    struct Work
    {
        uint32_t n2, n8, depth;
    };
    std::deque<Work> queue;
    queue.push_back({ 0, 0, 1 });
    while (!queue.empty())
    {
        Work w = queue.front();
        queue.pop_front();
        out->depth = std::max(out->depth, w.depth);
        out->inner_nodes++;
        const Node2 &src = b.nodes[w.n2];
        std::vector<uint32_t> kids;
        if (src.count > 0 || ntris == 0)
        {
            if (ntris > 0)
                kids.push_back(w.n2);
        }
        else
        {
            kids.push_back(src.left);
            kids.push_back(src.right);
            while (kids.size() < 8)
            {
                int best = -1;
                double best_area = -1.0;
                for (size_t i = 0; i < kids.size(); ++i)
                {
                    const Node2 &k = b.nodes[kids[i]];
                    if (k.count == 0 && k.box.area() > best_area)
                    {
                        best_area = k.box.area();
                        best = (int)i;
                    }
                }
                if (best < 0)
                    break;
                uint32_t id = kids[best];
                kids[best] = b.nodes[id].left;
                kids.push_back(b.nodes[id].right);
            }
        }
        if (root_area > 0)
            sah += src.box.area() / root_area;

        // octant slot assignment (greedy max matching of centroid offsets)
        int slot_of[8];
        for (int i = 0; i < 8; ++i)
            slot_of[i] = -1;
        {
            float pc[3];
            for (int k = 0; k < 3; ++k)
                pc[k] = 0.5f * (src.box.lo[k] + src.box.hi[k]);
            double cost[8][8];
            for (size_t i = 0; i < kids.size(); ++i)
            {
                const Aabb &cb = b.nodes[kids[i]].box;
                double off[3];
                for (int k = 0; k < 3; ++k)
                    off[k] = 0.5 * ((double)cb.lo[k] + cb.hi[k]) - pc[k];
                for (int s = 0; s < 8; ++s)
                    cost[i][s] = ((s & 1) ? -off[0] : off[0]) +
                                 ((s & 2) ? -off[1] : off[1]) +
                                 ((s & 4) ? -off[2] : off[2]);
            }
            bool slot_used[8] = {};
            for (size_t round = 0; round < kids.size(); ++round)
            {
                double bestc = -INFINITY;
                int bi = -1, bs = -1;
                for (size_t i = 0; i < kids.size(); ++i)
                {
                    if (slot_of[i] >= 0)
                        continue;
                    for (int s = 0; s < 8; ++s)
                        if (!slot_used[s] && cost[i][s] > bestc)
                        {
                            bestc = cost[i][s];
                            bi = (int)i;
                            bs = s;
                        }
                }
                slot_of[bi] = bs;
                slot_used[bs] = true;
            }
        }

        WlBvh8Node node;
        std::memset(&node, 0, sizeof node);
        for (int s = 0; s < 8; ++s)
        {
            node.meta[s] = 0xFF;
            for (int k = 0; k < 3; ++k)
            {
                node.q[k][s] = kHalfMaxFinite;
                node.q[3 + k][s] = 0;
            }
        }
        // quantisation frame = this node's box
        const Aabb &nb = src.box;
        float scale[3];
        for (int k = 0; k < 3; ++k)
        {
            float lo = ntris ? nb.lo[k] : 0.0f;
            float hi = ntris ? nb.hi[k] : 0.0f;
            node.p[k] = lo;
            int e = choose_exponent(lo, hi);
            node.e[k] = (uint8_t)(e + 127);
            scale[k] = exp2i(e);
            (void)hi;
        }
        bool any_inner = false;
        for (size_t i = 0; i < kids.size(); ++i)
            if (b.nodes[kids[i]].count == 0)
                any_inner = true;
        if (any_inner)
        {
            node.child_base = (uint32_t)out->nodes.size();
            out->nodes.resize(out->nodes.size() + 8);
        }
        node.tri_base = (uint32_t)out->tris.size();
        // leaves in slot order
        uint32_t tri_off = 0;
        for (int s = 0; s < 8; ++s)
        {
            int i = -1;
            for (size_t j = 0; j < kids.size(); ++j)
                if (slot_of[j] == s)
                    i = (int)j;
            if (i < 0)
                continue;
            const Node2 &k = b.nodes[kids[i]];
            for (int a = 0; a < 3; ++a)
            {
                double qlo = std::floor(((double)k.box.lo[a] - node.p[a]) / scale[a]) - 1.0;
                double qhi = std::ceil(((double)k.box.hi[a] - node.p[a]) / scale[a]) + 1.0;
                qlo = std::max(qlo, 0.0);
                qhi = std::min(qhi, (double)kQMax);
                node.q[a][s] = half_of_int((int)qlo);
                node.q[3 + a][s] = half_of_int((int)qhi);
            }
            if (k.count == 0)
            {
                node.imask |= (uint8_t)(1u << s);
                node.meta[s] = (uint8_t)s;
                queue.push_back({ kids[i], node.child_base + (uint32_t)s, w.depth + 1 });
            }
            else
            {
                node.meta[s] = (uint8_t)((tri_off << 2) | (k.count - 1));
                for (uint32_t t = 0; t < k.count; ++t)
                {
                    uint32_t p = b.prims[k.first + t];
                    WlTriRecord r;
                    std::memcpy(r.v0, tris + 9 * (size_t)p, 36);
                    r.prim = p;
                    r.input = 0;
                    r.local = p;
                    out->tris.push_back(r);
                }
                tri_off += k.count;
                out->leaves++;
                if (root_area > 0)
                    sah += k.box.area() / root_area * k.count;
            }
        }
        out->nodes[w.n8] = node;
    }
    out->sah_cost = sah;
    auto t1 = std::chrono::steady_clock::now();
    out->build_seconds = std::chrono::duration<double>(t1 - t0).count();
    if (out->depth > WL_BVH8_STACK_DEPTH)
    {
        if (err)
            *err = "BVH8 depth " + std::to_string(out->depth) +
                   " exceeds the traversal stack depth " +
                   std::to_string(WL_BVH8_STACK_DEPTH);
        return false;
    }
    if (out->nodes.size() > (size_t)0xFFFFFF * 8)
    {
        if (err)
            *err = "BVH8 node count exceeds the 24-bit node-group index";
        return false;
    }
    return true;
}

} // namespace wl
