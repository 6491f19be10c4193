// hostemu.cpp — CPU emulation of the BVH8 builder + traversal kernels and of
// the batched-serving index arithmetic, for tests only (libwl_hostemu.so).
// It compiles the exact traversal code of traverse_device.h for the host,
// with every node index, triangle index and stack slot the kernel would touch
// checked against its allocation, and the serving schedule (serve_plan.h:
// plan, find_item) with every path, pixel and film index the multi-request
// raygen and film-update kernels derive from it checked, so the indexing is
// verified on the CPU before any kernel runs on a GPU.
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "bvh_builder.h"
#include "serve_plan.h"

static thread_local size_t wl_trav_nodes_bound = 0;
static thread_local size_t wl_trav_tris_bound = 0;
static std::atomic<uint64_t> g_faults{ 0 };
static thread_local const char *g_fault_what = nullptr;

#define WL_TRAV_CHECK(cond, what)                                            \
    do                                                                       \
    {                                                                        \
        if (!(cond))                                                         \
        {                                                                    \
            g_faults++;                                                      \
            g_fault_what = what;                                             \
            return true;                                                     \
        }                                                                    \
    } while (0)

#include "traverse_device.h"

struct HostEmu
{
    wl::Bvh8Host bvh;
};

extern "C" {

int wlh_build(const float *tris, uint32_t ntris, void **out, char *err, int errlen)
{
    HostEmu *h = new HostEmu;
    std::string e;
    if (!wl::build_bvh8(tris, ntris, 8, &h->bvh, &e))
    {
        if (err && errlen > 0)
        {
            std::strncpy(err, e.c_str(), errlen - 1);
            err[errlen - 1] = 0;
        }
        delete h;
        return 1;
    }
    *out = h;
    return 0;
}

void wlh_free(void *p) { delete (HostEmu *)p; }

void wlh_info(void *p, uint32_t *out6)
{
    HostEmu *h = (HostEmu *)p;
    out6[0] = (uint32_t)h->bvh.nodes.size();
    out6[1] = (uint32_t)h->bvh.tris.size();
    out6[2] = h->bvh.depth;
    out6[3] = h->bvh.inner_nodes;
    out6[4] = h->bvh.leaves;
    out6[5] = (uint32_t)(h->bvh.sah_cost);
}

// rays: n x 8 floats; hits: n x 4 (t, prim bits, b1, b2); counts: n x 2 or NULL
uint64_t wlh_trace(void *p, const float *rays, uint32_t n, float *hits,
                   int any, uint32_t *counts, int nthreads)
{
    HostEmu *h = (HostEmu *)p;
    g_faults = 0;
    auto work = [&](uint32_t b, uint32_t e) {
        wl_trav_nodes_bound = h->bvh.nodes.size();
        wl_trav_tris_bound = h->bvh.tris.size();
        uint32_t stack[WL_BVH8_STACK_DEPTH];
        for (uint32_t i = b; i < e; ++i)
        {
            const float *r = rays + 8 * (size_t)i;
            wl::TravRay tr;
            wl::trav_setup(tr, r[0], r[1], r[2], r[4], r[5], r[6], r[3], r[7]);
            uint32_t nn = 0, nt = 0;
            wl::TravResult res;
            if (any)
                res = wl::trav_run<true, true, 1>(h->bvh.nodes.data(), h->bvh.tris.data(),
                                                  tr, true, stack, nn, nt);
            else
            {
                // the closest-hit kernels' step: one node or one triangle
                wl::TravState st;
                wl::trav_init(st, tr, true);
                for (;;)
                {
                    const bool tri = st.trimask != 0;
                    if (wl::trav_step_one<1>(h->bvh.nodes.data(), h->bvh.tris.data(), tr, st,
                                             stack))
                        break;
                    if (tri)
                        ++nt;
                    else
                        ++nn;
                }
                wl::trav_finish(st.res);
                res = st.res;
            }
            float *o = hits + 4 * (size_t)i;
            o[0] = res.t;
            std::memcpy(&o[1], &res.prim, 4);
            o[2] = res.b1;
            o[3] = res.b2;
            if (counts)
            {
                counts[2 * (size_t)i] = nn;
                counts[2 * (size_t)i + 1] = nt;
            }
        }
    };
    if (nthreads < 1)
        nthreads = 1;
    std::vector<std::thread> th;
    uint32_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t)
    {
        uint32_t b = t * chunk, e = std::min<uint64_t>(n, (uint64_t)b + chunk);
        if (b >= e)
            break;
        th.emplace_back(work, b, e);
    }
    for (auto &x : th)
        x.join();
    return g_faults.load();
}

// Serving schedule of n requests (req4: width, height, spp, frame_offset
// each) in megabatches of `capacity` paths, walked as the multi-request
// kernels walk it: every path of every megabatch through find_item to its
// item, pixel and sample (k_raygen_multi), every pixel of the megabatch's
// pixel list to its item and samples (k_accumulate_multi).  Counts each
// violation: an index outside the batch buffers, the item or the request's
// film; a path or pixel mapped to an item that does not contain it; a
// request sample (pixel, frame) covered other than exactly once, or a
// pixel's samples reaching its film out of ascending frame order.
// out4: megabatches, items, paths, largest megabatch.
uint64_t wlh_serve_check(const uint32_t *req4, uint32_t n, uint32_t capacity, uint64_t *out4)
{
    std::vector<wl::PlanRequest> r(n);
    for (uint32_t k = 0; k < n; ++k)
        r[k] = { req4[4 * k], req4[4 * k + 1], req4[4 * k + 2], req4[4 * k + 3] };
    wl::Plan P;
    wl::plan(r.data(), n, capacity, P);
    uint64_t bad = 0, maxp = 0;
    // per request: samples seen per pixel (counts), last frame added per pixel
    std::vector<std::vector<uint32_t>> seen(n), last(n);
    for (uint32_t k = 0; k < n; ++k)
    {
        seen[k].assign((size_t)r[k].width * r[k].height, 0);
        last[k].assign((size_t)r[k].width * r[k].height, 0xFFFFFFFFu);
    }
    std::vector<uint32_t> items_of_req(n);
    for (size_t b = 0; b < P.batches.size(); ++b)
    {
        const wl::PtBatchHeader &h = P.batches[b];
        maxp = std::max<uint64_t>(maxp, h.paths);
        if (h.paths > capacity || h.pixels > capacity ||
            (uint64_t)h.item_offset + h.items > P.items.size())
        {
            ++bad;
            continue;
        }
        const wl::PtItem *items = P.items.data() + h.item_offset;
        std::fill(items_of_req.begin(), items_of_req.end(), 0u);
        for (uint32_t i = 0; i < h.items; ++i)
        {
            if (items[i].req >= n)
                ++bad;
            else if (++items_of_req[items[i].req] > 1) // one item per request
                ++bad;
        }
        // k_raygen_multi: path p -> item, pixel, sample
        for (uint32_t p = 0; p < h.paths; ++p)
        {
            const uint32_t j = wl::find_item(items, h.items, p, false);
            if (j >= h.items)
            {
                ++bad;
                continue;
            }
            const wl::PtItem &it = items[j];
            if (p < it.path0 || p >= it.path0 + it.npix * it.nsamples)
            {
                ++bad;
                continue;
            }
            const uint32_t local = p - it.path0, pixel = local % it.npix, s = local / it.npix;
            const wl::PlanRequest &q = r[it.req];
            if (pixel >= q.width * q.height || pixel % q.width >= q.width ||
                pixel / q.width >= q.height || s >= it.nsamples ||
                it.frame0 + s >= q.frame_offset + q.spp || it.frame0 + s < q.frame_offset)
                ++bad;
        }
        // k_accumulate_multi: megabatch pixel x -> item, film pixel, its samples
        for (uint32_t x = 0; x < h.pixels; ++x)
        {
            const uint32_t j = wl::find_item(items, h.items, x, true);
            if (j >= h.items)
            {
                ++bad;
                continue;
            }
            const wl::PtItem &it = items[j];
            if (x < it.px0 || x >= it.px0 + it.npix)
            {
                ++bad;
                continue;
            }
            const uint32_t px = x - it.px0;
            const wl::PlanRequest &q = r[it.req];
            if (px >= (uint64_t)q.width * q.height)
            {
                ++bad;
                continue;
            }
            for (uint32_t s = 0; s < it.nsamples; ++s)
            {
                const uint64_t slot = it.path0 + (uint64_t)s * it.npix + px;
                if (slot >= h.paths || slot >= capacity)
                    ++bad;
                const uint32_t f = it.frame0 + s;
                uint32_t &l = last[it.req][px];
                if (l != 0xFFFFFFFFu && f <= l)
                    ++bad; // not ascending
                l = f;
                ++seen[it.req][px];
            }
        }
    }
    for (uint32_t k = 0; k < n; ++k)
    {
        for (size_t px = 0; px < seen[k].size(); ++px)
            if (seen[k][px] != r[k].spp ||
                (r[k].spp && last[k][px] != r[k].frame_offset + r[k].spp - 1))
                ++bad;
        if (r[k].spp && P.last_batch[k] >= P.batches.size())
            ++bad;
    }
    if (out4)
    {
        out4[0] = P.batches.size();
        out4[1] = P.items.size();
        out4[2] = P.paths;
        out4[3] = maxp;
    }
    return bad;
}

// The counter advance of the run-time chunk schedule (dyn_xcd_chunks).
uint32_t wlh_dyn_chunks(uint32_t n, uint32_t xcd, uint32_t chunk)
{
    return wl::dyn_xcd_chunks(n, xcd, chunk);
}

} // extern "C"
