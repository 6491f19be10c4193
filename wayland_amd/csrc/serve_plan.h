// serve_plan.h — the batched-serving schedule: how requests' whole-image
// sample passes are packed into megabatches (plan, host), and how a
// megabatch's path or pixel index finds its work item (find_item, host and
// device).  Shared by wl_serve.cpp, the multi-request kernels in
// pathtrace.hip and the CPU emulation (hostemu.cpp), which checks every
// index the kernels derive from them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "wl_pt.h"

namespace wl
{

// One request as the planner sees it.
struct PlanRequest
{
    uint32_t width, height, spp, frame_offset;
};

struct Plan
{
    std::vector<PtItem> items;
    std::vector<PtBatchHeader> batches;
    std::vector<uint32_t> last_batch; // per request
    uint64_t paths = 0;
};

// FIFO packing: each megabatch takes, in request order, as many whole-image
// sample passes of every unfinished request as still fit.  (Items stay in
// request order: grouping a megabatch's items by BVH was measured slower,
// since the traversal's static per-wave ray ranges then get uneven work.)
// Every request's pixel count must be at most `capacity`.
inline void plan(const PlanRequest *r, uint32_t n, uint32_t capacity, Plan &P)
{
    std::vector<uint32_t> rem(n), next(n, 0);
    P.items.clear();
    P.batches.clear();
    P.paths = 0;
    P.last_batch.assign(n, 0);
    uint32_t pending = 0;
    for (uint32_t k = 0; k < n; ++k)
    {
        rem[k] = r[k].spp;
        pending += rem[k] > 0;
    }
    while (pending)
    {
        PtBatchHeader h{ 0, 0, 0, (uint32_t)P.items.size() };
        uint32_t free_paths = capacity;
        for (uint32_t k = 0; k < n && free_paths; ++k)
        {
            if (!rem[k])
                continue;
            const uint32_t npix = r[k].width * r[k].height;
            if (npix > free_paths)
                continue;
            const uint32_t ns = std::min(rem[k], free_paths / npix);
            PtItem it{};
            it.req = k;
            it.npix = npix;
            it.frame0 = r[k].frame_offset + next[k];
            it.nsamples = ns;
            it.path0 = h.paths;
            it.px0 = h.pixels;
            P.items.push_back(it);
            h.paths += ns * npix;
            h.pixels += npix;
            h.items += 1;
            free_paths -= ns * npix;
            next[k] += ns;
            rem[k] -= ns;
            P.last_batch[k] = (uint32_t)P.batches.size();
            if (!rem[k])
                --pending;
        }
        P.paths += h.paths;
        P.batches.push_back(h);
    }
}

// Item of megabatch entry x (a path index, or with by_pixel a pixel index of
// the megabatch's pixel list) given the items' start offsets: binary search
// for items[lo].start <= x < items[lo + 1].start.
__host__ __device__ __forceinline__ uint32_t find_item(const PtItem *items, uint32_t nitems,
                                                       uint32_t x, bool by_pixel)
{
    uint32_t lo = 0, hi = nitems;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) / 2;
        const uint32_t st = by_pixel ? items[mid].px0 : items[mid].path0;
        if (st <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

} // namespace wl
