// hostemu.cpp — CPU emulation of the BVH8 builder + traversal kernels, for
// tests only (libwl_hostemu.so).  It compiles the exact traversal code of
// traverse_device.h for the host, with every node index, triangle index and
// stack slot the kernel would touch checked against its allocation, so the
// indexing is verified on the CPU before any kernel runs on a GPU.
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "bvh_builder.h"

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

} // extern "C"
