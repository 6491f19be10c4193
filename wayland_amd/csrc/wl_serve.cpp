// wl_serve.cpp — batched serving: many render requests (each a scene, an
// image size, a sample count and path-tracer parameters) in shared
// megabatches.  The reference renders one launch per request and frame
// (DeviceManager::Launch, src/Frontend/Core/DeviceManager.cpp:231-241);
// here the scheduler packs whole requests x sample ranges into megabatches
// of up to batch_paths paths, and a render call's whole megabatch sequence
// (per megabatch: raygen, max_depth x extend / shade / shadow, film update)
// is captured into one hipGraph and launched once: the batch descriptors live
// in device memory and each megabatch's first kernel picks the next one, so
// the host enqueues one graph launch per lane (one per kGraphBatches
// megabatches beyond that).  With two lanes the requests are
// split into two groups (each request whole in one group) whose megabatch
// sequences run concurrently on two streams with their own buffers, so each
// lane's kernels fill the other's tails.  An autotuner times candidate
// megabatch sizes, shade grids and lane counts on the workload itself.
//
// Each request's film equals its standalone wl_pt_render result bit for bit:
// a path's arithmetic depends only on its own state and its request's scene,
// and every pixel receives its samples in ascending order.
#include "wl_internal.h"
#include "serve_plan.h"
#include "wl_pt.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

using wl::set_error;

struct wl_server_s
{
    wl_context ctx = nullptr;
    wl_serve_config cfg{};
    hipStream_t stream = nullptr; // capture and replay stream
    // batch buffers
    void *d_buf = nullptr;
    uint32_t capacity = 0;
    uint32_t *d_counters = nullptr;
    unsigned long long *d_stats = nullptr;
    // request / batch descriptors
    wl::PtReq *d_reqs = nullptr;
    uint32_t reqs_cap = 0;
    wl::PtItem *d_items = nullptr;
    uint32_t items_cap = 0;
    wl::PtBatchHeader *d_batches = nullptr;
    uint32_t batches_cap = 0;
    uint32_t *d_batch_index = nullptr;
    wl::PtBatchHeader *d_header = nullptr;
    // the requests' BVHs back to back (traversal arena), refilled every call
    WlBvh8Node *d_arena_nodes = nullptr;
    uint32_t arena_nodes_cap = 0;
    WlTriRecord *d_arena_tris = nullptr;
    uint32_t arena_tris_cap = 0;
    // captured megabatch sequences and everything they bake in (compared field
    // by field): `exec` runs graph_batches megabatches, `exec_rest` the
    // remainder of a call with more than kGraphBatches
    hipGraphExec_t exec = nullptr, exec_rest = nullptr;
    std::vector<uint64_t> graph_key, rest_key;
    // device wall clock at the end of each megabatch, and at the call's start
    unsigned long long *d_times = nullptr;
    uint32_t times_cap = 0;
    unsigned long long *d_clock = nullptr;
    double clock_hz = 1.0e8;
    // second lane (created on first use): with cfg.lanes = 2 the requests are
    // split into two groups whose megabatches run concurrently, the second
    // group through this server's twin (own buffers, stream and graph)
    wl_server_s *twin = nullptr;
};

namespace
{

constexpr uint32_t kDefaultServeBatch = 1u << 22;
// megabatches per captured graph (about 50 kernel nodes each at depth 16)
constexpr uint32_t kGraphBatches = 64;

template <class T>
wl_status grow(T **p, uint32_t &cap, uint32_t need)
{
    if (need <= cap)
        return WL_SUCCESS;
    if (*p)
    {
        WL_HIP_CHECK(hipDeviceSynchronize());
        WL_HIP_CHECK(hipFree(*p));
        *p = nullptr;
        cap = 0;
    }
    const uint32_t n = std::max<uint32_t>(need, 64);
    WL_HIP_CHECK(hipMalloc((void **)p, sizeof(T) * (size_t)n));
    cap = n;
    return WL_SUCCESS;
}

void drop_graph(wl_server s)
{
    if (s->exec)
    {
        hipGraphExecDestroy(s->exec);
        s->exec = nullptr;
    }
    if (s->exec_rest)
    {
        hipGraphExecDestroy(s->exec_rest);
        s->exec_rest = nullptr;
    }
    s->graph_key.clear();
    s->rest_key.clear();
}

// Captures `count` consecutive megabatches into *exec (replacing it) unless
// it already holds them for `key`.
wl_status capture_batches(wl_server s, const wl::PtMultiLaunch &L, const wl::PtBuffers &B,
                          uint32_t count, const std::vector<uint64_t> &key,
                          hipGraphExec_t *exec, std::vector<uint64_t> *held)
{
    if (*exec && *held == key)
        return WL_SUCCESS;
    if (*exec)
    {
        hipGraphExecDestroy(*exec);
        *exec = nullptr;
        held->clear();
    }
    wl::pt_multi_init();
    hipGraph_t g = nullptr;
    WL_HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = hipSuccess;
    for (uint32_t b = 0; b < count && e1 == hipSuccess; ++b)
        e1 = wl::pt_run_multi(L, B, s->ctx->num_cus, s->stream);
    const hipError_t e2 = hipStreamEndCapture(s->stream, &g);
    if (e1 != hipSuccess || e2 != hipSuccess)
    {
        if (g)
            hipGraphDestroy(g);
        return wl::hip_error(e1 != hipSuccess ? e1 : e2, "megabatch capture");
    }
    const hipError_t e3 = hipGraphInstantiateWithFlags(exec, g, 0);
    hipGraphDestroy(g);
    if (e3 != hipSuccess)
    {
        *exec = nullptr;
        return wl::hip_error(e3, "megabatch graph instantiate");
    }
    *held = key;
    return WL_SUCCESS;
}

wl_status ensure_buffers(wl_server s, uint32_t capacity)
{
    if (capacity == s->capacity && s->d_buf)
        return WL_SUCCESS;
    drop_graph(s);
    if (s->d_buf)
    {
        WL_HIP_CHECK(hipDeviceSynchronize());
        WL_HIP_CHECK(hipFree(s->d_buf));
        s->d_buf = nullptr;
        s->capacity = 0;
    }
    WL_HIP_CHECK(hipMalloc(&s->d_buf, wl::pt_batch_bytes(capacity)));
    s->capacity = capacity;
    return WL_SUCCESS;
}

wl_pt_config request_config(const wl_render_request &r)
{
    wl_pt_config c{};
    c.width = r.width;
    c.height = r.height;
    c.spp = r.spp;
    c.frame_offset = r.frame_offset;
    c.max_depth = r.max_depth;
    c.rr_depth = r.rr_depth;
    c.rr_prob = r.rr_prob;
    c.epsilon = r.epsilon;
    c.strategy = r.strategy;
    return c;
}

wl_status normalise_config(wl_serve_config &c)
{
    if (c.batch_paths == 0)
        c.batch_paths = kDefaultServeBatch;
    if (c.shade_blocks_per_cu == 0)
        c.shade_blocks_per_cu = 8;
    if (c.batch_paths > (1u << 26))
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "batch_paths above 2^26");
    if (c.shade_blocks_per_cu > 64)
        return set_error(WL_ERROR_INVALID_VALUE, "shade_blocks_per_cu above 64");
    if (c.lanes == 0)
        c.lanes = 1;
    if (c.lanes > 2)
        return set_error(WL_ERROR_INVALID_VALUE, "lanes above 2");
    return WL_SUCCESS;
}

// What a lane enqueued, for finish_lane.
struct LaneRun
{
    wl::Plan pl;
    uint32_t nb = 0;
    uint32_t launches = 0; // graph launches (or megabatch sequences) enqueued
};

// Checks every request before any lane enqueues work, so an invalid request
// leaves every film untouched.
wl_status validate_requests(wl_server s, const wl_serve_config &cfg,
                            const wl_render_request *reqs, uint32_t n)
{
    for (uint32_t k = 0; k < n; ++k)
    {
        const wl_render_request &r = reqs[k];
        if (!r.scene || !r.film)
            return set_error(WL_ERROR_INVALID_VALUE, "request without scene or film");
        if (r.scene->ctx->device != s->ctx->device)
            return set_error(WL_ERROR_INVALID_VALUE, "request scene on another device");
        if (r.scene->guide && r.scene->guide_mode)
            return set_error(WL_ERROR_NOT_SUPPORTED, "batched serving of a guided scene");
        const wl_pt_config c = request_config(r);
        const wl_status st = wl::pt_check_config(r.scene, &c);
        if (st != WL_SUCCESS)
            return st;
        if (r.width * r.height > cfg.batch_paths)
            return set_error(WL_ERROR_LIMIT_EXCEEDED,
                             "batch_paths must hold the largest request's pixel count");
    }
    // Films must not overlap: one megabatch (or two lanes at once) updates
    // every request's film, so two requests sharing pixels would race.
    std::vector<std::pair<uintptr_t, uintptr_t>> span(n);
    for (uint32_t k = 0; k < n; ++k)
    {
        const uintptr_t b = reinterpret_cast<uintptr_t>(reqs[k].film);
        span[k] = { b, b + sizeof(float) * 6u * (uintptr_t)reqs[k].width * reqs[k].height };
    }
    std::sort(span.begin(), span.end());
    for (uint32_t k = 1; k < n; ++k)
        if (span[k].first < span[k - 1].second)
            return set_error(WL_ERROR_INVALID_VALUE, "two requests' films overlap");
    return WL_SUCCESS;
}

// Builds the lane's plan and arena and enqueues its megabatches on s->stream
// after the work already on `user` (the requests are validated); returns
// without waiting.
wl_status run_lane(wl_server s, const wl_serve_config &cfg, const wl_render_request *reqs,
                   uint32_t n, hipStream_t user, LaneRun &out)
{
    wl_status st = WL_SUCCESS;
    std::vector<wl::PtReq> hreq(n);
    uint32_t max_npix = 0, max_depth = 0, max_bvh = 0;
    for (uint32_t k = 0; k < n; ++k)
    {
        const wl_render_request &r = reqs[k];
        if (!r.scene || !r.film)
            return set_error(WL_ERROR_INVALID_VALUE, "request without scene or film");
        if (r.scene->ctx->device != s->ctx->device)
            return set_error(WL_ERROR_INVALID_VALUE, "request scene on another device");
        if (r.scene->guide && r.scene->guide_mode)
            return set_error(WL_ERROR_NOT_SUPPORTED, "batched serving of a guided scene");
        const wl_pt_config c = request_config(r);
        st = wl::pt_check_config(r.scene, &c);
        if (st != WL_SUCCESS)
            return st;
        hreq[k].P = wl::pt_params_for(r.scene, &c);
        hreq[k].S = wl::pt_scene_view(r.scene);
        hreq[k].film = r.film;
        max_npix = std::max(max_npix, r.width * r.height);
        max_depth = std::max(max_depth, r.max_depth);
        max_bvh = std::max(max_bvh, r.scene->accel->info.depth);
    }
    // traversal arena: each distinct BVH once, its child and triangle indices
    // rebased, so the multi-request traversal starts a ray at its request's
    // root and otherwise runs the single-BVH kernel
    std::vector<const wl_accel_s *> accs;
    std::vector<uint32_t> acc_node0, acc_tri0;
    uint64_t arena_nodes = 0, arena_tris = 0;
    for (uint32_t k = 0; k < n; ++k)
    {
        const wl_accel_s *a = reqs[k].scene->accel;
        size_t j = 0;
        while (j < accs.size() && accs[j] != a)
            ++j;
        if (j == accs.size())
        {
            accs.push_back(a);
            acc_node0.push_back((uint32_t)arena_nodes);
            acc_tri0.push_back((uint32_t)arena_tris);
            arena_nodes += a->num_nodes;
            arena_tris += std::max<uint32_t>(1, a->num_tri_records);
            // stack entries hold child group indices / 8 in 24 bits
            if (arena_nodes >= (1ull << 27) || arena_tris >= (1ull << 32))
                return set_error(WL_ERROR_LIMIT_EXCEEDED,
                                 "the requests' BVHs exceed the traversal arena (2^27 nodes)");
        }
        hreq[k].root = acc_node0[j];
    }
    if (max_npix > cfg.batch_paths)
        return set_error(WL_ERROR_LIMIT_EXCEEDED,
                         "batch_paths must hold the largest request's pixel count");
    wl::Plan &pl = out.pl;
    std::vector<wl::PlanRequest> preq(n);
    for (uint32_t k = 0; k < n; ++k)
        preq[k] = { reqs[k].width, reqs[k].height, reqs[k].spp, reqs[k].frame_offset };
    wl::plan(preq.data(), n, cfg.batch_paths, pl);
    const uint32_t nb = (uint32_t)pl.batches.size();
    out.nb = nb;
    st = ensure_buffers(s, cfg.batch_paths);
    if (st == WL_SUCCESS)
        st = grow(&s->d_reqs, s->reqs_cap, n);
    if (st == WL_SUCCESS)
        st = grow(&s->d_items, s->items_cap, (uint32_t)pl.items.size());
    if (st == WL_SUCCESS)
        st = grow(&s->d_batches, s->batches_cap, nb);
    if (st == WL_SUCCESS)
        st = grow(&s->d_times, s->times_cap, std::max<uint32_t>(nb, 1));
    if (st == WL_SUCCESS)
        st = grow(&s->d_arena_nodes, s->arena_nodes_cap, (uint32_t)arena_nodes);
    if (st == WL_SUCCESS)
        st = grow(&s->d_arena_tris, s->arena_tris_cap, (uint32_t)arena_tris);
    if (st != WL_SUCCESS)
        return st;
    // the descriptors of this call; the graph reads them through fixed pointers
    WL_HIP_CHECK(hipMemcpy(s->d_reqs, hreq.data(), sizeof(wl::PtReq) * n, hipMemcpyHostToDevice));
    WL_HIP_CHECK(hipMemcpy(s->d_items, pl.items.data(), sizeof(wl::PtItem) * pl.items.size(),
                           hipMemcpyHostToDevice));
    WL_HIP_CHECK(hipMemcpy(s->d_batches, pl.batches.data(), sizeof(wl::PtBatchHeader) * nb,
                           hipMemcpyHostToDevice));

    wl::PtBuffers B = wl::pt_carve(s->d_buf, s->capacity, s->d_counters, s->d_stats);
    wl::PtMultiLaunch L{};
    L.reqs = s->d_reqs;
    L.items = s->d_items;
    L.batches = s->d_batches;
    L.batch_index = s->d_batch_index;
    L.header = s->d_header;
    L.times = s->d_times;
    L.nodes = s->d_arena_nodes;
    L.tris = s->d_arena_tris;
    L.max_depth = max_depth;
    L.stack16 = max_bvh <= 16 ? 1u : 0u;
    L.capacity = s->capacity;
    L.shade_blocks_per_cu = cfg.shade_blocks_per_cu;

    hipEvent_t ready;
    WL_HIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    WL_HIP_CHECK(hipEventRecord(ready, user));
    WL_HIP_CHECK(hipStreamWaitEvent(s->stream, ready, 0));
    WL_HIP_CHECK(hipEventDestroy(ready));
    for (size_t j = 0; j < accs.size(); ++j)
    {
        const wl_accel_s *a = accs[j];
        WL_HIP_CHECK(wl::pt_arena_copy_nodes(s->d_arena_nodes, a->d_nodes, a->num_nodes,
                                             acc_node0[j], acc_tri0[j], s->stream));
        if (a->num_tri_records)
            WL_HIP_CHECK(hipMemcpyAsync(s->d_arena_tris + acc_tri0[j], a->d_tris,
                                        sizeof(WlTriRecord) * (size_t)a->num_tri_records,
                                        hipMemcpyDeviceToDevice, s->stream));
    }

    // the graphs bake in buffer pointers, grids, the bounce count and the
    // number of megabatches they run
    const uint32_t per_graph = std::min(nb, kGraphBatches);
    const uint32_t full = per_graph ? nb / per_graph : 0, rest = per_graph ? nb % per_graph : 0;
    if (cfg.use_graph && nb)
    {
        std::vector<uint64_t> key = {
            max_depth,
            L.stack16,
            cfg.shade_blocks_per_cu,
            s->capacity,
            (uint64_t)(uintptr_t)s->d_buf,
            (uint64_t)(uintptr_t)s->d_reqs,
            (uint64_t)(uintptr_t)s->d_items,
            (uint64_t)(uintptr_t)s->d_batches,
            (uint64_t)(uintptr_t)s->d_arena_nodes,
            (uint64_t)(uintptr_t)s->d_arena_tris,
            (uint64_t)(uintptr_t)s->d_times,
            per_graph,
        };
        st = capture_batches(s, L, B, per_graph, key, &s->exec, &s->graph_key);
        if (st == WL_SUCCESS && rest)
        {
            key.back() = rest;
            st = capture_batches(s, L, B, rest, key, &s->exec_rest, &s->rest_key);
        }
        if (st != WL_SUCCESS)
            return st;
    }
    // counters: zeroed by a plain copy before the launch (never inside the graph)
    const uint32_t zeros[16] = {};
    WL_HIP_CHECK(hipMemcpyAsync(s->d_batch_index, zeros, 4, hipMemcpyHostToDevice, s->stream));
    WL_HIP_CHECK(hipMemcpyAsync(s->d_stats, zeros, 16, hipMemcpyHostToDevice, s->stream));
    if (cfg.use_graph)
    {
        for (uint32_t k = 0; k < full; ++k)
            WL_HIP_CHECK(hipGraphLaunch(s->exec, s->stream));
        if (rest)
            WL_HIP_CHECK(hipGraphLaunch(s->exec_rest, s->stream));
    }
    else
    {
        for (uint32_t b = 0; b < nb; ++b)
            WL_HIP_CHECK(wl::pt_run_multi(L, B, s->ctx->num_cus, s->stream));
    }
    out.launches = cfg.use_graph ? full + (rest ? 1u : 0u) : nb;
    return WL_SUCCESS;
}

// Waits for a lane; its megabatch end times relative to `origin` (the device
// clock stamped on the caller's stream before any lane started), per request
// in `done` (lane order), and its ray counts added to `rays`.
wl_status finish_lane(wl_server s, const LaneRun &lr, const wl_render_request *reqs, uint32_t n,
                      unsigned long long origin, double hz, double *done, double &end,
                      unsigned long long rays[2])
{
    WL_HIP_CHECK(hipStreamSynchronize(s->stream));
    std::vector<unsigned long long> tb(lr.nb);
    if (lr.nb)
        WL_HIP_CHECK(hipMemcpy(tb.data(), s->d_times, sizeof(unsigned long long) * lr.nb,
                               hipMemcpyDeviceToHost));
    auto secs = [&](unsigned long long t) { return t > origin ? (double)(t - origin) / hz : 0.0; };
    end = lr.nb ? secs(tb[lr.nb - 1]) : 0.0;
    for (uint32_t k = 0; k < n; ++k)
        done[k] = reqs[k].spp ? secs(tb[lr.pl.last_batch[k]]) : 0.0;
    unsigned long long h[2] = { 0, 0 };
    WL_HIP_CHECK(hipMemcpy(h, s->d_stats, 16, hipMemcpyDeviceToHost));
    rays[0] += h[0];
    rays[1] += h[1];
    return WL_SUCCESS;
}

// Creates the twin lane of `s` (same device and configuration); servers are
// created and freed by create_server / destroy_server.
wl_status make_twin(wl_server s);
wl_status create_server(wl_context ctx, const wl_serve_config &c, wl_server *out);
wl_status destroy_server(wl_server s);

wl_status run(wl_server s, const wl_render_request *reqs, uint32_t n, wl_serve_stats *stats,
              double *done_seconds, hipStream_t user)
{
    if (!s || (n && !reqs))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (n == 0)
        return WL_SUCCESS;
    wl_serve_config cfg = s->cfg;
    wl_status st = normalise_config(cfg);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    st = validate_requests(s, cfg, reqs, n);
    if (st != WL_SUCCESS)
        return st;
    // requests -> lanes: each request to the lane with fewer paths so far
    // (a request stays in one lane, so its film gets its samples in order)
    std::vector<uint32_t> idx[2];
    uint64_t load[2] = { 0, 0 };
    const uint32_t lanes = (cfg.lanes == 2 && n >= 2) ? 2u : 1u;
    for (uint32_t k = 0; k < n; ++k)
    {
        const uint32_t l = (lanes == 2 && load[1] < load[0]) ? 1u : 0u;
        idx[l].push_back(k);
        load[l] += (uint64_t)reqs[k].width * reqs[k].height * reqs[k].spp;
    }
    const uint32_t used = idx[1].empty() ? 1u : 2u;
    if (used == 2 && !s->twin)
    {
        st = make_twin(s);
        if (st != WL_SUCCESS)
            return st;
    }
    std::vector<wl_render_request> sub[2];
    for (uint32_t l = 0; l < used; ++l)
        for (uint32_t k : idx[l])
            sub[l].push_back(reqs[k]);
    wl_server lane[2] = { s, s->twin };
    // the call's start on the device clock, once the work already queued on
    // `user` is done (each lane starts after it)
    WL_HIP_CHECK(wl::pt_clock_stamp(s->d_clock, user));
    LaneRun lr[2];
    st = WL_SUCCESS;
    for (uint32_t l = 0; l < used && st == WL_SUCCESS; ++l)
        st = run_lane(lane[l], cfg, sub[l].data(), (uint32_t)sub[l].size(), user, lr[l]);
    unsigned long long origin = 0;
    if (st == WL_SUCCESS)
    {
        // the lanes waited for the stamp: it is written once a lane is done
        WL_HIP_CHECK(hipStreamSynchronize(s->stream));
        WL_HIP_CHECK(hipMemcpy(&origin, s->d_clock, sizeof origin, hipMemcpyDeviceToHost));
    }
    std::vector<double> done[2];
    double end = 0.0;
    unsigned long long rays[2] = { 0, 0 };
    for (uint32_t l = 0; l < used && st == WL_SUCCESS; ++l)
    {
        done[l].assign(sub[l].size(), 0.0);
        double e = 0.0;
        st = finish_lane(lane[l], lr[l], sub[l].data(), (uint32_t)sub[l].size(), origin,
                         s->clock_hz, done[l].data(), e, rays);
        end = std::max(end, e);
    }
    if (st != WL_SUCCESS)
    {
        // a lane may still hold enqueued work: let it drain before returning
        for (uint32_t l = 0; l < used; ++l)
            if (lane[l])
                hipStreamSynchronize(lane[l]->stream);
    }
    if (st != WL_SUCCESS)
        return st;
    if (done_seconds)
        for (uint32_t l = 0; l < used; ++l)
            for (size_t j = 0; j < idx[l].size(); ++j)
                done_seconds[idx[l][j]] = done[l][j];
    if (stats)
    {
        std::memset(stats, 0, sizeof *stats);
        // from the call's start (after the work already queued on `user`)
        // to the end of the last lane, on the device clock
        stats->seconds = end;
        for (uint32_t l = 0; l < used; ++l)
        {
            stats->paths += lr[l].pl.paths;
            stats->batches += lr[l].nb;
            stats->launches += lr[l].launches;
        }
        stats->extension_rays = rays[0];
        stats->shadow_rays = rays[1];
    }
    return WL_SUCCESS;
}

wl_status create_server(wl_context ctx, const wl_serve_config &c, wl_server *out)
{
    wl_server s = new (std::nothrow) wl_server_s;
    if (!s)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "server allocation");
    s->ctx = ctx;
    s->cfg = c;
    wl::ctx_retain(ctx);
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_counters, 64);
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_stats, 64);
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_batch_index, 64);
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_header, 64);
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_clock, 64);
    int khz = 0;
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device);
    if (e == hipSuccess && khz > 0)
        s->clock_hz = 1e3 * khz;
    if (e != hipSuccess)
    {
        destroy_server(s);
        return wl::hip_error(e, "server set-up");
    }
    *out = s;
    return WL_SUCCESS;
}

wl_status make_twin(wl_server s)
{
    wl_serve_config c = s->cfg;
    c.lanes = 1;
    return create_server(s->ctx, c, &s->twin);
}

wl_status destroy_server(wl_server s)
{
    hipSetDevice(s->ctx->device);
    hipDeviceSynchronize();
    drop_graph(s);
    hipFree(s->d_buf);
    hipFree(s->d_times);
    hipFree(s->d_clock);
    hipFree(s->d_counters);
    hipFree(s->d_stats);
    hipFree(s->d_reqs);
    hipFree(s->d_items);
    hipFree(s->d_batches);
    hipFree(s->d_batch_index);
    hipFree(s->d_header);
    hipFree(s->d_arena_nodes);
    hipFree(s->d_arena_tris);
    if (s->stream)
        hipStreamDestroy(s->stream);
    wl_context ctx = s->ctx;
    delete s;
    wl::ctx_release(ctx);
    return WL_SUCCESS;
}

} // namespace

extern "C" {

wl_status wl_serve_default_config(wl_serve_config *cfg)
{
    if (!cfg)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    std::memset(cfg, 0, sizeof *cfg);
    cfg->batch_paths = kDefaultServeBatch;
    cfg->shade_blocks_per_cu = 8;
    cfg->use_graph = 1;
    cfg->lanes = 2;
    return WL_SUCCESS;
}

wl_status wl_server_create(wl_context ctx, const wl_serve_config *cfg, wl_server *out)
{
    if (!ctx || !out)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    wl_serve_config c;
    if (cfg)
        c = *cfg;
    else
        wl_serve_default_config(&c);
    wl_status st = normalise_config(c);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(hipSetDevice(ctx->device));
    return create_server(ctx, c, out);
}

wl_status wl_server_destroy(wl_server s)
{
    if (!s)
        return set_error(WL_ERROR_INVALID_VALUE, "server is NULL");
    if (s->twin)
        wl_server_destroy(s->twin);
    s->twin = nullptr;
    return destroy_server(s);
}

wl_status wl_server_get_config(wl_server s, wl_serve_config *cfg)
{
    if (!s || !cfg)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *cfg = s->cfg;
    return WL_SUCCESS;
}

wl_status wl_server_set_config(wl_server s, const wl_serve_config *cfg)
{
    if (!s || !cfg)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    wl_serve_config c = *cfg;
    wl_status st = normalise_config(c);
    if (st != WL_SUCCESS)
        return st;
    s->cfg = c;
    return WL_SUCCESS;
}

wl_status wl_server_render(wl_server s, const wl_render_request *reqs, uint32_t n,
                           wl_serve_stats *stats, double *done_seconds, void *stream)
{
    return run(s, reqs, n, stats, done_seconds, (hipStream_t)stream);
}

wl_status wl_server_autotune(wl_server s, const wl_render_request *reqs, uint32_t n,
                             uint32_t max_spp, wl_serve_config *chosen)
{
    if (!s || !reqs || n == 0)
        return set_error(WL_ERROR_INVALID_VALUE, "bad argument");
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    // the workload with every request cut to max_spp samples, into scratch films
    std::vector<wl_render_request> probe(reqs, reqs + n);
    size_t film_floats = 0;
    uint32_t max_npix = 0;
    for (auto &r : probe)
    {
        r.spp = std::min(r.spp, std::max<uint32_t>(1, max_spp));
        film_floats += 6 * (size_t)r.width * r.height;
        max_npix = std::max(max_npix, r.width * r.height);
    }
    float *scratch = nullptr;
    WL_HIP_CHECK(hipMalloc((void **)&scratch, 4 * film_floats));
    size_t off = 0;
    for (auto &r : probe)
    {
        r.film = scratch + off;
        off += 6 * (size_t)r.width * r.height;
    }
    const wl_serve_config saved = s->cfg;
    wl_serve_config best = saved;
    double best_rate = -1.0;
    const uint32_t batches[] = { 1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25,
                                 1u << 26 };
    wl_status st = WL_SUCCESS;
    // times one candidate (a warm run that captures its graph, then a timed
    // one); returns its rate in paths per second (0 on failure)
    auto trial = [&](const wl_serve_config &c) {
        s->cfg = c;
        wl_serve_stats a{}, b{};
        st = run(s, probe.data(), n, &a, nullptr, nullptr);
        if (st == WL_SUCCESS)
            st = run(s, probe.data(), n, &b, nullptr, nullptr);
        if (st != WL_SUCCESS)
            return 0.0;
        const double rate = (double)b.paths / std::max(1e-9, b.seconds);
        if (rate > best_rate)
        {
            best_rate = rate;
            best = c;
        }
        return rate;
    };
    // coordinate search: the megabatch capacity first (8 shade workgroups per
    // CU, two lanes when there are two requests), then the shade grid at the
    // best capacity, then the other lane count -- at most 7 + 2 + 1
    // candidates instead of the 42 of the full grid
    wl_serve_config c = saved;
    c.use_graph = 1;
    c.shade_blocks_per_cu = 8;
    c.lanes = n >= 2 ? 2u : 1u;
    // the probe's spp is capped, so capacities beyond its size time alike; the
    // full workload has more samples and fills larger megabatches: take the
    // largest capacity within 2 % of the fastest
    double rate_of[7] = {};
    for (int k = 0; k < 7; ++k)
    {
        if (batches[k] < max_npix || st != WL_SUCCESS)
            continue;
        c.batch_paths = batches[k];
        rate_of[k] = trial(c);
    }
    for (int k = 6; k >= 0 && st == WL_SUCCESS && best_rate > 0; --k)
        if (rate_of[k] >= 0.98 * best_rate)
        {
            best.batch_paths = batches[k];
            break;
        }
    for (uint32_t sh : { 4u, 16u })
    {
        if (st != WL_SUCCESS || best_rate <= 0)
            break;
        c = best;
        c.shade_blocks_per_cu = sh;
        trial(c);
    }
    if (st == WL_SUCCESS && best_rate > 0 && n >= 2)
    {
        c = best;
        c.lanes = best.lanes == 2 ? 1u : 2u;
        trial(c);
    }
    hipFree(scratch);
    s->cfg = st == WL_SUCCESS && best_rate > 0 ? best : saved;
    // the candidates left the largest batch buffers behind: keep only what
    // the chosen configuration renders with (the twin lane only for lanes 2)
    if (s->twin && s->cfg.lanes != 2)
    {
        wl_server_destroy(s->twin);
        s->twin = nullptr;
    }
    if (st == WL_SUCCESS)
        st = ensure_buffers(s, s->cfg.batch_paths);
    if (st == WL_SUCCESS && s->twin)
        st = ensure_buffers(s->twin, s->cfg.batch_paths);
    if (chosen)
        *chosen = s->cfg;
    return st;
}

} // extern "C"
