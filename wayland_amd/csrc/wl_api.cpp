// wl_api.cpp — C ABI entry points: errors, contexts, acceleration structures
// and raw traversal.  The path-tracer entry points live in wl_pt_api.cpp.
#include "wl_internal.h"
#include "bvh_builder.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <thread>

namespace
{
thread_local std::string g_last_error;
}

namespace wl
{
wl_status set_error(wl_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

wl_status hip_error(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? WL_ERROR_OUT_OF_MEMORY : WL_ERROR_DEVICE;
}

void ctx_retain(wl_context ctx) { ctx->refs.fetch_add(1); }

void ctx_release(wl_context ctx)
{
    if (ctx->refs.fetch_sub(1) != 1)
        return;
    hipSetDevice(ctx->device);
    if (ctx->stream)
    {
        hipStreamSynchronize(ctx->stream);
        hipStreamDestroy(ctx->stream);
    }
    if (!ctx->sched.empty())
    {
        hipDeviceSynchronize();
        for (auto &kv : ctx->sched)
            hipFree(kv.second.dev);
    }
    delete ctx;
}

hipError_t sched_block(wl_context ctx, hipStream_t s, wl_context_s::SchedBlock **out)
{
    auto it = ctx->sched.find(s);
    if (it != ctx->sched.end())
    {
        *out = &it->second;
        return hipSuccess;
    }
    wl_context_s::SchedBlock b;
    hipError_t e = hipMalloc((void **)&b.dev, 256);
    if (e == hipSuccess)
        e = hipMemset(b.dev, 0, 256);
    if (e == hipSuccess) // the zeros land before any stream's next launch
        e = hipDeviceSynchronize();
    if (e != hipSuccess)
    {
        if (b.dev)
            hipFree(b.dev);
        return e;
    }
    *out = &ctx->sched.emplace(s, b).first->second;
    return hipSuccess;
}
} // namespace wl

using wl::set_error;

extern "C" {

const char *wl_status_string(wl_status s)
{
    switch (s)
    {
    case WL_SUCCESS: return "WL_SUCCESS";
    case WL_ERROR_INVALID_VALUE: return "WL_ERROR_INVALID_VALUE";
    case WL_ERROR_OUT_OF_MEMORY: return "WL_ERROR_OUT_OF_MEMORY";
    case WL_ERROR_DEVICE: return "WL_ERROR_DEVICE";
    case WL_ERROR_NOT_SUPPORTED: return "WL_ERROR_NOT_SUPPORTED";
    case WL_ERROR_LIMIT_EXCEEDED: return "WL_ERROR_LIMIT_EXCEEDED";
    case WL_ERROR_NO_DEVICE: return "WL_ERROR_NO_DEVICE";
    case WL_ERROR_UNKNOWN_PROGRAM: return "WL_ERROR_UNKNOWN_PROGRAM";
    case WL_ERROR_QUEUE_OVERFLOW: return "WL_ERROR_QUEUE_OVERFLOW";
    }
    return "WL_ERROR_UNKNOWN_STATUS";
}

const char *wl_last_error_message(void) { return g_last_error.c_str(); }

wl_status wl_device_count(int *count)
{
    if (!count)
        return set_error(WL_ERROR_INVALID_VALUE, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess)
    {
        *count = 0;
        return wl::hip_error(e, "hipGetDeviceCount");
    }
    *count = n;
    return WL_SUCCESS;
}

wl_status wl_context_create(int device, wl_context *out)
{
    if (!out)
        return set_error(WL_ERROR_INVALID_VALUE, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(WL_ERROR_NO_DEVICE, "no HIP device is visible");
    if (device < 0 || device >= n)
        return set_error(WL_ERROR_INVALID_VALUE, "device index out of range");
    WL_HIP_CHECK(hipSetDevice(device));
    wl_context ctx = new (std::nothrow) wl_context_s;
    if (!ctx)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "context allocation");
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
    {
        delete ctx;
        return wl::hip_error(e, "hipStreamCreateWithFlags");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        ctx->num_cus = prop.multiProcessorCount;
    // per-device set-up of wl_pt_render_noalloc's fixed scratch array
    const wl_status st = wl::noalloc_setup(device);
    if (st != WL_SUCCESS)
    {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return st;
    }
    *out = ctx;
    return WL_SUCCESS;
}

wl_status wl_context_destroy(wl_context ctx)
{
    if (!ctx)
        return set_error(WL_ERROR_INVALID_VALUE, "context is NULL");
    wl::ctx_release(ctx);
    return WL_SUCCESS;
}

wl_status wl_context_get_stream(wl_context ctx, void **stream)
{
    if (!ctx || !stream)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *stream = ctx->stream;
    return WL_SUCCESS;
}

wl_status wl_context_synchronize(wl_context ctx)
{
    if (!ctx)
        return set_error(WL_ERROR_INVALID_VALUE, "context is NULL");
    WL_HIP_CHECK(hipSetDevice(ctx->device));
    WL_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return WL_SUCCESS;
}

wl_status wl_set_cache_path(wl_context ctx, const char *path)
{
    if (!ctx || !path)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    ctx->cache_path = path;
    return WL_SUCCESS;
}

// ---- acceleration structures -------------------------------------------

// Checks every index the traversal kernels derive from the node data against
// the allocation sizes, so no kernel can read outside the node and triangle
// arrays, and the LDS stack depth bound holds.
static bool validate_bvh8(const wl::Bvh8Host &b, std::string *err)
{
    const size_t nn = b.nodes.size(), nt = b.tris.size();
    std::vector<uint32_t> depth(nn, 0);
    depth[0] = 1;
    std::vector<uint32_t> queue{ 0 };
    std::vector<uint8_t> seen(nn, 0);
    seen[0] = 1;
    for (size_t qi = 0; qi < queue.size(); ++qi)
    {
        const WlBvh8Node &n = b.nodes[queue[qi]];
        for (int s = 0; s < 8; ++s)
        {
            if (n.imask & (1u << s))
            {
                size_t c = (size_t)n.child_base + s;
                if (c >= nn || (n.child_base & 7u) != 0 || seen[c])
                {
                    *err = "BVH8 inner child index out of range";
                    return false;
                }
                seen[c] = 1;
                depth[c] = depth[queue[qi]] + 1;
                if (depth[c] > WL_BVH8_STACK_DEPTH)
                {
                    *err = "BVH8 deeper than the traversal stack";
                    return false;
                }
                queue.push_back((uint32_t)c);
            }
            else if (n.meta[s] != 0xFF)
            {
                uint32_t off = n.meta[s] >> 2, cnt = (n.meta[s] & 3u) + 1u;
                if (off + cnt > 32 || (size_t)n.tri_base + off + cnt > nt)
                {
                    *err = "BVH8 leaf triangle range out of range";
                    return false;
                }
            }
        }
    }
    return true;
}

wl_status wl_accel_build_triangles(wl_context ctx, const wl_tri_input *inputs,
                                   uint32_t num_inputs, uint32_t build_flags,
                                   wl_accel *out)
{
    (void)build_flags;
    if (!ctx || !out || (num_inputs && !inputs))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    if (num_inputs == 0)
        return set_error(WL_ERROR_INVALID_VALUE,
                         "an acceleration structure needs at least one build input");
    // gather world-space triangles in global primitive order
    uint64_t total = 0;
    for (uint32_t i = 0; i < num_inputs; ++i)
    {
        const wl_tri_input &in = inputs[i];
        if (in.num_triangles && (!in.vertices || !in.indices))
            return set_error(WL_ERROR_INVALID_VALUE, "build input without data");
        if (in.num_sbt_records > 1)
            return set_error(WL_ERROR_NOT_SUPPORTED,
                             "per-primitive SBT offsets are not supported");
        total += in.num_triangles;
    }
    if (total > 0x7FFFFFFFull)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "too many triangles");
    std::vector<float> tris;
    try
    {
        tris.resize(9 * (size_t)total);
    }
    catch (...)
    {
        return set_error(WL_ERROR_OUT_OF_MEMORY, "host triangle buffer");
    }
    wl_accel a = new (std::nothrow) wl_accel_s;
    if (!a)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "accel allocation");
    a->ctx = ctx;
    uint64_t off = 0;
    for (uint32_t i = 0; i < num_inputs; ++i)
    {
        const wl_tri_input &in = inputs[i];
        a->input_offsets.push_back((uint32_t)off);
        a->input_tri_counts.push_back(in.num_triangles);
        a->input_num_vertices.push_back(in.num_vertices);
        a->h_indices.insert(a->h_indices.end(), in.indices,
                            in.indices + 3 * (size_t)in.num_triangles);
        for (uint32_t t = 0; t < in.num_triangles; ++t)
        {
            for (int k = 0; k < 3; ++k)
            {
                uint32_t v = in.indices[3 * (size_t)t + k];
                if (v >= in.num_vertices)
                {
                    delete a;
                    return set_error(WL_ERROR_INVALID_VALUE,
                                     "triangle index out of range of the vertex array");
                }
                std::memcpy(&tris[9 * (off + t) + 3 * k], in.vertices + 3 * (size_t)v, 12);
            }
        }
        off += in.num_triangles;
    }
    wl::Bvh8Host host;
    std::string err;
    int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (!wl::build_bvh8(tris.data(), (uint32_t)total, nthreads, &host, &err))
    {
        delete a;
        return set_error(WL_ERROR_LIMIT_EXCEEDED, err);
    }
    if (!validate_bvh8(host, &err))
    {
        delete a;
        return set_error(WL_ERROR_LIMIT_EXCEEDED, err);
    }
    // record the build input and input-local index of every triangle record
    for (WlTriRecord &r : host.tris)
    {
        auto it = std::upper_bound(a->input_offsets.begin(), a->input_offsets.end(), r.prim);
        uint32_t in = (uint32_t)(it - a->input_offsets.begin()) - 1;
        // skip empty inputs that share an offset with the next one
        while (in + 1 < a->input_offsets.size() && a->input_offsets[in + 1] <= r.prim)
            ++in;
        r.input = in;
        r.local = r.prim - a->input_offsets[in];
    }
    hipError_t e = hipSetDevice(ctx->device);
    size_t nbytes = host.nodes.size() * sizeof(WlBvh8Node);
    size_t tbytes = std::max<size_t>(1, host.tris.size()) * sizeof(WlTriRecord);
    if (e == hipSuccess)
        e = hipMalloc(&a->d_nodes, nbytes);
    if (e == hipSuccess)
        e = hipMalloc(&a->d_tris, tbytes);
    if (e == hipSuccess)
        e = hipMemcpyAsync(a->d_nodes, host.nodes.data(), nbytes,
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !host.tris.empty())
        e = hipMemcpyAsync(a->d_tris, host.tris.data(),
                           host.tris.size() * sizeof(WlTriRecord),
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
        if (a->d_nodes)
            hipFree(a->d_nodes);
        if (a->d_tris)
            hipFree(a->d_tris);
        delete a;
        return wl::hip_error(e, "accel upload");
    }
    a->num_nodes = (uint32_t)host.nodes.size();
    a->num_tris = total;
    a->num_tri_records = (uint32_t)host.tris.size();
    a->h_tris = std::move(tris);
    wl_accel_info &inf = a->info;
    inf.num_triangles = total;
    inf.num_inputs = num_inputs;
    inf.num_nodes = (uint32_t)host.nodes.size();
    inf.inner_nodes = host.inner_nodes;
    inf.leaves = host.leaves;
    inf.depth = host.depth;
    inf.device_bytes = nbytes + tbytes;
    inf.sah_cost = host.sah_cost;
    inf.build_seconds = host.build_seconds;
    wl::register_accel(a);
    wl::ctx_retain(ctx);
    *out = a;
    return WL_SUCCESS;
}

wl_status wl_accel_destroy(wl_accel a)
{
    if (!a)
        return set_error(WL_ERROR_INVALID_VALUE, "accel is NULL");
    wl::unregister_accel(a);
    hipSetDevice(a->ctx->device);
    hipDeviceSynchronize();
    hipFree(a->d_nodes);
    hipFree(a->d_tris);
    wl_context ctx = a->ctx;
    delete a;
    wl::ctx_release(ctx);
    return WL_SUCCESS;
}

wl_status wl_accel_get_info(wl_accel a, wl_accel_info *info)
{
    if (!a || !info)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *info = a->info;
    return WL_SUCCESS;
}

wl_status wl_accel_depth(wl_accel a, uint32_t *depth)
{
    if (!a || !depth)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *depth = 1;
    return WL_SUCCESS;
}

wl_status wl_accel_input_offsets(wl_accel a, uint32_t *offsets)
{
    if (!a || !offsets)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    std::copy(a->input_offsets.begin(), a->input_offsets.end(), offsets);
    return WL_SUCCESS;
}

// ---- raw traversal --------------------------------------------------------
// A NULL stream is HIP's null stream, as in every HIP API; the context's own
// stream is available through wl_context_get_stream.
static hipStream_t pick_stream(wl_context ctx, void *stream)
{
    (void)ctx;
    return (hipStream_t)stream;
}

wl_status wl_trace_closest(wl_accel a, const wl_ray *rays, uint32_t n,
                           wl_hit *hits, void *stream)
{
    if (!a || (n && (!rays || !hits)))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(a->ctx->device));
    WL_HIP_CHECK(wl::launch_trace_closest(a, rays, n, hits, nullptr,
                                          pick_stream(a->ctx, stream)));
    return WL_SUCCESS;
}

wl_status wl_trace_closest_counted(wl_accel a, const wl_ray *rays, uint32_t n,
                                   wl_hit *hits, uint32_t *counts, void *stream)
{
    if (!a || (n && (!rays || !hits || !counts)))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(a->ctx->device));
    WL_HIP_CHECK(wl::launch_trace_closest(a, rays, n, hits, counts,
                                          pick_stream(a->ctx, stream)));
    return WL_SUCCESS;
}

wl_status wl_trace_any(wl_accel a, const wl_ray *rays, uint32_t n,
                       uint32_t *bits, void *stream)
{
    if (!a || (n && (!rays || !bits)))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(a->ctx->device));
    WL_HIP_CHECK(wl::launch_trace_any(a, rays, n, bits, pick_stream(a->ctx, stream)));
    return WL_SUCCESS;
}

} // extern "C"
