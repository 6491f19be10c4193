// wl_pt_noalloc.hip — wl_pt_render_noalloc: the path tracer's render with
// its batch buffers (path state, ray / hit / shadow queues, queue counters)
// in one fixed device array, `scratch`, instead of buffers the scene grows
// with hipMalloc.  For callers that cannot allocate: code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation.
//
// Each GPU holds its own copy of `scratch` at its own address, from the
// library's first use on that device until the process ends.  Its address on
// a device is looked up once, when a context is created on that device
// (wl_noalloc_setup, called from wl_context_create), and kept per device; host
// code passes only that address to kernels, never the array's name (which in
// host code is a stand-in, not a device address).  Every call on a device
// shares that device's array: calls that overlap on one GPU race.
#include "wl_internal.h"
#include "wl_pt.h"

#include <algorithm>
#include <mutex>

namespace
{
constexpr uint32_t kMaxPaths = WL_PT_NOALLOC_MAX_BATCH_PATHS;
// header: queue counters at byte 0 (16 words), path / ray statistics at 128
constexpr size_t kHeaderBytes = 256;
constexpr size_t r256(size_t b) { return (b + 255) & ~(size_t)255; }
// the carve of wl::pt_carve for kMaxPaths paths (checked against
// wl::pt_batch_bytes at set-up)
constexpr size_t kBatchBytes = r256(16ull * kMaxPaths) * 3 + r256(32ull * kMaxPaths) * 3 +
                               r256(4ull * kMaxPaths) * 8 + r256(16ull * kMaxPaths) * 2 +
                               r256(4ull * 2048);
constexpr size_t kScratchBytes = kHeaderBytes + kBatchBytes;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Zeroes the counter / statistics header (a kernel, not a memset: the call may
// be captured into a graph).
__global__ void k_noalloc_clear(uint32_t *header)
{
    const uint32_t i = threadIdx.x;
    if (i < kHeaderBytes / 4)
        header[i] = 0u;
}

namespace wl
{
// Looks up this device's copy of `scratch` (the device must be current).
wl_status noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device])
        return WL_SUCCESS;
    if (pt_batch_bytes(kMaxPaths) > kBatchBytes)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "no-allocation scratch smaller than one batch");
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p)
        return set_error(WL_ERROR_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
    g_ptr[device] = p;
    return WL_SUCCESS;
}
} // namespace wl

extern "C" wl_status wl_pt_render_noalloc(wl_scene s, const wl_pt_config *c, float *film,
                                          void *stream)
{
    wl_status st = wl::pt_check_config(s, c);
    if (st != WL_SUCCESS)
        return st;
    if (!film)
        return wl::set_error(WL_ERROR_INVALID_VALUE, "film is NULL");
    if (s->guide && s->guide_mode)
        return wl::set_error(WL_ERROR_NOT_SUPPORTED,
                             "a guided render needs per-path guide buffers: detach the guide");
    const uint32_t cap = c->max_batch_paths ? c->max_batch_paths : kMaxPaths;
    if (cap > kMaxPaths)
        return wl::set_error(WL_ERROR_LIMIT_EXCEEDED,
                             "max_batch_paths exceeds WL_PT_NOALLOC_MAX_BATCH_PATHS");
    const int dev = s->ctx->device;
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = dev >= 0 && dev < kMaxDevices ? g_ptr[dev] : nullptr;
    }
    if (!base)
        return wl::set_error(WL_ERROR_DEVICE, "no-allocation scratch not set up on this device");
    if (c->spp == 0)
        return WL_SUCCESS;
    WL_HIP_CHECK(hipSetDevice(dev));
    hipStream_t hs = (hipStream_t)stream;
    const uint32_t npix_total = c->width * c->height;
    // batch = a pixel range x a sample range, as wl_pt_render forms them
    const uint32_t tile = std::min(npix_total, cap);
    const uint32_t ns = std::max<uint32_t>(1, std::min<uint32_t>(c->spp, cap / tile));
    char *b = (char *)base;
    uint32_t *counters = (uint32_t *)b;
    unsigned long long *stats = (unsigned long long *)(b + 128);
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(kHeaderBytes / 4), 0, hs, counters);
    const wl::PtBuffers B = wl::pt_carve(b + kHeaderBytes, (size_t)tile * ns, counters, stats);
    const wl::PtScene S = wl::pt_scene_view(s);
    for (uint32_t s0 = 0; s0 < c->spp; s0 += ns)
    {
        const uint32_t nss = std::min(ns, c->spp - s0);
        for (uint32_t p0 = 0; p0 < npix_total; p0 += tile)
        {
            wl::PtParams P = wl::pt_params_for(s, c);
            if (c->flags & WL_PT_SORT_BY_MATERIAL)
                P.sort_bins = std::min<uint32_t>(s->num_meshes + 1, 1024);
            P.pix0 = p0;
            P.npix = std::min(tile, npix_total - p0);
            P.nsamples = nss;
            P.frame0 = c->frame_offset + s0;
            WL_HIP_CHECK(wl::pt_run_batch(P, S, B, s->ctx->num_cus, hs));
            WL_HIP_CHECK(wl::pt_accumulate(P, B, film, s->ctx->num_cus, hs));
        }
    }
    return WL_SUCCESS;
}
