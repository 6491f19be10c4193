// traverse.hip — raw closest-hit / any-hit traversal kernels for gfx950.
// One ray per lane, 256 lanes per workgroup, LDS stack per lane (see
// traverse_device.h).  These are the kernels behind wl_trace_closest and
// wl_trace_any, which stand in for optixTrace (PathTracing.cu:72-75, :92-102).
//
// The LDS stack is sized per launch: a BVH at most 16 levels deep needs at
// most 15 entries, so its kernels reserve 16 (16 KiB per workgroup) instead
// of 32, which doubles the workgroups that fit a CU's LDS.
#include "traverse_device.h"
#include "wl_internal.h"

namespace wl
{

constexpr int kTraceBlock = 256;

template <bool COUNT, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(
    const WlBvh8Node *__restrict__ nodes, const WlTriRecord *__restrict__ tris,
    const wl_ray *__restrict__ rays, uint32_t n, wl_hit *__restrict__ hits,
    uint32_t *__restrict__ counts)
{
    __shared__ uint32_t stack[STACK * kTraceBlock];
    const uint32_t rid = blockIdx.x * kTraceBlock + threadIdx.x;
    const bool active = rid < n;
    TravRay r;
    if (active)
    {
        const float4 *rp = reinterpret_cast<const float4 *>(rays + rid);
        const float4 a = rp[0], b = rp[1];
        trav_setup(r, a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w);
    }
    else
    {
        trav_setup(r, 0, 0, 0, 1, 0, 0, 0, 0);
    }
    uint32_t nn = 0, nt = 0;
    TravResult h = trav_run<false, COUNT, kTraceBlock>(nodes, tris, r, active,
                                                       stack + threadIdx.x, nn, nt);
    if (active)
    {
        float4 o;
        o.x = h.t;
        o.y = __uint_as_float(h.prim);
        o.z = h.b1;
        o.w = h.b2;
        reinterpret_cast<float4 *>(hits)[rid] = o;
        if (COUNT)
        {
            counts[2 * (size_t)rid] = nn;
            counts[2 * (size_t)rid + 1] = nt;
        }
    }
}

template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace_any(
    const WlBvh8Node *__restrict__ nodes, const WlTriRecord *__restrict__ tris,
    const wl_ray *__restrict__ rays, uint32_t n, uint32_t *__restrict__ bits)
{
    __shared__ uint32_t stack[STACK * kTraceBlock];
    const uint32_t rid = blockIdx.x * kTraceBlock + threadIdx.x;
    const bool active = rid < n;
    TravRay r;
    if (active)
    {
        const float4 *rp = reinterpret_cast<const float4 *>(rays + rid);
        const float4 a = rp[0], b = rp[1];
        trav_setup(r, a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w);
    }
    else
    {
        trav_setup(r, 0, 0, 0, 1, 0, 0, 0, 0);
    }
    uint32_t nn = 0, nt = 0;
    TravResult h = trav_run<true, false, kTraceBlock>(nodes, tris, r, active,
                                                      stack + threadIdx.x, nn, nt);
    const bool occ = active && h.prim != WL_INVALID;
    // wave64 ballot: one 64-bit word of occlusion bits per wave
    const uint64_t m = __ballot(occ);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave_first = rid - lane;
    if (lane == 0)
    {
        const uint32_t w0 = wave_first >> 5;
        if ((size_t)w0 * 32 < n)
            bits[w0] = (uint32_t)m;
        if ((size_t)(w0 + 1) * 32 < n)
            bits[w0 + 1] = (uint32_t)(m >> 32);
    }
}

// Stack entries a traversal of this BVH can hold at once: depth - 1.
static inline bool shallow(const wl_accel_s *a) { return a->info.depth <= 16; }

hipError_t launch_trace_closest(const wl_accel_s *a, const wl_ray *rays,
                                uint32_t n, wl_hit *hits, uint32_t *counts,
                                hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    const bool sh = shallow(a);
    if (counts)
    {
        if (sh)
            hipLaunchKernelGGL((k_trace_closest<true, 16>), grid, block, 0, s, a->d_nodes,
                               a->d_tris, rays, n, hits, counts);
        else
            hipLaunchKernelGGL((k_trace_closest<true, WL_BVH8_STACK_DEPTH>), grid, block, 0, s,
                               a->d_nodes, a->d_tris, rays, n, hits, counts);
    }
    else
    {
        if (sh)
            hipLaunchKernelGGL((k_trace_closest<false, 16>), grid, block, 0, s, a->d_nodes,
                               a->d_tris, rays, n, hits, counts);
        else
            hipLaunchKernelGGL((k_trace_closest<false, WL_BVH8_STACK_DEPTH>), grid, block, 0, s,
                               a->d_nodes, a->d_tris, rays, n, hits, counts);
    }
    return hipGetLastError();
}

hipError_t launch_trace_any(const wl_accel_s *a, const wl_ray *rays,
                            uint32_t n, uint32_t *bits, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    if (shallow(a))
        hipLaunchKernelGGL((k_trace_any<16>), grid, block, 0, s, a->d_nodes, a->d_tris, rays,
                           n, bits);
    else
        hipLaunchKernelGGL((k_trace_any<WL_BVH8_STACK_DEPTH>), grid, block, 0, s, a->d_nodes,
                           a->d_tris, rays, n, bits);
    return hipGetLastError();
}

} // namespace wl
