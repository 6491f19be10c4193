// traverse.hip — raw closest-hit / any-hit traversal kernels for gfx950.
// These are the kernels behind wl_trace_closest and wl_trace_any, which stand
// in for optixTrace (PathTracing.cu:72-75, :92-102).
//
// One ray per lane, 256 lanes per workgroup, an LDS stack per lane, and a
// persistent grid whose lanes refill from their wave's range of rays
// (traverse_persistent.h).  The LDS stack is sized per launch: a BVH at most
// 16 levels deep needs at most 15 entries, so its kernels reserve 16
// (16 KiB per workgroup) instead of 32.  The instrumented closest-hit kernel
// (node / triangle counts) keeps one ray per lane without refill.
#include "traverse_persistent.h"
#include "wl_internal.h"

namespace wl
{

constexpr int kTraceBlock = 256;
// The first nodes of the BVH (laid out breadth first: the root, its group of
// 8 and their groups, levels 0-2 of a full tree) are copied to LDS per
// workgroup; every ray starts there, so about half of a ray's node visits
// read LDS instead of the L1/L2 path.  9,344 bytes beside the 16-entry stack
// keep 6 workgroups (24 waves) per CU; the 32-entry-stack kernels for deep
// BVHs have no room for it.
constexpr uint32_t kTopNodes = 73;

// Chunk scheduling (trav_persistent's DYN = 2): chunks of this many rays,
// taken at run time from one counter per XCD.  The static split it replaced
// (interleaved chunks per wave) left the kernel waiting for its slowest
// waves: 8.68-8.75 -> 9.32-9.41 Grays/s on 2^23 incoherent closest-hit rays.
constexpr uint32_t kTraceChunk = 256;

// The chunk counters' values at launch, one per XCD.
struct TraceSched
{
    uint32_t base[8];
};

template <int STACK>
constexpr uint32_t top_nodes() { return STACK == 16 ? kTopNodes : 0u; }

// Copies nodes [0, ntop) to LDS (all threads of the workgroup, then a barrier).
__device__ __forceinline__ void load_top(uint4 *top, const WlBvh8Node *__restrict__ nodes,
                                         uint32_t ntop)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(nodes);
    for (uint32_t i = threadIdx.x; i < 8u * ntop; i += kTraceBlock)
        top[i] = src[i];
    __syncthreads();
}

__device__ __forceinline__ void load_ray(const wl_ray *__restrict__ rays, uint32_t i, TravRay &r)
{
    load_ray_nt(rays, i, r);
}

// ntop: nodes copied to LDS, min(node count, top_nodes<STACK>()); sched /
// base: the launching stream's chunk counters and their values at launch.
// The grid is a multiple of 8.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6, 8)))
void k_trace_closest_p(
    const WlBvh8Node *__restrict__ nodes, const WlTriRecord *__restrict__ tris,
    const wl_ray *__restrict__ rays, uint32_t n, wl_hit *__restrict__ hits, uint32_t ntop,
    uint32_t *__restrict__ sched, TraceSched base)
{
    __shared__ uint32_t stack[STACK * kTraceBlock];
    __shared__ uint4 top[8 * (top_nodes<STACK>() ? top_nodes<STACK>() : 1)];
    if (top_nodes<STACK>())
        load_top(top, nodes, ntop);
    trav_persistent<false, kTraceBlock, false, false, true, 0, kRefillLanes, 2, kTraceChunk>(
        nodes, tris, n, stack + threadIdx.x,
        [&](uint32_t i, TravRay &r) { load_ray(rays, i, r); },
        [&](uint32_t i, const TravResult &h) { store_hit_nt(hits, i, h); },
        reinterpret_cast<const WlBvh8Node *>(top), top_nodes<STACK>() ? ntop : 0u, sched,
        base.base[blockIdx.x % 8u]);
}

template <int STACK>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7, 8)))
void k_trace_any_p(const WlBvh8Node *__restrict__ nodes, const WlTriRecord *__restrict__ tris,
                   const wl_ray *__restrict__ rays, uint32_t n, uint32_t *__restrict__ bits,
                   uint32_t ntop, uint32_t *__restrict__ sched, TraceSched base)
{
    __shared__ uint32_t stack[STACK * kTraceBlock];
    __shared__ uint4 top[8 * (top_nodes<STACK>() ? top_nodes<STACK>() : 1)];
    if (top_nodes<STACK>())
        load_top(top, nodes, ntop);
    trav_persistent<true, kTraceBlock, false, false, true, 0, kRefillLanes, 2, kTraceChunk>(
        nodes, tris, n, stack + threadIdx.x,
        [&](uint32_t i, TravRay &r) { load_ray(rays, i, r); },
        [&](uint32_t i, const TravResult &h) {
            if (h.prim != WL_INVALID)
                atomicOr(&bits[i >> 5], 1u << (i & 31u));
        },
        reinterpret_cast<const WlBvh8Node *>(top), top_nodes<STACK>() ? ntop : 0u, sched,
        base.base[blockIdx.x % 8u]);
}

// Instrumented closest hit: one ray per lane, per-ray node/triangle counts.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest_counted(
    const WlBvh8Node *__restrict__ nodes, const WlTriRecord *__restrict__ tris,
    const wl_ray *__restrict__ rays, uint32_t n, wl_hit *__restrict__ hits,
    uint32_t *__restrict__ counts)
{
    __shared__ uint32_t stack[STACK * kTraceBlock];
    const uint32_t rid = blockIdx.x * kTraceBlock + threadIdx.x;
    const bool active = rid < n;
    TravRay r;
    if (active)
        load_ray(rays, rid, r);
    else
        trav_setup(r, 0, 0, 0, 1, 0, 0, 0, 0);
    uint32_t nn = 0, nt = 0;
    TravResult h = trav_run<false, true, kTraceBlock>(nodes, tris, r, active,
                                                     stack + threadIdx.x, nn, nt);
    if (active)
    {
        reinterpret_cast<float4 *>(hits)[rid] =
            make_float4(h.t, __uint_as_float(h.prim), h.b1, h.b2);
        counts[2 * (size_t)rid] = nn;
        counts[2 * (size_t)rid + 1] = nt;
    }
}

// Zeroes the any-hit occlusion words (the kernel ORs bits in).
__global__ void k_zero_words(uint32_t *words, uint32_t nwords)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = i; k < nwords; k += gridDim.x * blockDim.x)
        words[k] = 0;
}

static inline bool shallow(const wl_accel_s *a) { return a->info.depth <= 16; }

// Persistent grid: the workgroups that fit the GPU at once (the launch then
// takes at most one workgroup per 256 rays).
template <class K>
static uint32_t persistent_grid(K kernel, uint32_t /*unused*/, int num_cus)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTraceBlock, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    return (uint32_t)(per_cu * std::max(1, num_cus));
}

// The counters' values for a launch over n rays, advanced past it.
static TraceSched take_bases(wl_context_s::SchedBlock *sb, uint32_t n)
{
    TraceSched t;
    for (uint32_t x = 0; x < 8; ++x)
    {
        t.base[x] = sb->base[x];
        sb->base[x] += dyn_xcd_chunks(n, x, kTraceChunk);
    }
    return t;
}

// The persistent grid for n rays, rounded up to a multiple of 8 so that each
// XCD's eighth of the rays has workgroups (chunk scheduling DYN = 2).
static uint32_t xcd_grid(uint32_t grid_cache, uint32_t n)
{
    const uint32_t g = (uint32_t)std::min<uint64_t>(
        grid_cache, ((uint64_t)n + kTraceBlock - 1) / kTraceBlock);
    return (g + 7u) & ~7u;
}

template <int STACK>
static hipError_t closest_impl(const wl_accel_s *a, const wl_ray *rays, uint32_t n,
                               wl_hit *hits, hipStream_t s)
{
    static uint32_t grid_cache = 0; // depends only on the kernel and the device
    static int cus_cache = -1;
    if (cus_cache != a->ctx->num_cus)
    {
        grid_cache = persistent_grid(k_trace_closest_p<STACK>, 0xFFFFFFFFu, a->ctx->num_cus);
        cus_cache = a->ctx->num_cus;
    }
    const uint32_t grid = xcd_grid(grid_cache, n);
    std::lock_guard<std::mutex> lock(a->ctx->sched_mu);
    wl_context_s::SchedBlock *sb = nullptr;
    const hipError_t e = sched_block(a->ctx, s, &sb);
    if (e != hipSuccess)
        return e;
    const TraceSched base = take_bases(sb, n);
    hipLaunchKernelGGL(k_trace_closest_p<STACK>, dim3(grid), dim3(kTraceBlock), 0, s,
                       a->d_nodes, a->d_tris, rays, n, hits,
                       std::min<uint32_t>(a->num_nodes, top_nodes<STACK>()), sb->dev, base);
    return hipGetLastError();
}

template <int STACK>
static hipError_t any_impl(const wl_accel_s *a, const wl_ray *rays, uint32_t n,
                           uint32_t *bits, hipStream_t s)
{
    static uint32_t grid_cache = 0;
    static int cus_cache = -1;
    if (cus_cache != a->ctx->num_cus)
    {
        grid_cache = persistent_grid(k_trace_any_p<STACK>, 0xFFFFFFFFu, a->ctx->num_cus);
        cus_cache = a->ctx->num_cus;
    }
    const uint32_t grid = xcd_grid(grid_cache, n);
    std::lock_guard<std::mutex> lock(a->ctx->sched_mu);
    wl_context_s::SchedBlock *sb = nullptr;
    const hipError_t e = sched_block(a->ctx, s, &sb);
    if (e != hipSuccess)
        return e;
    const TraceSched base = take_bases(sb, n);
    const uint32_t words = (n + 31) / 32;
    hipLaunchKernelGGL(k_zero_words, dim3(std::min<uint32_t>((words + 255) / 256, 4096)),
                       dim3(256), 0, s, bits, words);
    hipLaunchKernelGGL(k_trace_any_p<STACK>, dim3(grid), dim3(kTraceBlock), 0, s, a->d_nodes,
                       a->d_tris, rays, n, bits,
                       std::min<uint32_t>(a->num_nodes, top_nodes<STACK>()), sb->dev, base);
    return hipGetLastError();
}

hipError_t launch_trace_closest(const wl_accel_s *a, const wl_ray *rays,
                                uint32_t n, wl_hit *hits, uint32_t *counts,
                                hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const bool sh = shallow(a);
    if (counts)
    {
        dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
        if (sh)
            hipLaunchKernelGGL(k_trace_closest_counted<16>, grid, block, 0, s, a->d_nodes,
                               a->d_tris, rays, n, hits, counts);
        else
            hipLaunchKernelGGL(k_trace_closest_counted<WL_BVH8_STACK_DEPTH>, grid, block, 0, s,
                               a->d_nodes, a->d_tris, rays, n, hits, counts);
        return hipGetLastError();
    }
    return sh ? closest_impl<16>(a, rays, n, hits, s)
              : closest_impl<WL_BVH8_STACK_DEPTH>(a, rays, n, hits, s);
}

hipError_t launch_trace_any(const wl_accel_s *a, const wl_ray *rays,
                            uint32_t n, uint32_t *bits, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    return shallow(a) ? any_impl<16>(a, rays, n, bits, s)
                      : any_impl<WL_BVH8_STACK_DEPTH>(a, rays, n, bits, s);
}

} // namespace wl
