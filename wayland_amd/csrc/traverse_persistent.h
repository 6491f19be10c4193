// traverse_persistent.h — persistent traversal with lane refill (gfx950).
//
// Incoherent rays finish after very different numbers of steps, so a wave
// that runs one ray per lane to completion idles most of its lanes near the
// end.  Here a fixed grid of waves loops: every lane advances its ray by one
// step (trav_step_one: one pending triangle, then the next node once none is
// pending); once at least REFILL lanes of a wave are idle, the idle lanes
// start the next rays of the wave's own ray indices (after Aila & Laine,
// "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009).
//
// Each wave owns ceil(n / waves) rays (one range, or interleaved chunks) and
// refills from a wave-uniform register, so no atomics are needed.  The
// remapped index lb is a permutation of the workgroup indices and the chunks
// of all waves tile [0, n), so every ray is owned by exactly one wave.  A
// wave leaves the loop when its rays are used up and none of its lanes holds
// a ray, so the grid always drains.
#pragma once
#include "../../include/wayland_hip.h"
#include "traverse_device.h"

namespace wl
{

// Ray and hit streams are read once and written once: non-temporal accesses
// keep them from displacing BVH nodes in the caches.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load_ray_nt(const wl_ray *__restrict__ rays, uint32_t i,
                                            TravRay &r)
{
    const f4v *rp = reinterpret_cast<const f4v *>(rays + i);
    const f4v a = __builtin_nontemporal_load(rp), b = __builtin_nontemporal_load(rp + 1);
    trav_setup(r, a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w);
}

__device__ __forceinline__ void store_hit_nt(wl_hit *__restrict__ hits, uint32_t i,
                                             const TravResult &h)
{
    const f4v v = { h.t, __uint_as_float(h.prim), h.b1, h.b2 };
    __builtin_nontemporal_store(v, reinterpret_cast<f4v *>(hits) + i);
}

// Default refill threshold: a wave refills once at least this many of its
// lanes are idle (REFILL below overrides it per kernel; measured best: 16 for
// 2^23 incoherent closest-hit and any-hit rays, 8 to 16 within noise for
// closest hit; 16 to 28 equal for the path tracer's queues, which use 20).
constexpr int kRefillLanes = 16;

__device__ __forceinline__ uint32_t lane_prefix(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// fetch(i, TravRay&) sets up ray i (i < n); emit(i, const TravResult&)
// consumes its result.  BLOCK must be a multiple of 64.  With ROOTED the
// rays of one launch may belong to different BVHs stored back to back in one
// node and triangle array (child and triangle indices rebased): fetch(i,
// TravRay&, uint32_t &root) also gives the index of the ray's root node.
// XCDMAP: the hardware deals workgroups round-robin to the 8 XCDs, each with
// its own L2; the wave's ray range is taken from a remapped workgroup index so
// that XCD x works on one contiguous eighth of the rays and their hits.  On
// 2^23 incoherent rays this is 1-3 % faster for closest hit and 4 % for
// any-hit; the path tracer's queues gain nothing from it and keep the plain
// order.
// CHUNKS > 0: instead of one contiguous range, wave w takes chunks w,
// w + waves, w + 2 waves, ... of about 1/CHUNKS of its share each (at least
// 64 rays), so waves whose share is expensive (a stretch of a multi-scene
// queue that hits the largest BVH) get cheaper chunks too.  It suits queues
// whose cost varies along the queue (measured best: 4 for 2^23 incoherent
// rays, 8 for multi-request queues, 2 for the path tracer's shadow rays); the
// path tracer's extension rays keep contiguous ranges, which keep its
// coherent queues together.
// top / ntop (optional): nodes [0, ntop) copied to LDS by the caller (see
// node_visit).
// DYN = 2: chunks of DCHUNK rays are handed out at run time instead, so a
// wave slowed by expensive rays takes fewer chunks.  One global counter per
// XCD (`dyn`, 8 words): XCD x owns the x-th eighth of the rays, its waves
// start on chunks 0 .. W-1 of that eighth (W waves per XCD; the grid is a
// multiple of 8) and then take chunk W + (counter - base) each, base =
// `dyn_base`, the counter's value at launch.  Every wave that had a chunk
// makes exactly one fetch past the eighth's last chunk, so a launch advances
// counter x by the eighth's chunk count: the host keeps the bases and nothing
// is reset.
template <bool ANY, int BLOCK, bool AUX, bool ROOTED = false, bool XCDMAP = false,
          uint32_t CHUNKS = 0, int REFILL = kRefillLanes, int DYN = 0,
          uint32_t DCHUNK = 256, class Fetch, class Emit>
__device__ __forceinline__ void trav_persistent(const WlBvh8Node *__restrict__ nodes,
                                                const WlTriRecord *__restrict__ tris,
                                                uint32_t n, uint32_t *stack_lane,
                                                Fetch fetch, Emit emit,
                                                const WlBvh8Node *top = nullptr,
                                                uint32_t ntop = 0, uint32_t *dyn = nullptr,
                                                uint32_t dyn_base = 0)
{
    static_assert(BLOCK % 64 == 0, "whole waves only");
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    const uint32_t G = gridDim.x;
    const uint32_t lb = (XCDMAP && G % 8u == 0u) ? (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u
                                                 : blockIdx.x;
    const uint32_t wave = lb * (BLOCK / 64) + threadIdx.x / 64;
    // the wave's current range [next, end) of chunk cidx (wave-uniform)
    const uint32_t share = (uint32_t)(((uint64_t)n + waves - 1) / waves);
    const uint32_t C = CHUNKS == 0 ? share
                       : share / CHUNKS > 64u ? (share + CHUNKS - 1) / CHUNKS
                                              : 64u;
    uint64_t cidx = wave;
    uint32_t next, end;
    // DYN: this wave's XCD, its eighth [xlo, xhi) of the rays and its index
    // among the XCD's waves
    const uint32_t xcd = blockIdx.x % 8u;
    const uint32_t n8 = (n + 7u) / 8u;
    const uint32_t xlo = xcd * n8 < n ? xcd * n8 : n;
    const uint32_t xhi = xlo + n8 < n ? xlo + n8 : n;
    const uint32_t wx = (blockIdx.x / 8u) * (BLOCK / 64) + threadIdx.x / 64;
    const uint32_t wavesx = (G / 8u) * (BLOCK / 64);
    // the next chunk of the DYN schedule into [next, end)
    auto take_chunk = [&](bool first) {
        static_assert(DYN == 0 || DYN == 2, "DYN is 0 or 2");
        uint32_t c = wx;
        if (!first)
        {
            uint32_t k = 0;
            if (__lane_id() == 0)
                k = atomicAdd(dyn + xcd, 1u) - dyn_base;
            c = wavesx + __shfl(k, 0);
        }
        const uint64_t s0 = xlo + (uint64_t)c * DCHUNK;
        next = (uint32_t)(s0 < xhi ? s0 : xhi);
        end = (uint32_t)(s0 + DCHUNK < xhi ? s0 + DCHUNK : xhi);
    };
    if constexpr (DYN != 0)
        take_chunk(true);
    else
    {
        const uint64_t s0 = cidx * C;
        next = (uint32_t)(s0 < n ? s0 : n);
        end = (uint32_t)(s0 + C < n ? s0 + C : n);
    }
    TravRay r;
    trav_setup(r, 0, 0, 0, 1, 0, 0, 0, 0);
    TravState st;
    trav_init(st, r, false);
    uint32_t rid = 0;
    bool alive = false;
    for (;;)
    {
        const uint64_t idle = __ballot(!alive);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (next < end && nidle >= (uint32_t)REFILL)
        {
            if (!alive)
            {
                const uint32_t k = next + lane_prefix(idle);
                if (k < end)
                {
                    rid = k;
                    uint32_t root = 0;
                    if constexpr (ROOTED)
                        fetch(rid, r, root);
                    else
                        fetch(rid, r);
                    trav_init(st, r, true);
                    st.node = root;
                    alive = true;
                }
            }
            next = next + nidle < end ? next + nidle : end;
            if (DYN != 0 && next >= end)
                take_chunk(false);
            else if (DYN == 0 && CHUNKS != 0 && next >= end)
            {
                cidx += waves;
                const uint64_t s0 = cidx * C;
                next = (uint32_t)(s0 < n ? s0 : n);
                end = (uint32_t)(s0 + C < n ? s0 + C : n);
            }
        }
        if (__ballot(alive) == 0)
        {
            if (next >= end)
                break;
            continue;
        }
        bool fin = false;
        if (alive)
        {
            fin = trav_step_one<BLOCK, AUX, ANY>(nodes, tris, r, st, stack_lane, top, ntop);
        }
        if (fin)
        {
            trav_finish(st.res);
            emit(rid, st.res);
            alive = false;
        }
    }
}


} // namespace wl
