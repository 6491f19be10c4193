// traverse_persistent.h — persistent traversal with lane refill (gfx950).
//
// Incoherent rays finish after very different numbers of steps, so a wave
// that runs one ray per lane to completion idles most of its lanes near the
// end.  Here a fixed grid of waves loops: every lane advances its ray by one
// step (any-hit: a node and its triangles, trav_step; closest hit: one node
// or one triangle, trav_step_one); once at least kRefillLanes lanes of a
// wave are idle, the idle lanes start the next rays of the wave's own
// contiguous range of ray
// indices (after Aila & Laine, "Understanding the Efficiency of Ray Traversal
// on GPUs", HPG 2009).
//
// Each wave owns ceil(n / waves) consecutive rays and refills from a
// wave-uniform register, so no atomics are needed; with >1000 random rays per
// wave, the per-wave totals differ by a few percent at most.  A wave leaves
// the loop when its range is used up and none of its lanes holds a ray, so
// the grid always drains.
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

constexpr int kRefillLanes = 16;

__device__ __forceinline__ uint32_t lane_prefix(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// fetch(i, TravRay&) sets up ray i (i < n); emit(i, const TravResult&)
// consumes its result.  BLOCK must be a multiple of 64.  With PERLANE the
// rays of one launch may belong to different BVHs: fetch(i, TravRay&,
// nodes&, tris&) also sets the lane's node and triangle arrays.
template <bool ANY, int BLOCK, bool AUX, bool PERLANE = false, class Fetch, class Emit>
__device__ __forceinline__ void trav_persistent(const WlBvh8Node *__restrict__ nodes,
                                                const WlTriRecord *__restrict__ tris,
                                                uint32_t n, uint32_t *stack_lane,
                                                Fetch fetch, Emit emit)
{
    const WlBvh8Node *ln = nodes;
    const WlTriRecord *lt = tris;
    static_assert(BLOCK % 64 == 0, "whole waves only");
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64;
    const uint32_t chunk = (uint32_t)(((uint64_t)n + waves - 1) / waves);
    const uint64_t begin64 = (uint64_t)wave * chunk;
    const uint32_t end = (uint32_t)(begin64 + chunk < n ? begin64 + chunk : n);
    uint32_t next = (uint32_t)(begin64 < n ? begin64 : n); // wave-uniform
    TravRay r;
    trav_setup(r, 0, 0, 0, 1, 0, 0, 0, 0);
    TravState st;
    trav_init(st, r, false);
    uint32_t rid = 0;
    bool alive = false;
    uint32_t nn = 0, nt = 0;
    for (;;)
    {
        const uint64_t idle = __ballot(!alive);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (next < end && nidle >= (uint32_t)kRefillLanes)
        {
            if (!alive)
            {
                const uint32_t k = next + lane_prefix(idle);
                if (k < end)
                {
                    rid = k;
                    if constexpr (PERLANE)
                        fetch(rid, r, ln, lt);
                    else
                        fetch(rid, r);
                    trav_init(st, r, true);
                    alive = true;
                }
            }
            next = next + nidle < end ? next + nidle : end;
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
            if constexpr (ANY)
                fin = trav_step<ANY, false, BLOCK, AUX>(ln, lt, r, st, stack_lane, nn, nt);
            else
                fin = trav_step_one<BLOCK, AUX>(ln, lt, r, st, stack_lane);
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
