// traverse_persistent.h — persistent traversal with lane refill (gfx950).
//
// Incoherent rays finish after very different numbers of steps, so a wave
// that runs one ray per lane to completion idles most of its lanes near the
// end.  Here a fixed grid of waves loops: every lane advances its ray by one
// node (trav_step); once at least kRefillLanes lanes of a wave are idle, the
// wave takes that many new ray indices from a global counter with one atomic
// (ballot + mbcnt prefix) and the idle lanes start them (after Aila & Laine,
// "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009).
// Every wave leaves the loop once the counter has passed n and none of its
// lanes holds a ray, so the grid always drains.
#pragma once
#include "traverse_device.h"

namespace wl
{

constexpr int kRefillLanes = 16;

__device__ __forceinline__ uint32_t lane_prefix(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// fetch(i, TravRay&) sets up ray i (i < n); emit(i, const TravResult&)
// consumes its result.  `counter` must be zero when the grid starts.
template <bool ANY, int BLOCK, bool AUX, class Fetch, class Emit>
__device__ __forceinline__ void trav_persistent(const WlBvh8Node *__restrict__ nodes,
                                                const WlTriRecord *__restrict__ tris,
                                                uint32_t n, uint32_t *counter,
                                                uint32_t *stack_lane, Fetch fetch,
                                                Emit emit)
{
    TravRay r;
    trav_setup(r, 0, 0, 0, 1, 0, 0, 0, 0);
    TravState st;
    trav_init(st, r, false);
    uint32_t rid = 0;
    bool alive = false;
    bool exhausted = false; // wave-uniform
    const uint32_t lane = __lane_id();
    uint32_t nn = 0, nt = 0;
    for (;;)
    {
        const uint64_t idle = __ballot(!alive);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && nidle >= (uint32_t)kRefillLanes)
        {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0;
            if ((int)lane == leader)
                base = atomicAdd(counter, nidle);
            base = __shfl(base, leader);
            if ((uint64_t)base + nidle >= n)
                exhausted = true;
            if (!alive)
            {
                rid = base + lane_prefix(idle);
                if (rid < n)
                {
                    fetch(rid, r);
                    trav_init(st, r, true);
                    alive = true;
                }
            }
        }
        if (__ballot(alive) == 0)
        {
            if (exhausted)
                break;
            continue;
        }
        if (alive && trav_step<ANY, false, BLOCK, AUX>(nodes, tris, r, st, stack_lane, nn, nt))
        {
            trav_finish(st.res);
            emit(rid, st.res);
            alive = false;
        }
    }
}

} // namespace wl
