// bvh_builder.h — host SAH builder for the compressed BVH8 (see bvh8.h).
#pragma once
#include "bvh8.h"

#include <stdint.h>
#include <string>
#include <vector>

namespace wl
{

struct Bvh8Host
{
    std::vector<WlBvh8Node> nodes;   // node 0 is the root
    std::vector<WlTriRecord> tris;   // leaf triangle records
    uint32_t depth = 0;              // levels, root = 1
    uint32_t inner_nodes = 0;        // nodes actually used
    uint32_t leaves = 0;
    double sah_cost = 0.0;           // SAH cost (node 1, triangle 1 per test)
    double build_seconds = 0.0;
};

// Build over `ntris` world-space triangles (9 floats each, global primitive
// order).  Binned SAH (32 bins) BVH2 with at most WL_BVH8_MAX_LEAF_TRIS
// triangles per leaf, collapsed greedily (largest surface area first) to
// 8-wide nodes, children slot-ordered by octant and boxes quantised to 16
// bits.  Returns false with *err set on failure (e.g. depth limit).
bool build_bvh8(const float *tris, uint32_t ntris, int nthreads, Bvh8Host *out,
                std::string *err);

} // namespace wl
