// bvh8.h — in-memory layout of the 8-wide compressed BVH shared by the host
// builder (bvh_builder.cpp) and the gfx950 traversal kernels (traverse.hip).
//
// It replaces the opaque acceleration structure that optixAccelBuild produces
// (src/Backend/Optix/Core/AccelStructure.cpp:133-162).
//
// Node: 128 bytes, one L2 cache line on gfx950, fetched by one lane as eight
// 16-byte loads:
//   bytes   0..11  p[3]        box origin (float)
//   bytes  12..14  e[3]        per-axis scale exponent, scale = 2^(e-127)
//   byte   15      imask       bit s set: slot s is an inner node
//   bytes  16..19  child_base  inner child of slot s lives at child_base + s
//   bytes  20..23  tri_base    first triangle record of this node's leaves
//   bytes  24..31  meta[8]     leaf slot: (offset << 2) | (count - 1);
//                              empty slot: 0xFF
//   bytes  32..127 q[6][8]     quantised child boxes: lo_x, lo_y, lo_z,
//                              hi_x, hi_y, hi_z; plane = p + scale * q, q an
//                              integer 0..2047 held as binary16 (exact), so
//                              gfx950's v_fma_mix_f32 reads it directly.
// Child boxes are rounded outward and padded by one quantum, so the decoded
// box always contains the exact child box.  Empty slots have lo = 65504 (the
// largest finite binary16) and hi = 0, which a near/far-plane slab test never
// hits.
//
// Inner children are addressed as child_base + slot: every node that has an
// inner child owns a group of eight node slots.  That costs node memory for
// unused slots (never fetched) and buys a 4-byte traversal stack entry,
// (child_base / 8) << 8 | remaining-hit-mask, with no node reload on pop.
//
// Slots are assigned per node so that, for a ray whose direction sign bits
// are o (bit k set: d_k < 0), visiting slots in increasing (slot ^ o ^ 7)
// order is approximately front to back (child centroids matched to octant
// directions, Ylitie, Karras, Laine, HPG 2017).
//
// Triangle record: 48 bytes, world-space vertices gathered once at build time
// plus the global primitive id (mesh triangle offset + local index), the
// build input it came from and its index within that input.
#pragma once
#include <stdint.h>

#define WL_BVH8_STACK_DEPTH 32
#define WL_BVH8_MAX_LEAF_TRIS 4

struct alignas(16) WlBvh8Node
{
    float p[3];
    uint8_t e[3];
    uint8_t imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint16_t q[6][8];
};

struct alignas(16) WlTriRecord
{
    float v0[3];
    float v1[3];
    float v2[3];
    uint32_t prim;
    uint32_t input;
    uint32_t local;
};

static_assert(sizeof(WlBvh8Node) == 128, "BVH8 node must be 128 bytes");
static_assert(sizeof(WlTriRecord) == 48, "triangle record must be 48 bytes");
