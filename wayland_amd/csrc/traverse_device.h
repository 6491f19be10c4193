// traverse_device.h — gfx950 BVH8 traversal, one ray per lane.
//
// Watertight FP32 ray/triangle test (Woop, Benthin, Wald, JCGT 2013) with the
// same operation order as the host oracle: this translation unit is compiled
// with -ffp-contract=off, so every + - * / below rounds separately and the
// division and square root are IEEE correctly rounded.  Box tests use
// explicit fmaf (only the triangle test must match the host bit for bit: the
// box test decides what is visited, never what is hit, because boxes are
// conservative).
//
// Traversal state per lane: a current node group (child_base, hit mask in
// key space), a triangle mask of the node just visited, and an LDS stack of
// 4-byte group entries, WL_BVH8_STACK_DEPTH deep, laid out [depth][lane] so
// the 64 lanes of a wave touch 64 consecutive words.
//
// Two step shapes share the node and triangle code.  trav_step visits a node
// and then tests all of its triangles (the instrumented kernel and the
// one-ray-to-completion loop).  The persistent kernels, closest and any-hit,
// step with trav_step_one instead: a lane tests one pending triangle per
// step, and visits its next node once none is pending (in the same step as
// its last triangle), so lanes with several triangles do not hold the whole
// wave in a triangle loop.  Both visit the same nodes and test the same
// triangles in the same order, so the results are the same.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh8.h"

#define WL_INVALID 0xFFFFFFFFu

// Compiled for the device by hipcc and, for the CPU emulation used by the
// tests (hostemu.cpp), for the host.  WL_TRAV_CHECK is empty on the device;
// the host emulation defines it to record out-of-range accesses.
#define WL_HD __host__ __device__ __forceinline__
#ifndef WL_TRAV_CHECK
#define WL_TRAV_CHECK(cond, what) ((void)0)
#endif

namespace wl
{

WL_HD float wl_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
WL_HD uint32_t wl_f2u(float f) { return __builtin_bit_cast(uint32_t, f); }

struct TravRay
{
    float o[3];
    float tmin, tmax;
    float inv[3];
    float Sx, Sy, Sz;
    // bits 0-2: bit k set for d_k < 0 (slot ^ neg ^ 7 = visit order key, near
    // first); bits 3-8: kx | ky << 2 | kz << 4 (the triangle test's axis order)
    uint32_t code;
};

WL_HD float sel3(int k, float a, float b, float c)
{
    return k == 0 ? a : (k == 1 ? b : c);
}

WL_HD void trav_setup(TravRay &r, float ox, float oy,
                                           float oz, float dx, float dy,
                                           float dz, float tmin, float tmax)
{
    r.o[0] = ox;
    r.o[1] = oy;
    r.o[2] = oz;
    r.tmin = tmin;
    r.tmax = tmax;
    float d[3] = { dx, dy, dz };
    int kz = 0;
    if (fabsf(d[1]) > fabsf(d[kz]))
        kz = 1;
    if (fabsf(d[2]) > fabsf(d[kz]))
        kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dkz = sel3(kz, dx, dy, dz);
    if (dkz < 0.0f)
    {
        int t = kx;
        kx = ky;
        ky = t;
    }
    const uint32_t kxyz = (uint32_t)kx | (uint32_t)ky << 2 | (uint32_t)kz << 4;
    r.Sx = sel3(kx, dx, dy, dz) / dkz;
    r.Sy = sel3(ky, dx, dy, dz) / dkz;
    r.Sz = 1.0f / dkz;
    uint32_t neg = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        float di = d[i];
        if (fabsf(di) < 1e-30f)
            di = copysignf(1e-30f, di);
        // the reciprocals only place box planes, which are padded; the
        // triangle test uses the correctly rounded Sx, Sy, Sz
#if defined(__HIP_DEVICE_COMPILE__)
        r.inv[i] = __builtin_amdgcn_rcpf(di);
#else
        r.inv[i] = 1.0f / di;
#endif
        if (r.inv[i] < 0.0f)
            neg |= 1u << i;
    }
    r.code = neg | kxyz << 3;
}

// Watertight test; returns true and (t, b1, b2) for tmin <= t <= tmax.
WL_HD bool tri_test(const TravRay &r, const float4 w0,
                                         const float4 w1, const float4 w2,
                                         float tmax, float &t_out, float &b1_out,
                                         float &b2_out)
{
    // A = v0 - o, B = v1 - o, C = v2 - o, components permuted (kx, ky, kz)
    const int kx = (int)((r.code >> 3) & 3u), ky = (int)((r.code >> 5) & 3u),
              kz = (int)((r.code >> 7) & 3u);
    // subtract in world axes, then permute (the same differences, fewer selects)
    const float A0 = w0.x - r.o[0], A1 = w0.y - r.o[1], A2 = w0.z - r.o[2];
    const float B0 = w0.w - r.o[0], B1 = w1.x - r.o[1], B2 = w1.y - r.o[2];
    const float C0 = w1.z - r.o[0], C1 = w1.w - r.o[1], C2 = w2.x - r.o[2];
    const float Akx = sel3(kx, A0, A1, A2), Aky = sel3(ky, A0, A1, A2), Akz = sel3(kz, A0, A1, A2);
    const float Bkx = sel3(kx, B0, B1, B2), Bky = sel3(ky, B0, B1, B2), Bkz = sel3(kz, B0, B1, B2);
    const float Ckx = sel3(kx, C0, C1, C2), Cky = sel3(ky, C0, C1, C2), Ckz = sel3(kz, C0, C1, C2);
    const float Ax = Akx - r.Sx * Akz;
    const float Ay = Aky - r.Sy * Akz;
    const float Bx = Bkx - r.Sx * Bkz;
    const float By = Bky - r.Sy * Bkz;
    const float Cx = Ckx - r.Sx * Ckz;
    const float Cy = Cky - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f)
    {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    const float det = U + V + W;
    if (det == 0.0f)
        return false;
    const float Az = r.Sz * Akz;
    const float Bz = r.Sz * Bkz;
    const float Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= r.tmin && t <= tmax))
        return false;
    t_out = t;
    b1_out = V * rcp;
    b2_out = W * rcp;
    return true;
}

// Quantised planes are binary16 halves of a word; the fma that consumes the
// widened value compiles to one v_fma_mix_f32 (conversion folded in).
WL_HD float q_lo16(uint32_t w)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu));
}
WL_HD float q_hi16(uint32_t w)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}

struct TravResult
{
    float t;
    uint32_t prim;
    float b1, b2;
    uint32_t input, local; // build input and input-local index (AUX only)
};

// Resumable traversal state of one ray: the current node group (child_base,
// hit mask in key space), the node to visit next, the stack depth and the
// best hit so far.  The stack itself lives in LDS.
struct TravState
{
    TravResult res;
    uint32_t grp_base, grp_mask, node;
    int sp;
    bool visit;
    uint32_t trimask, tri_base; // pending triangles (trav_step_one)
};

WL_HD void trav_init(TravState &st, const TravRay &r, bool active)
{
    st.res.t = r.tmax;
    st.res.prim = WL_INVALID;
    st.res.b1 = 0.0f;
    st.res.b2 = 0.0f;
    st.res.input = WL_INVALID;
    st.res.local = WL_INVALID;
    st.sp = 0;
    st.grp_base = 0;
    st.grp_mask = 0;
    st.node = 0;
    st.visit = active;
    st.trimask = 0;
    st.tri_base = 0;
}

// The eight child boxes of `node` against [tmin, min(t_best, ...)]: inner
// hits in key space (bit s ^ neg ^ 7), and the triangle bits of the hit leaves.
struct NodeHits
{
    uint32_t child_base, tri_base, inner, trimask;
};

// top (optional): a copy of nodes [0, ntop) in LDS, read instead of memory
// for those nodes (the builder lays the top levels out first).
WL_HD NodeHits node_visit(const WlBvh8Node *__restrict__ nodes, uint32_t node,
                          const TravRay &r, float t_best, const WlBvh8Node *top = nullptr,
                          uint32_t ntop = 0)
{
    // near/far quantised planes chosen by the ray's direction signs
    // (lo planes at words 2-4, hi planes at 5-7: a per-lane offset picks them)
    const uint32_t nx = r.code & 1u, ny = (r.code >> 1) & 1u, nz = (r.code >> 2) & 1u;
    uint4 h0, h1, nqx, fqx, nqy, fqy, nqz, fqz;
    auto fetch = [&](auto np) {
        h0 = np[0];
        h1 = np[1];
        nqx = np[2 + 3 * nx];
        fqx = np[5 - 3 * nx];
        nqy = np[3 + 3 * ny];
        fqy = np[6 - 3 * ny];
        nqz = np[4 + 3 * nz];
        fqz = np[7 - 3 * nz];
    };
#if defined(__HIP_DEVICE_COMPILE__)
    // distinct address spaces keep the two paths apart: LDS reads for the
    // cached nodes, global loads for the rest (not one generic-pointer load)
    typedef __attribute__((address_space(3))) const uint4 lds_u4;
    typedef __attribute__((address_space(1))) const uint4 glb_u4;
    if (node < ntop)
        fetch((lds_u4 *)(reinterpret_cast<const uint4 *>(top + node)));
    else
        fetch((glb_u4 *)(reinterpret_cast<const uint4 *>(nodes + node)));
#else
    fetch(reinterpret_cast<const uint4 *>(node < ntop ? top + node : nodes + node));
#endif
    const float px = wl_u2f(h0.x), py = wl_u2f(h0.y), pz = wl_u2f(h0.z);
    const uint32_t ew = h0.w;
    const float sx = wl_u2f((ew & 0xFFu) << 23);
    const float sy = wl_u2f(((ew >> 8) & 0xFFu) << 23);
    const float sz = wl_u2f(((ew >> 16) & 0xFFu) << 23);
    const uint32_t imask = ew >> 24;
    const uint32_t meta_lo = h1.z, meta_hi = h1.w;
    // per-axis plane parameters: t = q * sinv + adj
    const float ax = (px - r.o[0]) * r.inv[0];
    const float ay = (py - r.o[1]) * r.inv[1];
    const float az = (pz - r.o[2]) * r.inv[2];
    const float six = sx * r.inv[0], siy = sy * r.inv[1], siz = sz * r.inv[2];
    const float cull = t_best * 1.00001f;
    const float tmin = r.tmin;
    const uint32_t nxw[4] = { nqx.x, nqx.y, nqx.z, nqx.w };
    const uint32_t fxw[4] = { fqx.x, fqx.y, fqx.z, fqx.w };
    const uint32_t nyw[4] = { nqy.x, nqy.y, nqy.z, nqy.w };
    const uint32_t fyw[4] = { fqy.x, fqy.y, fqy.z, fqy.w };
    const uint32_t nzw[4] = { nqz.x, nqz.y, nqz.z, nqz.w };
    const uint32_t fzw[4] = { fqz.x, fqz.y, fqz.z, fqz.w };
    // all eight tests first, without branches: the hit mask
    uint32_t hitmask = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s)
    {
        const int wi = s >> 1;
        float qnx, qfx, qny, qfy, qnz, qfz;
        if (s & 1)
        {
            qnx = q_hi16(nxw[wi]); qfx = q_hi16(fxw[wi]);
            qny = q_hi16(nyw[wi]); qfy = q_hi16(fyw[wi]);
            qnz = q_hi16(nzw[wi]); qfz = q_hi16(fzw[wi]);
        }
        else
        {
            qnx = q_lo16(nxw[wi]); qfx = q_lo16(fxw[wi]);
            qny = q_lo16(nyw[wi]); qfy = q_lo16(fyw[wi]);
            qnz = q_lo16(nzw[wi]); qfz = q_lo16(fzw[wi]);
        }
        const float tnx = __builtin_fmaf(qnx, six, ax);
        const float tfx = __builtin_fmaf(qfx, six, ax);
        const float tny = __builtin_fmaf(qny, siy, ay);
        const float tfy = __builtin_fmaf(qfy, siy, ay);
        const float tnz = __builtin_fmaf(qnz, siz, az);
        const float tfz = __builtin_fmaf(qfz, siz, az);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        float tf = fminf(fminf(tfx, tfy), tfz);
        tf = __builtin_fmaf(fabsf(tf), 4e-6f, tf);
        hitmask |= (tn <= fminf(tf, cull)) ? (1u << s) : 0u;
    }
    NodeHits h;
    h.child_base = h1.x;
    h.tri_base = h1.y;
    // inner hits to key space: bit s -> bit s ^ k, k = neg ^ 7 (xor of a 3-bit index
    // = swaps of single bits, pairs and nibbles)
    uint32_t m = hitmask & imask;
    const uint32_t k = (r.code & 7u) ^ 7u;
    m = (k & 1u) ? (((m & 0x55u) << 1) | ((m >> 1) & 0x55u)) : m;
    m = (k & 2u) ? (((m & 0x33u) << 2) | ((m >> 2) & 0x33u)) : m;
    m = (k & 4u) ? (((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu)) : m;
    h.inner = m;
    // hit leaves: their triangle ranges (few per node, so a loop)
    uint32_t trimask = 0;
    uint32_t leaf = hitmask & ~imask;
    while (leaf)
    {
        const uint32_t s = __builtin_ctz(leaf);
        leaf &= leaf - 1u;
        const uint32_t mm = ((s < 4 ? meta_lo : meta_hi) >> (8 * (s & 3))) & 0xFFu;
        // 0xFF marks an empty slot (its inverted box can still pass the
        // padded test on a tiny node)
        if (mm != 0xFFu)
        {
            const uint32_t off = mm >> 2, cnt = (mm & 3u) + 1u;
            trimask |= ((1u << cnt) - 1u) << off;
        }
    }
    h.trimask = trimask;
    return h;
}

// Test triangle tri_base + i; keeps the closest hit (ties to the smaller
// primitive id).  Returns true for an accepted hit.
template <bool ANY, bool AUX>
WL_HD bool tri_visit(const WlTriRecord *__restrict__ tris, uint32_t tri_base, uint32_t i,
                     const TravRay &r, TravState &st)
{
    const float4 *tp = reinterpret_cast<const float4 *>(tris + tri_base + i);
    const float4 w0 = tp[0], w1 = tp[1], w2 = tp[2];
    float t, b1, b2;
    // the best hit so far bounds t: closest hit keeps only t <= res.t (ties to
    // the smaller primitive id below); any-hit holds res.t = tmax until its hit
    if (tri_test(r, w0, w1, w2, st.res.t, t, b1, b2))
    {
        const uint32_t prim = wl_f2u(w2.y);
        if (ANY || t < st.res.t || (t == st.res.t && prim < st.res.prim))
        {
            st.res.t = t;
            st.res.prim = prim;
            st.res.b1 = b1;
            st.res.b2 = b2;
            if (AUX)
            {
                st.res.input = wl_f2u(w2.z);
                st.res.local = wl_f2u(w2.w);
            }
            return true;
        }
    }
    return false;
}

// Chunks of XCD xcd's eighth of n rays under the run-time chunk schedule of
// the persistent trace kernels (traverse_persistent.h, DYN = 2): eighth x is
// [x * ceil(n / 8), (x + 1) * ceil(n / 8)) clipped to n, and a launch
// advances XCD x's counter by exactly this many.
WL_HD uint32_t dyn_xcd_chunks(uint32_t n, uint32_t xcd, uint32_t chunk)
{
    const uint32_t n8 = (n + 7u) / 8u;
    const uint32_t lo = xcd * n8 < n ? xcd * n8 : n;
    const uint32_t hi = lo + n8 < n ? lo + n8 : n;
    return (hi - lo + chunk - 1u) / chunk;
}

// Next node from the current group, else from the stack; false when none.
template <int BLOCK>
WL_HD bool next_node(const TravRay &r, TravState &st, uint32_t *lds_stack_lane)
{
    if (st.grp_mask == 0)
    {
        if (st.sp == 0)
            return false;
        --st.sp;
        const uint32_t e = lds_stack_lane[st.sp * BLOCK];
        st.grp_base = (e >> 8) << 3;
        st.grp_mask = e & 0xFFu;
    }
    const uint32_t pos = __builtin_ctz(st.grp_mask);
    st.grp_mask &= st.grp_mask - 1u;
    st.node = st.grp_base + (pos ^ (r.code & 7u) ^ 7u);
    if (st.grp_mask)
    {
        lds_stack_lane[st.sp * BLOCK] = ((st.grp_base >> 3) << 8) | st.grp_mask;
        ++st.sp;
    }
    return true;
}

// One step: visit the pending node (8 child boxes), test the triangles of its
// leaves, then pick the next node from the current group or the stack.
// Returns true when the ray is finished (stack empty, or first hit for ANY).
template <bool ANY, bool COUNT, int BLOCK, bool AUX = false>
WL_HD bool trav_step(const WlBvh8Node *__restrict__ nodes,
                     const WlTriRecord *__restrict__ tris, const TravRay &r,
                     TravState &st, uint32_t *lds_stack_lane, uint32_t &n_nodes,
                     uint32_t &n_tris, const WlBvh8Node *top = nullptr, uint32_t ntop = 0)
{
    uint32_t trimask = 0, tri_base = 0;
    if (st.visit)
    {
        WL_TRAV_CHECK(st.node < wl_trav_nodes_bound, "node index");
        const NodeHits h = node_visit(nodes, st.node, r, st.res.t, top, ntop);
        if (COUNT)
            ++n_nodes;
        trimask = h.trimask;
        tri_base = h.tri_base;
        st.grp_base = h.child_base;
        st.grp_mask = h.inner;
        st.visit = false;
    }
    // ---- triangles of the node just visited
    while (trimask)
    {
        const uint32_t i = __builtin_ctz(trimask);
        trimask &= trimask - 1u;
        WL_TRAV_CHECK((size_t)tri_base + i < wl_trav_tris_bound, "triangle index");
        if (COUNT)
            ++n_tris;
        if (tri_visit<ANY, AUX>(tris, tri_base, i, r, st) && ANY)
            return true;
    }
    // ---- next node: from the current group, else from the stack
    WL_TRAV_CHECK((st.grp_mask & (st.grp_mask - 1u)) == 0u || st.sp < WL_BVH8_STACK_DEPTH,
                  "stack depth");
    if (!next_node<BLOCK>(r, st, lds_stack_lane))
        return true;
    st.visit = true;
    return false;
}

// Persistent-kernel step: one pending triangle, then, once no triangle is
// pending, one node (the next from the group or the stack, then its box
// test).  A lane that tests its node's last triangle goes on to the next node
// in the same step: the order of tests is that of doing the two one after the
// other, and the wave's node branch runs with more of its lanes.  Returns
// true when the ray is finished (no node left, or with ANY its first hit).
template <int BLOCK, bool AUX = false, bool ANY = false>
WL_HD bool trav_step_one(const WlBvh8Node *__restrict__ nodes,
                         const WlTriRecord *__restrict__ tris, const TravRay &r,
                         TravState &st, uint32_t *lds_stack_lane,
                         const WlBvh8Node *top = nullptr, uint32_t ntop = 0)
{
    if (st.trimask)
    {
        const uint32_t i = __builtin_ctz(st.trimask);
        st.trimask &= st.trimask - 1u;
        WL_TRAV_CHECK((size_t)st.tri_base + i < wl_trav_tris_bound, "triangle index");
        if (tri_visit<ANY, AUX>(tris, st.tri_base, i, r, st) && ANY)
            return true;
        if (st.trimask)
            return false;
    }
    if (!st.visit)
    {
        WL_TRAV_CHECK((st.grp_mask & (st.grp_mask - 1u)) == 0u || st.sp < WL_BVH8_STACK_DEPTH,
                  "stack depth");
        if (!next_node<BLOCK>(r, st, lds_stack_lane))
            return true;
    }
    WL_TRAV_CHECK(st.node < wl_trav_nodes_bound, "node index");
    const NodeHits h = node_visit(nodes, st.node, r, st.res.t, top, ntop);
    st.trimask = h.trimask;
    st.tri_base = h.tri_base;
    st.grp_base = h.child_base;
    st.grp_mask = h.inner;
    st.visit = false;
    return false;
}

WL_HD void trav_finish(TravResult &res)
{
    if (res.prim == WL_INVALID)
        res.t = wl_u2f(0x7F800000u); // +inf
}

// Traverse one ray to completion.
template <bool ANY, bool COUNT, int BLOCK, bool AUX = false>
WL_HD TravResult
trav_run(const WlBvh8Node *__restrict__ nodes,
         const WlTriRecord *__restrict__ tris, const TravRay &r, bool active,
         uint32_t *lds_stack_lane, uint32_t &n_nodes, uint32_t &n_tris)
{
    TravState st;
    trav_init(st, r, active);
    if (active)
        while (!trav_step<ANY, COUNT, BLOCK, AUX>(nodes, tris, r, st, lds_stack_lane,
                                                 n_nodes, n_tris))
        {
        }
    trav_finish(st.res);
    return st.res;
}

} // namespace wl
