// pathtrace.hip — wavefront path tracer for gfx950.
//
// Restates the per-pixel megakernel of the reference (raygen loop
// PathTracing.cu:40-139, __closesthit__radiance :181-289, __miss__radiance
// :163-171, NEE :81-128, RR :130-138) as queue kernels:
//
//   [extend -> shade -> shadow] x max_depth -> accumulate
//
//   extend     closest-hit BVH8 traversal of the active ray queue.  The
//              first pass of a single-scene batch traces path i's camera
//              ray (TEA seed (pixel, frame), pinhole ray) computed in place,
//              and the first shade pass recomputes that vertex's state and
//              starts the path's radiance: nothing is written per path
//              before the first traversal.  Multi-request batches keep a
//              raygen pass (k_raygen_multi).
//   shade      miss / emitter hit with MIS / depth cap / Disney BSDF
//              sample-pdf-eval / NEE light sample + BSDF eval / RR, in the
//              reference's RNG draw order; appends the continuing paths and
//              the shadow rays to compacted queues (wave64 ballot + mbcnt
//              prefix, one atomic per queue per workgroup).
//   shadow     any-hit traversal of the shadow queue; adds the precomputed
//              NEE contribution of unoccluded rays.
//   accumulate compensated (Kahan) FP32 film: sum + compensation per pixel.
//
// Radiance contributions reach a path in the reference's order (one per
// bounce at most), so per-sample radiance matches a serial evaluation.
// Every kernel loops over a device-side queue count with a fixed grid, so the
// host enqueues a whole batch without synchronising.
#include "guide_device.h"
#include "shading_device.h"
#include "traverse_persistent.h"
#include "serve_plan.h"
#include "wl_pt.h"

namespace wl
{

constexpr int kPtBlock = 256;

// ---- workgroup-level append to two queues -----------------------------------
// Every wave of the workgroup calls this in the same loop pass.  Per wave a
// ballot and mbcnt prefix; per workgroup one atomic per queue, issued by one
// thread.  All waves appending to one counter serialise at one L2 channel,
// so a per-wave atomic makes the counter the bottleneck of the shade kernel;
// a per-workgroup one divides that by the waves per workgroup.  The LDS
// slots alternate between passes (pass parity), so two barriers per pass
// suffice.
struct BlockAppendLds
{
    uint32_t cnt[2][2][kPtBlock / 64]; // [pass parity][queue][wave]
    uint32_t base[2][2];               // [pass parity][queue]
};

__device__ __forceinline__ void block_append2(BlockAppendLds &L, uint32_t pass,
                                              uint32_t *c0, bool p0, uint32_t *c1, bool p1,
                                              uint32_t &slot0, uint32_t &slot1)
{
    const uint32_t par = pass & 1u;
    const uint32_t wave = threadIdx.x / 64u, lane = __lane_id();
    const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
    if (lane == 0)
    {
        L.cnt[par][0][wave] = (uint32_t)__popcll(m0);
        L.cnt[par][1][wave] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t t0 = 0, t1 = 0;
        for (uint32_t w = 0; w < kPtBlock / 64; ++w)
        {
            t0 += L.cnt[par][0][w];
            t1 += L.cnt[par][1][w];
        }
        L.base[par][0] = t0 ? atomicAdd(c0, t0) : 0u;
        L.base[par][1] = t1 ? atomicAdd(c1, t1) : 0u;
    }
    __syncthreads();
    uint32_t o0 = L.base[par][0], o1 = L.base[par][1];
    for (uint32_t w = 0; w < wave; ++w)
    {
        o0 += L.cnt[par][0][w];
        o1 += L.cnt[par][1][w];
    }
    slot0 = o0 + __builtin_amdgcn_mbcnt_hi((uint32_t)(m0 >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m0, 0u));
    slot1 = o1 + __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
}

__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 v3of(float4 a) { return mk(a.x, a.y, a.z); }

// ---- camera rays ----------------------------------------------------------
// Camera ray of path p of a batch: its pixel and sample, the seed after
// PinholeGenerateRay's two draws (Camera.h:28-35: u, then v), the direction.
__device__ __forceinline__ void camera_ray(const PtParams &P, uint32_t p, V3 &o, V3 &d,
                                           uint32_t &seed)
{
    const uint32_t pixel = P.pix0 + p % P.npix;
    const uint32_t s = p / P.npix;
    const uint32_t x = pixel % P.width, y = pixel / P.width;
    seed = tea4(pixel, P.frame0 + s);
    const float u = 2.f * ((float)(int)x + rnd(seed)) / (float)(int)P.width - 1.f;
    const float v = 2.f * ((float)(int)y + rnd(seed)) / (float)(int)P.height - 1.f;
    d = normalize(ld3(P.cam.lookat) + ld3(P.cam.right) * u + ld3(P.cam.up) * v);
    o = ld3(P.cam.pos);
}

// Start of a single-scene batch: the first queue's length and the material
// bins.  No per-path state is written: the first extend pass computes each
// path's camera ray with camera_ray, and the first shade pass recomputes the
// first vertex's state (origin, direction and seed, throughput 1, depth 0,
// lastTraceTerm 1, lastPdf 1) and writes the path's radiance (and, when
// training a guide, its record accumulator), so a path costs no ray, queue
// or radiance write and read before its first traversal.
__global__ void k_batch_start(PtParams P, PtBuffers B)
{
    if (threadIdx.x == 0)
        B.counters[0] = P.npix * P.nsamples;
    for (uint32_t b = threadIdx.x; b < P.sort_bins; b += blockDim.x)
        B.sort_bins[b] = 0;
}

// ---- top BVH levels in LDS ---------------------------------------------------
// The single-scene traversal kernels with the 16-entry stack copy the first
// 73 nodes (levels 0-2 of the breadth-first node layout: root, its group and
// their groups) to LDS; every ray starts there.  9,344 bytes beside the
// stack keep 6 workgroups per CU.  Multi-request launches start rays at other
// roots of the node arena and keep the plain loads.
constexpr uint32_t kPtTopNodes = 73;

template <int STACK, bool MULTI>
constexpr uint32_t pt_top_nodes() { return (STACK == 16 && !MULTI) ? kPtTopNodes : 0u; }

template <int STACK, bool MULTI>
__device__ __forceinline__ uint32_t pt_load_top(uint4 *top, const PtScene &S)
{
    if (!pt_top_nodes<STACK, MULTI>())
        return 0u;
    const uint32_t ntop = min(S.num_nodes, pt_top_nodes<STACK, MULTI>());
    const uint4 *src = reinterpret_cast<const uint4 *>(S.nodes);
    for (uint32_t i = threadIdx.x; i < 8u * ntop; i += kPtBlock)
        top[i] = src[i];
    __syncthreads();
    return ntop;
}

// ---- extend: closest hit for the active queue -----------------------------
template <int STACK, bool MULTI>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(STACK == 16 ? 6 : 5)))
void k_extend(PtScene S, PtBuffers B, int parity, const PtReq *__restrict__ reqs, float eps)
{
    __shared__ uint32_t stack[STACK * kPtBlock];
    __shared__ uint4 top[8 * (pt_top_nodes<STACK, MULTI>() ? pt_top_nodes<STACK, MULTI>() : 1)];
    const uint32_t ntop = pt_load_top<STACK, MULTI>(top, S);
    const uint32_t count = B.counters[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        B.counters[parity ^ 1] = 0; // next queue
        B.counters[2] = 0;          // shadow queue
        B.stats[0] += count;        // extension rays (single writer)
    }
    const wl_ray *rays = parity ? B.rays1 : B.rays0;
    const uint32_t *qreq = parity ? B.qreq1 : B.qreq0; // multi-request batches
    // the queues carry the path's lastPdf and seed in the rays' w words (the
    // shade pass reads its state back from the ray): every extension ray is
    // traced on [epsilon, 1e30], epsilon the batch's or the ray's request's
    auto load_eps = [&](uint32_t i, TravRay &r, float e) {
        const f4v *rp = reinterpret_cast<const f4v *>(rays + i);
        const f4v a = __builtin_nontemporal_load(rp), b = __builtin_nontemporal_load(rp + 1);
        trav_setup(r, a.x, a.y, a.z, b.x, b.y, b.z, e, 1e30f);
    };
    auto load = [&](uint32_t i, TravRay &r) { load_eps(i, r, eps); };
    auto store = [&](uint32_t i, const TravResult &h) { store_hit_nt(B.hits, i, h); };
    if constexpr (MULTI) // S.nodes / S.tris: the arena of the requests' BVHs
        trav_persistent<false, kPtBlock, false, true, false, 8, 20>(
            S.nodes, S.tris, count, stack + threadIdx.x,
            [&](uint32_t i, TravRay &r, uint32_t &root) {
                const PtReq &rq = reqs[qreq[i]];
                root = rq.root;
                load_eps(i, r, rq.P.epsilon);
            },
            store);
    else
        trav_persistent<false, kPtBlock, false, false, false, 0, 20>(
            S.nodes, S.tris, count, stack + threadIdx.x, load, store,
            reinterpret_cast<const WlBvh8Node *>(top), ntop);
}

// The first bounce of a single-scene batch: queue entry i is path i, and its
// camera ray is computed here (the same camera_ray as the shade pass uses),
// not read from a ray queue: nothing is written before the first traversal.
template <int STACK>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(5)))
void k_extend_first(PtParams P, PtScene S, PtBuffers B)
{
    __shared__ uint32_t stack[STACK * kPtBlock];
    __shared__ uint4 top[8 * (pt_top_nodes<STACK, false>() ? pt_top_nodes<STACK, false>() : 1)];
    const uint32_t ntop = pt_load_top<STACK, false>(top, S);
    const uint32_t count = P.npix * P.nsamples;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        B.counters[1] = 0; // next queue
        B.counters[2] = 0; // shadow queue
        B.stats[0] += count;
    }
    auto load = [&](uint32_t i, TravRay &r) {
        V3 o, d;
        uint32_t seed;
        camera_ray(P, i, o, d, seed);
        trav_setup(r, o.x, o.y, o.z, d.x, d.y, d.z, P.epsilon, 1e30f);
    };
    auto store = [&](uint32_t i, const TravResult &h) { store_hit_nt(B.hits, i, h); };
    trav_persistent<false, kPtBlock, false, false, false, 0, 20>(
        S.nodes, S.tris, count, stack + threadIdx.x, load, store,
        reinterpret_cast<const WlBvh8Node *>(top), ntop);
}

// ---- material sort of the shade queue ------------------------------------------
// Counting sort of the queue entries by the mesh (= material slot) they hit;
// misses take the last bin.  Per block: an LDS histogram, one global atomic per
// bin; the scatter reserves each block's range per bin the same way, so the
// global atomics per bounce are (blocks x bins), not one per entry.
constexpr uint32_t kSortMaxBins = 1024;

__device__ __forceinline__ uint32_t sort_key(const PtScene &S, uint32_t prim, uint32_t nbins)
{
    if (prim == WL_INVALID)
        return nbins - 1;
    const float4 *sr = reinterpret_cast<const float4 *>(S.prims) + 3 * (size_t)prim;
    return min(__float_as_uint(sr[2].y), nbins - 2);
}

__global__ __launch_bounds__(kPtBlock) void k_sort_count(PtParams P, PtScene S, PtBuffers B,
                                                         int parity)
{
    __shared__ uint32_t hist[kSortMaxBins];
    const uint32_t nb = P.sort_bins;
    for (uint32_t b = threadIdx.x; b < nb; b += kPtBlock)
        hist[b] = 0;
    __syncthreads();
    const uint32_t count = B.counters[parity];
    for (uint32_t i = blockIdx.x * kPtBlock + threadIdx.x; i < count; i += gridDim.x * kPtBlock)
    {
        const uint32_t key =
            sort_key(S, __float_as_uint(reinterpret_cast<const float4 *>(B.hits)[i].y), nb);
        B.sort_keys[i] = key;
        atomicAdd(&hist[key], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += kPtBlock)
        if (hist[b])
            atomicAdd(&B.sort_bins[b], hist[b]);
}

// One block: bin counts -> running bin offsets; clears the counts.
__global__ void k_sort_scan(PtParams P, PtBuffers B)
{
    __shared__ uint32_t part[kPtBlock];
    const uint32_t nb = P.sort_bins;
    const uint32_t per = (nb + kPtBlock - 1) / kPtBlock;
    const uint32_t b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    uint32_t s = 0;
    for (uint32_t b = b0; b < b1; ++b)
        s += B.sort_bins[b];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t run = 0;
        for (int t = 0; t < kPtBlock; ++t)
        {
            const uint32_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (uint32_t b = b0; b < b1; ++b)
    {
        const uint32_t c = B.sort_bins[b];
        B.sort_bins[kSortMaxBins + b] = run;
        B.sort_bins[b] = 0;
        run += c;
    }
}

__global__ __launch_bounds__(kPtBlock) void k_sort_scatter(PtParams P, PtBuffers B, int parity)
{
    __shared__ uint32_t hist[kSortMaxBins];
    __shared__ uint32_t base[kSortMaxBins];
    const uint32_t nb = P.sort_bins;
    const uint32_t count = B.counters[parity];
    for (uint32_t start = blockIdx.x * kPtBlock; start < count; start += gridDim.x * kPtBlock)
    {
        for (uint32_t b = threadIdx.x; b < nb; b += kPtBlock)
            hist[b] = 0;
        __syncthreads();
        const uint32_t i = start + threadIdx.x;
        uint32_t key = 0, local = 0;
        if (i < count)
        {
            key = B.sort_keys[i];
            local = atomicAdd(&hist[key], 1u);
        }
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < nb; b += kPtBlock)
            if (hist[b])
                base[b] = atomicAdd(&B.sort_bins[kSortMaxBins + b], hist[b]);
        __syncthreads();
        if (i < count)
            B.sort_perm[base[key] + local] = i;
        __syncthreads();
    }
}

// ---- guide inference for the active queue ------------------------------------
// Query of a hit: position normalised to the scene box, geometric normal
// facing the incoming ray, outgoing direction.  The shade kernel builds the
// same query from the same values for its training records.
__device__ __forceinline__ void guide_query(const PtParams &P, V3 hitPos, V3 Ng, V3 rayDir,
                                            float q[9])
{
    q[0] = (hitPos.x - P.gbox_lo[0]) * P.gbox_inv[0];
    q[1] = (hitPos.y - P.gbox_lo[1]) * P.gbox_inv[1];
    q[2] = (hitPos.z - P.gbox_lo[2]) * P.gbox_inv[2];
    q[3] = Ng.x;
    q[4] = Ng.y;
    q[5] = Ng.z;
    q[6] = -rayDir.x;
    q[7] = -rayDir.y;
    q[8] = -rayDir.z;
}

// Path state word thr.w: bits 0-15 the path's depth, bits 16-30 the number of
// guided vertices it has passed, bit 31 set once its guide training record
// is taken.
constexpr uint32_t kDepthMask = 0xFFFFu;
constexpr uint32_t kGvisMask = 0x7FFFu;
constexpr uint32_t kGrecBit = 1u << 31;

// Whether the shade pass samples the guide at the vertex of queue item i: a
// hit on a mesh that emits nothing, below the depth cap, whose material has
// trans = 0 and clearcoat = 0, on a path that has passed fewer than gverts
// guided vertices (gverts = 0: no limit).  The shade kernel decides from the
// same values (its emitter, depth-cap and material branches, then gverts).
// SEL 0: every hit (no path state: diagnostics), 1: a path's first vertex
// (depth 0, none guided), 2: any later vertex (state read from thr).
template <int SEL>
__device__ __forceinline__ bool guide_wanted(const PtParams &P, const PtScene &S,
                                             const PtBuffers &B, const uint32_t *qin,
                                             uint32_t prim, uint32_t i)
{
    if (prim == WL_INVALID)
        return false;
    if (SEL == 0)
        return true;
    uint32_t depth = 0, gvis = 0;
    if (SEL == 2)
    {
        const uint32_t raw = __float_as_uint(B.thr[qin[i]].w);
        depth = raw & kDepthMask;
        gvis = (raw >> 16) & kGvisMask;
    }
    if (depth + 1 >= P.max_depth || (P.gverts != 0 && gvis >= P.gverts))
        return false;
    const uint32_t mesh = __float_as_uint(S.prims[12 * (size_t)prim + 9]);
    const PtMesh M = S.meshes[mesh];
    if (M.light != WL_INVALID)
        return false;
    const wl_disney_material &m = S.materials[M.material];
    return m.trans == 0.0f && m.clearcoat == 0.0f;
}

// Guide inference for the active queue.  Each wave walks its 64-item chunks
// of the queue, appends the items whose vertex the shade pass will guide to a
// ring of item indices in LDS, and runs the network on full 16-query tiles
// from the ring (the remainder at the end), so the MFMA work follows the
// guided vertices, not the queue length.  Results go to gout at the item's
// queue index; other items' slots are not written.
template <int SEL>
__global__ __launch_bounds__(kPtBlock)
void k_guide_infer_pt(PtParams P, PtScene S, PtBuffers B, int parity)
{
    constexpr int kRing = 128; // holds < 16 left over + 64 appended
    __shared__ bf16x8 sA[kGuideFwdVecs];
    __shared__ float sb[kGuideBiases];
    __shared__ float sq[9 * kPtBlock];
    __shared__ uint32_t sring[(kPtBlock / 64) * kRing];
    float *q = sq + 9 * threadIdx.x;
    uint32_t *ring = sring + (threadIdx.x / 64) * kRing;
    const bf16x8 *fwd = reinterpret_cast<const bf16x8 *>(P.gfwd);
    guide_stage<kPtBlock, kGuideFwdVecs>(sA, fwd);
    guide_stage_biases<kPtBlock>(sb, P.gparams);
    __syncthreads();
    const uint32_t count = B.counters[parity];
    const uint32_t *qin = parity ? B.qpath1 : B.qpath0;
    const wl_ray *rays = parity ? B.rays1 : B.rays0;
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t nwaves = gridDim.x * (kPtBlock / 64);
    const uint32_t wave = blockIdx.x * (kPtBlock / 64) + threadIdx.x / 64;
    // one tile: ring entries [head, head + n), n <= 16 (wave-uniform)
    auto run_tile = [&](uint32_t head, uint32_t n) {
        const uint32_t k = (uint32_t)(lane & 15);
        const bool valid = k < n;
        const uint32_t i = valid ? ring[(head + k) % kRing] : 0u;
#pragma unroll
        for (int t = 0; t < 9; ++t)
            q[t] = 0.f;
        if (valid)
        {
            const float4 hit = reinterpret_cast<const float4 *>(B.hits)[i];
            float4 r0, r1;
            if (SEL == 1)
            {
                // the first queue: path i's camera ray (no ray queue is written)
                V3 co, cd;
                uint32_t sd;
                camera_ray(P, i, co, cd, sd);
                r0 = f4(co, P.epsilon);
                r1 = f4(cd, 1e30f);
            }
            else
            {
                const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
                r0 = rp[0];
                r1 = rp[1];
            }
            const uint32_t prim = __float_as_uint(hit.y);
            const float4 *sr = reinterpret_cast<const float4 *>(S.prims) + 3 * (size_t)prim;
            const float4 w0 = sr[0], w1 = sr[1], w2 = sr[2];
            const V3 o = mk(r0.x, r0.y, r0.z), d = mk(r1.x, r1.y, r1.z);
            const V3 va = mk(w0.x, w0.y, w0.z), vb = mk(w0.w, w1.x, w1.y),
                     vc = mk(w1.z, w1.w, w2.x);
            V3 Ng = normalize(cross(vb - va, vc - va));
            if (dot(Ng, d) > 0)
                Ng = -Ng;
            guide_query(P, o + d * hit.x, Ng, d, q);
        }
        bf16x8 x[2];
        guide_encode(q, g, x);
        floatx4 o[2];
        guide_forward<false>(sA, sb, lane, x, o, nullptr);
        guide_decode_lanes(o, g, B.gout + kGuideDecoded * (size_t)i, valid);
    };
    uint32_t head = 0, tail = 0; // ring positions, wave-uniform
    for (uint32_t base = wave * 64u; base < count; base += nwaves * 64u)
    {
        const uint32_t i = base + (uint32_t)lane;
        bool want = false;
        if (i < count)
        {
            const uint32_t prim = reinterpret_cast<const uint32_t *>(B.hits)[4 * (size_t)i + 1];
            want = guide_wanted<SEL>(P, S, B, qin, prim, i);
        }
        const uint64_t m = __ballot(want);
        if (want)
            ring[(tail + lane_prefix(m)) % kRing] = i;
        tail += (uint32_t)__popcll(m);
        // the ring is read by other lanes of this wave: order the LDS
        // accesses at wave scope
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        while (tail - head >= 16u)
        {
            run_tile(head, 16u);
            head += 16u;
        }
    }
    if (tail != head)
        run_tile(head, tail - head);
}

// ---- shade ------------------------------------------------------------------
// GSAMPLE: one-sample MIS of the learned guide and the BSDF at vertices whose
// material has trans = 0 and clearcoat = 0.  There the reference's pdf equals
// the density its sampler draws from wherever the BSDF is non-zero, so
// mixing in the guide keeps the estimator's expectation unchanged.  The mixed
// density replaces the BSDF pdf everywhere the reference uses it (the path
// weight, lastPdf for emitter-hit MIS, the NEE MIS weight).
// MULTI: paths of many requests; each lane shades with its request's scene
// and parameters.
// GTRAIN: paths carry guide training records (radiance after the recorded
// vertex is summed; GSAMPLE vertices record themselves).
// TRACE: also writes the per-vertex record of PtBuffers::trace.
template <bool GSAMPLE, bool GTRAIN, bool MULTI, bool TRACE = false>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(3)))
void k_shade(PtParams P0, PtScene S0, PtBuffers B, int parity, const PtReq *__restrict__ reqs,
             int first)
{
    const uint32_t count = B.counters[parity];
    const uint32_t *qin = parity ? B.qpath1 : B.qpath0;
    uint32_t *qout = parity ? B.qpath0 : B.qpath1;
    wl_ray *rout = parity ? B.rays0 : B.rays1;
    __shared__ BlockAppendLds app;
    uint32_t pass = 0;
    // every thread of the workgroup runs the same passes (base is per
    // workgroup), as the workgroup barriers of the append require
    for (uint32_t base = blockIdx.x * kPtBlock; base < count; base += gridDim.x * kPtBlock)
    {
        const uint32_t j = base + threadIdx.x;
        const bool in_range = j < count;
        // material-sorted order: the j-th entry of the sorted queue
        const uint32_t i = (!MULTI && P0.sort_bins && in_range) ? B.sort_perm[j] : j;
        bool cont = false, shadow = false, shadow_gacc = false;
        uint32_t path = 0, req = 0, root = 0;
        wl_ray next_ray{}, shadow_ray{};
        V3 contrib = mk(0, 0, 0);
        if (in_range)
        {
            // the first queue is the identity: entry i is path i
            const bool ident = first != 0;
            path = ident ? i : qin[i];
            if (MULTI)
                req = (parity ? B.qreq1 : B.qreq0)[i];
            const PtReq *rq = MULTI ? reqs + req : nullptr;
            if (MULTI)
                root = rq->root;
            const PtParams &P = MULTI ? rq->P : P0;
            const PtScene &S = MULTI ? rq->S : S0;
            const float4 hit = reinterpret_cast<const float4 *>(B.hits)[i];
            float4 s_pos, s_dir, s_thr, s_ltt;
            if (first)
            {
                // the first vertex: single-scene batches recompute the camera
                // ray; multi-request ones read it (with lastPdf 1 and the
                // seed in its w words) from the queue the raygen pass wrote,
                // whose item lookup the shade pass would otherwise repeat
                if (MULTI)
                {
                    const float4 *rp = reinterpret_cast<const float4 *>(B.rays0 + i);
                    s_pos = rp[0];
                    s_dir = rp[1];
                }
                else
                {
                    V3 o, d;
                    uint32_t sd;
                    camera_ray(P, path, o, d, sd);
                    s_pos = f4(o, 1.0f);                // w: lastPdf
                    s_dir = f4(d, __uint_as_float(sd)); // w: seed
                }
                s_thr = f4(mk(1, 1, 1), __uint_as_float(0u)); // w: depth
                s_ltt = f4(mk(1, 1, 1), 1.0f); // w: miss weight
            }
            else
            {
                // the ray this pass shades: origin and lastPdf, then the
                // direction and the seed (written so by the last pass)
                const float4 *rp =
                    reinterpret_cast<const float4 *>((parity ? B.rays1 : B.rays0) + i);
                s_pos = rp[0];
                s_dir = rp[1];
                s_thr = B.thr[path];
                s_ltt = B.ltt[path];
            }
            V3 rayPos = v3of(s_pos);
            float lastPdf = s_pos.w;
            V3 rayDir = v3of(s_dir);
            uint32_t seed = __float_as_uint(s_dir.w);
            V3 throughput = v3of(s_thr);
            const uint32_t sw = __float_as_uint(s_thr.w);
            uint32_t depth = sw & kDepthMask;
            uint32_t gvis = (sw >> 16) & kGvisMask; // guided vertices passed
            V3 lastTraceTerm = v3of(s_ltt);
            float missw = 1.0f; // miss weight of the ray this vertex samples
            // radiance changes only at a miss or an emitter hit: read and
            // written back only then
            const bool gtrain = GTRAIN && (P.gmode & 2u);
            V3 radiance = mk(0, 0, 0), rad_in = mk(0, 0, 0);
            bool rad_dirty = false;
            const uint32_t prim = __float_as_uint(hit.y);
            bool done = false;
            // guiding: the path's training record is taken (its first guided
            // vertex); the radiance added after it goes to gacc
            const bool grec_done = gtrain && (sw & kGrecBit) != 0;
            bool grec_now = false;
            if (prim == WL_INVALID)
            {
                // __miss__radiance: radiance += throughput * bg.  The
                // reference leaves the last vertex's weight f |cos| / pdf
                // out of this term, so its expectation is throughput * bg
                // times the probability that the last vertex's sampler
                // sends the ray to the background.  s_ltt.w keeps that
                // expectation when a guided vertex sampled the ray: the
                // reference sampler's density over the mixed density
                // there, and exactly 1 after every other vertex.
                radiance = ident ? mk(0, 0, 0) : v3of(B.rad[path]);
                rad_in = radiance;
                radiance = radiance + throughput * ld3(P.bg) * s_ltt.w;
                rad_dirty = true;
                done = true;
            }
            else
            {
                depth += 1;
                const float t = hit.x, b1 = hit.z, b2 = hit.w;
                const float4 *sr = reinterpret_cast<const float4 *>(S.prims) + 3 * (size_t)prim;
                const float4 w0 = sr[0], w1 = sr[1], w2 = sr[2];
                const V3 va = mk(w0.x, w0.y, w0.z), vb = mk(w0.w, w1.x, w1.y),
                         vc = mk(w1.z, w1.w, w2.x);
                const uint32_t mesh = __float_as_uint(w2.y);
                const PtMesh M = S.meshes[mesh];
                V3 Ng = normalize(cross(vb - va, vc - va));
                V3 Ns;
                if (M.has_normal)
                {
                    const float4 *nr = reinterpret_cast<const float4 *>(S.normals) + 3 * (size_t)prim;
                    const float4 n0 = nr[0], n1 = nr[1], n2 = nr[2];
                    const V3 na = mk(n0.x, n0.y, n0.z), nb = mk(n0.w, n1.x, n1.y),
                             nc = mk(n1.z, n1.w, n2.x);
                    Ns = normalize(na * (1 - b1 - b2) + nb * b1 + nc * b2);
                }
                else
                    Ns = Ng;
                const V3 hitPos = rayPos + rayDir * t;
                if (M.light != WL_INVALID)
                {
                    const PtLight lt = S.lights[M.light];
                    if (depth == 1 || (P.strategy != WL_STRATEGY_NEE &&
                                       (lt.two_sided || dot(Ns, rayDir) < 0)))
                    {
                        float neePdf = 0.f;
                        if (P.strategy != WL_STRATEGY_UPT && depth > 1)
                        {
                            // PdfAreaLightPos (Pdf.h:20-31)
                            float pdf = 1.0f / (float)P.num_lights;
                            pdf /= (float)lt.tri_count;
                            pdf /= tri_area(va, vb, vc);
                            const float dist = t;
                            neePdf = pdf * dist * dist / fabsf(dot(Ns, rayDir));
                        }
                        const float uptPdf = lastPdf;
                        radiance = ident ? mk(0, 0, 0) : v3of(B.rad[path]);
                        rad_in = radiance;
                        rad_dirty = true;
                        radiance = radiance + throughput * ld3(lt.L) * lastTraceTerm *
                                                  (uptPdf / (uptPdf + neePdf));
                    }
                    done = true;
                }
                else if (depth >= P.max_depth)
                {
                    throughput = mk(0, 0, 0);
                    done = true;
                }
                else
                {
                    const Mat mat = load_mat(S.materials + M.material);
                    // texture colour (PathTracing.cu:261-268)
                    V3 tex = mk(1, 1, 1);
                    if (M.texture != WL_INVALID)
                    {
                        const float4 *ur = reinterpret_cast<const float4 *>(S.uvs) + 2 * (size_t)prim;
                        const float4 u01 = ur[0], u2 = ur[1];
                        const float w0 = 1 - b1 - b2;
                        const float tu = (u01.x * w0 + u01.z * b1) + u2.x * b2;
                        const float tv = (u01.y * w0 + u01.w * b1) + u2.y * b2;
                        const PtTexture T = S.textures[M.texture];
                        tex = tex_bilinear(S.texels + T.offset, T.width, T.height, tu, tv);
                    }
                    if (mat.trans < 0.1f)
                    {
                        if (dot(Ng, rayDir) > 0)
                            Ng = -Ng;
                        if (dot(Ns, rayDir) > 0)
                            Ns = -Ns;
                    }
                    const bool guided = GSAMPLE && mat.trans == 0.0f &&
                                        mat.clearcoat == 0.0f && (P.gmode & 1u) &&
                                        (P.gverts == 0 || gvis < P.gverts);
                    if (GSAMPLE && guided)
                        gvis += 1;
                    GuideMix gm;
                    bool took_guide = false;
                    V3 nd;
                    if (GSAMPLE && guided)
                    {
                        const float *gsrc = B.gout + kGuideDecoded * (size_t)i;
                        guide_load(gsrc, gm);
                        if (rnd(seed) < P.galpha)
                        {
                            const float u0 = rnd(seed);
                            const float u1 = rnd(seed);
                            const float u2 = rnd(seed);
                            nd = guide_sample(gm, gsrc + 40, u0, u1, u2);
                            took_guide = true;
                        }
                    }
                    if (!took_guide)
                    {
                        nd = sample_disney(mat, Ns, -rayDir, seed);
                        nd = normalize(nd);
                    }
                    float pdf = pdf_disney(mat, Ns, -rayDir, nd);
                    if (GSAMPLE && guided && !(pdf < 0.0f || __builtin_isnan(pdf)))
                    {
                        // the reference sampler's own density: its pdf less
                        // the diffuse term below the surface, where the
                        // sampler draws nothing (trans = 0, clearcoat = 0)
                        const float cl = dot(nd, Ns);
                        const float pb = pdf;
                        float sdens = pdf;
                        if (cl < 0.0f)
                            sdens -= 0.5f * (1.0f - mat.metallic) * (-cl * (1.0f / kPi));
                        pdf = P.galpha * guide_pdf(gm, nd) + (1.0f - P.galpha) * pdf;
                        if (!(pdf > 0.0f))
                            pdf = -1.0f; // a zero mixed density ends the path
                        else if (__builtin_isinf(pb))
                            // the reference pdf is infinite where its GGX
                            // reflection folds back onto -V (1 / |L.H|, and
                            // its sampler does land there): the limit of
                            // s / p_mix as p_bsdf grows, not inf / inf
                            missw = 1.0f / (1.0f - P.galpha);
                        else
                            missw = fmaxf(sdens, 0.0f) / pdf; // 0 if only q is infinite
                    }
                    if (pdf < 0.0f || __builtin_isnan(pdf))
                    {
                        done = true;
                    }
                    else
                    {
                        throughput = throughput * lastTraceTerm;
                        const float cosWeight = dot(nd, Ns);
                        V3 bsdf = eval_disney(mat, Ns, Ng, -rayDir, nd, tex);
                        bsdf = clamp3(bsdf, 0.f, 1e30f);
                        lastTraceTerm = bsdf * fabsf(cosWeight) / pdf;
                        if (TRACE && depth - 1 < B.trace_max)
                        {
                            float4 *tr = reinterpret_cast<float4 *>(
                                B.trace + 8 * ((size_t)path * B.trace_max + depth - 1));
                            tr[0] = make_float4((float)depth, (float)mesh, pdf,
                                                (lastTraceTerm.x + lastTraceTerm.y +
                                                 lastTraceTerm.z) / 3.0f);
                            tr[1] = make_float4((throughput.x + throughput.y + throughput.z) /
                                                    3.0f,
                                                0.0f, 0.0f, mat.trans);
                        }
                        const V3 lastRayDir = rayDir;
                        if (GSAMPLE && guided && gtrain && !grec_done)
                        {
                            // training record: query, sampled direction and
                            // the throughput reaching this vertex
                            float q[9];
                            guide_query(P, hitPos, Ng, rayDir, q);
                            float4 *r = reinterpret_cast<float4 *>(B.grec + 16 * (size_t)path);
                            r[0] = make_float4(q[0], q[1], q[2], q[3]);
                            r[1] = make_float4(q[4], q[5], q[6], q[7]);
                            r[2] = make_float4(q[8], nd.x, nd.y, nd.z);
                            r[3] = make_float4(0.0f, (throughput.x + throughput.y + throughput.z) *
                                                         (1.0f / 3.0f),
                                               0.0f, 0.0f);
                            grec_now = true;
                        }
                        rayPos = hitPos;
                        rayDir = nd;
                        lastPdf = pdf;
                        // ---- NEE at this vertex (PathTracing.cu:81-128)
                        if (P.strategy != WL_STRATEGY_UPT && P.num_lights > 0)
                        {
                            uint32_t lid = (uint32_t)((float)P.num_lights * rnd(seed));
                            lid = lid < P.num_lights ? lid : P.num_lights - 1;
                            const PtLight lt = S.lights[lid];
                            float lpdf = 1.0f / (float)P.num_lights;
                            uint32_t triIdx = (uint32_t)(rnd(seed) * (float)lt.tri_count);
                            triIdx = triIdx < lt.tri_count ? triIdx : lt.tri_count - 1;
                            lpdf = lpdf / (float)lt.tri_count;
                            const uint32_t lprim = lt.tri_offset + triIdx;
                            const float4 *lr = reinterpret_cast<const float4 *>(S.prims) + 3 * (size_t)lprim;
                            const float4 l0 = lr[0], l1 = lr[1], l2 = lr[2];
                            const V3 p0 = mk(l0.x, l0.y, l0.z), p1 = mk(l0.w, l1.x, l1.y),
                                     p2 = mk(l1.z, l1.w, l2.x);
                            float u = rnd(seed), v = rnd(seed);
                            if (u + v > 1)
                            {
                                u = 1 - u;
                                v = 1 - v;
                            }
                            const float cx = 1 - u - v, cy = u;
                            const V3 lpos = p0 + u * (p1 - p0) + v * (p2 - p0);
                            lpdf = lpdf / tri_area(p0, p1, p2);
                            V3 lN;
                            if (lt.has_normal)
                            {
                                const float4 *nr = reinterpret_cast<const float4 *>(S.normals) + 3 * (size_t)lprim;
                                const float4 n0 = nr[0], n1 = nr[1], n2 = nr[2];
                                const V3 na = mk(n0.x, n0.y, n0.z), nb = mk(n0.w, n1.x, n1.y),
                                         nc = mk(n1.z, n1.w, n2.x);
                                // weights (1 - cx - cy, cx, cy), as Sample.h:121 applies them
                                lN = normalize(na * (1 - cx - cy) + nb * cx + nc * cy);
                            }
                            else
                                lN = normalize(cross(p1 - p0, p2 - p0));
                            V3 visRay = lpos - rayPos;
                            const float dist = length(visRay);
                            visRay = visRay / dist;
                            const float cosTheta1 = dot(lN, visRay);
                            if (lt.two_sided || cosTheta1 < 0)
                            {
                                const float cosTheta2 = dot(Ns, visRay);
                                const float neePdf = lpdf * dist * dist / fabsf(cosTheta1);
                                float uptPdf = 0.f;
                                if (P.strategy != WL_STRATEGY_NEE)
                                {
                                    uptPdf = pdf_disney(mat, Ns, -lastRayDir, visRay);
                                    if (uptPdf < 0 || __builtin_isnan(uptPdf))
                                        uptPdf = 0.f;
                                    if (GSAMPLE && guided)
                                        uptPdf = P.galpha * guide_pdf(gm, visRay) +
                                                 (1.0f - P.galpha) * uptPdf;
                                }
                                V3 bv = eval_disney(mat, Ns, Ng, -lastRayDir, visRay, tex);
                                bv = clamp3(bv, 0.f, 1e30f);
                                contrib = throughput * bv * fabsf(cosTheta2) * ld3(lt.L) / neePdf *
                                          (1 - uptPdf / (uptPdf + neePdf));
                                shadow = true;
                                shadow_ray = wl_ray{ rayPos.x, rayPos.y, rayPos.z, P.epsilon,
                                                     visRay.x, visRay.y, visRay.z,
                                                     dist * (1 - P.epsilon) };
                            }
                        }
                        // ---- Russian roulette (PathTracing.cu:130-138)
                        if (depth > P.rr_depth)
                        {
                            if (rnd(seed) > P.rr_prob)
                            {
                                throughput = mk(0, 0, 0);
                                done = true;
                            }
                            else
                                throughput = throughput / P.rr_prob;
                        }
                    }
                }
            }
            if (rad_dirty || ident) // the first pass starts every path's radiance
                B.rad[path] = f4(radiance, 0.0f);
            if (gtrain)
            {
                // radiance reaching the path after its first guided vertex
                if (grec_done && rad_dirty)
                {
                    float4 gacc = B.gacc[path];
                    gacc.x += radiance.x - rad_in.x;
                    gacc.y += radiance.y - rad_in.y;
                    gacc.z += radiance.z - rad_in.z;
                    B.gacc[path] = gacc;
                }
                if (grec_now) // nothing gathered yet; w marks the record
                    B.gacc[path] = make_float4(0.f, 0.f, 0.f, 1.0f);
                else if (ident) // no record yet
                    B.gacc[path] = make_float4(0.f, 0.f, 0.f, 0.f);
                shadow_gacc = grec_done;
            }
            if (!done)
            {
                cont = true;
                B.thr[path] = f4(throughput,
                                 __uint_as_float(depth | (gvis << 16) |
                                                 (grec_done || grec_now ? kGrecBit : 0u)));
                B.ltt[path] = f4(lastTraceTerm, missw);
                // the path's lastPdf and seed ride in the ray's w words
                next_ray = wl_ray{ rayPos.x, rayPos.y, rayPos.z, lastPdf,
                                   rayDir.x, rayDir.y, rayDir.z, __uint_as_float(seed) };
            }
        }
        // ---- compaction (whole wave participates)
        uint32_t slot, sslot;
        block_append2(app, pass++, &B.counters[parity ^ 1], cont, &B.counters[2], shadow, slot,
                      sslot);
        if (cont)
        {
            qout[slot] = path;
            rout[slot] = next_ray;
            if (MULTI)
                (parity ? B.qreq0 : B.qreq1)[slot] = req;
        }
        if (shadow)
        {
            B.spath[sslot] = path;
            if (MULTI)
                B.sreq[sslot] = root;
            B.srays[sslot] = shadow_ray;
            B.scontrib[sslot] = f4(contrib, shadow_gacc ? 1.0f : 0.0f);
        }
    }
}

// ---- shadow: any hit; unoccluded rays add their NEE contribution ----------
template <int STACK, bool MULTI>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(STACK == 16 ? 6 : 5)))
void k_shadow(PtScene S, PtBuffers B, const PtReq *__restrict__ reqs)
{
    __shared__ uint32_t stack[STACK * kPtBlock];
    __shared__ uint4 top[8 * (pt_top_nodes<STACK, MULTI>() ? pt_top_nodes<STACK, MULTI>() : 1)];
    const uint32_t ntop = pt_load_top<STACK, MULTI>(top, S);
    const uint32_t count = B.counters[2];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        B.stats[1] += count;
    auto load = [&](uint32_t i, TravRay &r) { load_ray_nt(B.srays, i, r); };
    auto add = [&](uint32_t i, const TravResult &h) {
        if (h.prim != WL_INVALID)
            return;
        // unoccluded: add the NEE contribution (one shadow ray per path)
        const uint32_t path = B.spath[i];
        const float4 c = B.scontrib[i];
        float4 rr = B.rad[path];
        rr.x = rr.x + c.x;
        rr.y = rr.y + c.y;
        rr.z = rr.z + c.z;
        B.rad[path] = rr;
        if (c.w != 0.0f) // the path records guide training radiance
        {
            float4 ga = B.gacc[path];
            ga.x += c.x;
            ga.y += c.y;
            ga.z += c.z;
            B.gacc[path] = ga;
        }
    };
    if constexpr (MULTI) // S.nodes / S.tris: the arena of the requests' BVHs
        trav_persistent<true, kPtBlock, false, true, false, 8, 20>(
            S.nodes, S.tris, count, stack + threadIdx.x,
            [&](uint32_t i, TravRay &r, uint32_t &root) {
                root = B.sreq[i];
                load(i, r);
            },
            add);
    else
        trav_persistent<true, kPtBlock, false, false, false, 2, 20>(
            S.nodes, S.tris, count, stack + threadIdx.x, load, add,
            reinterpret_cast<const WlBvh8Node *>(top), ntop);
}

// ---- batched serving: raygen and film update of a multi-request megabatch ----
__global__ void k_batch_begin(PtMultiLaunch L)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *L.header = L.batches[*L.batch_index];
}

__global__ void k_batch_end(PtMultiLaunch L)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        const uint32_t b = *L.batch_index;
        L.times[b] = (unsigned long long)wall_clock64(); // the batch's film update is done
        *L.batch_index = b + 1;
    }
}

__global__ void k_clock_stamp(unsigned long long *dst)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *dst = (unsigned long long)wall_clock64();
}

__global__ __launch_bounds__(kPtBlock) void k_raygen_multi(PtMultiLaunch L, PtBuffers B)
{
    const PtBatchHeader h = *L.header;
    const uint32_t npaths = h.paths, nitems = h.items;
    const PtItem *items = L.items + h.item_offset;
    for (uint32_t p = blockIdx.x * kPtBlock + threadIdx.x; p < npaths;
         p += gridDim.x * kPtBlock)
    {
        const PtItem it = items[find_item(items, nitems, p, false)];
        const PtParams &P = L.reqs[it.req].P;
        const uint32_t local = p - it.path0;
        const uint32_t pixel = local % it.npix;
        const uint32_t s = local / it.npix;
        const uint32_t x = pixel % P.width, y = pixel / P.width;
        uint32_t seed = tea4(pixel, it.frame0 + s);
        // PinholeGenerateRay (Camera.h:28-35): two draws, u then v
        const float u = 2.f * ((float)(int)x + rnd(seed)) / (float)(int)P.width - 1.f;
        const float v = 2.f * ((float)(int)y + rnd(seed)) / (float)(int)P.height - 1.f;
        const V3 d = normalize(ld3(P.cam.lookat) + ld3(P.cam.right) * u + ld3(P.cam.up) * v);
        const V3 o = ld3(P.cam.pos);
        // the ray carries lastPdf 1 and the seed in its w words, as the
        // shade pass writes its rays; the first shade pass takes path = p
        // and starts the path's radiance and state
        B.rays0[p] = wl_ray{ o.x, o.y, o.z, 1.0f, d.x, d.y, d.z, __uint_as_float(seed) };
        B.qreq0[p] = it.req;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        B.counters[0] = npaths;
}

// Kahan update of each item's film, samples in ascending order (the order
// a single-request render adds them in).
__global__ __launch_bounds__(kPtBlock) void k_accumulate_multi(PtMultiLaunch L, PtBuffers B)
{
    const PtBatchHeader h = *L.header;
    const uint32_t nitems = h.items, npix = h.pixels;
    const PtItem *items = L.items + h.item_offset;
    for (uint32_t x = blockIdx.x * kPtBlock + threadIdx.x; x < npix; x += gridDim.x * kPtBlock)
    {
        const PtItem it = items[find_item(items, nitems, x, true)];
        const uint32_t px = x - it.px0;
        float *f = L.reqs[it.req].film + 6 * (size_t)px;
        float sum[3] = { f[0], f[1], f[2] }, c[3] = { f[3], f[4], f[5] };
        for (uint32_t s = 0; s < it.nsamples; ++s)
        {
            const float4 r = B.rad[it.path0 + (size_t)s * it.npix + px];
            const float v[3] = { r.x, r.y, r.z };
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                const float y = v[k] - c[k];
                const float t = sum[k] + y;
                c[k] = (t - sum[k]) - y;
                sum[k] = t;
            }
        }
        for (int k = 0; k < 3; ++k)
        {
            f[k] = sum[k];
            f[3 + k] = c[k];
        }
    }
}

// ---- guide training samples -----------------------------------------------------
// A recorded vertex's weight is F/p of its sampled direction: the radiance
// the path gathered through that direction divided by the throughput that
// reached the vertex (channel means).  Zero-weight records carry no gradient
// and are dropped.
// Each thread takes kCollectPer paths per pass (coalesced: path base + k
// kPtBlock + thread) and the workgroup claims the training slots of the whole
// pass with one atomic: one per 2048 paths, where one per wave serialised
// the kernel on the counter.
constexpr uint32_t kCollectPer = 8;

__global__ __launch_bounds__(kPtBlock) void k_guide_collect(PtParams P, PtBuffers B)
{
    __shared__ uint32_t s_cnt[kPtBlock / 64];
    __shared__ uint32_t s_base;
    const uint32_t npaths = P.npix * P.nsamples;
    const uint32_t wave = threadIdx.x / 64u, lane = __lane_id();
    constexpr uint32_t span = kPtBlock * kCollectPer;
    // block-uniform loop: every thread runs the same passes (barriers)
    for (uint32_t base = blockIdx.x * span; base < npaths; base += gridDim.x * span)
    {
        float w[kCollectPer];
        uint64_t m[kCollectPer];
        uint32_t wtotal = 0;
#pragma unroll
        for (uint32_t k = 0; k < kCollectPer; ++k)
        {
            const uint32_t p = base + k * kPtBlock + threadIdx.x;
            bool keep = false;
            w[k] = 0.f;
            if (p < npaths)
            {
                const float4 ga = B.gacc[p];
                if (ga.w != 0.0f)
                {
                    const float thr = B.grec[16 * (size_t)p + 13];
                    const float wk = (ga.x + ga.y + ga.z) * (1.0f / 3.0f) / fmaxf(thr, 1e-30f);
                    keep = wk > 0.0f; // false for NaN
                    w[k] = fminf(wk, B.gwmax);
                }
            }
            m[k] = __ballot(keep);
            wtotal += (uint32_t)__popcll(m[k]);
        }
        if (lane == 0)
            s_cnt[wave] = wtotal;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            uint32_t t = 0;
            for (uint32_t v = 0; v < kPtBlock / 64; ++v)
                t += s_cnt[v];
            s_base = t ? atomicAdd(B.gcount, t) : 0u;
        }
        __syncthreads();
        uint32_t o = s_base;
        for (uint32_t v = 0; v < wave; ++v)
            o += s_cnt[v];
#pragma unroll
        for (uint32_t k = 0; k < kCollectPer; ++k)
        {
            if ((m[k] >> lane) & 1u)
            {
                const uint32_t slot = o + lane_prefix(m[k]);
                const uint32_t p = base + k * kPtBlock + threadIdx.x;
                if (slot < B.gtrain_cap)
                {
                    const float4 *src = reinterpret_cast<const float4 *>(B.grec + 16 * (size_t)p);
                    float4 *dst = reinterpret_cast<float4 *>(B.gtrain + 16 * (size_t)slot);
                    dst[0] = src[0];
                    dst[1] = src[1];
                    dst[2] = src[2];
                    dst[3] = make_float4(w[k], 0.0f, 0.0f, 0.0f);
                }
            }
            o += (uint32_t)__popcll(m[k]);
        }
        __syncthreads(); // s_cnt and s_base are rewritten by the next pass
    }
}

// ---- film -----------------------------------------------------------------
// Kahan-compensated sum over the batch's samples of each pixel, in sample
// order: film = [sum RGB, compensation RGB] per pixel.
__global__ __launch_bounds__(kPtBlock) void k_accumulate(PtParams P, PtBuffers B, float *film)
{
    for (uint32_t px = blockIdx.x * kPtBlock + threadIdx.x; px < P.npix;
         px += gridDim.x * kPtBlock)
    {
        float *f = film + 6 * ((size_t)P.pix0 + px);
        float sum[3] = { f[0], f[1], f[2] }, c[3] = { f[3], f[4], f[5] };
        for (uint32_t s = 0; s < P.nsamples; ++s)
        {
            const float4 r = B.rad[(size_t)s * P.npix + px];
            const float v[3] = { r.x, r.y, r.z };
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                const float y = v[k] - c[k];
                const float t = sum[k] + y;
                c[k] = (t - sum[k]) - y;
                sum[k] = t;
            }
        }
        for (int k = 0; k < 3; ++k)
        {
            f[k] = sum[k];
            f[3 + k] = c[k];
        }
    }
}

__global__ void k_store_samples(PtParams P, PtBuffers B, float *rgb)
{
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < P.npix;
         px += gridDim.x * blockDim.x)
    {
        const float4 r = B.rad[px];
        const size_t o = 3 * ((size_t)P.pix0 + px);
        rgb[o] = r.x;
        rgb[o + 1] = r.y;
        rgb[o + 2] = r.z;
    }
}

// Reference progressive film (PathTracing.cu:140-158): running mean with
// float factors and double temporaries in an RGBA float buffer, clamp,
// ACES, x255, truncate to u8.
__global__ void k_frame_update(PtParams P, PtBuffers B, uint32_t frame, float *acc,
                               uint8_t *color)
{
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < P.npix;
         px += gridDim.x * blockDim.x)
    {
        const float4 r = B.rad[px];
        const double tf[4] = { 0.0 + (double)(r.x / 1), 0.0 + (double)(r.y / 1),
                               0.0 + (double)(r.z / 1), 1.0 };
        float *a = acc + 4 * ((size_t)P.pix0 + px);
        if (frame == 0)
        {
            for (int k = 0; k < 4; ++k)
                a[k] = (float)tf[k];
        }
        else
        {
            const double f0 = (double)((float)frame / ((float)frame + 1.0f));
            const double f1 = (double)(1.0f / ((float)frame + 1.0f));
            for (int k = 0; k < 4; ++k)
                a[k] = (float)((double)a[k] * f0 + tf[k] * f1);
        }
        uint8_t *c = color + 4 * ((size_t)P.pix0 + px);
        for (int k = 0; k < 3; ++k)
            c[k] = (uint8_t)(aces1(clampf(a[k], 0.f, 1e30f)) * 255.0f);
        c[3] = (uint8_t)(1.0f * 255.0f);
    }
}

__global__ void k_resolve(const float *film, uint32_t npix, uint32_t nsamples, float *rgb)
{
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix;
         px += gridDim.x * blockDim.x)
    {
        for (int k = 0; k < 3; ++k)
            rgb[3 * (size_t)px + k] =
                (film[6 * (size_t)px + k] - film[6 * (size_t)px + 3 + k]) / (float)nsamples;
    }
}

// ---- host side: one batch ---------------------------------------------------
hipError_t pt_run_batch(const PtParams &P, const PtScene &S, const PtBuffers &B,
                        int num_cus, hipStream_t st)
{
    const uint32_t npaths = P.npix * P.nsamples;
    const uint32_t blocks_all = (npaths + kPtBlock - 1) / kPtBlock;
    // a BVH at most 16 deep needs at most 15 stack entries per lane
    const bool shallow = S.bvh_depth <= 16;
    // persistent traversal grids: the workgroups that fit the GPU at once
    static int occ16 = 0, occ32 = 0, occf16 = 0, occf32 = 0;
    if (!occ16)
    {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occf16, k_extend_first<16>, kPtBlock,
                                                         0) != hipSuccess ||
            occf16 < 1)
            occf16 = 4;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occf32, k_extend_first<WL_BVH8_STACK_DEPTH>, kPtBlock, 0) != hipSuccess ||
            occf32 < 1)
            occf32 = 4;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ16, k_extend<16, false>, kPtBlock,
                                                         0) != hipSuccess ||
            occ16 < 1)
            occ16 = 4;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ32, k_extend<WL_BVH8_STACK_DEPTH, false>, kPtBlock, 0) != hipSuccess ||
            occ32 < 1)
            occ32 = 4;
    }
    const uint32_t g_trav =
        std::min<uint32_t>(blocks_all, (uint32_t)num_cus * (uint32_t)(shallow ? occ16 : occ32));
    const uint32_t g_first =
        std::min<uint32_t>(blocks_all, (uint32_t)num_cus * (uint32_t)(shallow ? occf16 : occf32));
    const uint32_t g_shade = std::min<uint32_t>(blocks_all, (uint32_t)num_cus * 8);
    const uint32_t g_sort = std::min<uint32_t>(blocks_all, (uint32_t)num_cus * 4);
    // guide inference: 64 queries per workgroup pass, up to 4 workgroups per CU
    const uint32_t g_guide =
        std::max<uint32_t>(1, std::min<uint32_t>((npaths + 63) / 64, (uint32_t)num_cus * 4));
    hipLaunchKernelGGL(k_batch_start, dim3(1), dim3(kPtBlock), 0, st, P, B);
    for (uint32_t b = 0; b < P.max_depth; ++b)
    {
        const int parity = (int)(b & 1u);
        if (b == 0 && shallow)
            hipLaunchKernelGGL((k_extend_first<16>), dim3(g_first), dim3(kPtBlock), 0, st, P, S,
                               B);
        else if (b == 0)
            hipLaunchKernelGGL((k_extend_first<WL_BVH8_STACK_DEPTH>), dim3(g_first),
                               dim3(kPtBlock), 0, st, P, S, B);
        else if (shallow)
            hipLaunchKernelGGL((k_extend<16, false>), dim3(g_trav), dim3(kPtBlock), 0, st, S, B,
                               parity, nullptr, P.epsilon);
        else
            hipLaunchKernelGGL((k_extend<WL_BVH8_STACK_DEPTH, false>), dim3(g_trav),
                               dim3(kPtBlock), 0, st, S, B, parity, nullptr, P.epsilon);
        if (P.sort_bins)
        {
            hipLaunchKernelGGL(k_sort_count, dim3(g_sort), dim3(kPtBlock), 0, st, P, S, B, parity);
            hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kPtBlock), 0, st, P, B);
            hipLaunchKernelGGL(k_sort_scatter, dim3(g_sort), dim3(kPtBlock), 0, st, P, B, parity);
        }
        // guided sampling on the first gbounces bounces (0 = all)
        const bool gsample = (P.gmode & 1u) && (P.gbounces == 0 || b < P.gbounces);
        const bool gtrain = (P.gmode & 2u) != 0;
        if (gsample)
        {
            if (b == 0)
                hipLaunchKernelGGL(k_guide_infer_pt<1>, dim3(g_guide), dim3(kPtBlock), 0, st, P,
                                   S, B, parity);
            else
                hipLaunchKernelGGL(k_guide_infer_pt<2>, dim3(g_guide), dim3(kPtBlock), 0, st, P,
                                   S, B, parity);
            if (gtrain)
                hipLaunchKernelGGL((k_shade<true, true, false>), dim3(g_shade), dim3(kPtBlock), 0,
                                   st, P, S, B, parity, nullptr, b == 0 ? 1 : 0);
            else
                hipLaunchKernelGGL((k_shade<true, false, false>), dim3(g_shade), dim3(kPtBlock),
                                   0, st, P, S, B, parity, nullptr, b == 0 ? 1 : 0);
        }
        else if (gtrain)
            hipLaunchKernelGGL((k_shade<false, true, false>), dim3(g_shade), dim3(kPtBlock), 0, st,
                               P, S, B, parity, nullptr, b == 0 ? 1 : 0);
        else if (B.trace)
            hipLaunchKernelGGL((k_shade<false, false, false, true>), dim3(g_shade),
                               dim3(kPtBlock), 0, st, P, S, B, parity, nullptr, b == 0 ? 1 : 0);
        else
            hipLaunchKernelGGL((k_shade<false, false, false>), dim3(g_shade), dim3(kPtBlock), 0,
                               st, P, S, B, parity, nullptr, b == 0 ? 1 : 0);
        if (shallow)
            hipLaunchKernelGGL((k_shadow<16, false>), dim3(g_trav), dim3(kPtBlock), 0, st, S, B,
                               nullptr);
        else
            hipLaunchKernelGGL((k_shadow<WL_BVH8_STACK_DEPTH, false>), dim3(g_trav),
                               dim3(kPtBlock), 0, st, S, B, nullptr);
    }
    return hipGetLastError();
}

static int g_multi_occ16 = 0, g_multi_occ32 = 0;

void pt_multi_init()
{
    int &occ16 = g_multi_occ16, &occ32 = g_multi_occ32;
    if (!occ16)
    {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ16, k_extend<16, true>, kPtBlock,
                                                         0) != hipSuccess ||
            occ16 < 1)
            occ16 = 4;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ32, k_extend<WL_BVH8_STACK_DEPTH, true>, kPtBlock, 0) != hipSuccess ||
            occ32 < 1)
            occ32 = 4;
    }
}

hipError_t pt_clock_stamp(unsigned long long *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_clock_stamp, dim3(1), dim3(64), 0, st, dst);
    return hipGetLastError();
}

hipError_t pt_run_multi(const PtMultiLaunch &L, const PtBuffers &B, int num_cus, hipStream_t st)
{
    pt_multi_init();
    const int occ16 = g_multi_occ16, occ32 = g_multi_occ32;
    const uint32_t blocks_all = (L.capacity + kPtBlock - 1) / kPtBlock;
    const uint32_t cus = (uint32_t)std::max(1, num_cus);
    const uint32_t g_gen = std::min<uint32_t>(blocks_all, cus * 8);
    const uint32_t g_trav = std::min<uint32_t>(blocks_all, cus * (uint32_t)(L.stack16 ? occ16 : occ32));
    const uint32_t g_shade =
        std::min<uint32_t>(blocks_all, cus * std::max<uint32_t>(1, L.shade_blocks_per_cu));
    const PtParams P0{};
    const PtScene S0{};
    PtScene SA{}; // traversal: the node/triangle arena of the requests' BVHs
    SA.nodes = L.nodes;
    SA.tris = L.tris;
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, L);
    hipLaunchKernelGGL(k_raygen_multi, dim3(g_gen), dim3(kPtBlock), 0, st, L, B);
    for (uint32_t b = 0; b < L.max_depth; ++b)
    {
        const int parity = (int)(b & 1u);
        if (L.stack16)
            hipLaunchKernelGGL((k_extend<16, true>), dim3(g_trav), dim3(kPtBlock), 0, st, SA, B,
                               parity, L.reqs, 0.0f);
        else
            hipLaunchKernelGGL((k_extend<WL_BVH8_STACK_DEPTH, true>), dim3(g_trav),
                               dim3(kPtBlock), 0, st, SA, B, parity, L.reqs, 0.0f);
        hipLaunchKernelGGL((k_shade<false, false, true>), dim3(g_shade), dim3(kPtBlock), 0, st,
                           P0, S0, B, parity, L.reqs, b == 0 ? 1 : 0);
        if (L.stack16)
            hipLaunchKernelGGL((k_shadow<16, true>), dim3(g_trav), dim3(kPtBlock), 0, st, SA, B,
                               L.reqs);
        else
            hipLaunchKernelGGL((k_shadow<WL_BVH8_STACK_DEPTH, true>), dim3(g_trav),
                               dim3(kPtBlock), 0, st, SA, B, L.reqs);
    }
    hipLaunchKernelGGL(k_accumulate_multi, dim3(g_gen), dim3(kPtBlock), 0, st, L, B);
    hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, st, L);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_arena_copy_nodes(WlBvh8Node *__restrict__ dst,
                                                          const WlBvh8Node *__restrict__ src,
                                                          uint32_t n, uint32_t node0,
                                                          uint32_t tri0)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    WlBvh8Node v = src[i];
    v.child_base += node0;
    v.tri_base += tri0;
    dst[node0 + i] = v;
}

hipError_t pt_arena_copy_nodes(WlBvh8Node *dst, const WlBvh8Node *src, uint32_t n,
                               uint32_t node0, uint32_t tri0, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_arena_copy_nodes, dim3((n + 255) / 256), dim3(256), 0, st, dst, src, n,
                       node0, tri0);
    return hipGetLastError();
}

hipError_t pt_guide_infer_queue(const PtParams &P, const PtScene &S, const PtBuffers &B,
                                uint32_t n, int num_cus, hipStream_t st)
{
    const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((n + 63) / 64, (uint32_t)num_cus * 4));
    hipLaunchKernelGGL(k_guide_infer_pt<0>, dim3(g), dim3(kPtBlock), 0, st, P, S, B, 0);
    return hipGetLastError();
}

hipError_t pt_guide_collect(const PtParams &P, const PtBuffers &B, int num_cus, hipStream_t st)
{
    const uint32_t npaths = P.npix * P.nsamples;
    const uint32_t span = kPtBlock * kCollectPer;
    const uint32_t g =
        std::max<uint32_t>(1, std::min<uint32_t>((npaths + span - 1) / span, (uint32_t)num_cus * 8));
    hipLaunchKernelGGL(k_guide_collect, dim3(g), dim3(kPtBlock), 0, st, P, B);
    return hipGetLastError();
}

hipError_t pt_accumulate(const PtParams &P, const PtBuffers &B, float *film, int num_cus,
                         hipStream_t st)
{
    const uint32_t g = std::min<uint32_t>((P.npix + kPtBlock - 1) / kPtBlock, (uint32_t)num_cus * 8);
    hipLaunchKernelGGL(k_accumulate, dim3(g), dim3(kPtBlock), 0, st, P, B, film);
    return hipGetLastError();
}

hipError_t pt_store_samples(const PtParams &P, const PtBuffers &B, float *rgb, hipStream_t st)
{
    const uint32_t g = std::min<uint32_t>((P.npix + 255) / 256, 4096);
    hipLaunchKernelGGL(k_store_samples, dim3(g), dim3(256), 0, st, P, B, rgb);
    return hipGetLastError();
}

hipError_t pt_frame_update(const PtParams &P, const PtBuffers &B, uint32_t frame, float *acc,
                           uint8_t *color, hipStream_t st)
{
    const uint32_t g = std::min<uint32_t>((P.npix + 255) / 256, 4096);
    hipLaunchKernelGGL(k_frame_update, dim3(g), dim3(256), 0, st, P, B, frame, acc, color);
    return hipGetLastError();
}

hipError_t pt_resolve(const float *film, uint32_t npix, uint32_t nsamples, float *rgb,
                      hipStream_t st)
{
    const uint32_t g = std::min<uint32_t>((npix + 255) / 256, 4096);
    hipLaunchKernelGGL(k_resolve, dim3(g), dim3(256), 0, st, film, npix, nsamples, rgb);
    return hipGetLastError();
}

} // namespace wl
