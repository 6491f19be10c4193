// wl_pt.h — device data of the wavefront path tracer (pathtrace.hip) and the
// host functions that enqueue it (used by wl_pt_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wayland_hip.h"
#include "bvh8.h"

namespace wl
{

struct PtMesh
{
    uint32_t has_normal;
    uint32_t light;     // area light id or WL_INVALID_INDEX
    uint32_t material;  // index into materials
    uint32_t tri_offset;
    uint32_t texture;   // index into textures or WL_INVALID_INDEX
};

// One RGBA8 image in PtScene::texels (rows top to bottom)
struct PtTexture
{
    uint32_t offset;    // first texel
    uint32_t width, height;
    uint32_t pad;
};

struct PtLight
{
    uint32_t tri_offset; // global primitive id of the light mesh's triangle 0
    uint32_t tri_count;
    uint32_t two_sided;
    uint32_t has_normal;
    float L[3];
    uint32_t mesh;
};

// Scene as the kernels read it.
//   prims:   per global primitive, 3 x float4 = v0.xyz v1.x | v1.yz v2.xy |
//            v2.z, mesh, local index, 0   (48 bytes, primitive order)
//   normals: per global primitive, 3 x float4 = n0.xyz n1.x | n1.yz n2.xy |
//            n2.z, 0, 0, 0 (only read for meshes with normals)
struct PtScene
{
    const WlBvh8Node *nodes;
    const WlTriRecord *tris;
    const float *prims;
    const float *normals;
    const PtMesh *meshes;
    const wl_disney_material *materials;
    const PtLight *lights;
    const float *uvs;           // 8 floats per primitive: uv of v0, v1, v2, pad
    const PtTexture *textures;
    const uint32_t *texels;
    uint32_t bvh_depth;
    uint32_t num_nodes; // BVH8 nodes (the top ones are copied to LDS by traversal)
};

// One batch = pixels [pix0, pix0 + npix) x samples [frame0, frame0 + nsamples)
struct PtParams
{
    uint32_t width, height;
    uint32_t pix0, npix, nsamples, frame0;
    uint32_t max_depth, rr_depth;
    float rr_prob, epsilon;
    uint32_t strategy, num_lights;
    float bg[3];
    wl_pinhole_camera cam;
    // learned guiding (gmode 0 = off): bit 0 samples the guide at vertices
    // whose material is guidable, bit 1 also records training samples
    uint32_t sort_bins;        // material-sorted shade queue: bins (0 = off)
    uint32_t gmode;
    uint32_t gbounces;         // guided sampling on bounces [0, gbounces), 0 = all
    uint32_t gverts;           // guided vertices per path at most (0 = no limit)
    float galpha;              // probability of sampling the guide
    const void *gfwd;          // packed BF16 forward operands
    const float *gparams;      // FP32 master parameters (biases)
    float gbox_lo[3], gbox_inv[3]; // position normalisation to [0,1]^3
};

// Path state (structure of float4 arrays, one entry per path) and queues.
struct PtBuffers
{
    float4 *thr;      // throughput, w = depth (bits 0-15), guided vertices (16-31)
    float4 *ltt;      // lastTraceTerm, w = weight of a miss of the next ray
    float4 *rad;      // radiance
    wl_ray *rays0, *rays1;
    uint32_t *qpath0, *qpath1;
    wl_hit *hits;
    wl_ray *srays;
    uint32_t *spath;
    float4 *scontrib;
    uint32_t *sort_keys, *sort_perm; // material sort: key and order per queue slot
    uint32_t *sort_bins;             // [0, 1024) counts, [1024, 2048) offsets
    uint32_t *qreq0, *qreq1, *sreq; // multi-request batches: request of each
                                    // ray queue slot / BVH root (node arena
                                    // index) of each shadow queue slot
    uint32_t *counters;       // [0], [1] active queues, [2] shadow queue
    unsigned long long *stats; // [0] extension rays, [1] shadow rays
    // learned guiding (only with PtParams::gmode != 0)
    float *gout;      // decoded mixture, 48 floats per queue slot
    float4 *gacc;     // per path: radiance added after its first guided
                      // vertex; w = 1 once that vertex is recorded
    float *grec;      // per path: training record of that vertex (16 floats)
    float *gtrain;    // training samples (16 floats each)
    uint32_t *gcount; // number of training samples
    uint32_t gtrain_cap;
    float gwmax;      // training weights are clamped to this
    // per-vertex record of each path (wl_pt_trace_samples only, else NULL):
    // 8 floats per surface vertex that sampled a direction, at
    // trace[(path * trace_max + depth - 1) * 8]: depth, mesh, pdf, mean
    // f|cos|/pdf, mean throughput reaching the vertex, 0, 0, trans
    float *trace;
    uint32_t trace_max;
};

hipError_t pt_run_batch(const PtParams &P, const PtScene &S, const PtBuffers &B,
                        int num_cus, hipStream_t st);

// ---- batched serving: one megabatch mixes paths of many requests ----------
// A request's scene, parameters and film; the ray queues of a multi-request
// batch carry each entry's request id (qreq0/qreq1), the shadow queue the
// root of the request's BVH in the node arena (sreq), the one thing the
// shadow kernel needs.
struct PtReq
{
    PtParams P; // pix0 / npix / nsamples / frame0 unused; gmode 0
    PtScene S;  // S.nodes / S.tris: the request's own BVH (traversal uses the arena)
    float *film; // 6 floats per pixel (Kahan sum + compensation)
    uint32_t root; // the request's BVH root in the launch's node arena
    uint32_t pad_;
};

// Work item: all pixels of request `req` for samples [frame0, frame0 +
// nsamples); its paths are [path0, path0 + npix * nsamples) of the megabatch
// (sample-major like a single-request batch), its pixels [px0, px0 + npix)
// of the megabatch's pixel list.  A megabatch holds each request at most
// once, so the film update of a pixel never races.
struct PtItem
{
    uint32_t req, npix, frame0, nsamples;
    uint32_t path0, px0, pad_[2];
};

// Megabatch headers: PtBatchHeader per batch; the graph's first kernel
// copies the header of batch *batch_index into `header`, its last kernel
// advances *batch_index, so consecutive replays walk the batch list.
struct PtBatchHeader
{
    uint32_t paths, items, pixels, item_offset;
};

struct PtMultiLaunch
{
    const PtReq *reqs;
    const PtItem *items;          // all batches' items
    const PtBatchHeader *batches; // all batches' headers
    uint32_t *batch_index;        // device counter of the batch being run
    PtBatchHeader *header;        // the current batch's header
    unsigned long long *times;    // device wall clock at each batch's end
    // the requests' BVHs back to back, child and triangle indices rebased, so
    // multi-request traversal keeps the single-BVH kernel's registers and
    // starts each ray at its request's root
    const WlBvh8Node *nodes;
    const WlTriRecord *tris;
    uint32_t max_depth;  // bounces enqueued (maximum over the requests)
    uint32_t stack16;    // every BVH at most 16 levels deep
    uint32_t capacity;   // paths the buffers hold
    uint32_t shade_blocks_per_cu;
};
// Enqueues raygen, max_depth x (extend, shade, shadow) and the film update
// of one megabatch; every count comes from device memory, so the sequence
// can be captured once into a hipGraph and replayed.
hipError_t pt_run_multi(const PtMultiLaunch &L, const PtBuffers &B, int num_cus,
                        hipStream_t st);
// One-time occupancy queries of the multi-request kernels (outside capture).
void pt_multi_init();
// Writes the device wall clock (wall_clock64) to *dst once the stream gets there.
hipError_t pt_clock_stamp(unsigned long long *dst, hipStream_t st);
// Copies n BVH nodes into an arena at node offset node0, adding node0 to the
// child indices and tri0 to the triangle indices (the triangle records go to
// tri0 unchanged).
hipError_t pt_arena_copy_nodes(WlBvh8Node *dst, const WlBvh8Node *src, uint32_t n,
                               uint32_t node0, uint32_t tri0, hipStream_t st);
// The path tracer's guide inference (k_guide_infer_pt) over queue 0 of B:
// n = B.counters[0] entries of rays0 / hits, decoded mixtures of every hit
// to B.gout (misses are not written).
hipError_t pt_guide_infer_queue(const PtParams &P, const PtScene &S, const PtBuffers &B,
                                uint32_t n, int num_cus, hipStream_t st);
// Moves the finished batch's training records with a positive weight into
// the training buffer.
hipError_t pt_guide_collect(const PtParams &P, const PtBuffers &B, int num_cus,
                            hipStream_t st);
hipError_t pt_accumulate(const PtParams &P, const PtBuffers &B, float *film,
                         int num_cus, hipStream_t st);
hipError_t pt_store_samples(const PtParams &P, const PtBuffers &B, float *rgb,
                            hipStream_t st);
hipError_t pt_frame_update(const PtParams &P, const PtBuffers &B, uint32_t frame,
                           float *acc, uint8_t *color, hipStream_t st);
hipError_t pt_resolve(const float *film, uint32_t npix, uint32_t nsamples,
                      float *rgb, hipStream_t st);

} // namespace wl
