// wl_internal.h — library-internal types behind the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <atomic>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/wayland_hip.h"
#include "bvh8.h"
#include "wl_guide.h"

// Reference counted: the caller's handle plus one reference per accel, scene
// and pipeline made from it, so children may be released in any order.
struct wl_context_s
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::string cache_path;
    int num_cus = 0;
    std::atomic<int> refs{ 1 };
    // Chunk counters of the raw persistent trace kernels: one block of eight
    // device counters per stream that launches them, with the counters'
    // values at the next launch on that stream (kept by the host: a launch
    // advances them by known amounts).  Launches on one stream use the block
    // in stream order; launches on different streams never share one.
    struct SchedBlock
    {
        uint32_t *dev = nullptr;
        uint32_t base[8] = {};
    };
    std::mutex sched_mu; // held from reading the bases to the launch
    std::unordered_map<hipStream_t, SchedBlock> sched;
};

namespace wl
{
void ctx_retain(wl_context ctx);
void ctx_release(wl_context ctx);
// The chunk counter block of stream s (allocated on first use); the caller
// holds ctx->sched_mu.
hipError_t sched_block(wl_context ctx, hipStream_t s, wl_context_s::SchedBlock **out);
} // namespace wl

struct wl_accel_s
{
    wl_context ctx = nullptr;
    WlBvh8Node *d_nodes = nullptr;
    WlTriRecord *d_tris = nullptr;
    uint32_t num_nodes = 0;
    uint64_t num_tris = 0;
    uint32_t num_tri_records = 0; // entries of d_tris (leaf triangle records)
    std::vector<uint32_t> input_offsets;
    std::vector<uint32_t> input_tri_counts;
    std::vector<uint32_t> input_num_vertices;
    std::vector<float> h_tris;        // 9 floats per global primitive
    std::vector<uint32_t> h_indices;  // 3 input-local vertex ids per primitive
    wl_accel_info info{};
};

// Path-tracer scene: device copies of the shading data plus the batch
// buffers, grown on demand and kept for later calls.
struct wl_scene_s
{
    wl_context ctx = nullptr;
    wl_accel accel = nullptr;
    uint32_t num_meshes = 0, num_lights = 0;
    float *d_prims = nullptr;                 // 12 floats per primitive
    float *d_normals = nullptr;               // 12 floats per primitive
    void *d_meshes = nullptr;                 // wl::PtMesh[num_meshes]
    wl_disney_material *d_materials = nullptr;
    void *d_lights = nullptr;                 // wl::PtLight[num_lights]
    wl_pinhole_camera camera{};
    float bg[3] = { 0, 0, 0 };
    // batch buffers
    uint32_t capacity = 0;
    void *d_batch = nullptr;                  // one allocation, carved up
    uint32_t *d_counters = nullptr;
    unsigned long long *d_stats = nullptr;
    // second batch lane: a large unguided wl_pt_render splits each batch into
    // two pixel halves that run concurrently, the second on stream2
    uint32_t capacity2 = 0;
    void *d_batch2 = nullptr;
    uint32_t *d_counters2 = nullptr;
    unsigned long long *d_stats2 = nullptr;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // learned guiding
    wl_guide guide = nullptr;
    uint32_t guide_mode = 0;
    float box_lo[3] = { 0, 0, 0 }, box_inv[3] = { 1, 1, 1 }; // scene bounds
    uint32_t gcapacity = 0;                   // paths covered by d_gbatch
    void *d_gbatch = nullptr;                 // gout, gacc, grec
    // textures (wl_scene_set_textures)
    std::vector<uint32_t> mesh_texture;       // image index per mesh or WL_INVALID_INDEX
    float *d_uvs = nullptr;                   // 8 floats per primitive (3 x uv + pad)
    void *d_textures = nullptr;               // wl::PtTexture[num_images]
    uint32_t *d_texels = nullptr;             // RGBA8 texels, images back to back
};

struct wl_guide_s
{
    wl_context ctx = nullptr;
    wl_guide_config cfg{};
    wl::GuideDev dev{};
    void *d_alloc = nullptr;
    uint64_t steps = 0;
    hipStream_t last_stream = nullptr;
    // The network the path tracer samples from: a copy of the packed forward
    // operands and the FP32 parameters (biases), published from dev.fwd /
    // dev.params.  A training render trains on `tstream` while its next
    // batch renders from this copy; `pending` means a training is enqueued
    // there (ending in ev_trained) whose result is not yet published.
    void *pub_fwd = nullptr;
    float *pub_params = nullptr;
    hipStream_t tstream = nullptr;
    hipEvent_t ev_pub = nullptr, ev_trained = nullptr;
    bool pending = false;
};

namespace wl
{
// thread-local error reporting
wl_status set_error(wl_status s, const std::string &msg);
wl_status hip_error(hipError_t e, const char *what);

// live accels, for resolving LaunchParams `traversable` handles
void register_accel(wl_accel a);
void unregister_accel(wl_accel a);
wl_accel lookup_accel(uint64_t handle);

// kernel launchers (traverse.hip)
hipError_t launch_trace_closest(const wl_accel_s *a, const wl_ray *rays,
                                uint32_t n, wl_hit *hits, uint32_t *counts,
                                hipStream_t s);
hipError_t launch_trace_any(const wl_accel_s *a, const wl_ray *rays,
                            uint32_t n, uint32_t *bits, hipStream_t s);
// path-tracer helpers shared by wl_pt_api.cpp and wl_serve.cpp
struct PtBuffers;
struct PtScene;
struct PtParams;
PtBuffers pt_carve(void *base, size_t cap, uint32_t *counters, unsigned long long *stats);
size_t pt_batch_bytes(size_t cap);
PtScene pt_scene_view(wl_scene s);
wl_status pt_check_config(wl_scene s, const wl_pt_config *c);
PtParams pt_params_for(wl_scene s, const wl_pt_config *c);
// looks up this device's no-allocation scratch array (wl_pt_noalloc.hip)
wl_status noalloc_setup(int device);
// Adam steps of a guide over samples[0, min(*n_dev, n_cap)) (wl_guide_api.cpp)
wl_status guide_train_device(wl_guide g, const float *samples, const uint32_t *n_dev,
                             uint32_t n_cap, uint32_t steps, hipStream_t s);
// copies the guide's current network to the path tracer's copy, on s
wl_status guide_publish(wl_guide g, hipStream_t s);
// orders s after the guide's pending training and publishes its result
wl_status guide_settle(wl_guide g, hipStream_t s);
} // namespace wl

#define WL_HIP_CHECK(expr)                                                   \
    do                                                                       \
    {                                                                        \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            return wl::hip_error(e_, #expr);                                 \
    } while (0)
