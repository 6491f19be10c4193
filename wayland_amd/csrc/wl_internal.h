// wl_internal.h — library-internal types behind the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/wayland_hip.h"
#include "bvh8.h"

struct wl_context_s
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::string cache_path;
    int num_cus = 0;
};

struct wl_accel_s
{
    wl_context ctx = nullptr;
    WlBvh8Node *d_nodes = nullptr;
    WlTriRecord *d_tris = nullptr;
    uint32_t num_nodes = 0;
    uint64_t num_tris = 0;
    std::vector<uint32_t> input_offsets;
    std::vector<uint32_t> input_tri_counts;
    wl_accel_info info{};
};

// Device-side scene for the path tracer (structure of arrays).
struct WlDeviceMesh
{
    uint32_t vertex_offset;   // into d_vertices / d_normals
    uint32_t tri_offset;      // global primitive id of local triangle 0
    uint32_t tri_count;
    uint32_t has_normal;
    uint32_t light;           // area light id or WL_INVALID_INDEX
    uint32_t material;        // == mesh index
    uint32_t pad[2];
};

struct WlDeviceLight
{
    uint32_t mesh;
    uint32_t tri_count;
    uint32_t two_sided;
    uint32_t has_normal;
    float L[3];
    uint32_t pad;
};

struct wl_scene_s
{
    wl_context ctx = nullptr;
    wl_accel accel = nullptr;
    uint32_t num_meshes = 0, num_lights = 0;
    float *d_vertices = nullptr;       // float3, mesh-local order
    float *d_normals = nullptr;        // float3
    uint32_t *d_indices = nullptr;     // uint3 per global triangle (mesh-local)
    uint32_t *d_prim_mesh = nullptr;   // global triangle -> mesh
    WlDeviceMesh *d_meshes = nullptr;
    wl_disney_material *d_materials = nullptr;
    WlDeviceLight *d_lights = nullptr;
    wl_pinhole_camera camera{};
    float bg[3] = { 0, 0, 0 };
    std::vector<WlDeviceMesh> h_meshes;
};

namespace wl
{
// thread-local error reporting
wl_status set_error(wl_status s, const std::string &msg);
wl_status hip_error(hipError_t e, const char *what);

// kernel launchers (traverse.hip)
hipError_t launch_trace_closest(const wl_accel_s *a, const wl_ray *rays,
                                uint32_t n, wl_hit *hits, uint32_t *counts,
                                hipStream_t s);
hipError_t launch_trace_any(const wl_accel_s *a, const wl_ray *rays,
                            uint32_t n, uint32_t *bits, hipStream_t s);
} // namespace wl

#define WL_HIP_CHECK(expr)                                                   \
    do                                                                       \
    {                                                                        \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            return wl::hip_error(e_, #expr);                                 \
    } while (0)
