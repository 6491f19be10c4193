// wl_pt_api.cpp — C ABI of the path tracer: scene set-up (the device buffer
// layout of src/Frontend/Core/DeviceManager.cpp:77-181 + the hit records of
// PathTracingProgramManager.cpp:57-132, pre-gathered per primitive) and the
// render entry points.
#include "wl_internal.h"
#include "wl_pt.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>

using wl::set_error;

namespace
{

} // namespace

namespace wl
{
// Carves the batch buffers for `cap` paths out of one allocation.
PtBuffers pt_carve(void *base, size_t cap, uint32_t *counters, unsigned long long *stats)
{
    PtBuffers B{};
    char *p = (char *)base;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += (bytes + 255) & ~(size_t)255;
        return (void *)q;
    };
    B.thr = (float4 *)take(16 * cap);
    B.ltt = (float4 *)take(16 * cap);
    B.rad = (float4 *)take(16 * cap);
    B.rays0 = (wl_ray *)take(32 * cap);
    B.rays1 = (wl_ray *)take(32 * cap);
    B.qpath0 = (uint32_t *)take(4 * cap);
    B.qpath1 = (uint32_t *)take(4 * cap);
    B.hits = (wl_hit *)take(16 * cap);
    B.srays = (wl_ray *)take(32 * cap);
    B.spath = (uint32_t *)take(4 * cap);
    B.scontrib = (float4 *)take(16 * cap);
    B.sort_keys = (uint32_t *)take(4 * cap);
    B.sort_perm = (uint32_t *)take(4 * cap);
    B.sort_bins = (uint32_t *)take(4 * 2048);
    B.qreq0 = (uint32_t *)take(4 * cap);
    B.qreq1 = (uint32_t *)take(4 * cap);
    B.sreq = (uint32_t *)take(4 * cap);
    B.counters = counters;
    B.stats = stats;
    return B;
}

size_t pt_batch_bytes(size_t cap)
{
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return r(16 * cap) * 3 + r(32 * cap) * 3 + r(4 * cap) * 8 + r(16 * cap) * 2 + r(4 * 2048);
}

PtScene pt_scene_view(wl_scene s)
{
    PtScene S;
    S.nodes = s->accel->d_nodes;
    S.tris = s->accel->d_tris;
    S.prims = s->d_prims;
    S.normals = s->d_normals;
    S.meshes = (const PtMesh *)s->d_meshes;
    S.materials = s->d_materials;
    S.lights = (const PtLight *)s->d_lights;
    S.uvs = s->d_uvs;
    S.textures = (const PtTexture *)s->d_textures;
    S.texels = s->d_texels;
    S.bvh_depth = s->accel->info.depth;
    S.num_nodes = s->accel->num_nodes;
    return S;
}

wl_status pt_check_config(wl_scene s, const wl_pt_config *c)
{
    if (!s || !c)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (c->width == 0 || c->height == 0)
        return set_error(WL_ERROR_INVALID_VALUE, "empty image");
    if ((uint64_t)c->width * c->height > 0x7FFFFFFFull)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "image too large");
    if (c->max_depth == 0 || c->max_depth > 1024)
        return set_error(WL_ERROR_INVALID_VALUE, "max_depth must be in [1, 1024]");
    if (!(c->rr_prob > 0.0f && c->rr_prob <= 1.0f))
        return set_error(WL_ERROR_INVALID_VALUE, "rr_prob must be in (0, 1]");
    if (c->strategy > WL_STRATEGY_MIS)
        return set_error(WL_ERROR_INVALID_VALUE, "unknown strategy");
    return WL_SUCCESS;
}

// Kernel parameters of a render of `s` with `c` (no guide).
PtParams pt_params_for(wl_scene s, const wl_pt_config *c)
{
    PtParams P{};
    P.width = c->width;
    P.height = c->height;
    P.max_depth = c->max_depth;
    P.rr_depth = c->rr_depth;
    P.rr_prob = c->rr_prob;
    P.epsilon = c->epsilon;
    P.strategy = c->strategy;
    P.num_lights = s->num_lights;
    std::memcpy(P.bg, s->bg, sizeof P.bg);
    P.cam = s->camera;
    return P;
}
} // namespace wl

namespace
{

wl::PtBuffers carve(wl_scene s)
{
    wl::PtBuffers B = wl::pt_carve(s->d_batch, s->capacity, s->d_counters, s->d_stats);
    const size_t cap = s->capacity;
    if (s->guide && s->guide_mode && s->d_gbatch)
    {
        char *g = (char *)s->d_gbatch;
        B.gout = (float *)g;
        g += (192 * cap + 255) & ~(size_t)255;
        B.gacc = (float4 *)g;
        g += (16 * cap + 255) & ~(size_t)255;
        B.grec = (float *)g;
        wl::GuideDev &D = s->guide->dev;
        B.gtrain = D.train;
        B.gcount = D.count;
        B.gtrain_cap = D.train_cap;
        B.gwmax = s->guide->cfg.max_weight;
    }
    return B;
}

size_t batch_bytes(size_t cap) { return wl::pt_batch_bytes(cap); }

size_t guide_batch_bytes(size_t cap)
{
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return r(192 * cap) + r(16 * cap) + r(64 * cap);
}

wl_status ensure_capacity(wl_scene s, uint32_t paths)
{
    if (s->guide && s->guide_mode && s->gcapacity < std::max(paths, s->capacity))
    {
        const uint32_t gcap = std::max(paths, s->capacity);
        if (s->d_gbatch)
        {
            WL_HIP_CHECK(hipDeviceSynchronize());
            WL_HIP_CHECK(hipFree(s->d_gbatch));
            s->d_gbatch = nullptr;
            s->gcapacity = 0;
        }
        WL_HIP_CHECK(hipMalloc(&s->d_gbatch, guide_batch_bytes(gcap)));
        s->gcapacity = gcap;
    }
    if (paths <= s->capacity)
        return WL_SUCCESS;
    if (s->d_batch)
    {
        WL_HIP_CHECK(hipDeviceSynchronize());
        WL_HIP_CHECK(hipFree(s->d_batch));
        s->d_batch = nullptr;
        s->capacity = 0;
    }
    WL_HIP_CHECK(hipMalloc(&s->d_batch, batch_bytes(paths)));
    s->capacity = paths;
    return WL_SUCCESS;
}

// The second batch lane (buffers, counters, stream, fork/join events),
// created on first use and grown like the first.
wl_status ensure_lane2(wl_scene s, uint32_t paths)
{
    if (!s->stream2)
        WL_HIP_CHECK(hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking));
    if (!s->ev_fork)
        WL_HIP_CHECK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    if (!s->ev_join)
        WL_HIP_CHECK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    if (!s->d_counters2)
        WL_HIP_CHECK(hipMalloc((void **)&s->d_counters2, 64));
    if (!s->d_stats2)
        WL_HIP_CHECK(hipMalloc((void **)&s->d_stats2, 64));
    if (paths <= s->capacity2)
        return WL_SUCCESS;
    if (s->d_batch2)
    {
        WL_HIP_CHECK(hipDeviceSynchronize());
        WL_HIP_CHECK(hipFree(s->d_batch2));
        s->d_batch2 = nullptr;
        s->capacity2 = 0;
    }
    WL_HIP_CHECK(hipMalloc(&s->d_batch2, batch_bytes(paths)));
    s->capacity2 = paths;
    return WL_SUCCESS;
}

// Guided training batches: after the bounces, wait for the training of the
// previous batch (it reads the guide's sample buffer and writes the network)
// and publish its network to the path tracer's copy, move this batch's
// training records into the sample buffer, then train on them on the guide's
// own stream, where the training overlaps the next batch's kernels.  Batch k
// thus samples from the network trained through batch k - 2, which depends
// on none of its own paths.
wl_status after_batch(wl_scene s, const wl::PtParams &P, const wl::PtBuffers &B, hipStream_t hs)
{
    if (!(P.gmode & WL_GUIDE_TRAIN))
        return WL_SUCCESS;
    wl_guide g = s->guide;
    wl_status st = wl::guide_settle(g, hs);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(wl::pt_guide_collect(P, B, s->ctx->num_cus, hs));
    WL_HIP_CHECK(hipEventRecord(g->ev_pub, hs));
    WL_HIP_CHECK(hipStreamWaitEvent(g->tstream, g->ev_pub, 0));
    st = wl::guide_train_device(g, g->dev.train, g->dev.count, g->dev.train_cap,
                                g->cfg.steps_per_batch, g->tstream);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(wl::guide_roll_count(g->dev.count, g->dev.train_cap, g->tstream));
    WL_HIP_CHECK(hipEventRecord(g->ev_trained, g->tstream));
    g->pending = true;
    return WL_SUCCESS;
}

// A guided render that does not train samples from the guide's latest
// network: the render's stream waits for a pending training.
wl_status before_render(wl_scene s, hipStream_t hs)
{
    if (s->guide && s->guide_mode && !(s->guide_mode & WL_GUIDE_TRAIN))
        return wl::guide_settle(s->guide, hs);
    return WL_SUCCESS;
}

wl::PtScene scene_view(wl_scene s) { return wl::pt_scene_view(s); }

wl_status check_config(wl_scene s, const wl_pt_config *c) { return wl::pt_check_config(s, c); }

wl::PtParams params_for(wl_scene s, const wl_pt_config *c)
{
    wl::PtParams P = wl::pt_params_for(s, c);
    if (c->flags & WL_PT_SORT_BY_MATERIAL)
        P.sort_bins = std::min<uint32_t>(s->num_meshes + 1, 1024);
    if (s->guide && s->guide_mode)
    {
        P.gmode = s->guide_mode;
        P.galpha = s->guide->cfg.mix_alpha;
        P.gbounces = s->guide->cfg.guided_bounces;
        P.gverts = s->guide->cfg.guided_vertices;
        P.gfwd = s->guide->pub_fwd;
        P.gparams = s->guide->pub_params;
        std::memcpy(P.gbox_lo, s->box_lo, sizeof P.gbox_lo);
        std::memcpy(P.gbox_inv, s->box_inv, sizeof P.gbox_inv);
    }
    return P;
}

// Paths per batch when the caller gives none: 2^26 paths hold 16 GB of path
// state and queues, so the later, shorter bounces still fill the GPU.
constexpr uint32_t kDefaultBatchPaths = 1u << 26;
// batches of at least this many paths run as two concurrent halves
constexpr uint64_t kSplitMinPaths = 1u << 22;

} // namespace

extern "C" {

wl_status wl_scene_create(wl_context ctx, wl_accel accel,
                          const wl_mesh_shading *meshes, uint32_t num_meshes,
                          const wl_area_light *lights, uint32_t num_lights,
                          const wl_pinhole_camera *camera, const float bg[3],
                          wl_scene *out)
{
    if (!ctx || !accel || !out || (num_meshes && !meshes) || (num_lights && !lights))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    if (num_meshes != accel->input_offsets.size())
        return set_error(WL_ERROR_INVALID_VALUE,
                         "one wl_mesh_shading per build input is required");
    for (uint32_t l = 0; l < num_lights; ++l)
    {
        if (lights[l].mesh >= num_meshes)
            return set_error(WL_ERROR_INVALID_VALUE, "light mesh out of range");
        if (accel->input_tri_counts[lights[l].mesh] == 0)
            return set_error(WL_ERROR_INVALID_VALUE, "light mesh has no triangles");
    }
    for (uint32_t m = 0; m < num_meshes; ++m)
        if (meshes[m].area_light != WL_INVALID_INDEX && meshes[m].area_light >= num_lights)
            return set_error(WL_ERROR_INVALID_VALUE, "mesh area_light out of range");
    const size_t T = accel->num_tris;
    std::vector<float> prims(12 * std::max<size_t>(1, T), 0.0f);
    std::vector<float> normals;
    bool any_normals = false;
    for (uint32_t m = 0; m < num_meshes; ++m)
        any_normals |= meshes[m].normals != nullptr;
    if (any_normals)
        normals.assign(12 * std::max<size_t>(1, T), 0.0f);
    std::vector<wl::PtMesh> hm(num_meshes);
    for (uint32_t m = 0; m < num_meshes; ++m)
    {
        const uint32_t off = accel->input_offsets[m], cnt = accel->input_tri_counts[m];
        hm[m].has_normal = meshes[m].normals ? 1u : 0u;
        hm[m].light = meshes[m].area_light;
        hm[m].material = m;
        hm[m].tri_offset = off;
        hm[m].texture = WL_INVALID_INDEX;
        for (uint32_t t = 0; t < cnt; ++t)
        {
            const size_t g = (size_t)off + t;
            float *r = &prims[12 * g];
            std::memcpy(r, &accel->h_tris[9 * g], 36);
            std::memcpy(&r[9], &m, 4);
            std::memcpy(&r[10], &t, 4);
            if (meshes[m].normals)
            {
                float *n = &normals[12 * g];
                for (int k = 0; k < 3; ++k)
                {
                    uint32_t v = accel->h_indices[3 * g + k];
                    std::memcpy(n + 3 * k, meshes[m].normals + 3 * (size_t)v, 12);
                }
            }
        }
    }
    // scene bounds (position normalisation of the guide's queries)
    float lo[3] = { 0, 0, 0 }, hi[3] = { 1, 1, 1 };
    if (T > 0)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = 3.0e38f;
            hi[k] = -3.0e38f;
        }
        for (size_t v = 0; v < 3 * T; ++v)
            for (int k = 0; k < 3; ++k)
            {
                const float x = accel->h_tris[3 * v + k];
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
    }
    std::vector<wl::PtLight> hl(std::max<uint32_t>(1, num_lights));
    for (uint32_t l = 0; l < num_lights; ++l)
    {
        const uint32_t m = lights[l].mesh;
        hl[l].tri_offset = accel->input_offsets[m];
        hl[l].tri_count = accel->input_tri_counts[m];
        hl[l].two_sided = lights[l].two_sided ? 1u : 0u;
        hl[l].has_normal = meshes[m].normals ? 1u : 0u;
        std::memcpy(hl[l].L, lights[l].L, 12);
        hl[l].mesh = m;
    }
    std::vector<wl_disney_material> mats(std::max<uint32_t>(1, num_meshes));
    for (uint32_t m = 0; m < num_meshes; ++m)
        mats[m] = meshes[m].material;

    wl_scene s = new (std::nothrow) wl_scene_s;
    if (!s)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "scene allocation");
    s->ctx = ctx;
    wl::ctx_retain(ctx);
    s->accel = accel;
    s->num_meshes = num_meshes;
    s->num_lights = num_lights;
    s->mesh_texture.assign(num_meshes, WL_INVALID_INDEX);
    if (camera)
        s->camera = *camera;
    if (bg)
        std::memcpy(s->bg, bg, sizeof s->bg);
    for (int k = 0; k < 3; ++k)
    {
        s->box_lo[k] = lo[k];
        const float ext = hi[k] - lo[k];
        s->box_inv[k] = ext > 0.f ? 1.0f / ext : 1.0f;
    }
    hipError_t e = hipSetDevice(ctx->device);
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e != hipSuccess)
            return;
        e = hipMalloc(dst, bytes);
        if (e == hipSuccess)
            e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    up((void **)&s->d_prims, prims.data(), prims.size() * 4);
    if (any_normals)
        up((void **)&s->d_normals, normals.data(), normals.size() * 4);
    up(&s->d_meshes, hm.data(), std::max<size_t>(1, hm.size()) * sizeof(wl::PtMesh));
    up((void **)&s->d_materials, mats.data(), mats.size() * sizeof(wl_disney_material));
    up(&s->d_lights, hl.data(), hl.size() * sizeof(wl::PtLight));
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_counters, 64);
    if (e == hipSuccess)
        e = hipMalloc((void **)&s->d_stats, 64);
    if (e != hipSuccess)
    {
        wl_scene_destroy(s);
        return wl::hip_error(e, "scene upload");
    }
    *out = s;
    return WL_SUCCESS;
}

wl_status wl_scene_destroy(wl_scene s)
{
    if (!s)
        return set_error(WL_ERROR_INVALID_VALUE, "scene is NULL");
    hipSetDevice(s->ctx->device);
    hipDeviceSynchronize();
    hipFree(s->d_prims);
    hipFree(s->d_normals);
    hipFree(s->d_meshes);
    hipFree(s->d_materials);
    hipFree(s->d_lights);
    hipFree(s->d_batch);
    hipFree(s->d_counters);
    hipFree(s->d_stats);
    hipFree(s->d_batch2);
    hipFree(s->d_counters2);
    hipFree(s->d_stats2);
    if (s->ev_fork)
        hipEventDestroy(s->ev_fork);
    if (s->ev_join)
        hipEventDestroy(s->ev_join);
    if (s->stream2)
        hipStreamDestroy(s->stream2);
    hipFree(s->d_gbatch);
    hipFree(s->d_uvs);
    hipFree(s->d_textures);
    hipFree(s->d_texels);
    wl_context ctx = s->ctx;
    delete s;
    wl::ctx_release(ctx);
    return WL_SUCCESS;
}

wl_status wl_scene_set_textures(wl_scene s, const wl_texture_image *images, uint32_t num_images,
                                const wl_mesh_texture *meshes, uint32_t num_meshes)
{
    if (!s)
        return set_error(WL_ERROR_INVALID_VALUE, "scene is NULL");
    if (num_images && !images)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL images");
    if (num_images && (!meshes || num_meshes != s->num_meshes))
        return set_error(WL_ERROR_INVALID_VALUE, "one wl_mesh_texture per build input is required");
    const wl_accel a = s->accel;
    std::vector<wl::PtTexture> tab(std::max<uint32_t>(1, num_images));
    uint64_t texels = 0;
    for (uint32_t i = 0; i < num_images; ++i)
    {
        const wl_texture_image &im = images[i];
        if (!im.rgba || im.width == 0 || im.height == 0)
            return set_error(WL_ERROR_INVALID_VALUE, "empty texture image");
        if (im.width > 65536 || im.height > 65536)
            return set_error(WL_ERROR_LIMIT_EXCEEDED, "texture side above 65536");
        tab[i] = wl::PtTexture{ (uint32_t)texels, im.width, im.height, 0u };
        texels += (uint64_t)im.width * im.height;
        if (texels > 0xFFFFFFFFull)
            return set_error(WL_ERROR_LIMIT_EXCEEDED, "more than 2^32 texels");
    }
    std::vector<uint32_t> tex_of(s->num_meshes, WL_INVALID_INDEX);
    for (uint32_t m = 0; num_images && m < s->num_meshes; ++m)
    {
        if (meshes[m].texture != WL_INVALID_INDEX && meshes[m].texture >= num_images)
            return set_error(WL_ERROR_INVALID_VALUE, "mesh texture index out of range");
        tex_of[m] = meshes[m].texture;
    }
    // texcoords gathered per primitive, as normals are (zeros for a mesh without)
    std::vector<float> uvs;
    std::vector<uint32_t> pix;
    if (num_images)
    {
        uvs.assign(8 * std::max<uint64_t>(1, a->num_tris), 0.0f);
        for (uint32_t m = 0; m < s->num_meshes; ++m)
        {
            if (!meshes[m].texcoords)
                continue;
            const uint32_t off = a->input_offsets[m], cnt = a->input_tri_counts[m];
            for (uint32_t t = 0; t < cnt; ++t)
            {
                const size_t g = (size_t)off + t;
                for (int k = 0; k < 3; ++k)
                {
                    const uint32_t v = a->h_indices[3 * g + k];
                    uvs[8 * g + 2 * k] = meshes[m].texcoords[2 * (size_t)v];
                    uvs[8 * g + 2 * k + 1] = meshes[m].texcoords[2 * (size_t)v + 1];
                }
            }
        }
        pix.resize((size_t)texels);
        for (uint32_t i = 0; i < num_images; ++i)
            std::memcpy(pix.data() + tab[i].offset, images[i].rgba,
                        4 * (size_t)images[i].width * images[i].height);
    }
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    // running renders may still read the old textures
    WL_HIP_CHECK(hipDeviceSynchronize());
    std::vector<wl::PtMesh> hm(std::max<uint32_t>(1, s->num_meshes));
    WL_HIP_CHECK(hipMemcpy(hm.data(), s->d_meshes, hm.size() * sizeof(wl::PtMesh),
                           hipMemcpyDeviceToHost));
    hipFree(s->d_uvs);
    hipFree(s->d_textures);
    hipFree(s->d_texels);
    s->d_uvs = nullptr;
    s->d_textures = nullptr;
    s->d_texels = nullptr;
    hipError_t e = hipSuccess;
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e != hipSuccess)
            return;
        e = hipMalloc(dst, bytes);
        if (e == hipSuccess)
            e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    if (num_images)
    {
        up((void **)&s->d_uvs, uvs.data(), uvs.size() * 4);
        up(&s->d_textures, tab.data(), tab.size() * sizeof(wl::PtTexture));
        up((void **)&s->d_texels, pix.data(), pix.size() * 4);
    }
    // the texture field of the per-mesh records; after a failed upload the
    // scene is left without textures
    if (e != hipSuccess)
        std::fill(tex_of.begin(), tex_of.end(), WL_INVALID_INDEX);
    for (uint32_t m = 0; m < s->num_meshes; ++m)
        hm[m].texture = tex_of[m];
    const hipError_t e2 = hipMemcpy(s->d_meshes, hm.data(), hm.size() * sizeof(wl::PtMesh),
                                    hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = e2;
    if (e != hipSuccess)
    {
        hipFree(s->d_uvs);
        hipFree(s->d_textures);
        hipFree(s->d_texels);
        s->d_uvs = nullptr;
        s->d_textures = nullptr;
        s->d_texels = nullptr;
        s->mesh_texture.assign(s->num_meshes, WL_INVALID_INDEX);
        return wl::hip_error(e, "texture upload");
    }
    s->mesh_texture = tex_of;
    return WL_SUCCESS;
}

wl_status wl_scene_set_camera(wl_scene s, const wl_pinhole_camera *camera)
{
    if (!s || !camera)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    s->camera = *camera;
    return WL_SUCCESS;
}

// Splits each batch of wl_pt_render into two pixel halves, [p0, p0 + h)
// on the caller's stream and [p0 + h, p0 + np) on stream2.
static wl_status render_two_lanes(wl_scene s, const wl_pt_config *c, float *film,
                                  wl_pt_stats *stats, hipStream_t hs, uint32_t tile, uint32_t ns,
                                  std::chrono::steady_clock::time_point t0)
{
    const uint32_t npix_total = c->width * c->height;
    const uint32_t half = (tile + 1) / 2;
    wl_status st = ensure_capacity(s, half * ns);
    if (st == WL_SUCCESS)
        st = ensure_lane2(s, half * ns);
    if (st != WL_SUCCESS)
        return st;
    wl::PtBuffers B = carve(s);
    wl::PtBuffers B2 = wl::pt_carve(s->d_batch2, s->capacity2, s->d_counters2, s->d_stats2);
    wl::PtScene S = scene_view(s);
    hipStream_t h2 = s->stream2;
    // fork: the second lane starts after the caller's earlier work
    WL_HIP_CHECK(hipEventRecord(s->ev_fork, hs));
    WL_HIP_CHECK(hipStreamWaitEvent(h2, s->ev_fork, 0));
    if (stats)
    {
        WL_HIP_CHECK(hipMemsetAsync(s->d_stats, 0, 16, hs));
        WL_HIP_CHECK(hipMemsetAsync(s->d_stats2, 0, 16, h2));
    }
    for (uint32_t s0 = 0; s0 < c->spp; s0 += ns)
    {
        const uint32_t nss = std::min(ns, c->spp - s0);
        for (uint32_t p0 = 0; p0 < npix_total; p0 += tile)
        {
            const uint32_t np = std::min(tile, npix_total - p0);
            const uint32_t na = (np + 1) / 2, nb = np - na;
            wl::PtParams P = params_for(s, c);
            P.nsamples = nss;
            P.frame0 = c->frame_offset + s0;
            P.pix0 = p0;
            P.npix = na;
            WL_HIP_CHECK(wl::pt_run_batch(P, S, B, s->ctx->num_cus, hs));
            WL_HIP_CHECK(wl::pt_accumulate(P, B, film, s->ctx->num_cus, hs));
            if (nb)
            {
                P.pix0 = p0 + na;
                P.npix = nb;
                WL_HIP_CHECK(wl::pt_run_batch(P, S, B2, s->ctx->num_cus, h2));
                WL_HIP_CHECK(wl::pt_accumulate(P, B2, film, s->ctx->num_cus, h2));
            }
        }
    }
    // join: the caller's stream continues after both lanes
    WL_HIP_CHECK(hipEventRecord(s->ev_join, h2));
    WL_HIP_CHECK(hipStreamWaitEvent(hs, s->ev_join, 0));
    if (stats)
    {
        unsigned long long h[2] = { 0, 0 }, h2s[2] = { 0, 0 };
        WL_HIP_CHECK(hipMemcpyAsync(h, s->d_stats, 16, hipMemcpyDeviceToHost, hs));
        WL_HIP_CHECK(hipMemcpyAsync(h2s, s->d_stats2, 16, hipMemcpyDeviceToHost, hs));
        WL_HIP_CHECK(hipStreamSynchronize(hs));
        auto t1 = std::chrono::steady_clock::now();
        stats->paths = (uint64_t)npix_total * c->spp;
        stats->extension_rays = h[0] + h2s[0];
        stats->shadow_rays = h[1] + h2s[1];
        stats->max_bounce = c->max_depth;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return WL_SUCCESS;
}

wl_status wl_pt_render(wl_scene s, const wl_pt_config *c, float *film,
                       wl_pt_stats *stats, void *stream)
{
    wl_status st = check_config(s, c);
    if (st != WL_SUCCESS)
        return st;
    if (!film)
        return set_error(WL_ERROR_INVALID_VALUE, "film is NULL");
    if (c->spp == 0)
        return WL_SUCCESS;
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t hs = (hipStream_t)stream;
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t npix_total = c->width * c->height;
    const uint32_t max_paths = c->max_batch_paths ? c->max_batch_paths : kDefaultBatchPaths;
    // batch = a pixel range x a sample range
    uint32_t tile = std::min(npix_total, max_paths);
    uint32_t ns = std::max<uint32_t>(1, std::min<uint32_t>(c->spp, max_paths / tile));
    // Large unguided batches run as two concurrent pixel halves (the second
    // on the scene's second stream, with its own buffers): each half's
    // kernels fill the other's tails.  A pixel's samples stay in one half and
    // in order, so the film is the same as from one lane.
    const bool guided = s->guide && s->guide_mode;
    if (!guided && (uint64_t)tile * ns >= kSplitMinPaths && tile >= 2)
        return render_two_lanes(s, c, film, stats, hs, tile, ns, t0);
    st = ensure_capacity(s, tile * ns);
    if (st != WL_SUCCESS)
        return st;
    wl::PtBuffers B = carve(s);
    st = before_render(s, hs);
    if (st != WL_SUCCESS)
        return st;
    wl::PtScene S = scene_view(s);
    if (stats)
        WL_HIP_CHECK(hipMemsetAsync(s->d_stats, 0, 16, hs));
    for (uint32_t s0 = 0; s0 < c->spp; s0 += ns)
    {
        const uint32_t nss = std::min(ns, c->spp - s0);
        for (uint32_t p0 = 0; p0 < npix_total; p0 += tile)
        {
            wl::PtParams P = params_for(s, c);
            P.pix0 = p0;
            P.npix = std::min(tile, npix_total - p0);
            P.nsamples = nss;
            P.frame0 = c->frame_offset + s0;
            WL_HIP_CHECK(wl::pt_run_batch(P, S, B, s->ctx->num_cus, hs));
            WL_HIP_CHECK(wl::pt_accumulate(P, B, film, s->ctx->num_cus, hs));
            st = after_batch(s, P, B, hs);
            if (st != WL_SUCCESS)
                return st;
        }
    }
    if (stats)
    {
        unsigned long long h[2] = { 0, 0 };
        WL_HIP_CHECK(hipMemcpyAsync(h, s->d_stats, 16, hipMemcpyDeviceToHost, hs));
        WL_HIP_CHECK(hipStreamSynchronize(hs));
        auto t1 = std::chrono::steady_clock::now();
        stats->paths = (uint64_t)npix_total * c->spp;
        stats->extension_rays = h[0];
        stats->shadow_rays = h[1];
        stats->max_bounce = c->max_depth;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return WL_SUCCESS;
}

static wl_status one_frame(wl_scene s, const wl_pt_config *c, uint32_t frame,
                           wl::PtBuffers &B, hipStream_t hs, uint32_t p0, uint32_t np)
{
    wl::PtParams P = params_for(s, c);
    P.pix0 = p0;
    P.npix = np;
    P.nsamples = 1;
    P.frame0 = frame;
    WL_HIP_CHECK(wl::pt_run_batch(P, scene_view(s), B, s->ctx->num_cus, hs));
    return after_batch(s, P, B, hs);
}

wl_status wl_pt_render_samples(wl_scene s, const wl_pt_config *c, uint32_t frame,
                               float *rgb, void *stream)
{
    wl_status st = check_config(s, c);
    if (st != WL_SUCCESS)
        return st;
    if (!rgb)
        return set_error(WL_ERROR_INVALID_VALUE, "rgb is NULL");
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t hs = (hipStream_t)stream;
    const uint32_t npix_total = c->width * c->height;
    const uint32_t max_paths = c->max_batch_paths ? c->max_batch_paths : kDefaultBatchPaths;
    const uint32_t tile = std::min(npix_total, max_paths);
    st = ensure_capacity(s, tile);
    if (st != WL_SUCCESS)
        return st;
    wl::PtBuffers B = carve(s);
    st = before_render(s, hs);
    if (st != WL_SUCCESS)
        return st;
    for (uint32_t p0 = 0; p0 < npix_total; p0 += tile)
    {
        const uint32_t np = std::min(tile, npix_total - p0);
        st = one_frame(s, c, frame, B, hs, p0, np);
        if (st != WL_SUCCESS)
            return st;
        wl::PtParams P = params_for(s, c);
        P.pix0 = p0;
        P.npix = np;
        P.nsamples = 1;
        WL_HIP_CHECK(wl::pt_store_samples(P, B, rgb, hs));
    }
    return WL_SUCCESS;
}

wl_status wl_pt_trace_samples(wl_scene s, const wl_pt_config *c, uint32_t frame, float *rgb,
                              float *trace, uint32_t max_vertices, void *stream)
{
    wl_status st = check_config(s, c);
    if (st != WL_SUCCESS)
        return st;
    if (!rgb || !trace)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL buffer");
    if (max_vertices == 0 || max_vertices > 1024)
        return set_error(WL_ERROR_INVALID_VALUE, "max_vertices must be in [1, 1024]");
    if (s->guide && s->guide_mode)
        return set_error(WL_ERROR_NOT_SUPPORTED, "traced renders are unguided");
    const uint32_t npix = c->width * c->height;
    const uint32_t max_paths = c->max_batch_paths ? c->max_batch_paths : kDefaultBatchPaths;
    if (npix > max_paths)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "a traced render is one batch");
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t hs = (hipStream_t)stream;
    st = ensure_capacity(s, npix);
    if (st != WL_SUCCESS)
        return st;
    wl::PtBuffers B = carve(s);
    B.trace = trace;
    B.trace_max = max_vertices;
    st = one_frame(s, c, frame, B, hs, 0, npix);
    if (st != WL_SUCCESS)
        return st;
    wl::PtParams P = params_for(s, c);
    P.pix0 = 0;
    P.npix = npix;
    P.nsamples = 1;
    WL_HIP_CHECK(wl::pt_store_samples(P, B, rgb, hs));
    return WL_SUCCESS;
}

wl_status wl_pt_launch_frame(wl_scene s, const wl_pt_config *c, uint32_t frame,
                             float *radiance_rgba, uint8_t *color_rgba, void *stream)
{
    wl_status st = check_config(s, c);
    if (st != WL_SUCCESS)
        return st;
    if (!radiance_rgba || !color_rgba)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL buffer");
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t hs = (hipStream_t)stream;
    const uint32_t npix_total = c->width * c->height;
    const uint32_t max_paths = c->max_batch_paths ? c->max_batch_paths : kDefaultBatchPaths;
    const uint32_t tile = std::min(npix_total, max_paths);
    st = ensure_capacity(s, tile);
    if (st != WL_SUCCESS)
        return st;
    wl::PtBuffers B = carve(s);
    st = before_render(s, hs);
    if (st != WL_SUCCESS)
        return st;
    for (uint32_t p0 = 0; p0 < npix_total; p0 += tile)
    {
        const uint32_t np = std::min(tile, npix_total - p0);
        st = one_frame(s, c, frame, B, hs, p0, np);
        if (st != WL_SUCCESS)
            return st;
        wl::PtParams P = params_for(s, c);
        P.pix0 = p0;
        P.npix = np;
        P.nsamples = 1;
        WL_HIP_CHECK(wl::pt_frame_update(P, B, frame, radiance_rgba, color_rgba, hs));
    }
    return WL_SUCCESS;
}

wl_status wl_guide_infer_hits(wl_scene s, const wl_ray *rays, const wl_hit *hits, uint32_t n,
                              float *decoded, void *stream)
{
    if (!s || !rays || !hits || !decoded)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (!s->guide || !s->guide_mode)
        return set_error(WL_ERROR_INVALID_VALUE, "the scene has no guide attached");
    if (n == 0)
        return WL_SUCCESS;
    WL_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t hs = (hipStream_t)stream;
    wl_status st = wl::guide_settle(s->guide, hs); // the guide's latest network
    if (st != WL_SUCCESS)
        return st;
    wl_pt_config c{};
    c.width = 1;
    c.height = 1;
    c.max_depth = 1;
    c.rr_prob = 1.0f;
    wl::PtParams P = params_for(s, &c);
    wl::PtBuffers B{};
    B.rays0 = const_cast<wl_ray *>(rays);
    B.hits = const_cast<wl_hit *>(hits);
    B.counters = s->d_counters; // the kernel reads the queue length from here
    B.gout = decoded;
    WL_HIP_CHECK(hipMemcpyAsync(s->d_counters, &n, sizeof n, hipMemcpyHostToDevice, hs));
    WL_HIP_CHECK(wl::pt_guide_infer_queue(P, scene_view(s), B, n, s->ctx->num_cus, hs));
    WL_HIP_CHECK(hipStreamSynchronize(hs)); // n is read from host memory
    return WL_SUCCESS;
}

wl_status wl_scene_set_guide(wl_scene s, wl_guide g, uint32_t mode)
{
    if (!s)
        return set_error(WL_ERROR_INVALID_VALUE, "scene is NULL");
    if (mode & ~(uint32_t)(WL_GUIDE_SAMPLE | WL_GUIDE_TRAIN))
        return set_error(WL_ERROR_INVALID_VALUE, "unknown guide mode bits");
    if (g && g->ctx->device != s->ctx->device)
        return set_error(WL_ERROR_INVALID_VALUE, "guide and scene on different devices");
    if (!g)
        mode = 0;
    s->guide = mode ? g : nullptr;
    s->guide_mode = mode;
    return WL_SUCCESS;
}

wl_status wl_film_resolve(wl_context ctx, const float *film, uint32_t npix,
                          uint32_t nsamples, float *rgb, void *stream)
{
    if (!ctx || !film || !rgb || nsamples == 0)
        return set_error(WL_ERROR_INVALID_VALUE, "bad argument");
    WL_HIP_CHECK(hipSetDevice(ctx->device));
    WL_HIP_CHECK(wl::pt_resolve(film, npix, nsamples, rgb, (hipStream_t)stream));
    return WL_SUCCESS;
}

} // extern "C"
