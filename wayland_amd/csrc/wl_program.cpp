// wl_program.cpp — the OptiX program interface on top of the built-in gfx950
// integrators: program discovery by name (Module::IdentifyPrograms,
// Module-AutoConfig.cpp:76-178), pipelines (Pipeline.cpp:93-136), SBTs
// (ShaderBindingTable.h:42-109) and launches (Launcher.cpp:12-20).
#include "wl_internal.h"
#include "wl_program.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <vector>

using wl::set_error;

struct wl_pipeline_s
{
    wl_context ctx = nullptr;
    uint32_t builtin = 0;
    std::vector<wl_program_group> groups;
    uint32_t max_trace_depth = 0;
};

struct wl_sbt_s
{
    wl_pipeline pipeline = nullptr;
    std::vector<uint8_t> raygen, miss, hit;
    size_t miss_stride = 0, hit_stride = 0;
    uint32_t miss_count = 0, hit_count = 0;
    uint8_t *d_hit = nullptr;
    // PathTracing program: scene derived from the records on first launch
    wl_scene scene = nullptr;
    wl_accel scene_accel = nullptr;
    const void *scene_lights = nullptr;
    uint32_t scene_light_count = 0;
};

namespace
{
// Heap objects that are never destroyed: handles may be released from
// finalizers that run after the library's static destructors.
std::mutex &accel_mu()
{
    static std::mutex *m = new std::mutex;
    return *m;
}
std::set<wl_accel> &accels()
{
    static std::set<wl_accel> *s = new std::set<wl_accel>;
    return *s;
}

// texture objects (wl_texture_create): an image copy per handle
struct TextureObject
{
    std::vector<uint8_t> rgba;
    uint32_t width = 0, height = 0;
};
std::mutex &texture_mu()
{
    static std::mutex *m = new std::mutex;
    return *m;
}
std::set<TextureObject *> &textures()
{
    static std::set<TextureObject *> *s = new std::set<TextureObject *>;
    return *s;
}

constexpr uint32_t kHeaderMagic = 0x574C4847u; // "WLHG"

bool is_name_char(char c)
{
    return (c >= '0' && c <= '9') || (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') ||
           c == '_' || c == '-';
}
bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

// Matches `extern` WS+ `"C"` WS+ `__global__` WS+ `void` WS+ NAME WS* `()`
// starting at s[i]; NAME = "__" [_0-9A-Za-z-]+.  Returns the name or "".
std::string match_signature(const char *s, size_t n, size_t i, size_t *end)
{
    auto lit = [&](size_t &k, const char *w) {
        size_t L = std::strlen(w);
        if (k + L > n || std::memcmp(s + k, w, L) != 0)
            return false;
        k += L;
        return true;
    };
    auto ws1 = [&](size_t &k) {
        size_t k0 = k;
        while (k < n && is_space(s[k]))
            ++k;
        return k > k0;
    };
    size_t k = i;
    if (!lit(k, "extern") || !ws1(k) || !lit(k, "\"C\"") || !ws1(k) ||
        !lit(k, "__global__") || !ws1(k) || !lit(k, "void") || !ws1(k))
        return "";
    size_t n0 = k;
    if (!lit(k, "__"))
        return "";
    while (k < n && is_name_char(s[k]))
        ++k;
    if (k - n0 <= 2)
        return "";
    std::string name(s + n0, k - n0);
    while (k < n && is_space(s[k]))
        ++k;
    if (!lit(k, "()"))
        return "";
    *end = k;
    return name;
}

struct HitInfo
{
    std::string suffix;
    bool exists[3] = { false, false, false }; // intersection, anyhit, closesthit
};

void copy_name(char dst[64], const std::string &s)
{
    std::memset(dst, 0, 64);
    std::strncpy(dst, s.c_str(), 63);
}

bool has_program(const std::vector<wl_program_group> &g, uint32_t kind, int slot,
                 const char *name)
{
    for (const auto &x : g)
        if (x.kind == kind && std::strcmp(x.entry[slot], name) == 0)
            return true;
    return false;
}

} // namespace

namespace wl
{
void register_accel(wl_accel a)
{
    std::lock_guard<std::mutex> g(accel_mu());
    accels().insert(a);
}
void unregister_accel(wl_accel a)
{
    std::lock_guard<std::mutex> g(accel_mu());
    accels().erase(a);
}
wl_accel lookup_accel(uint64_t handle)
{
    std::lock_guard<std::mutex> g(accel_mu());
    auto it = accels().find(reinterpret_cast<wl_accel>(handle));
    return it == accels().end() ? nullptr : *it;
}
} // namespace wl

extern "C" {

wl_status wl_identify_programs(const char *src, size_t len, wl_program_group *groups,
                               uint32_t capacity, uint32_t *count)
{
    if (!src || !count || (capacity && !groups))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    static const char *prefixes[8] = { "raygen__", "intersection__", "anyhit__",
                                       "closesthit__", "miss__", "direct_callable__",
                                       "continuation_callable__", "exception__" };
    std::vector<wl_program_group> out;
    std::vector<HitInfo> hits;
    for (size_t i = 0; i < len; ++i)
    {
        if (src[i] != 'e')
            continue;
        size_t end = 0;
        std::string name = match_signature(src, len, i, &end);
        if (name.empty())
            continue;
        i = end - 1;
        std::string body = name.substr(2);
        int type = -1;
        for (int p = 0; p < 8; ++p)
            if (body.compare(0, std::strlen(prefixes[p]), prefixes[p]) == 0)
            {
                type = p;
                break;
            }
        if (type == 0 || type == 4)
        {
            wl_program_group g{};
            g.kind = type == 0 ? WL_PROGRAM_RAYGEN : WL_PROGRAM_MISS;
            copy_name(g.entry[0], name);
            out.push_back(g);
        }
        else if (type >= 1 && type <= 3)
        {
            // a hit program joins the first group whose suffix ends its name
            auto it = std::find_if(hits.begin(), hits.end(), [&](const HitInfo &h) {
                return name.size() >= h.suffix.size() &&
                       name.compare(name.size() - h.suffix.size(), h.suffix.size(),
                                    h.suffix) == 0;
            });
            if (it == hits.end())
            {
                HitInfo h;
                h.suffix = body.substr(std::strlen(prefixes[type]));
                hits.push_back(h);
                it = hits.end() - 1;
            }
            it->exists[type - 1] = true;
        }
        // other kinds (callables, exception) are skipped, as the reference does
    }
    for (const HitInfo &h : hits)
    {
        wl_program_group g{};
        g.kind = WL_PROGRAM_HITGROUP;
        for (int k = 0; k < 3; ++k)
            copy_name(g.entry[k], h.exists[k] ? "__" + std::string(prefixes[k + 1]) + h.suffix : "");
        out.push_back(g);
    }
    *count = (uint32_t)out.size();
    for (uint32_t k = 0; k < out.size() && k < capacity; ++k)
        groups[k] = out[k];
    if (out.size() > capacity)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "more program groups than capacity");
    return WL_SUCCESS;
}

wl_status wl_pipeline_create(wl_context ctx, const wl_program_group *groups,
                             uint32_t n, uint32_t max_trace_depth, wl_pipeline *out,
                             uint32_t *builtin)
{
    if (!ctx || !out || (n && !groups))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    std::vector<wl_program_group> g(groups, groups + n);
    uint32_t which = 0;
    if (has_program(g, WL_PROGRAM_RAYGEN, 0, "__raygen__RenderFrame") &&
        has_program(g, WL_PROGRAM_MISS, 0, "__miss__radiance") &&
        has_program(g, WL_PROGRAM_MISS, 0, "__miss__shadow") &&
        has_program(g, WL_PROGRAM_HITGROUP, 2, "__closesthit__radiance"))
        which = WL_BUILTIN_PATHTRACING;
    else if (has_program(g, WL_PROGRAM_RAYGEN, 0, "__raygen__RenderFrame") &&
             has_program(g, WL_PROGRAM_MISS, 0, "__miss__Empty") &&
             has_program(g, WL_PROGRAM_HITGROUP, 2, "__closesthit__PT"))
        which = WL_BUILTIN_EXAMPLE_PATHTRACER;
    if (!which)
        return set_error(WL_ERROR_UNKNOWN_PROGRAM,
                         "the program set matches no built-in integrator "
                         "(PathTracing or examples/PathTracer)");
    wl_pipeline p = new (std::nothrow) wl_pipeline_s;
    if (!p)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "pipeline allocation");
    p->ctx = ctx;
    wl::ctx_retain(ctx);
    p->builtin = which;
    p->groups = g;
    p->max_trace_depth = max_trace_depth;
    *out = p;
    if (builtin)
        *builtin = which;
    return WL_SUCCESS;
}

wl_status wl_pipeline_destroy(wl_pipeline p)
{
    if (!p)
        return set_error(WL_ERROR_INVALID_VALUE, "pipeline is NULL");
    wl_context ctx = p->ctx;
    delete p;
    wl::ctx_release(ctx);
    return WL_SUCCESS;
}

wl_status wl_sbt_pack_header(wl_pipeline p, uint32_t group_index, void *record)
{
    if (!p || !record)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (group_index >= p->groups.size())
        return set_error(WL_ERROR_INVALID_VALUE, "program group index out of range");
    uint8_t h[WL_SBT_RECORD_HEADER_SIZE] = {};
    std::memcpy(h, &group_index, 4);
    std::memcpy(h + 4, &kHeaderMagic, 4);
    std::memcpy(h + 8, &p->builtin, 4);
    std::memcpy(record, h, sizeof h);
    return WL_SUCCESS;
}

wl_status wl_sbt_create(wl_pipeline p, const void *raygen, size_t raygen_size,
                        const void *miss, uint32_t miss_count, size_t miss_stride,
                        const void *hit, uint32_t hit_count, size_t hit_stride,
                        wl_sbt *out)
{
    if (!p || !out || (raygen_size && !raygen) || (miss_count && !miss) ||
        (hit_count && !hit))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    if ((miss_count && (miss_stride < WL_SBT_RECORD_HEADER_SIZE ||
                        miss_stride % WL_SBT_RECORD_ALIGNMENT)) ||
        (hit_count && (hit_stride < WL_SBT_RECORD_HEADER_SIZE ||
                       hit_stride % WL_SBT_RECORD_ALIGNMENT)))
        return set_error(WL_ERROR_INVALID_VALUE,
                         "record strides must hold the 32-byte header and be 16-byte aligned");
    wl_sbt s = new (std::nothrow) wl_sbt_s;
    if (!s)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "sbt allocation");
    s->pipeline = p;
    s->raygen.assign((const uint8_t *)raygen, (const uint8_t *)raygen + raygen_size);
    s->miss.assign((const uint8_t *)miss, (const uint8_t *)miss + miss_count * miss_stride);
    s->hit.assign((const uint8_t *)hit, (const uint8_t *)hit + hit_count * hit_stride);
    s->miss_stride = miss_stride;
    s->hit_stride = hit_stride;
    s->miss_count = miss_count;
    s->hit_count = hit_count;
    hipError_t e = hipSetDevice(p->ctx->device);
    if (e == hipSuccess && !s->hit.empty())
        e = hipMalloc((void **)&s->d_hit, s->hit.size());
    if (e == hipSuccess && !s->hit.empty())
        e = hipMemcpy(s->d_hit, s->hit.data(), s->hit.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        hipFree(s->d_hit);
        delete s;
        return wl::hip_error(e, "sbt upload");
    }
    *out = s;
    return WL_SUCCESS;
}

wl_status wl_sbt_destroy(wl_sbt s)
{
    if (!s)
        return set_error(WL_ERROR_INVALID_VALUE, "sbt is NULL");
    if (s->scene)
        wl_scene_destroy(s->scene);
    hipFree(s->d_hit);
    delete s;
    return WL_SUCCESS;
}

wl_status wl_texture_create(wl_context ctx, const wl_texture_image *image, uint64_t *texture)
{
    if (!ctx || !image || !texture || !image->rgba)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (image->width == 0 || image->height == 0)
        return set_error(WL_ERROR_INVALID_VALUE, "empty texture image");
    if (image->width > 65536 || image->height > 65536)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "texture side above 65536");
    TextureObject *t = new (std::nothrow) TextureObject;
    if (!t)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "texture object");
    t->width = image->width;
    t->height = image->height;
    t->rgba.assign(image->rgba, image->rgba + 4 * (size_t)image->width * image->height);
    {
        std::lock_guard<std::mutex> g(texture_mu());
        textures().insert(t);
    }
    *texture = reinterpret_cast<uint64_t>(t);
    return WL_SUCCESS;
}

wl_status wl_texture_destroy(uint64_t texture)
{
    TextureObject *t = reinterpret_cast<TextureObject *>(texture);
    {
        std::lock_guard<std::mutex> g(texture_mu());
        if (!textures().erase(t))
            return set_error(WL_ERROR_INVALID_VALUE, "unknown texture object");
    }
    delete t;
    return WL_SUCCESS;
}

wl_status wl_accel_traversable(wl_accel a, uint64_t *handle)
{
    if (!a || !handle)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *handle = reinterpret_cast<uint64_t>(a);
    return WL_SUCCESS;
}

static wl_status launch_example(wl_pipeline p, hipStream_t st, const wl_example_params &prm,
                                wl_sbt sbt, uint32_t w, uint32_t h)
{
    wl_accel a = wl::lookup_accel(prm.traversable);
    if (!a)
        return set_error(WL_ERROR_INVALID_VALUE, "params.traversable is not a live accel");
    if (w != h)
        return set_error(WL_ERROR_NOT_SUPPORTED,
                         "the examples program indexes pixels as width * row + col "
                         "with row < width (PathTracer.cu:51-67); only square launches "
                         "stay inside the colour buffer");
    if (!prm.colorBuffer)
        return set_error(WL_ERROR_INVALID_VALUE, "params.colorBuffer is NULL");
    if (sbt->hit_count < a->input_offsets.size() ||
        sbt->hit_stride < WL_SBT_RECORD_HEADER_SIZE + sizeof(wl_example_hit_data))
        return set_error(WL_ERROR_INVALID_VALUE,
                         "one hit record with wl_example_hit_data per build input is required");
    for (size_t i = 0; i < a->input_offsets.size(); ++i)
    {
        wl_example_hit_data d;
        std::memcpy(&d, sbt->hit.data() + i * sbt->hit_stride + WL_SBT_RECORD_HEADER_SIZE,
                    sizeof d);
        if (a->input_tri_counts[i] && (!d.indices || !d.normals))
            return set_error(WL_ERROR_INVALID_VALUE, "hit record without indices/normals");
    }
    wl::ExampleLaunch L{};
    L.params = prm;
    L.nodes = a->d_nodes;
    L.tris = a->d_tris;
    L.hit_records = sbt->d_hit;
    L.hit_stride = sbt->hit_stride;
    L.width = w;
    L.height = h;
    WL_HIP_CHECK(wl::launch_example_pt(L, p->ctx->num_cus, st));
    return WL_SUCCESS;
}

// Builds (once per SBT + accel + light table) the path tracer's own scene
// from the reference's hit records, light table and miss record.
static wl_status derive_scene(wl_pipeline p, wl_sbt sbt, wl_accel a,
                              const wl_pathtracing_params &prm)
{
    if (sbt->scene && sbt->scene_accel == a && sbt->scene_lights == prm.areaLights &&
        sbt->scene_light_count == prm.areaLightCount)
        return WL_SUCCESS;
    if (sbt->scene)
    {
        wl_scene_destroy(sbt->scene);
        sbt->scene = nullptr;
    }
    const size_t M = a->input_offsets.size();
    const size_t rtc = 2; // ray types: radiance, shadow
    if (sbt->hit_count < M * rtc ||
        sbt->hit_stride < WL_SBT_RECORD_HEADER_SIZE + sizeof(wl_pathtracing_hit_data))
        return set_error(WL_ERROR_INVALID_VALUE,
                         "two hit records (radiance, shadow) per build input are required");
    std::vector<wl_mesh_shading> sh(M);
    std::vector<std::vector<float>> normals(M), texcoords(M);
    // textures: the objects the records name, each once (held under the
    // registry lock while copied)
    std::vector<std::vector<uint8_t>> images;
    std::vector<wl_texture_image> image_desc;
    std::vector<wl_mesh_texture> mt(M);
    std::vector<uint64_t> handles;
    for (size_t m = 0; m < M; ++m)
    {
        wl_pathtracing_hit_data d;
        std::memcpy(&d, sbt->hit.data() + (m * rtc) * sbt->hit_stride + WL_SBT_RECORD_HEADER_SIZE,
                    sizeof d);
        mt[m].texcoords = nullptr;
        mt[m].texture = WL_INVALID_INDEX;
        mt[m].pad_ = 0;
        if (d.hasTexture)
        {
            size_t k = 0;
            while (k < handles.size() && handles[k] != d.texture)
                ++k;
            if (k == handles.size())
            {
                std::lock_guard<std::mutex> g(texture_mu());
                TextureObject *t = reinterpret_cast<TextureObject *>(d.texture);
                if (!textures().count(t))
                    return set_error(WL_ERROR_INVALID_VALUE, "hit record names an unknown texture object");
                handles.push_back(d.texture);
                images.push_back(t->rgba);
                image_desc.push_back(wl_texture_image{ nullptr, t->width, t->height });
            }
            mt[m].texture = (uint32_t)k;
            if (d.texcoords && a->input_num_vertices[m])
            {
                texcoords[m].resize(2 * (size_t)a->input_num_vertices[m]);
                WL_HIP_CHECK(hipMemcpy(texcoords[m].data(), d.texcoords, texcoords[m].size() * 4,
                                       hipMemcpyDeviceToHost));
                mt[m].texcoords = texcoords[m].data();
            }
        }
        std::memset(&sh[m], 0, sizeof sh[m]);
        std::memcpy(&sh[m].material, d.disneyMat, sizeof d.disneyMat);
        sh[m].area_light = d.areaLightID;
        if (d.hasNormal && a->input_num_vertices[m])
        {
            if (!d.normals)
                return set_error(WL_ERROR_INVALID_VALUE, "hasNormal without a normals buffer");
            normals[m].resize(3 * (size_t)a->input_num_vertices[m]);
            WL_HIP_CHECK(hipMemcpy(normals[m].data(), d.normals, normals[m].size() * 4,
                                   hipMemcpyDeviceToHost));
            sh[m].normals = normals[m].data();
        }
    }
    std::vector<wl_device_area_light> dl(prm.areaLightCount);
    if (prm.areaLightCount)
    {
        if (!prm.areaLights)
            return set_error(WL_ERROR_INVALID_VALUE, "areaLights is NULL");
        WL_HIP_CHECK(hipMemcpy(dl.data(), prm.areaLights,
                               dl.size() * sizeof(wl_device_area_light), hipMemcpyDeviceToHost));
    }
    std::vector<wl_area_light> lights(prm.areaLightCount);
    for (uint32_t l = 0; l < prm.areaLightCount; ++l)
    {
        uint32_t mesh = WL_INVALID_INDEX;
        for (size_t m = 0; m < M; ++m)
            if (sh[m].area_light == l)
                mesh = (uint32_t)m;
        if (mesh == WL_INVALID_INDEX)
            return set_error(WL_ERROR_INVALID_VALUE, "an area light has no mesh");
        lights[l].mesh = mesh;
        std::memcpy(lights[l].L, dl[l].L, 12);
        lights[l].two_sided = dl[l].twoSided;
    }
    float bg[3] = { 0, 0, 0 };
    if (sbt->miss_count && sbt->miss_stride >= WL_SBT_RECORD_HEADER_SIZE + 12)
        std::memcpy(bg, sbt->miss.data() + WL_SBT_RECORD_HEADER_SIZE, 12);
    wl_scene s = nullptr;
    wl_status st = wl_scene_create(p->ctx, a, sh.data(), (uint32_t)M, lights.data(),
                                   (uint32_t)lights.size(), &prm.camera, bg, &s);
    if (st != WL_SUCCESS)
        return st;
    if (!images.empty())
    {
        for (size_t k = 0; k < images.size(); ++k)
            image_desc[k].rgba = images[k].data();
        st = wl_scene_set_textures(s, image_desc.data(), (uint32_t)image_desc.size(), mt.data(),
                                   (uint32_t)M);
        if (st != WL_SUCCESS)
        {
            wl_scene_destroy(s);
            return st;
        }
    }
    sbt->scene = s;
    sbt->scene_accel = a;
    sbt->scene_lights = prm.areaLights;
    sbt->scene_light_count = prm.areaLightCount;
    return WL_SUCCESS;
}

wl_status wl_launch(wl_pipeline p, void *stream, const void *params, size_t params_size,
                    wl_sbt sbt, uint32_t w, uint32_t h, uint32_t d)
{
    if (!p || !params || !sbt)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (sbt->pipeline != p)
        return set_error(WL_ERROR_INVALID_VALUE, "sbt belongs to another pipeline");
    if (d != 1)
        return set_error(WL_ERROR_NOT_SUPPORTED, "launch depth must be 1");
    if (w == 0 || h == 0)
        return WL_SUCCESS;
    WL_HIP_CHECK(hipSetDevice(p->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    if (p->builtin == WL_BUILTIN_EXAMPLE_PATHTRACER)
    {
        if (params_size != sizeof(wl_example_params))
            return set_error(WL_ERROR_INVALID_VALUE, "params size is not sizeof(LaunchParam)");
        wl_example_params prm;
        std::memcpy(&prm, params, sizeof prm);
        return launch_example(p, st, prm, sbt, w, h);
    }
    if (params_size != sizeof(wl_pathtracing_params))
        return set_error(WL_ERROR_INVALID_VALUE, "params size is not sizeof(LaunchParams)");
    wl_pathtracing_params prm;
    std::memcpy(&prm, params, sizeof prm);
    wl_accel a = wl::lookup_accel(prm.traversable);
    if (!a)
        return set_error(WL_ERROR_INVALID_VALUE, "params.traversable is not a live accel");
    if ((uint32_t)prm.fbSize[0] != w || (uint32_t)prm.fbSize[1] != h)
        return set_error(WL_ERROR_NOT_SUPPORTED, "fbSize must equal the launch size");
    wl_status s = derive_scene(p, sbt, a, prm);
    if (s != WL_SUCCESS)
        return s;
    wl_scene_set_camera(sbt->scene, &prm.camera);
    wl_pt_config cfg{};
    cfg.width = w;
    cfg.height = h;
    cfg.spp = 1;
    cfg.max_depth = 25;
    cfg.rr_depth = 5;
    cfg.rr_prob = 0.9f;
    cfg.epsilon = 1e-3f;
    cfg.strategy = WL_STRATEGY_MIS;
    return wl_pt_launch_frame(sbt->scene, &cfg, prm.frameID, prm.radianceBuffer,
                              prm.colorBuffer, stream);
}

wl_status wl_float_to_u8(const float *src, size_t n, uint8_t *dst, void *stream)
{
    if (n && (!src || !dst))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(wl::launch_float_to_u8(src, n, dst, (hipStream_t)stream));
    return WL_SUCCESS;
}

} // extern "C"
