/*
 * wayland_hip.h — C ABI of the MI355X (gfx950) rendering backend for
 * EasyRender (Extra-Creativity/Wayland).
 *
 * This is the boundary that replaces the OptiX backend
 * (src/Backend/Optix/Core) under the EasyRender path tracer.  Every entry
 * point names the reference interface it stands in for.  Plain C: opaque
 * handles, plain pointers and sizes; a stream is passed as `void *` holding a
 * hipStream_t (NULL = HIP's null stream, as everywhere in HIP; the context's
 * own stream is returned by wl_context_get_stream).
 *
 * Conventions
 *  - Every call returns a wl_status (0 = WL_SUCCESS) and never throws or
 *    aborts across the boundary.  wl_last_error_message() gives detail for
 *    the calling thread's last failure.  (Reference: ErrorCheck.h:9-58 only
 *    reports errors in debug builds; here they are always reported.)
 *  - Handles are owned by the library and freed by the matching *_destroy.
 *    A handle is not internally synchronised: calls on one handle from
 *    several threads need external locking.
 *  - Ray, hit and film buffers are caller-owned DEVICE memory.  Geometry and
 *    scene description inputs are HOST memory and are copied, so callers may
 *    free them after the call (as TriangleDataBuffer copies its input,
 *    BuildInput.cpp:118-148).
 *  - Work is enqueued on the given stream; the call returns without
 *    synchronising unless stated.
 */
#ifndef WAYLAND_HIP_H
#define WAYLAND_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum wl_status
{
    WL_SUCCESS = 0,
    WL_ERROR_INVALID_VALUE = 1,
    WL_ERROR_OUT_OF_MEMORY = 2,
    WL_ERROR_DEVICE = 3,          /* a HIP runtime call failed            */
    WL_ERROR_NOT_SUPPORTED = 4,
    WL_ERROR_LIMIT_EXCEEDED = 5,  /* e.g. BVH deeper than the stack       */
    WL_ERROR_NO_DEVICE = 6,
    WL_ERROR_UNKNOWN_PROGRAM = 7,
    WL_ERROR_QUEUE_OVERFLOW = 8
} wl_status;

#define WL_INVALID_INDEX 0xFFFFFFFFu

typedef struct wl_context_s *wl_context;
typedef struct wl_accel_s *wl_accel;
typedef struct wl_scene_s *wl_scene;
typedef struct wl_guide_s *wl_guide;
typedef struct wl_server_s *wl_server;

/* ---- errors (reference: HostUtils/ErrorCheck.h:9-58, ErrorCheck.cpp:14-22) */
const char *wl_status_string(wl_status status);
const char *wl_last_error_message(void);

/* ---- contexts (reference: Core/ContextManager.cpp:43-56 InitAndGetDeviceCount,
 *      :99-145 CreateContext, ContextManager.h:84-123 LocalContextSetter) ---- */
wl_status wl_device_count(int *count);
/* One context = one device + one stream (default priority). */
wl_status wl_context_create(int device, wl_context *out);
wl_status wl_context_destroy(wl_context ctx);
wl_status wl_context_get_stream(wl_context ctx, void **stream);
wl_status wl_context_synchronize(wl_context ctx);
/* ContextManager::SetCachePath (ContextManager.h:58): kernels are compiled
 * ahead of time for gfx950, so the path is only recorded. */
wl_status wl_set_cache_path(wl_context ctx, const char *path);

/* ---- acceleration structures --------------------------------------------
 * wl_tri_input mirrors one TriangleBuildInputArray::AddBuildInput call
 * (BuildInput.cpp:193-207): FLOAT3 vertices, UINT3 indices, one geometry
 * flag and `num_sbt_records` SBT records for this input.                   */
typedef struct wl_tri_input
{
    const float *vertices;      /* host, 3 floats per vertex (12-byte stride) */
    uint32_t num_vertices;
    const uint32_t *indices;    /* host, 3 per triangle, input-local          */
    uint32_t num_triangles;
    uint32_t geometry_flags;    /* GeometryFlags (BuildInput.h:50-56)         */
    uint32_t num_sbt_records;   /* 1 (one record per build input)            */
} wl_tri_input;

typedef struct wl_accel_info
{
    uint64_t num_triangles;
    uint32_t num_inputs;
    uint32_t num_nodes;         /* BVH8 node slots allocated                  */
    uint32_t inner_nodes;       /* nodes in use                               */
    uint32_t leaves;
    uint32_t depth;             /* BVH8 levels (root = 1)                     */
    uint32_t pad_;
    uint64_t device_bytes;      /* nodes + triangle records on the device     */
    double sah_cost;
    double build_seconds;
} wl_accel_info;

/* StaticAccelStructure(buildInputs) (AccelStructure.cpp:133-162) → SAH BVH8.
 * Global primitive id of triangle j of input i = (sum of num_triangles of
 * inputs < i) + j. */
wl_status wl_accel_build_triangles(wl_context ctx, const wl_tri_input *inputs,
                                   uint32_t num_inputs, uint32_t build_flags,
                                   wl_accel *out);
wl_status wl_accel_destroy(wl_accel accel);
wl_status wl_accel_get_info(wl_accel accel, wl_accel_info *info);
/* TriangleBuildInputArray::GetDepth() (BuildInput.h:192) — always 1. */
wl_status wl_accel_depth(wl_accel accel, uint32_t *depth);
/* Triangle offset of each build input (num_inputs entries). */
wl_status wl_accel_input_offsets(wl_accel accel, uint32_t *offsets);

/* ---- raw traversal (optixTrace call sites PathTracing.cu:72-75, :92-102)
 * wl_ray is what optixTrace receives: origin, tmin, direction, tmax.      */
typedef struct wl_ray
{
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} wl_ray;

/* Closest hit: t, global primitive id, barycentrics (b1, b2) weighting v1, v2
 * as optixGetTriangleBarycentrics() does.  Miss: prim = WL_INVALID_INDEX,
 * t = +inf.  Ties in t go to the smaller primitive id.  Watertight FP32 test
 * (Woop, Benthin, Wald 2013); hits are accepted for tmin <= t <= tmax. */
typedef struct wl_hit
{
    float t;
    uint32_t prim;
    float b1, b2;
} wl_hit;

/* wl_trace_closest and wl_trace_any hand the rays out to the GPU's waves at
 * run time from a small block of counters kept per stream (the context's):
 * the first trace on a stream allocates that block (256 bytes, freed with the
 * context) and synchronises the device, so a caller that captures traces
 * into a graph makes one trace on that stream before capturing.  Traces on
 * one stream are ordered; traces on different streams share nothing. */
wl_status wl_trace_closest(wl_accel accel, const wl_ray *d_rays, uint32_t n,
                           wl_hit *d_hits, void *stream);
/* Any hit with OPTIX_RAY_FLAG_TERMINATE_ON_FIRST_HIT semantics: bit (i % 32)
 * of d_occluded[i / 32] is set when ray i hits anything in [tmin, tmax].
 * d_occluded holds ceil(n / 32) words. */
wl_status wl_trace_any(wl_accel accel, const wl_ray *d_rays, uint32_t n,
                       uint32_t *d_occluded, void *stream);
/* Instrumented closest hit: also writes, per ray, the number of BVH8 nodes
 * fetched and triangles tested (d_counts[2*i], d_counts[2*i+1]). */
wl_status wl_trace_closest_counted(wl_accel accel, const wl_ray *d_rays,
                                   uint32_t n, wl_hit *d_hits,
                                   uint32_t *d_counts, void *stream);

/* ---- scene for the path tracer ---------------------------------------- */
/* Device/Material.h:7-21 DisneyMaterial (64 bytes, last float padding). */
typedef struct wl_disney_material
{
    float color[3];
    float metallic, roughness, specular, specularTint, eta, trans, subsurface,
        anisotropic, sheen, sheenTint, clearcoat, clearcoatgloss;
    float pad_;
} wl_disney_material;

/* Device/Camera.h:9-15 PinholeCamFrame; `right` already aspect-scaled as
 * PathTracingProgramManager.cpp:22-28 does. */
typedef struct wl_pinhole_camera
{
    float pos[3];
    float lookat[3];
    float up[3];
    float right[3];
} wl_pinhole_camera;

/* Per build input shading data: the HitData of PathTracingLaunchParams.h:64-78
 * without device pointers (the library owns the copies). */
typedef struct wl_mesh_shading
{
    const float *normals;       /* host, 3 floats per vertex, or NULL          */
    wl_disney_material material;
    uint32_t area_light;        /* index into lights, or WL_INVALID_INDEX      */
    uint32_t pad_[3];
} wl_mesh_shading;

/* Device/Light.h:9-17 DeviceAreaLight: emitter = the triangles of `mesh`. */
typedef struct wl_area_light
{
    uint32_t mesh;
    float L[3];
    uint32_t two_sided;
} wl_area_light;

wl_status wl_scene_create(wl_context ctx, wl_accel accel,
                          const wl_mesh_shading *meshes, uint32_t num_meshes,
                          const wl_area_light *lights, uint32_t num_lights,
                          const wl_pinhole_camera *camera, const float bg[3],
                          wl_scene *out);
wl_status wl_scene_destroy(wl_scene scene);
wl_status wl_scene_set_camera(wl_scene scene, const wl_pinhole_camera *camera);

/* An RGBA8 image, rows top to bottom: what stbi_load(path, ..., 4) returns
 * in Texture::TexLoadImage (src/Frontend/Core/Texture.cpp:13-16). */
typedef struct wl_texture_image
{
    const uint8_t *rgba;        /* host, 4 * width * height bytes              */
    uint32_t width, height;
} wl_texture_image;

/* Texture binding of one build input (the hasTexture / texture / texcoords
 * fields GenerateSBT fills, PathTracingProgramManager.cpp:97-119). */
typedef struct wl_mesh_texture
{
    const float *texcoords;     /* host, 2 floats per vertex, or NULL: all zero,
                                   as the reference's zero-filled texcoord
                                   buffer gives a mesh without uv            */
    uint32_t texture;           /* image index, or WL_INVALID_INDEX            */
    uint32_t pad_;
} wl_mesh_texture;

/* Textures of a scene: UploadTextures (DeviceManager.cpp:183-229) plus the
 * per-mesh binding.  At a hit on a textured mesh the material's base colour
 * is multiplied by the filtered texel (PathTracing.cu:261-268, Evaluate.h:21,
 * 86), for the bounce and for the NEE at that vertex (PathTracing.cu:120-122,
 * 280).  Sampling follows the reference's texture object: normalized
 * coordinates, wrap addressing, bilinear filtering with 8-bit fractional
 * weights, texels read as u8 / 255.  One entry per build input;
 * num_images == 0 removes the scene's textures. */
wl_status wl_scene_set_textures(wl_scene scene, const wl_texture_image *images,
                                uint32_t num_images, const wl_mesh_texture *meshes,
                                uint32_t num_meshes);

/* ---- wavefront path tracer (PathTracing.cu:40-291) --------------------- */
enum
{
    WL_STRATEGY_UPT = 0,
    WL_STRATEGY_NEE = 1,
    WL_STRATEGY_MIS = 2
};
enum
{
    WL_PT_SORT_BY_MATERIAL = 1u << 0 /* shade queue binned by material id */
};

typedef struct wl_pt_config
{
    uint32_t width, height;
    uint32_t spp;          /* samples per pixel in this call                  */
    uint32_t frame_offset; /* sample s uses the reference's frame id offset+s */
    uint32_t max_depth;    /* PathTracing.cu:235 (25)                          */
    uint32_t rr_depth;     /* PathTracing.cu:130 RR when depth > rr_depth (5) */
    float rr_prob;         /* PathTracing.cu:68 (0.9)                          */
    float epsilon;         /* PathTracing.cu:19 (1e-3)                         */
    uint32_t strategy;     /* PathTracing.cu:21-29 (MIS)                       */
    uint32_t flags;
    uint32_t max_batch_paths; /* 0 = library default (2^26)                   */
    uint32_t pad_;
} wl_pt_config;

typedef struct wl_pt_stats
{
    uint64_t paths;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint32_t max_bounce;
    uint32_t pad_;
    double seconds;        /* host wall time of the call (synchronising)     */
} wl_pt_stats;

/* Adds cfg->spp samples per pixel into a compensated FP32 film:
 * d_film = 6 floats per pixel (sum RGB, Kahan compensation RGB), caller
 * zeroes it before the first call.  Mean = sum / total samples.  If stats is
 * non-NULL the call synchronises and fills it.  Without a guide, batches of
 * at least 2^22 paths run as two pixel halves at once, the second on a stream
 * of the scene; `stream` then waits on an event after both, so work enqueued
 * on it later is ordered after the whole render. */
wl_status wl_pt_render(wl_scene scene, const wl_pt_config *cfg, float *d_film,
                       wl_pt_stats *stats, void *stream);
/* The per-sample radiance of frame `frame` (one sample per pixel, TEA seed
 * (pixel, frame)): d_rgb = W*H*3 floats.  For parity with the reference. */
wl_status wl_pt_render_samples(wl_scene scene, const wl_pt_config *cfg,
                               uint32_t frame, float *d_rgb, void *stream);
/* wl_pt_render_samples of one batch (W*H <= max_batch_paths, no guide) with
 * a per-vertex record of every path, for parity diagnosis against the
 * oracle's per-vertex trace: d_trace = W*H*max_vertices*8 floats, zeroed by
 * the caller; path p's vertex of depth k (the k-th surface vertex, counted
 * from 1) that sampled a direction writes 8 floats at ((p*max_vertices)+k-1)*8:
 * depth, mesh, pdf, mean f|cos|/pdf, mean throughput reaching the vertex, 0,
 * 0, trans.  Vertices beyond max_vertices are not recorded. */
wl_status wl_pt_trace_samples(wl_scene scene, const wl_pt_config *cfg, uint32_t frame,
                              float *d_rgb, float *d_trace, uint32_t max_vertices,
                              void *stream);
/* One progressive frame exactly as the reference raygen ends
 * (PathTracing.cu:145-158): running mean into d_radiance (RGBA float) with
 * float factors and double temporaries, then ACES (Tonemap.h:34-42) -> u8. */
wl_status wl_pt_launch_frame(wl_scene scene, const wl_pt_config *cfg,
                             uint32_t frame, float *d_radiance_rgba,
                             uint8_t *d_color_rgba, void *stream);
/* wl_pt_render without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The batch
 * buffers (path state, ray / hit / shadow queues, queue counters) live in one
 * fixed device array of the library, sized for WL_PT_NOALLOC_MAX_BATCH_PATHS
 * paths per batch (about 55 MB, held on each device the library runs on
 * until the process ends); the call makes no hipMalloc / hipFree and no
 * memset node.  Batches are formed as in wl_pt_render (one lane, no stats):
 * cfg->max_batch_paths = 0 means WL_PT_NOALLOC_MAX_BATCH_PATHS, a larger value
 * is rejected with WL_ERROR_LIMIT_EXCEEDED, and a scene with a guide attached
 * with WL_ERROR_NOT_SUPPORTED.  The film is the one wl_pt_render writes for
 * the same cfg, bit for bit.  Every call on one device shares that array:
 * calls that overlap on one GPU (different streams or threads) race, so order
 * them on one stream.  Needs a context created on the device first. */
#define WL_PT_NOALLOC_MAX_BATCH_PATHS (1u << 18)
wl_status wl_pt_render_noalloc(wl_scene scene, const wl_pt_config *cfg, float *d_film,
                               void *stream);
/* Film resolve: d_rgb[i] = sum / nsamples (+ compensation). */
wl_status wl_film_resolve(wl_context ctx, const float *d_film, uint32_t npix,
                          uint32_t nsamples, float *d_rgb, void *stream);

/* ---- learned path guiding -------------------------------------------------
 * No counterpart in the reference, whose path tracer samples the BSDF only
 * (Sample.h:134-231).  A small MLP (64 -> 4 x 64 ReLU -> 32, BF16 MFMA
 * forward and backward, FP32 accumulation and FP32 master weights) maps a
 * query (position normalised to the scene box, normal, outgoing direction;
 * sin/cos frequency encoding of the position) to an 8-lobe von Mises-Fisher
 * mixture over directions.  With a guide attached to a scene the tracer picks
 * the guide with probability mix_alpha at vertices whose material has
 * trans = 0 and clearcoat = 0, else the reference BSDF sampler, and weights
 * by the mixed density (one-sample MIS), so the estimate stays unbiased.
 * Only vertices on bounces below guided_bounces (0: every bounce) that are
 * among the path's first guided_vertices such vertices (0: all) are guided.  In
 * training mode every path records its first guided vertex, and after each
 * batch the recorded samples train the network for steps_per_batch Adam
 * steps on the KL objective -F/p log q (Mueller et al., TOG 2019).  That
 * training runs on the guide's own stream while the next batch renders: the
 * render samples from a published copy of the network, updated at the end of
 * each training batch, so batch k samples from the network trained through
 * batch k - 2 (none of its own paths).  Every other guide call, and a render
 * that does not train, first waits for a pending training and publishes it.
 *
 * Parameter layout (WL_GUIDE_PARAMS floats): for layer l = 0..4 the weights
 * W_l[out][in] row-major then the biases b_l[out]; out = 64 (l < 4) or 32.
 * Output row r: lobe 2 ((r % 16) / 4) + r / 16, parameter r % 4 = logit,
 * raw concentration (kappa = softplus), theta, phi.
 * Query: 9 floats.  Training sample: 16 floats = query, sampled direction,
 * weight F/p, 3 unused. */
#define WL_GUIDE_PARAMS 18720u
enum
{
    WL_GUIDE_SAMPLE = 1u << 0, /* guided sampling                             */
    WL_GUIDE_TRAIN = 1u << 1   /* record training samples and train per batch */
};

typedef struct wl_guide_config
{
    uint32_t seed;            /* weight initialisation                        */
    uint32_t batch_size;      /* samples per Adam step, multiple of 32        */
    uint32_t steps_per_batch; /* Adam steps after each rendered batch         */
    uint32_t max_samples;     /* training samples kept per batch              */
    float learning_rate, beta1, beta2, epsilon;
    float mix_alpha;          /* probability of sampling the guide            */
    float max_weight;         /* training weights are clamped to this         */
    uint32_t guided_bounces;  /* guide the first bounces only (0 = every one) */
    uint32_t guided_vertices; /* guide a path's first guidable vertices only
                                 (0 = every one)                              */
} wl_guide_config;

typedef struct wl_guide_info
{
    uint64_t steps;           /* Adam steps taken                             */
    uint32_t last_samples;    /* training samples of the last batch           */
    uint32_t pad_;
} wl_guide_info;

wl_status wl_guide_default_config(wl_guide_config *cfg);
wl_status wl_guide_create(wl_context ctx, const wl_guide_config *cfg, wl_guide *out);
wl_status wl_guide_destroy(wl_guide guide);
/* Host copies of the FP32 master parameters; set also clears the Adam
 * state and the step count.  n must be WL_GUIDE_PARAMS. */
wl_status wl_guide_set_params(wl_guide guide, const float *params, uint32_t n);
wl_status wl_guide_get_params(wl_guide guide, float *params, uint32_t n);
wl_status wl_guide_set_mix(wl_guide guide, float mix_alpha);
/* Synchronises the device. */
wl_status wl_guide_get_info(wl_guide guide, wl_guide_info *info);
/* Network outputs of n queries: d_out = 32 floats per query (row order). */
wl_status wl_guide_infer(wl_guide guide, const float *d_queries, uint32_t n, float *d_out,
                         void *stream);
/* Mixture sampling on given outputs: d_u = 3 uniforms per query (lobe, then
 * the two direction numbers); writes the direction and its mixture density.
 * Both calls run the path tracer's own decode chain (the outputs decoded
 * across lane groups into the stored per-query form, read back, sampled with
 * the stored lobe weights), so they test what a guided vertex does. */
wl_status wl_guide_mixture_sample(wl_context ctx, const float *d_out, const float *d_u,
                                  uint32_t n, float *d_dirs, float *d_pdf, void *stream);
wl_status wl_guide_mixture_pdf(wl_context ctx, const float *d_out, const float *d_dirs,
                               uint32_t n, float *d_pdf, void *stream);
/* Mean loss and its gradient over n <= batch_size samples, no update. */
wl_status wl_guide_gradient(wl_guide guide, const float *d_samples, uint32_t n,
                            float *d_grad, float *d_loss, void *stream);
/* `steps` Adam steps over consecutive mini-batches of the n samples. */
wl_status wl_guide_train(wl_guide guide, const float *d_samples, uint32_t n,
                         uint32_t steps, void *stream);
/* The path tracer's own guide inference (the kernel a guided bounce runs)
 * on n rays and their closest hits in `scene`, whose guide must be attached:
 * per query the decoded mixture as a guided vertex stores it, 48 floats:
 * (log pi_j + log C(kappa_j), kappa_j, mu_j.x, mu_j.y) for lobes j = 0..7,
 * then mu_j.z, then pi_j (C: the vMF normalisation).  The query of a hit is
 * the hit position normalised to the scene box, the geometric normal facing
 * the ray, and the reversed ray direction.  Misses give no output.  For
 * parity testing; synchronises the stream. */
wl_status wl_guide_infer_hits(wl_scene scene, const wl_ray *d_rays, const wl_hit *d_hits,
                              uint32_t n, float *d_decoded, void *stream);
/* Attaches (mode != 0) or detaches (guide NULL or mode 0) a guide. */
wl_status wl_scene_set_guide(wl_scene scene, wl_guide guide, uint32_t mode);

/* ---- batched serving ---------------------------------------------------------
 * Replaces one DeviceManager::Launch per request and frame
 * (src/Frontend/Core/DeviceManager.cpp:231-241, Launcher.cpp:12-20) for a
 * set of independent render requests.  The server packs whole-image sample
 * passes of the requests (in order) into megabatches of up to batch_paths
 * paths; a call's whole megabatch sequence is captured into one hipGraph and
 * launched once per lane, each megabatch reading its work list from device
 * memory.
 * Each request's film equals wl_pt_render of the same request bit for bit.
 * Requests may use different scenes of the server's device; their scenes must
 * have no guide attached, and their films must not overlap (rejected with
 * WL_ERROR_INVALID_VALUE: the requests' films are updated concurrently). */
typedef struct wl_render_request
{
    wl_scene scene;
    float *film;             /* device, 6 floats per pixel (as wl_pt_render) */
    uint32_t width, height;
    uint32_t spp, frame_offset;
    uint32_t max_depth, rr_depth;
    float rr_prob, epsilon;
    uint32_t strategy, pad_;
} wl_render_request;

typedef struct wl_serve_config
{
    uint32_t batch_paths;          /* megabatch capacity in paths; must hold
                                      the largest request's pixel count     */
    uint32_t shade_blocks_per_cu;  /* shade kernel workgroups per CU        */
    uint32_t use_graph;            /* launch the megabatch sequence as one
                                      captured hipGraph (0: plain launches) */
    uint32_t lanes;                /* 2: requests split into two groups whose
                                      megabatches run concurrently (own
                                      buffers and stream each); 1: one      */
} wl_serve_config;

typedef struct wl_serve_stats
{
    double seconds;          /* device-clock time from the call's start (after
                                work already on the stream) to the last
                                megabatch's end */
    uint64_t paths;
    uint64_t extension_rays, shadow_rays;
    uint32_t batches;        /* megabatches over both lanes                */
    uint32_t launches;       /* graph launches over both lanes: one per lane
                                for up to 64 megabatches                    */
} wl_serve_stats;

wl_status wl_serve_default_config(wl_serve_config *cfg);
wl_status wl_server_create(wl_context ctx, const wl_serve_config *cfg, wl_server *out);
wl_status wl_server_destroy(wl_server server);
wl_status wl_server_get_config(wl_server server, wl_serve_config *cfg);
wl_status wl_server_set_config(wl_server server, const wl_serve_config *cfg);
/* Renders n requests; waits for work already on `stream` and returns when
 * every film is complete.  done_seconds (optional, n doubles): each request's
 * completion time (the end of its last megabatch) from the call's start, on
 * the device clock.  A lane's megabatches run as one captured graph launch
 * (up to 64 megabatches per graph). */
wl_status wl_server_render(wl_server server, const wl_render_request *reqs, uint32_t n,
                           wl_serve_stats *stats, double *done_seconds, void *stream);
/* Times candidate megabatch sizes, shade grids and lane counts on the
 * requests (each cut to max_spp samples, rendered into scratch films), one
 * parameter at a time (capacity, then shade grid, then lanes), and keeps
 * the fastest. */
wl_status wl_server_autotune(wl_server server, const wl_render_request *reqs, uint32_t n,
                             uint32_t max_spp, wl_serve_config *chosen);

/* ---- programs, pipeline, SBT, launch: the OptiX program interface -------
 *
 * Program discovery by name, as Module::IdentifyPrograms does it
 * (src/Backend/Optix/Core/Module-AutoConfig.cpp:76-178): every
 *   extern "C" __global__ void __<kind>__<name>()
 * in the source text is a program; raygen and miss programs become groups in
 * file order; intersection/any-hit/closest-hit programs whose names end with
 * the same suffix merge into one hit group; hit groups follow all raygen and
 * miss groups, in order of first appearance.  The device code itself is not
 * compiled: a pipeline maps the discovered program set to a built-in gfx950
 * integrator (the PathTracing program of src/Frontend/Programs/PathTracing,
 * or the Lambertian program of examples/PathTracer).                       */
enum
{
    WL_PROGRAM_RAYGEN = 0,
    WL_PROGRAM_MISS = 1,
    WL_PROGRAM_HITGROUP = 2
};

typedef struct wl_program_group
{
    uint32_t kind;
    /* raygen/miss: entry[0]; hit group: entry[0] intersection,
     * entry[1] any-hit, entry[2] closest-hit ("" when absent) */
    char entry[3][64];
} wl_program_group;

wl_status wl_identify_programs(const char *source, size_t length,
                               wl_program_group *groups, uint32_t capacity,
                               uint32_t *count);

typedef struct wl_pipeline_s *wl_pipeline;
typedef struct wl_sbt_s *wl_sbt;

enum
{
    WL_BUILTIN_PATHTRACING = 1,      /* PathTracing.cu program set         */
    WL_BUILTIN_EXAMPLE_PATHTRACER = 2 /* examples/PathTracer/PathTracer.cu */
};

/* Pipeline (Pipeline.cpp:93-136): WL_ERROR_UNKNOWN_PROGRAM when the program
 * set is not one of the built-in integrators.  *builtin gets its id. */
wl_status wl_pipeline_create(wl_context ctx, const wl_program_group *groups,
                             uint32_t num_groups, uint32_t max_trace_depth,
                             wl_pipeline *out, uint32_t *builtin);
wl_status wl_pipeline_destroy(wl_pipeline p);

/* SBT records keep the reference's layout: a 32-byte header followed by the
 * program's data, 16-byte aligned (SBTData.h:8-25).  wl_sbt_pack_header
 * writes the header for program group `group_index` (optixSbtRecordPackHeader).
 * Hit record of build input i and ray type r = hit_records[r + ray_type_count*i]
 * (the addressing of GetSBTHitRecordBuffer, Traversable.h:95-102). */
#define WL_SBT_RECORD_HEADER_SIZE 32
#define WL_SBT_RECORD_ALIGNMENT 16
wl_status wl_sbt_pack_header(wl_pipeline p, uint32_t group_index, void *record);
wl_status wl_sbt_create(wl_pipeline p, const void *raygen_record, size_t raygen_size,
                        const void *miss_records, uint32_t miss_count,
                        size_t miss_stride, const void *hit_records,
                        uint32_t hit_count, size_t hit_stride, wl_sbt *out);
wl_status wl_sbt_destroy(wl_sbt sbt);

/* A texture object for PathTracing hit records: cudaCreateTextureObject over
 * a uchar4 array with wrap addressing and linear filtering
 * (DeviceManager.cpp:192-227).  The handle goes into
 * wl_pathtracing_hit_data::texture with hasTexture = 1, and texcoords points
 * at 2 floats per vertex on the device (NULL reads (0, 0)).  The library
 * keeps its own copy of the image; a scene derived from an SBT copies it
 * again, so the object may be destroyed after the first launch. */
wl_status wl_texture_create(wl_context ctx, const wl_texture_image *image, uint64_t *texture);
wl_status wl_texture_destroy(uint64_t texture);

/* The value a LaunchParams `traversable` field holds for this accel. */
wl_status wl_accel_traversable(wl_accel accel, uint64_t *handle);

/* optixLaunch (Launcher.cpp:12-20): params is host memory laid out as the
 * program's LaunchParams (copied at launch, like Launcher's device buffer):
 *   PathTracing:      wl_pathtracing_params (PathTracingLaunchParams.h:44-57)
 *   examples program: wl_example_params      (examples/PathTracer/Param.h:11-18) */
wl_status wl_launch(wl_pipeline p, void *stream, const void *params,
                    size_t params_size, wl_sbt sbt, uint32_t width,
                    uint32_t height, uint32_t depth);

/* Reference device structs the two built-in programs read (same layouts). */
typedef struct wl_pathtracing_params
{
    uint32_t frameID;
    int32_t fbSize[2];
    uint32_t pad0_;
    float *radianceBuffer;         /* vec4 per pixel                        */
    uint8_t *colorBuffer;          /* u8vec4 per pixel                      */
    wl_pinhole_camera camera;
    uint32_t areaLightCount;
    uint32_t pad1_;
    const void *areaLights;        /* wl_device_area_light[areaLightCount]  */
    uint64_t traversable;
} wl_pathtracing_params;

typedef struct wl_device_area_light /* Device/Light.h:9-17 */
{
    uint8_t twoSided, hasNormal, pad0_[2];
    uint32_t triangleNum;
    float L[3];
    uint32_t pad1_;
    const float *vertices;
    const float *normals;
    const int32_t *indices;
} wl_device_area_light;

typedef struct wl_pathtracing_hit_data /* PathTracingLaunchParams.h:64-78 */
{
    uint8_t hasTexture, hasNormal, pad0_[6];
    uint64_t texture;
    uint32_t meshID, areaLightID, materialID;
    float disneyMat[15];           /* DisneyMaterial, 60 bytes               */
    const int32_t *indices;
    const float *vertices;
    const float *normals;
    const float *texcoords;
} wl_pathtracing_hit_data;

typedef struct wl_example_camera /* DeviceUtils/Camera.h:7-14 */
{
    float horizontalRatio, verticalRatio;
    float position[3], gaze[3], up[3];
} wl_example_camera;

typedef struct wl_example_params /* examples/PathTracer/Param.h:11-18 */
{
    uint32_t sampleNum;
    uint32_t maxDepth;
    wl_example_camera camera;
    uint32_t pad_;
    float *colorBuffer;            /* vec4 per pixel                        */
    uint64_t traversable;
} wl_example_params;

typedef struct wl_example_hit_data /* examples/PathTracer/Param.h:20-28 */
{
    float color[3];
    float emission[3];
    const int32_t *indices;        /* device, ivec3 per triangle            */
    const float *normals;          /* device, vec3 per vertex               */
} wl_example_hit_data;

/* examples/PathTracer/Conversion.cu:6-19 FromFloatToChar: v > 1 -> 1, then
 * (u8)(v * 255), on the device. */
wl_status wl_float_to_u8(const float *d_src, size_t n, uint8_t *d_dst, void *stream);

#ifdef __cplusplus
}
#endif
#endif
