/*
 * oracle.h — CPU oracle for the EasyRender path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library includes, links
 * or calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load liboracle.so, and only as the reference results the
 * GPU path is compared against / the timed CPU baseline.  It is a plain-C restatement of the reference algorithm
 * (Extra-Creativity/Wayland, "EasyRender"); each function cites the reference
 * file:line it follows.  The reference itself cannot be built here (it needs
 * the CUDA toolkit, the OptiX SDK, glm and spdlog), so there is no oracle/_ref.
 *
 * Every function is compiled twice: an FP32 build (suffix _f32) whose
 * operation order and rounding follow the reference's host compile
 * (-ffp-contract=off, correctly rounded + - * / sqrt, glibc sin/cos/log), and
 * an FP64 build (suffix _f64) used as the converged-image reference.
 *
 * Pinning status of each function (what checks it independently) is listed in
 * tests/test_oracle.py; functions with no independent check say
 * "parity unpinned" there.
 */
#ifndef WL_ORACLE_H
#define WL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_INVALID 0xFFFFFFFFu

/* ---- RNG: src/Frontend/Device/Random.h:35-71 -------------------------- */
uint32_t oracle_tea(uint32_t v0, uint32_t v1, const uint32_t key[4],
                    uint32_t rounds, uint32_t *v1_out);
uint32_t oracle_tea4(uint32_t v0, uint32_t v1);
uint32_t oracle_lcg(uint32_t *state);
float oracle_rnd(uint32_t *state);

/* ---- rays / hits --------------------------------------------------------
 * Ray layout matches what OptiX receives at PathTracing.cu:72-75 and :92-102:
 * origin, tmin, direction, tmax (32 bytes). */
typedef struct
{
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} oracle_ray;

/* Closest hit: t, global primitive id (mesh triangle offset + local index),
 * barycentrics (b1, b2) weighting v1 and v2 as optixGetTriangleBarycentrics
 * does (src/Frontend/Device/Common.h:40-51). */
typedef struct
{
    float t;
    uint32_t prim;
    float b1, b2;
} oracle_hit;

/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013) that
 * stands in for OptiX's closed triangle test.  tris: 9 floats per triangle. */
int oracle_intersect_f32(const oracle_ray *ray, const float v[9], float *t,
                         float *b1, float *b2);

void oracle_trace_closest_brute(const float *tris, uint32_t ntris,
                                const oracle_ray *rays, uint32_t nrays,
                                oracle_hit *hits, int nthreads);
void oracle_trace_any_brute(const float *tris, uint32_t ntris,
                            const oracle_ray *rays, uint32_t nrays,
                            uint8_t *occluded, int nthreads);

/* Binary SAH BVH over the same triangles, the oracle's fast path for large
 * inputs.  Results equal the brute force (same test, same tie-break). */
typedef struct oracle_bvh oracle_bvh;
oracle_bvh *oracle_bvh_build(const float *tris, uint32_t ntris);
void oracle_bvh_free(oracle_bvh *bvh);
void oracle_bvh_info(const oracle_bvh *bvh, uint32_t *nodes, uint32_t *depth);
void oracle_bvh_trace_closest(const oracle_bvh *bvh, const oracle_ray *rays,
                              uint32_t n, oracle_hit *hits, int nthreads);
void oracle_bvh_trace_any(const oracle_bvh *bvh, const oracle_ray *rays,
                          uint32_t n, uint8_t *occluded, int nthreads);

/* ---- materials: src/Frontend/Device/Material.h:7-21 (64 bytes) ---------- */
typedef struct
{
    float color[3];
    float metallic;
    float roughness;
    float specular;
    float specularTint;
    float eta;
    float trans;
    float subsurface;
    float anisotropic;
    float sheen;
    float sheenTint;
    float clearcoat;
    float clearcoatgloss;
    float pad_;
} oracle_material;

/* Disney BSDF: Evaluate.h:78-191, Pdf.h:42-118, Sample.h:134-231. */
void oracle_disney_eval_f32(const oracle_material *m, const float Ns[3],
                            const float Ng[3], const float V[3],
                            const float L[3], float out[3]);
float oracle_disney_pdf_f32(const oracle_material *m, const float Ns[3],
                            const float Ng[3], const float V[3],
                            const float L[3]);
void oracle_disney_sample_f32(const oracle_material *m, const float Ns[3],
                              const float Ng[3], const float V[3],
                              uint32_t *seed, float L[3]);
void oracle_disney_eval_f64(const oracle_material *m, const double Ns[3],
                            const double Ng[3], const double V[3],
                            const double L[3], double out[3]);
double oracle_disney_pdf_f64(const oracle_material *m, const double Ns[3],
                             const double Ng[3], const double V[3],
                             const double L[3]);
void oracle_disney_sample_f64(const oracle_material *m, const double Ns[3],
                              const double Ng[3], const double V[3],
                              uint32_t *seed, double L[3]);
/* Batched forms for one (material, Ns, Ng, V): seeds -> directions (3n),
 * directions (3n) -> pdfs (n), directions (3n) -> eval values (3n). */
void oracle_disney_sample_batch_f32(const oracle_material *m, const float Ns[3],
                                    const float Ng[3], const float V[3],
                                    const uint32_t *seeds, uint32_t n, float *L);
void oracle_disney_pdf_batch_f32(const oracle_material *m, const float Ns[3],
                                 const float Ng[3], const float V[3], const float *L,
                                 uint32_t n, float *pdf);
void oracle_disney_eval_batch_f32(const oracle_material *m, const float Ns[3],
                                  const float Ng[3], const float V[3], const float *L,
                                  uint32_t n, float *out);
void oracle_disney_sample_batch_f64(const oracle_material *m, const double Ns[3],
                                    const double Ng[3], const double V[3],
                                    const uint32_t *seeds, uint32_t n, double *L);
void oracle_disney_pdf_batch_f64(const oracle_material *m, const double Ns[3],
                                 const double Ng[3], const double V[3], const double *L,
                                 uint32_t n, double *pdf);
void oracle_disney_eval_batch_f64(const oracle_material *m, const double Ns[3],
                                  const double Ng[3], const double V[3], const double *L,
                                  uint32_t n, double *out);
/* MaterialMath (Common.h:119-181), one function at a time: which = 0
 * SchlickFresnel(u), 1 Fresnel(cos_i, cos_t, eta), 2 GTR1(NdotH, a),
 * 3 GTR2(NdotH, a), 4 SmithG_GGX(NdotV, alphaG), 5 Lambda(w[3], n[3]). */
float oracle_material_math_f32(int which, const float *args);
double oracle_material_math_f64(int which, const double *args);
/* Orthonormal basis, Common.h:102-116 (Duff et al. 2017). */
void oracle_onb_f32(const float n[3], float u[3], float v[3]);
/* Texture fetch of the reference's texture object (wrap, bilinear with 8-bit
 * fractional weights, u8 / 255) on an RGBA8 image, rows top to bottom. */
void oracle_tex_bilinear_f32(const uint8_t *rgba, uint32_t w, uint32_t h, float u, float v,
                             float out[3]);
/* ACES fit tonemap, Tonemap.h:34-42. */
void oracle_aces_f32(const float rad[3], float out[3]);

/* ---- scene + integrator ------------------------------------------------- */
typedef struct
{
    uint32_t num_meshes;
    const float *vertices;          /* float3, all meshes back to back       */
    const float *normals;           /* float3 per vertex (NULL: no normals)  */
    const uint32_t *indices;        /* uint3 per triangle, mesh-local        */
    const uint32_t *mesh_vertex_offset;
    const uint32_t *mesh_tri_offset;
    const uint32_t *mesh_tri_count;
    const uint32_t *mesh_has_normal;
    const uint32_t *mesh_light;     /* area light id or ORACLE_INVALID       */
    const oracle_material *mesh_material;
    uint32_t num_lights;
    const uint32_t *light_mesh;
    const float *light_L;           /* float3 per light                      */
    const uint32_t *light_two_sided;
    float cam_pos[3], cam_lookat[3], cam_up[3], cam_right[3];
    float bg[3];                    /* MissData::bg_color                    */
    /* textures (PathTracing.cu:261-268); num_textures 0 = none             */
    uint32_t num_textures;
    const float *texcoords;         /* float2 per vertex, indexed as vertices */
    const uint32_t *mesh_texture;   /* image per mesh or ORACLE_INVALID      */
    const uint32_t *texture_offset; /* first texel of each image             */
    const uint32_t *texture_size;   /* width, height per image               */
    const uint8_t *texels;          /* RGBA8, rows top to bottom             */
} oracle_scene_desc;

typedef struct
{
    uint32_t width, height;
    uint32_t max_depth;   /* PathTracing.cu:235 depth cap (25)             */
    uint32_t rr_depth;    /* PathTracing.cu:130 RR when depth > rr_depth (5)*/
    float rr_prob;        /* PathTracing.cu:68 survival probability (0.9)   */
    float epsilon;        /* PathTracing.cu:19 ray tmin / shadow shrink     */
    uint32_t strategy;    /* PathTracing.cu:21-29: 0 UPT, 1 NEE, 2 MIS      */
} oracle_pt_config;

typedef struct oracle_pt oracle_pt;
oracle_pt *oracle_pt_create(const oracle_scene_desc *desc);
void oracle_pt_free(oracle_pt *pt);
/* One sample of PathTracing.cu:40-139 for pixel (px, py) at frame `frame`.
 * Returns the number of extension rays traced; *shadow gets shadow rays. */
uint32_t oracle_pt_sample_f32(const oracle_pt *pt, const oracle_pt_config *c,
                              uint32_t px, uint32_t py, uint32_t frame,
                              float rgb[3], uint32_t *shadow);
/* One sample with a per-vertex record (ORACLE_TRACE_STRIDE floats per
 * surface vertex: depth, mesh, pdf, mean f|cos|/pdf, mean throughput at the
 * vertex, mean NEE contribution, mean radiance after the vertex, trans);
 * returns the number of vertices recorded (<= max_vertices). */
#define ORACLE_TRACE_STRIDE 8
uint32_t oracle_pt_trace_f32(const oracle_pt *pt, const oracle_pt_config *c, uint32_t px,
                             uint32_t py, uint32_t frame, float rgb[3], float *trace,
                             uint32_t max_vertices);
/* Per-sample radiance image of one frame (W*H*3 floats). */
void oracle_pt_render_frame_f32(const oracle_pt *pt, const oracle_pt_config *c,
                                uint32_t frame, float *rgb, int nthreads,
                                uint64_t *rays_out);
/* FP64 estimator averaged over frames [frame0, frame0+nframes): the
 * converged-image reference. */
void oracle_pt_render_f64(const oracle_pt *pt, const oracle_pt_config *c,
                          uint32_t frame0, uint32_t nframes, double *rgb_mean,
                          int nthreads);
/* The same over the ww x wh pixels (x0 + stride i, y0 + stride j) of the
 * W x H image (stride 1: a window; same per-pixel seeds as the whole
 * image); rgb_var (optional) gets the per-pixel variance of one sample
 * (unbiased over the nframes samples). */
void oracle_pt_render_f64_window(const oracle_pt *pt, const oracle_pt_config *c, uint32_t x0,
                                 uint32_t y0, uint32_t ww, uint32_t wh, uint32_t stride,
                                 uint32_t frame0, uint32_t nframes, double *rgb_mean,
                                 double *rgb_var, int nthreads);
/* Film update of PathTracing.cu:145-158: running mean in a float RGBA buffer
 * with float factors and double temporaries, then ACES -> u8. */
void oracle_film_update(float *radiance_rgba, uint8_t *color_rgba,
                        const float *this_frame_rgb, uint32_t npix,
                        uint32_t frame);

/* examples/PathTracer/PathTracer.cu:49-144 (Lambertian recursive PT).
 * Mesh arrays as above; color/emission per mesh (float3). Output RGBA. */
typedef struct
{
    uint32_t width, height;
    uint32_t sample_num;  /* main.cpp:101 (1024)   */
    uint32_t max_depth;   /* main.cpp:102 (9)      */
    float cam_pos[3];
    float cam_gaze[3];
    float cam_up[3];
    float horizontal_ratio, vertical_ratio;
} oracle_example_config;
void oracle_example_render(const oracle_pt *pt, const float *mesh_color,
                           const float *mesh_emission,
                           const oracle_example_config *c, uint32_t row0,
                           uint32_t row1, float *rgba, int nthreads);

/* ---- learned path guiding (no counterpart in the reference) -------------
 * FP64 restatement of the guiding network that the library trains on the GPU
 * (oracle_guide.c): a 64 -> 64 -> 64 -> 64 -> 64 -> 32 ReLU MLP whose 32
 * outputs parameterise an 8-lobe von Mises-Fisher mixture over directions.
 *
 * Parameters (float, ORACLE_GUIDE_NPARAMS): for layer l = 0..4 the weight
 * matrix W_l[out][in] row-major, then the bias b_l[out]; out = 64 for l < 4,
 * 32 for l = 4; in = 64.
 *
 * Query (9 floats): position normalised to [0,1]^3, normal, outgoing
 * direction.  Input encoding x[64], with g = (e % 16) / 4 and pair
 * p = 2 (e / 16) + (e % 4) / 2: for g < 3 and p < 6, sin (e even) or cos
 * (e odd) of 2^p * pi * position[g]; for g < 3 and p = 6, normal[g] (e even)
 * or outgoing direction[g] (e odd); every other x[e] = 0.
 *
 * Output row r belongs to lobe 2 * ((r % 16) / 4) + r / 16, parameter r % 4:
 * 0 = mixture logit a, 1 = raw concentration k (kappa = softplus(k)),
 * 2 = theta, 3 = phi (mean direction (sin t cos p, sin t sin p, cos t)).
 *
 * Training sample (16 floats): query[9], sampled direction w[3], weight
 * (= F / p of the sample, >= 0), then 3 unused.  Per-sample loss
 * -weight * log q(w); the batch loss is the mean over the samples. */
#define ORACLE_GUIDE_NPARAMS 18720
#define ORACLE_GUIDE_NOUT 32
#define ORACLE_GUIDE_LOBES 8

/* Round-to-nearest-even float -> bfloat16 -> float. */
float oracle_bf16_round(float x);
void oracle_guide_encode(const float q[9], double x[64]);
/* Network outputs.  emulate_bf16: round weights, layer inputs to bfloat16 as
 * the GPU kernels do (accumulation stays FP64 here). */
void oracle_guide_forward(const double *params, const float q[9], int emulate_bf16,
                          double out[32]);
/* Decoded mixture: weight pi, concentration kappa, mean direction mu[3]. */
void oracle_guide_decode(const double out[32], double pi[8], double kappa[8],
                         double mu[24]);
double oracle_guide_pdf(const double out[32], const double w[3]);
/* Sample: u0 picks the lobe by the cumulative weights, (u1, u2) the vMF
 * direction (Jakob 2012 inversion) in the Duff et al. frame around mu. */
void oracle_guide_sample(const double out[32], double u0, double u1, double u2,
                         double w[3]);
void oracle_guide_pdf_batch(const double out[32], const double *dirs, uint32_t n,
                            double *pdf);
void oracle_guide_sample_batch(const double out[32], const double *u, uint32_t n,
                               double *dirs);
/* Mean loss over n samples; grad (ORACLE_GUIDE_NPARAMS doubles) receives its
 * gradient with respect to params.  emulate_bf16 rounds weights, layer
 * inputs and back-propagated deltas to bfloat16 where the GPU kernels do. */
double oracle_guide_loss_grad(const double *params, const float *samples, uint32_t n,
                              int emulate_bf16, double *grad);
/* One Adam step (Kingma & Ba 2015, bias-corrected), step = 1, 2, ... */
void oracle_guide_adam(double *params, double *m, double *v, const double *grad,
                       uint32_t step, double lr, double beta1, double beta2, double eps);

#ifdef __cplusplus
}
#endif
#endif
