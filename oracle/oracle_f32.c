/*
 * oracle_f32.c — FP32 instantiation of the CPU oracle (TEST INFRASTRUCTURE
 * ONLY; see oracle.h).  Follows the reference's host-compile numerics:
 * IEEE single precision, no contraction, glibc sinf/cosf/logf.
 */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX(n) n##_f32
#define K(x) x##f
#define RSQRT sqrtf
#define RSIN sinf
#define RCOS cosf
#define RLOG logf
#define RABS fabsf
#define RCOPYSIGN copysignf
#define RFMAX fmaxf
#define RISNAN isnan
#define RISINF isinf
#define RFLOOR floorf

#include "oracle_impl.inc"

int oracle_intersect_f32(const oracle_ray *ray, const float v[9], float *t,
                         float *b1, float *b2)
{
    WRay w;
    wray_setup(&w, ray->o, ray->d, ray->tmin, ray->tmax);
    return wt_test(&w, v, t, b1, b2);
}

void oracle_onb_f32(const float n[3], float u[3], float v[3])
{
    V3 a, b;
    onb(v3(n[0], n[1], n[2]), &a, &b);
    vstorer(u, a);
    vstorer(v, b);
}

/* ---- brute force: every triangle, same test, same tie-break ------------ */
typedef struct
{
    const float *tris;
    uint32_t ntris;
    const oracle_ray *rays;
    oracle_hit *hits;
    uint8_t *occ;
} brute_ctx;

static void brute_closest_task(void *p, uint32_t b, uint32_t e)
{
    brute_ctx *c = (brute_ctx *)p;
    for (uint32_t r = b; r < e; ++r)
    {
        WRay w;
        const oracle_ray *ray = &c->rays[r];
        wray_setup(&w, ray->o, ray->d, ray->tmin, ray->tmax);
        float best_t = ray->tmax, bb1 = 0, bb2 = 0;
        uint32_t best = ORACLE_INVALID;
        for (uint32_t i = 0; i < c->ntris; ++i)
        {
            float t, b1, b2;
            if (wt_test(&w, c->tris + 9 * (size_t)i, &t, &b1, &b2))
            {
                if (t < best_t || (t == best_t && i < best))
                {
                    best_t = t;
                    best = i;
                    bb1 = b1;
                    bb2 = b2;
                }
            }
        }
        c->hits[r].t = best == ORACLE_INVALID ? INFINITY : best_t;
        c->hits[r].prim = best;
        c->hits[r].b1 = bb1;
        c->hits[r].b2 = bb2;
    }
}

static void brute_any_task(void *p, uint32_t b, uint32_t e)
{
    brute_ctx *c = (brute_ctx *)p;
    for (uint32_t r = b; r < e; ++r)
    {
        WRay w;
        const oracle_ray *ray = &c->rays[r];
        wray_setup(&w, ray->o, ray->d, ray->tmin, ray->tmax);
        uint8_t o = 0;
        for (uint32_t i = 0; i < c->ntris && !o; ++i)
        {
            float t, b1, b2;
            if (wt_test(&w, c->tris + 9 * (size_t)i, &t, &b1, &b2))
                o = 1;
        }
        c->occ[r] = o;
    }
}

void oracle_trace_closest_brute(const float *tris, uint32_t ntris,
                                const oracle_ray *rays, uint32_t nrays,
                                oracle_hit *hits, int nthreads)
{
    brute_ctx c = { tris, ntris, rays, hits, NULL };
    oracle_parallel_for(nrays, 64, nthreads, brute_closest_task, &c);
}

void oracle_trace_any_brute(const float *tris, uint32_t ntris,
                            const oracle_ray *rays, uint32_t nrays,
                            uint8_t *occluded, int nthreads)
{
    brute_ctx c = { tris, ntris, rays, NULL, occluded };
    oracle_parallel_for(nrays, 64, nthreads, brute_any_task, &c);
}

/* ---- BVH traversal ------------------------------------------------------ */
typedef struct
{
    const oracle_bvh *bvh;
    const oracle_ray *rays;
    oracle_hit *hits;
    uint8_t *occ;
} bvh_ctx;

static void bvh_closest_task(void *p, uint32_t b, uint32_t e)
{
    bvh_ctx *c = (bvh_ctx *)p;
    for (uint32_t r = b; r < e; ++r)
    {
        const oracle_ray *ray = &c->rays[r];
        oracle_hit *h = &c->hits[r];
        if (c->bvh->ntris == 0)
        {
            h->t = INFINITY;
            h->prim = ORACLE_INVALID;
            h->b1 = h->b2 = 0;
            continue;
        }
        trace_closest_bvh(c->bvh, ray->o, ray->d, ray->tmin, ray->tmax, &h->t,
                          &h->prim, &h->b1, &h->b2);
    }
}

static void bvh_any_task(void *p, uint32_t b, uint32_t e)
{
    bvh_ctx *c = (bvh_ctx *)p;
    for (uint32_t r = b; r < e; ++r)
    {
        const oracle_ray *ray = &c->rays[r];
        c->occ[r] = c->bvh->ntris == 0
                        ? 0
                        : (uint8_t)trace_any_bvh(c->bvh, ray->o, ray->d,
                                                 ray->tmin, ray->tmax);
    }
}

void oracle_bvh_trace_closest(const oracle_bvh *bvh, const oracle_ray *rays,
                              uint32_t n, oracle_hit *hits, int nthreads)
{
    bvh_ctx c = { bvh, rays, hits, NULL };
    oracle_parallel_for(n, 256, nthreads, bvh_closest_task, &c);
}

void oracle_bvh_trace_any(const oracle_bvh *bvh, const oracle_ray *rays,
                          uint32_t n, uint8_t *occluded, int nthreads)
{
    bvh_ctx c = { bvh, rays, NULL, occluded };
    oracle_parallel_for(n, 256, nthreads, bvh_any_task, &c);
}

/* ---- integrator -------------------------------------------------------- */
uint32_t oracle_pt_trace_f32(const oracle_pt *pt, const oracle_pt_config *c, uint32_t px,
                             uint32_t py, uint32_t frame, float rgb[3], float *trace,
                             uint32_t max_vertices)
{
    V3 r;
    uint32_t n = 0;
    pt_sample_traced(pt, c, px, py, frame, &r, NULL, trace, max_vertices, &n);
    rgb[0] = r.x;
    rgb[1] = r.y;
    rgb[2] = r.z;
    return n;
}

uint32_t oracle_pt_sample_f32(const oracle_pt *pt, const oracle_pt_config *c,
                              uint32_t px, uint32_t py, uint32_t frame,
                              float rgb[3], uint32_t *shadow)
{
    V3 r;
    uint32_t n = pt_sample(pt, c, px, py, frame, &r, shadow);
    rgb[0] = r.x;
    rgb[1] = r.y;
    rgb[2] = r.z;
    return n;
}

typedef struct
{
    const oracle_pt *pt;
    const oracle_pt_config *c;
    uint32_t frame;
    float *rgb;
    uint64_t rays[256];
    pthread_mutex_t *mu;
    uint64_t total;
} frame_ctx;

#include <pthread.h>

static void frame_task(void *p, uint32_t b, uint32_t e)
{
    frame_ctx *fc = (frame_ctx *)p;
    uint64_t rays = 0;
    for (uint32_t i = b; i < e; ++i)
    {
        uint32_t x = i % fc->c->width, y = i / fc->c->width;
        uint32_t sh = 0;
        rays += oracle_pt_sample_f32(fc->pt, fc->c, x, y, fc->frame,
                                     fc->rgb + 3 * (size_t)i, &sh);
        rays += sh;
    }
    pthread_mutex_lock(fc->mu);
    fc->total += rays;
    pthread_mutex_unlock(fc->mu);
}

void oracle_pt_render_frame_f32(const oracle_pt *pt, const oracle_pt_config *c,
                                uint32_t frame, float *rgb, int nthreads,
                                uint64_t *rays_out)
{
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    frame_ctx fc;
    memset(&fc, 0, sizeof fc);
    fc.pt = pt;
    fc.c = c;
    fc.frame = frame;
    fc.rgb = rgb;
    fc.mu = &mu;
    oracle_parallel_for(c->width * c->height, 256, nthreads, frame_task, &fc);
    pthread_mutex_destroy(&mu);
    if (rays_out)
        *rays_out = fc.total;
}

/* ---- examples/PathTracer/PathTracer.cu:49-144 --------------------------- */
typedef struct
{
    const oracle_pt *pt;
    const float *mcolor, *memit;
} ex_scene;

/* PathTracer.cu:39-47 RandomSampleDir: uniform hemisphere. */
static V3 ex_sample_dir(V3 n, uint32_t *seed)
{
    V3 x, y;
    onb(n, &x, &y);
    /* `rnd(seed) * 2 * pi` with pi a float constant */
    float angle = rndr(seed) * 2 * PI_, h = rndr(seed);
    return vadd(vscale(vadd(vscale(x, sinf(angle)), vscale(y, cosf(angle))),
                       sqrtf(1 - h * h)),
                vscale(n, h));
}

/* Recursive __closesthit__PT (:96-144) / __miss__Empty (:90-94); returns the
 * payload colour and updates the seed the way the payload registers do. */
static V3 ex_trace(const ex_scene *s, V3 o, V3 d, float tmin, float tmax,
                   uint32_t depth, uint32_t *seed)
{
    float oo[3] = { o.x, o.y, o.z }, dd[3] = { d.x, d.y, d.z };
    float t, b1, b2;
    uint32_t prim;
    trace_closest_bvh(s->pt->bvh, oo, dd, tmin, tmax, &t, &prim, &b1, &b2);
    if (prim == ORACLE_INVALID)
        return v3(0, 0, 0); /* miss: colour 0, seed untouched */
    uint32_t local;
    uint32_t mesh = prim_to_mesh(s->pt, prim, &local);
    V3 emission = vload(s->memit + 3 * (size_t)mesh);
    V3 color = vload(s->mcolor + 3 * (size_t)mesh);
    float RR_factor = 1.0f;
    if (depth == 0 || !(emission.x == 0 && emission.y == 0 && emission.z == 0))
        return emission;
    else if (depth <= 5)
    {
        if (rndr(seed) > 0.99f)
            return emission;
        RR_factor /= 0.99f;
    }
    const uint32_t *tri = scene_tri(s->pt, mesh, local);
    V3 normal = vnormalize(bary(scene_normal(s->pt, mesh, tri[0]),
                                scene_normal(s->pt, mesh, tri[1]),
                                scene_normal(s->pt, mesh, tri[2]), b1, b2));
    V3 hit = vadd(o, vscale(d, t));
    V3 nd = ex_sample_dir(normal, seed);
    float cw = vdot(nd, normal);
    float cosWeight = cw > 0.f ? cw : 0.f;
    float coeff = 2;
    V3 result = ex_trace(s, hit, nd, 1e-5f, 30.f, depth - 1, seed);
    return vscale(vscale(vscale(vmul(result, color), cosWeight), coeff), RR_factor);
}

typedef struct
{
    ex_scene s;
    const oracle_example_config *c;
    uint32_t row0;
    float *rgba;
} ex_ctx;

static void ex_task(void *p, uint32_t b, uint32_t e)
{
    ex_ctx *ec = (ex_ctx *)p;
    const oracle_example_config *c = ec->c;
    for (uint32_t i = b; i < e; ++i)
    {
        /* launch index .x is the row, .y the column (PathTracer.cu:51) */
        uint32_t row = ec->row0 + i / c->height, col = i % c->height;
        float csx = 1.f / (float)c->width, csy = 1.f / (float)c->height;
        float vpos = 1 - 2.f * (float)row * csx, hpos = 2.f * (float)col * csy - 1;
        V3 gaze = vload(c->cam_gaze), up = vload(c->cam_up);
        V3 hvec = vscale(vcross(gaze, up), c->horizontal_ratio);
        V3 vvec = vscale(up, c->vertical_ratio);
        V3 rayDir = vnormalize(vadd(vadd(gaze, vscale(hvec, hpos)), vscale(vvec, vpos)));
        uint32_t idx = c->width * row + col;
        uint32_t seed = oracle_tea4(idx, 10086);
        V3 result = v3(0, 0, 0);
        float w = 1.f / (float)c->sample_num;
        for (uint32_t k = 0; k < c->sample_num; ++k)
        {
            float r1 = rndr(&seed);
            V3 rd = vadd(rayDir, vscale(hvec, r1 * 2 * csy));
            float r2 = rndr(&seed);
            rd = vadd(rd, vscale(vvec, r2 * 2 * csx));
            V3 col3 = ex_trace(&ec->s, vload(c->cam_pos), rd, 0.f, 30.f,
                               c->max_depth, &seed);
            result = vadd(result, vscale(col3, w));
        }
        float *o = ec->rgba + 4 * (size_t)i;
        o[0] = result.x;
        o[1] = result.y;
        o[2] = result.z;
        o[3] = 1.0f;
    }
}

void oracle_example_render(const oracle_pt *pt, const float *mesh_color,
                           const float *mesh_emission,
                           const oracle_example_config *c, uint32_t row0,
                           uint32_t row1, float *rgba, int nthreads)
{
    ex_ctx ec;
    ec.s.pt = pt;
    ec.s.mcolor = mesh_color;
    ec.s.memit = mesh_emission;
    ec.c = c;
    ec.row0 = row0;
    ec.rgba = rgba;
    oracle_parallel_for((row1 - row0) * c->height, 64, nthreads, ex_task, &ec);
}

void oracle_tex_bilinear_f32(const uint8_t *rgba, uint32_t w, uint32_t h, float u, float v,
                             float out[3])
{
    V3 r = tex_bilinear_rgba(rgba, w, h, u, v);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
