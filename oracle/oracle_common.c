/*
 * oracle_common.c — RNG, threading, oracle BVH build, scene set-up and the
 * film update of the CPU oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 */
#include "oracle_internal.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- RNG: src/Frontend/Device/Random.h -------------------------------- */

/* Random.h:35-50 tea<N>: N rounds of TEA (Wheeler & Needham 1994) on the
 * pair (v0, v1) with the fixed key; returns v0.  `key` and `rounds` are
 * parameters here so the tests can check the cipher on published vectors. */
uint32_t oracle_tea(uint32_t v0, uint32_t v1, const uint32_t key[4],
                    uint32_t rounds, uint32_t *v1_out)
{
    uint32_t s0 = 0;
    for (uint32_t n = 0; n < rounds; n++)
    {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s0) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s0) ^ ((v0 >> 5) + key[3]);
    }
    if (v1_out)
        *v1_out = v1;
    return v0;
}

uint32_t oracle_tea4(uint32_t v0, uint32_t v1)
{
    static const uint32_t key[4] = { 0xa341316cu, 0xc8013ea4u, 0xad90777du,
                                     0x7e95761eu };
    return oracle_tea(v0, v1, key, 4, NULL);
}

/* Random.h:53-59 */
uint32_t oracle_lcg(uint32_t *state)
{
    *state = 1664525u * *state + 1013904223u;
    return *state & 0x00FFFFFFu;
}

/* Random.h:68-71 */
float oracle_rnd(uint32_t *state)
{
    return (float)oracle_lcg(state) / (float)0x01000000;
}

/* ---- threading -------------------------------------------------------- */
typedef struct
{
    uint32_t n, chunk;
    volatile uint32_t next;
    pthread_mutex_t mu;
    oracle_task_fn fn;
    void *ctx;
} pf_state;

static void *pf_worker(void *arg)
{
    pf_state *s = (pf_state *)arg;
    for (;;)
    {
        pthread_mutex_lock(&s->mu);
        uint32_t b = s->next;
        uint32_t e = b + s->chunk > s->n || b + s->chunk < b ? s->n : b + s->chunk;
        s->next = e;
        pthread_mutex_unlock(&s->mu);
        if (b >= s->n)
            break;
        s->fn(s->ctx, b, e);
    }
    return NULL;
}

void oracle_parallel_for(uint32_t n, uint32_t chunk, int nthreads,
                         oracle_task_fn fn, void *ctx)
{
    if (nthreads <= 1 || n <= chunk)
    {
        if (n)
            fn(ctx, 0, n);
        return;
    }
    if (nthreads > 256)
        nthreads = 256;
    pf_state s;
    s.n = n;
    s.chunk = chunk ? chunk : 1;
    s.next = 0;
    s.fn = fn;
    s.ctx = ctx;
    pthread_mutex_init(&s.mu, NULL);
    pthread_t th[256];
    for (int i = 0; i < nthreads; ++i)
        pthread_create(&th[i], NULL, pf_worker, &s);
    for (int i = 0; i < nthreads; ++i)
        pthread_join(th[i], NULL);
    pthread_mutex_destroy(&s.mu);
}

/* ---- oracle BVH: binned SAH over triangle centroids -------------------- */
#define NBINS 16
#define LEAF_MAX 4

typedef struct
{
    float lo[3], hi[3];
} aabb;

static void aabb_empty(aabb *b)
{
    for (int i = 0; i < 3; ++i)
    {
        b->lo[i] = INFINITY;
        b->hi[i] = -INFINITY;
    }
}
static void aabb_grow(aabb *b, const aabb *o)
{
    for (int i = 0; i < 3; ++i)
    {
        if (o->lo[i] < b->lo[i])
            b->lo[i] = o->lo[i];
        if (o->hi[i] > b->hi[i])
            b->hi[i] = o->hi[i];
    }
}
static double aabb_area(const aabb *b)
{
    double dx = (double)b->hi[0] - b->lo[0], dy = (double)b->hi[1] - b->lo[1],
           dz = (double)b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0)
        return 0;
    return dx * dy + dy * dz + dz * dx;
}

oracle_bvh *oracle_bvh_build(const float *tris, uint32_t ntris)
{
    oracle_bvh *bvh = (oracle_bvh *)calloc(1, sizeof(oracle_bvh));
    bvh->ntris = ntris;
    bvh->tris = (float *)malloc(sizeof(float) * 9 * (size_t)(ntris ? ntris : 1));
    memcpy(bvh->tris, tris, sizeof(float) * 9 * (size_t)ntris);
    bvh->prims = (uint32_t *)malloc(sizeof(uint32_t) * (ntris ? ntris : 1));
    aabb *pb = (aabb *)malloc(sizeof(aabb) * (ntris ? ntris : 1));
    float *cen = (float *)malloc(sizeof(float) * 3 * (ntris ? ntris : 1));
    for (uint32_t i = 0; i < ntris; ++i)
    {
        bvh->prims[i] = i;
        aabb_empty(&pb[i]);
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k)
            {
                float x = tris[9 * (size_t)i + 3 * v + k];
                if (x < pb[i].lo[k])
                    pb[i].lo[k] = x;
                if (x > pb[i].hi[k])
                    pb[i].hi[k] = x;
            }
        for (int k = 0; k < 3; ++k)
            cen[3 * (size_t)i + k] = 0.5f * (pb[i].lo[k] + pb[i].hi[k]);
    }
    uint32_t cap = 2 * (ntris ? ntris : 1);
    bvh->nodes = (obvh_node *)malloc(sizeof(obvh_node) * cap);
    bvh->nnodes = 1;
    typedef struct
    {
        uint32_t node, begin, end, depth;
    } item;
    item *stack = (item *)malloc(sizeof(item) * 256);
    int sp = 0;
    stack[sp++] = (item){ 0, 0, ntris, 1 };
    bvh->depth = 1;
    while (sp > 0)
    {
        item it = stack[--sp];
        obvh_node *n = &bvh->nodes[it.node];
        aabb nb, cb;
        aabb_empty(&nb);
        aabb_empty(&cb);
        for (uint32_t i = it.begin; i < it.end; ++i)
        {
            uint32_t p = bvh->prims[i];
            aabb_grow(&nb, &pb[p]);
            aabb c;
            for (int k = 0; k < 3; ++k)
                c.lo[k] = c.hi[k] = cen[3 * (size_t)p + k];
            aabb_grow(&cb, &c);
        }
        if (it.end == it.begin)
        {
            for (int k = 0; k < 3; ++k)
            {
                nb.lo[k] = 0;
                nb.hi[k] = -1; /* empty: never hit */
            }
        }
        memcpy(n->lo, nb.lo, sizeof nb.lo);
        memcpy(n->hi, nb.hi, sizeof nb.hi);
        if (it.depth > bvh->depth)
            bvh->depth = it.depth;
        uint32_t cnt = it.end - it.begin;
        int best_axis = -1;
        int best_split = 0;
        double best_cost = (double)cnt; /* leaf cost (intersection cost 1) */
        if (cnt > 1)
        {
            for (int axis = 0; axis < 3; ++axis)
            {
                float lo = cb.lo[axis], hi = cb.hi[axis];
                if (!(hi > lo))
                    continue;
                aabb bins[NBINS];
                uint32_t counts[NBINS] = { 0 };
                for (int b = 0; b < NBINS; ++b)
                    aabb_empty(&bins[b]);
                double scale = NBINS / ((double)hi - lo);
                for (uint32_t i = it.begin; i < it.end; ++i)
                {
                    uint32_t p = bvh->prims[i];
                    int b = (int)(((double)cen[3 * (size_t)p + axis] - lo) * scale);
                    if (b >= NBINS)
                        b = NBINS - 1;
                    if (b < 0)
                        b = 0;
                    counts[b]++;
                    aabb_grow(&bins[b], &pb[p]);
                }
                double rarea[NBINS];
                uint32_t rcount[NBINS];
                aabb acc;
                aabb_empty(&acc);
                uint32_t c = 0;
                for (int b = NBINS - 1; b > 0; --b)
                {
                    aabb_grow(&acc, &bins[b]);
                    c += counts[b];
                    rarea[b] = aabb_area(&acc);
                    rcount[b] = c;
                }
                aabb_empty(&acc);
                c = 0;
                double parea = aabb_area(&nb);
                for (int b = 0; b < NBINS - 1; ++b)
                {
                    aabb_grow(&acc, &bins[b]);
                    c += counts[b];
                    if (c == 0 || rcount[b + 1] == 0)
                        continue;
                    double cost = 1.0 + (aabb_area(&acc) * c +
                                         rarea[b + 1] * rcount[b + 1]) /
                                            (parea > 0 ? parea : 1.0);
                    if (cost < best_cost)
                    {
                        best_cost = cost;
                        best_axis = axis;
                        best_split = b;
                    }
                }
            }
        }
        uint32_t mid = it.begin;
        if (best_axis >= 0)
        {
            float lo = cb.lo[best_axis], hi = cb.hi[best_axis];
            double scale = NBINS / ((double)hi - lo);
            uint32_t i = it.begin, j = it.end;
            while (i < j)
            {
                uint32_t p = bvh->prims[i];
                int b = (int)(((double)cen[3 * (size_t)p + best_axis] - lo) * scale);
                if (b >= NBINS)
                    b = NBINS - 1;
                if (b <= best_split)
                    ++i;
                else
                {
                    --j;
                    bvh->prims[i] = bvh->prims[j];
                    bvh->prims[j] = p;
                }
            }
            mid = i;
        }
        else if (cnt > LEAF_MAX)
        {
            /* no useful SAH split: fall back to a median split on the widest
             * centroid axis so leaves stay small */
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis])
                    axis = k;
            mid = it.begin + cnt / 2;
            best_axis = axis;
        }
        if (best_axis < 0 || mid == it.begin || mid == it.end)
        {
            if (cnt > LEAF_MAX && mid != it.begin + cnt / 2)
            {
                mid = it.begin + cnt / 2;
                best_axis = 0;
            }
            else
            {
                n->first = it.begin;
                n->count = (uint16_t)cnt;
                n->axis = 0;
                continue;
            }
        }
        uint32_t l = bvh->nnodes;
        bvh->nnodes += 2;
        n = &bvh->nodes[it.node];
        n->first = l;
        n->count = 0;
        n->axis = (uint16_t)best_axis;
        stack[sp++] = (item){ l + 1, mid, it.end, it.depth + 1 };
        stack[sp++] = (item){ l, it.begin, mid, it.depth + 1 };
    }
    free(stack);
    free(pb);
    free(cen);
    return bvh;
}

void oracle_bvh_free(oracle_bvh *bvh)
{
    if (!bvh)
        return;
    free(bvh->nodes);
    free(bvh->prims);
    free(bvh->tris);
    free(bvh);
}

void oracle_bvh_info(const oracle_bvh *bvh, uint32_t *nodes, uint32_t *depth)
{
    if (nodes)
        *nodes = bvh->nnodes;
    if (depth)
        *depth = bvh->depth;
}

/* ---- scene ------------------------------------------------------------- */
static void *dupmem(const void *p, size_t bytes)
{
    void *q = malloc(bytes ? bytes : 1);
    if (p && bytes)
        memcpy(q, p, bytes);
    else
        memset(q, 0, bytes ? bytes : 1);
    return q;
}

oracle_pt *oracle_pt_create(const oracle_scene_desc *d)
{
    oracle_pt *pt = (oracle_pt *)calloc(1, sizeof(oracle_pt));
    uint32_t M = d->num_meshes;
    pt->num_meshes = M;
    uint32_t nv = 0, nt = 0;
    for (uint32_t m = 0; m < M; ++m)
    {
        uint32_t vend = d->mesh_vertex_offset[m];
        /* vertex count of a mesh = max index + 1 over its triangles */
        uint32_t maxi = 0;
        for (uint32_t t = 0; t < d->mesh_tri_count[m]; ++t)
            for (int k = 0; k < 3; ++k)
            {
                uint32_t v = d->indices[3 * ((size_t)d->mesh_tri_offset[m] + t) + k];
                if (v + 1 > maxi)
                    maxi = v + 1;
            }
        if (vend + maxi > nv)
            nv = vend + maxi;
        if (d->mesh_tri_offset[m] + d->mesh_tri_count[m] > nt)
            nt = d->mesh_tri_offset[m] + d->mesh_tri_count[m];
    }
    pt->vertices = (float *)dupmem(d->vertices, sizeof(float) * 3 * (size_t)nv);
    pt->normals = (float *)dupmem(d->normals, d->normals ? sizeof(float) * 3 * (size_t)nv : 0);
    pt->indices = (uint32_t *)dupmem(d->indices, sizeof(uint32_t) * 3 * (size_t)nt);
    pt->mesh_vertex_offset = (uint32_t *)dupmem(d->mesh_vertex_offset, 4 * (size_t)M);
    pt->mesh_tri_offset = (uint32_t *)dupmem(d->mesh_tri_offset, 4 * (size_t)M);
    pt->mesh_tri_count = (uint32_t *)dupmem(d->mesh_tri_count, 4 * (size_t)M);
    pt->mesh_has_normal = (uint32_t *)dupmem(d->mesh_has_normal, 4 * (size_t)M);
    pt->mesh_light = (uint32_t *)dupmem(d->mesh_light, 4 * (size_t)M);
    pt->mesh_material = (oracle_material *)dupmem(d->mesh_material, sizeof(oracle_material) * (size_t)M);
    pt->num_lights = d->num_lights;
    pt->light_mesh = (uint32_t *)dupmem(d->light_mesh, 4 * (size_t)d->num_lights);
    pt->light_L = (float *)dupmem(d->light_L, 12 * (size_t)d->num_lights);
    pt->light_two_sided = (uint32_t *)dupmem(d->light_two_sided, 4 * (size_t)d->num_lights);
    memcpy(pt->cam_pos, d->cam_pos, 12);
    memcpy(pt->cam_lookat, d->cam_lookat, 12);
    memcpy(pt->cam_up, d->cam_up, 12);
    memcpy(pt->cam_right, d->cam_right, 12);
    memcpy(pt->bg, d->bg, 12);
    if (d->num_textures)
    {
        uint64_t ntex = 0;
        for (uint32_t i = 0; i < d->num_textures; ++i)
        {
            uint64_t end = (uint64_t)d->texture_offset[i] +
                           (uint64_t)d->texture_size[2 * i] * d->texture_size[2 * i + 1];
            if (end > ntex)
                ntex = end;
        }
        pt->num_textures = d->num_textures;
        pt->texcoords = (float *)dupmem(d->texcoords, sizeof(float) * 2 * (size_t)nv);
        pt->mesh_texture = (uint32_t *)dupmem(d->mesh_texture, 4 * (size_t)M);
        pt->texture_offset = (uint32_t *)dupmem(d->texture_offset, 4 * (size_t)d->num_textures);
        pt->texture_size = (uint32_t *)dupmem(d->texture_size, 8 * (size_t)d->num_textures);
        pt->texels = (uint8_t *)dupmem(d->texels, 4 * (size_t)ntex);
    }
    /* gather world-space triangles in global primitive order */
    float *tris = (float *)malloc(sizeof(float) * 9 * (size_t)(nt ? nt : 1));
    for (uint32_t m = 0; m < M; ++m)
        for (uint32_t t = 0; t < d->mesh_tri_count[m]; ++t)
        {
            size_t g = (size_t)d->mesh_tri_offset[m] + t;
            for (int k = 0; k < 3; ++k)
            {
                uint32_t v = d->indices[3 * g + k] + d->mesh_vertex_offset[m];
                memcpy(tris + 9 * g + 3 * k, d->vertices + 3 * (size_t)v, 12);
            }
        }
    pt->bvh = oracle_bvh_build(tris, nt);
    free(tris);
    return pt;
}

void oracle_pt_free(oracle_pt *pt)
{
    if (!pt)
        return;
    free(pt->vertices);
    free(pt->normals);
    free(pt->indices);
    free(pt->mesh_vertex_offset);
    free(pt->mesh_tri_offset);
    free(pt->mesh_tri_count);
    free(pt->mesh_has_normal);
    free(pt->mesh_light);
    free(pt->mesh_material);
    free(pt->light_mesh);
    free(pt->light_L);
    free(pt->light_two_sided);
    free(pt->texcoords);
    free(pt->mesh_texture);
    free(pt->texture_offset);
    free(pt->texture_size);
    free(pt->texels);
    oracle_bvh_free(pt->bvh);
    free(pt);
}

/* ---- film: PathTracing.cu:140-158 + Tonemap.h:34-42 -------------------- */
void oracle_aces_f32(const float rad_in[3], float out[3])
{
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    for (int k = 0; k < 3; ++k)
    {
        float r = rad_in[k] * 0.6f;
        float v = (r * (a * r + b)) / (r * (c * r + d) + e);
        out[k] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    }
}

void oracle_film_update(float *acc, uint8_t *color, const float *this_frame,
                        uint32_t npix, uint32_t frame)
{
    for (uint32_t i = 0; i < npix; ++i)
    {
        /* thisFrame = dvec4(0,0,0,1) + radiance / 1 */
        double tf[4] = { 0.0 + (double)(this_frame[3 * (size_t)i] / 1),
                         0.0 + (double)(this_frame[3 * (size_t)i + 1] / 1),
                         0.0 + (double)(this_frame[3 * (size_t)i + 2] / 1), 1.0 };
        float *dst = acc + 4 * (size_t)i;
        if (frame == 0)
        {
            for (int k = 0; k < 4; ++k)
                dst[k] = (float)tf[k];
        }
        else
        {
            double f0 = (double)((float)frame / ((float)frame + 1.0f));
            double f1 = (double)(1.0f / ((float)frame + 1.0f));
            for (int k = 0; k < 4; ++k)
                dst[k] = (float)((double)dst[k] * f0 + tf[k] * f1);
        }
        float rad[3], tm[3];
        for (int k = 0; k < 3; ++k)
        {
            float v = dst[k];
            rad[k] = v < 0.f ? 0.f : (v > 1e30f ? 1e30f : v);
        }
        oracle_aces_f32(rad, tm);
        uint8_t *c = color + 4 * (size_t)i;
        for (int k = 0; k < 3; ++k)
            c[k] = (uint8_t)(tm[k] * 255.0f);
        c[3] = (uint8_t)(1.0f * 255.0f);
    }
}
