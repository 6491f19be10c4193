/*
 * sanitize_check.c — drives every entry point of the CPU oracle (TEST
 * INFRASTRUCTURE ONLY; see oracle.h) on a small procedural scene, for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`, run by tests/test_sanitize.py).  Any
 * out-of-bounds access, use after free, leak, misaligned access, signed
 * overflow or invalid shift aborts the program with the sanitizer's report.
 *
 * Scene: a closed room (floor, ceiling, three walls; the front is open), a
 * two-sided and a one-sided quad light, a tilted box with per-vertex
 * normals, one textured wall, and a glass slab (trans 1) so the refraction
 * and total-internal-reflection branches run.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXV 256
#define MAXT 256
#define MAXM 16

static float V[3 * MAXV], N[3 * MAXV], UV[2 * MAXV];
static uint32_t I[3 * MAXT];
static uint32_t mv_off[MAXM], mt_off[MAXM], mt_cnt[MAXM], m_hasn[MAXM], m_light[MAXM],
    m_tex[MAXM];
static oracle_material mats[MAXM];
static uint32_t nv, nt, nm;

static void begin_mesh(oracle_material m, int has_normal)
{
    mv_off[nm] = nv;
    mt_off[nm] = nt;
    mt_cnt[nm] = 0;
    m_hasn[nm] = (uint32_t)has_normal;
    m_light[nm] = 0xFFFFFFFFu;
    m_tex[nm] = 0xFFFFFFFFu;
    mats[nm] = m;
    ++nm;
}

static void quad(const float p[4][3], const float n[3])
{
    const uint32_t b = nv - mv_off[nm - 1];
    for (int k = 0; k < 4; ++k)
    {
        memcpy(V + 3 * nv, p[k], sizeof(float) * 3);
        memcpy(N + 3 * nv, n, sizeof(float) * 3);
        UV[2 * nv] = (float)(k == 1 || k == 2);
        UV[2 * nv + 1] = (float)(k >= 2) * 1.5f;
        ++nv;
    }
    const uint32_t t[6] = { b, b + 1, b + 2, b, b + 2, b + 3 };
    memcpy(I + 3 * nt, t, sizeof t);
    nt += 2;
    mt_cnt[nm - 1] += 2;
}

static oracle_material mat(float r, float g, float b, float metallic, float rough, float trans)
{
    oracle_material m;
    memset(&m, 0, sizeof m);
    m.color[0] = r;
    m.color[1] = g;
    m.color[2] = b;
    m.metallic = metallic;
    m.roughness = rough;
    m.eta = 1.5f;
    m.trans = trans;
    m.sheenTint = 0.5f;
    m.clearcoatgloss = 1.0f;
    m.specular = 0.5f;
    return m;
}

static void box(float x0, float y0, float z0, float x1, float y1, float z1)
{
    const float f[6][4][3] = {
        { { x0, y0, z1 }, { x1, y0, z1 }, { x1, y1, z1 }, { x0, y1, z1 } },
        { { x1, y0, z0 }, { x0, y0, z0 }, { x0, y1, z0 }, { x1, y1, z0 } },
        { { x0, y1, z1 }, { x1, y1, z1 }, { x1, y1, z0 }, { x0, y1, z0 } },
        { { x0, y0, z0 }, { x1, y0, z0 }, { x1, y0, z1 }, { x0, y0, z1 } },
        { { x1, y0, z1 }, { x1, y0, z0 }, { x1, y1, z0 }, { x1, y1, z1 } },
        { { x0, y0, z0 }, { x0, y0, z1 }, { x0, y1, z1 }, { x0, y1, z0 } },
    };
    const float n[6][3] = { { 0, 0, 1 }, { 0, 0, -1 }, { 0, 1, 0 },
                            { 0, -1, 0 }, { 1, 0, 0 }, { -1, 0, 0 } };
    for (int k = 0; k < 6; ++k)
        quad(f[k], n[k]);
}

static uint32_t rng_state = 12345u;
static float frand(void)
{
    rng_state = rng_state * 1664525u + 1013904223u;
    return (float)(rng_state >> 8) * (1.0f / 16777216.0f);
}

static int check_pt(void)
{
    /* room */
    begin_mesh(mat(0.7f, 0.7f, 0.7f, 0.f, 0.5f, 0.f), 0);
    const float fl[4][3] = { { -1, 0, 1 }, { 1, 0, 1 }, { 1, 0, -1 }, { -1, 0, -1 } };
    const float ce[4][3] = { { -1, 2, -1 }, { 1, 2, -1 }, { 1, 2, 1 }, { -1, 2, 1 } };
    const float up[3] = { 0, 1, 0 }, dn[3] = { 0, -1, 0 };
    quad(fl, up);
    quad(ce, dn);
    begin_mesh(mat(0.6f, 0.2f, 0.2f, 0.3f, 0.3f, 0.f), 0); /* textured back wall */
    const float bw[4][3] = { { -1, 0, -1 }, { 1, 0, -1 }, { 1, 2, -1 }, { -1, 2, -1 } };
    const float fz[3] = { 0, 0, 1 };
    quad(bw, fz);
    m_tex[nm - 1] = 0;
    begin_mesh(mat(0.2f, 0.6f, 0.2f, 1.f, 0.05f, 0.f), 0);
    const float lw[4][3] = { { -1, 0, 1 }, { -1, 0, -1 }, { -1, 2, -1 }, { -1, 2, 1 } };
    const float rw[4][3] = { { 1, 0, -1 }, { 1, 0, 1 }, { 1, 2, 1 }, { 1, 2, -1 } };
    const float px[3] = { 1, 0, 0 }, nx[3] = { -1, 0, 0 };
    quad(lw, px);
    quad(rw, nx);
    /* lights: a two-sided quad under the ceiling and a one-sided one on a wall */
    begin_mesh(mat(0, 0, 0, 0, 0.5f, 0), 0);
    const float l0[4][3] = { { -0.3f, 1.98f, -0.3f }, { 0.3f, 1.98f, -0.3f },
                             { 0.3f, 1.98f, 0.3f }, { -0.3f, 1.98f, 0.3f } };
    quad(l0, dn);
    m_light[nm - 1] = 0;
    begin_mesh(mat(0, 0, 0, 0, 0.5f, 0), 0);
    const float l1[4][3] = { { 0.99f, 0.5f, -0.2f }, { 0.99f, 0.5f, 0.2f },
                             { 0.99f, 0.9f, 0.2f }, { 0.99f, 0.9f, -0.2f } };
    quad(l1, nx);
    m_light[nm - 1] = 1;
    /* a box with per-vertex normals, and a glass slab */
    begin_mesh(mat(0.8f, 0.8f, 0.3f, 0.2f, 0.8f, 0.f), 1);
    box(-0.7f, 0.0f, -0.6f, -0.2f, 0.8f, -0.1f);
    for (uint32_t v = mv_off[nm - 1]; v < nv; ++v)
    {
        N[3 * v] += 0.1f;
        const float l = sqrtf(N[3 * v] * N[3 * v] + N[3 * v + 1] * N[3 * v + 1] +
                              N[3 * v + 2] * N[3 * v + 2]);
        for (int k = 0; k < 3; ++k)
            N[3 * v + k] /= l;
    }
    begin_mesh(mat(1, 1, 1, 0, 0.02f, 1.f), 0);
    box(0.1f, 0.0f, -0.2f, 0.7f, 0.6f, 0.4f);

    static uint8_t texels[4 * 5 * 3];
    for (int k = 0; k < 4 * 5 * 3; ++k)
        texels[k] = (uint8_t)(k * 37);
    const uint32_t toff[1] = { 0 }, tsize[2] = { 5, 3 };
    const uint32_t light_mesh[2] = { 4, 5 }, two_sided[2] = { 1, 0 };
    const float light_L[6] = { 17, 12, 4, 5, 5, 5 };
    oracle_scene_desc d;
    memset(&d, 0, sizeof d);
    d.num_meshes = nm;
    d.vertices = V;
    d.normals = N;
    d.indices = I;
    d.mesh_vertex_offset = mv_off;
    d.mesh_tri_offset = mt_off;
    d.mesh_tri_count = mt_cnt;
    d.mesh_has_normal = m_hasn;
    d.mesh_light = m_light;
    d.mesh_material = mats;
    d.num_lights = 2;
    d.light_mesh = light_mesh;
    d.light_L = light_L;
    d.light_two_sided = two_sided;
    const float cp[3] = { 0, 1, 3.4f }, cl[3] = { 0, 0, -1 }, cu[3] = { 0, 0.4f, 0 },
                cr[3] = { 0.4f, 0, 0 };
    memcpy(d.cam_pos, cp, sizeof cp);
    memcpy(d.cam_lookat, cl, sizeof cl);
    memcpy(d.cam_up, cu, sizeof cu);
    memcpy(d.cam_right, cr, sizeof cr);
    d.bg[0] = d.bg[1] = d.bg[2] = 0.1f;
    d.num_textures = 1;
    d.texcoords = UV;
    d.mesh_texture = m_tex;
    d.texture_offset = toff;
    d.texture_size = tsize;
    d.texels = texels;
    oracle_pt *pt = oracle_pt_create(&d);
    if (!pt)
        return 1;

    const uint32_t W = 24, H = 16;
    for (uint32_t strat = 0; strat < 3; ++strat)
    {
        oracle_pt_config c = { W, H, 12, 2, 0.8f, 1e-3f, strat };
        float *img = malloc(sizeof(float) * 3 * W * H);
        uint64_t rays = 0;
        oracle_pt_render_frame_f32(pt, &c, strat + 1, img, 3, &rays);
        double s = 0;
        for (uint32_t k = 0; k < 3 * W * H; ++k)
            s += img[k];
        printf("strategy %u: mean %.4f, %llu rays\n", strat, s / (3.0 * W * H),
               (unsigned long long)rays);
        free(img);
    }
    oracle_pt_config c = { W, H, 25, 5, 0.9f, 1e-3f, 2 };
    float tr[64 * ORACLE_TRACE_STRIDE];
    for (uint32_t k = 0; k < 200; ++k)
    {
        float rgb[3];
        uint32_t sh = 0;
        oracle_pt_sample_f32(pt, &c, k % W, (k / W) % H, k, rgb, &sh);
        oracle_pt_trace_f32(pt, &c, (k * 7) % W, (k * 3) % H, k + 99, rgb, tr, 64);
        oracle_pt_trace_f32(pt, &c, (k * 5) % W, (k * 11) % H, k + 7, rgb, tr, 2);
    }
    double *mean = malloc(sizeof(double) * 3 * 5 * 4), *var = malloc(sizeof(double) * 3 * 5 * 4);
    oracle_pt_render_f64_window(pt, &c, W - 5, H - 4, 5, 4, 1, 1000, 6, mean, var, 2);
    oracle_pt_render_f64_window(pt, &c, 1, 0, 5, 4, 3, 1000, 2, mean, NULL, 2);
    double *full = malloc(sizeof(double) * 3 * W * H);
    oracle_pt_render_f64(pt, &c, 5, 2, full, 4);
    /* film update + ACES on a frame */
    float *rad = calloc(4 * W * H, sizeof(float));
    uint8_t *col = calloc(4 * W * H, 1);
    float *frame = malloc(sizeof(float) * 3 * W * H);
    for (uint32_t k = 0; k < 3 * W * H; ++k)
        frame[k] = (float)full[k];
    for (uint32_t f = 0; f < 3; ++f)
        oracle_film_update(rad, col, frame, W * H, f);
    /* examples/PathTracer program on the same meshes */
    float mc[3 * MAXM], me[3 * MAXM];
    for (uint32_t m = 0; m < nm; ++m)
        for (int k = 0; k < 3; ++k)
        {
            mc[3 * m + k] = mats[m].color[k];
            me[3 * m + k] = m_light[m] != 0xFFFFFFFFu ? 10.0f : 0.0f;
        }
    oracle_example_config ec;
    memset(&ec, 0, sizeof ec);
    ec.width = 12;
    ec.height = 10;
    ec.sample_num = 3;
    ec.max_depth = 9;
    memcpy(ec.cam_pos, cp, sizeof cp);
    memcpy(ec.cam_gaze, cl, sizeof cl);
    memcpy(ec.cam_up, up, sizeof up);
    ec.horizontal_ratio = 0.6f;
    ec.vertical_ratio = 0.5f;
    float *ex = malloc(sizeof(float) * 4 * 10 * 12);
    oracle_example_render(pt, mc, me, &ec, 2, 12, ex, 2);
    free(ex);
    free(mean);
    free(var);
    free(full);
    free(rad);
    free(col);
    free(frame);
    oracle_pt_free(pt);
    return 0;
}

static int check_traversal(void)
{
    enum { T = 40, R = 500 };
    float *tris = malloc(sizeof(float) * 9 * T);
    for (int k = 0; k < 9 * T; ++k)
        tris[k] = frand();
    oracle_ray *rays = malloc(sizeof(oracle_ray) * R);
    for (int k = 0; k < R; ++k)
    {
        for (int j = 0; j < 3; ++j)
        {
            rays[k].o[j] = frand() * 1.4f - 0.2f;
            rays[k].d[j] = frand() * 2 - 1;
        }
        rays[k].tmin = k % 7 ? 0.f : 0.3f;
        rays[k].tmax = k % 5 ? INFINITY : 0.5f;
    }
    oracle_hit *h0 = malloc(sizeof(oracle_hit) * R), *h1 = malloc(sizeof(oracle_hit) * R);
    uint8_t *o0 = malloc(R), *o1 = malloc(R);
    oracle_trace_closest_brute(tris, T, rays, R, h0, 3);
    oracle_trace_any_brute(tris, T, rays, R, o0, 3);
    oracle_bvh *b = oracle_bvh_build(tris, T);
    uint32_t nodes, depth;
    oracle_bvh_info(b, &nodes, &depth);
    oracle_bvh_trace_closest(b, rays, R, h1, 3);
    oracle_bvh_trace_any(b, rays, R, o1, 3);
    int bad = 0;
    for (int k = 0; k < R; ++k)
        bad += h0[k].prim != h1[k].prim || o0[k] != o1[k];
    printf("traversal: %d rays, %u nodes, depth %u, %d mismatches\n", R, nodes, depth, bad);
    oracle_bvh_free(b);
    oracle_bvh *e = oracle_bvh_build(tris, 0);
    if (e)
        oracle_bvh_free(e);
    free(tris);
    free(rays);
    free(h0);
    free(h1);
    free(o0);
    free(o1);
    return bad != 0;
}

static int check_bsdf(void)
{
    const float Ns[3] = { 0, 0, 1 }, Ng[3] = { 0, 0, 1 };
    const double Nsd[3] = { 0, 0, 1 }, Ngd[3] = { 0, 0, 1 };
    for (int k = 0; k < 64; ++k)
    {
        oracle_material m = mat(frand(), frand(), frand(), frand(), 0.02f + frand(),
                                k % 3 == 0 ? 1.f : 0.f);
        m.clearcoat = k % 4 == 0 ? 1.f : 0.f;
        m.sheen = frand();
        m.subsurface = frand();
        m.eta = k % 2 ? 1.5f : 1.0f / 1.5f;
        const float z = k % 5 == 0 ? -0.4f : 0.6f;
        const float Vv[3] = { sqrtf(1 - z * z), 0, z };
        const double Vd[3] = { Vv[0], 0, z };
        uint32_t seeds[16];
        for (int j = 0; j < 16; ++j)
            seeds[j] = (uint32_t)(k * 977 + j);
        float L[48], pdf[16], f[48];
        double Ld[48], pdfd[16], fd[48];
        oracle_disney_sample_batch_f32(&m, Ns, Ng, Vv, seeds, 16, L);
        oracle_disney_pdf_batch_f32(&m, Ns, Ng, Vv, L, 16, pdf);
        oracle_disney_eval_batch_f32(&m, Ns, Ng, Vv, L, 16, f);
        oracle_disney_sample_batch_f64(&m, Nsd, Ngd, Vd, seeds, 16, Ld);
        oracle_disney_pdf_batch_f64(&m, Nsd, Ngd, Vd, Ld, 16, pdfd);
        oracle_disney_eval_batch_f64(&m, Nsd, Ngd, Vd, Ld, 16, fd);
        uint32_t s = (uint32_t)k;
        float l1[3], e1[3];
        oracle_disney_sample_f32(&m, Ns, Ng, Vv, &s, l1);
        oracle_disney_eval_f32(&m, Ns, Ng, Vv, l1, e1);
        (void)oracle_disney_pdf_f32(&m, Ns, Ng, Vv, l1);
    }
    const float a[6] = { 0.3f, 0.2f, 0.9f, 0.f, 0.f, 1.f };
    const double ad[6] = { 0.3, 0.2, 0.9, 0., 0., 1. };
    for (int w = 0; w < 6; ++w)
    {
        (void)oracle_material_math_f32(w, a);
        (void)oracle_material_math_f64(w, ad);
    }
    float u[3], v[3];
    const float nn[3] = { 0, 0, -1 };
    oracle_onb_f32(nn, u, v);
    const uint8_t img[4 * 2 * 2] = { 1, 2, 3, 4, 50, 60, 70, 80, 9, 9, 9, 9, 200, 100, 0, 255 };
    float t[3];
    oracle_tex_bilinear_f32(img, 2, 2, -3.25f, 7.5f, t);
    oracle_tex_bilinear_f32(img, 2, 2, NAN, INFINITY, t);
    const float r[3] = { 0.5f, 4.f, 1e6f };
    oracle_aces_f32(r, t);
    uint32_t st = 7;
    for (int k = 0; k < 10; ++k)
        (void)oracle_rnd(&st);
    uint32_t v1;
    const uint32_t key[4] = { 1, 2, 3, 4 };
    (void)oracle_tea(1, 2, key, 32, &v1);
    (void)oracle_tea4(5, 6);
    return 0;
}

static int check_guide(void)
{
    double *p = malloc(sizeof(double) * ORACLE_GUIDE_NPARAMS);
    double *m = calloc(ORACLE_GUIDE_NPARAMS, sizeof(double));
    double *v = calloc(ORACLE_GUIDE_NPARAMS, sizeof(double));
    double *g = malloc(sizeof(double) * ORACLE_GUIDE_NPARAMS);
    for (int k = 0; k < ORACLE_GUIDE_NPARAMS; ++k)
        p[k] = (frand() - 0.5f) * 0.2;
    enum { S = 40 };
    float *samples = malloc(sizeof(float) * 16 * S);
    for (int k = 0; k < 16 * S; ++k)
        samples[k] = frand();
    for (int k = 0; k < S; ++k)
    {
        float *w = samples + 16 * k + 9;
        const float l = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) + 1e-6f;
        w[0] /= l;
        w[1] /= l;
        w[2] /= l;
    }
    double out[32], x[64], pi[8], kap[8], mu[24], w[3];
    oracle_guide_encode(samples, x);
    oracle_guide_forward(p, samples, 1, out);
    oracle_guide_forward(p, samples, 0, out);
    oracle_guide_decode(out, pi, kap, mu);
    oracle_guide_sample(out, 0.3, 0.7, 0.2, w);
    (void)oracle_guide_pdf(out, w);
    double us[9] = { 0.1, 0.2, 0.3, 0.9, 0.5, 0.99, 0.0, 0.0, 0.0 }, dirs[9], pdfs[3];
    oracle_guide_sample_batch(out, us, 3, dirs);
    oracle_guide_pdf_batch(out, dirs, 3, pdfs);
    for (uint32_t step = 1; step <= 2; ++step)
    {
        (void)oracle_guide_loss_grad(p, samples, S, (int)step - 1, g);
        oracle_guide_adam(p, m, v, g, step, 1e-3, 0.9, 0.99, 1e-8);
    }
    (void)oracle_bf16_round(3.1415926f);
    free(p);
    free(m);
    free(v);
    free(g);
    free(samples);
    return 0;
}

int main(void)
{
    int bad = check_traversal();
    bad |= check_bsdf();
    bad |= check_pt();
    bad |= check_guide();
    printf(bad ? "SANITIZE_CHECK_FAILED\n" : "SANITIZE_CHECK_OK\n");
    return bad;
}
