/*
 * oracle_f64.c — FP64 instantiation of the CPU oracle (TEST INFRASTRUCTURE
 * ONLY; see oracle.h): the same estimator in double precision, averaged over
 * many frames, is the converged-image reference for relMSE.
 */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX(n) n##_f64
#define K(x) x
#define RSQRT sqrt
#define RSIN sin
#define RCOS cos
#define RLOG log
#define RABS fabs
#define RCOPYSIGN copysign
#define RFMAX fmax
#define RISNAN isnan
#define RISINF isinf
#define RFLOOR floor

#include "oracle_impl.inc"

typedef struct
{
    const oracle_pt *pt;
    const oracle_pt_config *c;
    uint32_t frame0, nframes;
    double *out;
    uint32_t x0, y0, ww; /* window origin and width (whole image: 0, 0, W) */
    uint32_t stride;     /* pixel step inside the window (1: every pixel)  */
    double *var;         /* optional per-pixel variance of one sample      */
} f64_ctx;

static void f64_task(void *p, uint32_t b, uint32_t e)
{
    f64_ctx *fc = (f64_ctx *)p;
    for (uint32_t i = b; i < e; ++i)
    {
        uint32_t x = fc->x0 + fc->stride * (i % fc->ww), y = fc->y0 + fc->stride * (i / fc->ww);
        double acc[3] = { 0, 0, 0 }, acc2[3] = { 0, 0, 0 };
        for (uint32_t f = 0; f < fc->nframes; ++f)
        {
            V3 r;
            pt_sample(fc->pt, fc->c, x, y, fc->frame0 + f, &r, NULL);
            acc[0] += r.x;
            acc[1] += r.y;
            acc[2] += r.z;
            acc2[0] += r.x * r.x;
            acc2[1] += r.y * r.y;
            acc2[2] += r.z * r.z;
        }
        const double n = (double)fc->nframes;
        for (int k = 0; k < 3; ++k)
        {
            const double m = acc[k] / n;
            fc->out[3 * (size_t)i + k] = m;
            if (fc->var)
                fc->var[3 * (size_t)i + k] = n > 1 ? (acc2[k] - n * m * m) / (n - 1) : 0.0;
        }
    }
}

void oracle_pt_render_f64(const oracle_pt *pt, const oracle_pt_config *c,
                          uint32_t frame0, uint32_t nframes, double *rgb_mean,
                          int nthreads)
{
    oracle_pt_render_f64_window(pt, c, 0, 0, c->width, c->height, 1, frame0, nframes,
                                rgb_mean, NULL, nthreads);
}

void oracle_pt_render_f64_window(const oracle_pt *pt, const oracle_pt_config *c, uint32_t x0,
                                 uint32_t y0, uint32_t ww, uint32_t wh, uint32_t stride,
                                 uint32_t frame0, uint32_t nframes, double *rgb_mean,
                                 double *rgb_var, int nthreads)
{
    f64_ctx fc = { pt, c, frame0, nframes, rgb_mean, x0, y0, ww, stride ? stride : 1, rgb_var };
    if (pt->bvh->ntris == 0)
    {
        memset(rgb_mean, 0, sizeof(double) * 3 * (size_t)ww * wh);
        if (rgb_var)
            memset(rgb_var, 0, sizeof(double) * 3 * (size_t)ww * wh);
        return;
    }
    oracle_parallel_for(ww * wh, 1, nthreads, f64_task, &fc);
}
