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
} f64_ctx;

static void f64_task(void *p, uint32_t b, uint32_t e)
{
    f64_ctx *fc = (f64_ctx *)p;
    for (uint32_t i = b; i < e; ++i)
    {
        uint32_t x = i % fc->c->width, y = i / fc->c->width;
        double acc[3] = { 0, 0, 0 };
        for (uint32_t f = 0; f < fc->nframes; ++f)
        {
            V3 r;
            pt_sample(fc->pt, fc->c, x, y, fc->frame0 + f, &r, NULL);
            acc[0] += r.x;
            acc[1] += r.y;
            acc[2] += r.z;
        }
        for (int k = 0; k < 3; ++k)
            fc->out[3 * (size_t)i + k] = acc[k] / (double)fc->nframes;
    }
}

void oracle_pt_render_f64(const oracle_pt *pt, const oracle_pt_config *c,
                          uint32_t frame0, uint32_t nframes, double *rgb_mean,
                          int nthreads)
{
    f64_ctx fc = { pt, c, frame0, nframes, rgb_mean };
    if (pt->bvh->ntris == 0)
    {
        memset(rgb_mean, 0, sizeof(double) * 3 * (size_t)c->width * c->height);
        return;
    }
    oracle_parallel_for(c->width * c->height, 4, nthreads, f64_task, &fc);
}
