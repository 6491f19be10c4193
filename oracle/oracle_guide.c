/*
 * oracle_guide.c — FP64 restatement of the learned path-guiding network.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no learned
 * component; this follows the published methods the library's GPU kernels
 * implement:
 *   - von Mises-Fisher lobes on S^2, density kappa / (2 pi (1 - e^{-2 kappa}))
 *     * exp(kappa (mu.w - 1)), and the numerically stable inversion
 *     w_z = 1 + log(u + (1 - u) e^{-2 kappa}) / kappa (W. Jakob, "Numerically
 *     stable sampling of the von Mises Fisher distribution on S^2", 2012);
 *   - the KL-divergence training objective of neural importance sampling
 *     (Mueller et al., "Neural Importance Sampling", TOG 2019): per sample
 *     -F/p * log q(w), F/p carried as the sample weight;
 *   - Adam (Kingma & Ba, ICLR 2015).
 * The frame around mu is the Duff et al. 2017 orthonormal basis that the
 * reference uses for its BSDF samples (src/Frontend/Device/Common.h:102-116).
 *
 * Pinning: the gradient is checked against central finite differences of the
 * loss, the mixture density against numerical integration over the sphere
 * and the sampler against a chi-square histogram of its density
 * (tests/test_guide_oracle.py).
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>

#include "oracle.h"

#define NIN 64
#define NH 64
#define NOUT 32
#define NL 5

static const double PI_D = 3.14159265358979323846;

static int layer_out(int l) { return l < 4 ? NH : NOUT; }
static size_t w_off(int l) { return (size_t)l * (NH * NIN + NH); }
static size_t b_off(int l) { return w_off(l) + (size_t)layer_out(l) * NIN; }

float oracle_bf16_round(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) /* inf / nan: truncate */
        u &= 0xFFFF0000u;
    else
        u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

/* emulate: the value as the GPU kernels hold it in a BF16 operand */
static double bf(double x, int emulate) { return emulate ? (double)oracle_bf16_round((float)x) : x; }
/* emulate: the value as an FP32 register holds it */
static double f32(double x, int emulate) { return emulate ? (double)(float)x : x; }

void oracle_guide_encode(const float q[9], double x[64])
{
    for (int e = 0; e < 64; ++e)
    {
        const int g = (e % 16) / 4, p = 2 * (e / 16) + (e % 4) / 2, odd = e % 2;
        x[e] = 0.0;
        if (g == 3)
            continue;
        if (p < 6)
        {
            const double a = ldexp(PI_D, p) * (double)q[g];
            x[e] = odd ? cos(a) : sin(a);
        }
        else if (p == 6)
            x[e] = odd ? (double)q[6 + g] : (double)q[3 + g];
    }
}

/* Forward pass keeping every layer input (act[l][64], the input of layer l)
 * and the outputs. */
static void forward_keep(const double *P, const float q[9], int em, double act[NL][NIN],
                         double out[NOUT])
{
    double x[NIN];
    oracle_guide_encode(q, x);
    for (int i = 0; i < NIN; ++i)
        act[0][i] = bf(x[i], em);
    for (int l = 0; l < NL; ++l)
    {
        const int no = layer_out(l);
        const double *W = P + w_off(l), *b = P + b_off(l);
        for (int o = 0; o < no; ++o)
        {
            double z = f32(b[o], em);
            for (int i = 0; i < NIN; ++i)
                z += bf(W[(size_t)o * NIN + i], em) * act[l][i];
            if (l < NL - 1)
                act[l + 1][o] = bf(z > 0.0 ? z : 0.0, em);
            else
                out[o] = z;
        }
    }
}

void oracle_guide_forward(const double *params, const float q[9], int emulate_bf16,
                          double out[32])
{
    double act[NL][NIN];
    forward_keep(params, q, emulate_bf16, act, out);
}

static int row_of(int lobe, int param) { return 16 * (lobe % 2) + 4 * (lobe / 2) + param; }

static double softplus(double k) { return k > 30.0 ? k : log1p(exp(k)); }
static double sigmoid(double k) { return 1.0 / (1.0 + exp(-k)); }

/* log of the vMF normalisation kappa / (2 pi (1 - e^{-2 kappa})). */
static double log_norm(double kappa)
{
    if (kappa < 1e-4)
        return log((1.0 + kappa + kappa * kappa / 3.0) / (4.0 * PI_D));
    return log(kappa) - log(2.0 * PI_D) - log(-expm1(-2.0 * kappa));
}

/* d log_norm / d kappa = 1/kappa - 2 / (e^{2 kappa} - 1). */
static double dlog_norm(double kappa)
{
    if (kappa < 1e-3)
        return 1.0 - kappa / 3.0 + kappa * kappa * kappa / 45.0;
    return 1.0 / kappa - 2.0 / expm1(2.0 * kappa);
}

void oracle_guide_decode(const double out[32], double pi[8], double kappa[8], double mu[24])
{
    double amax = -1e300;
    for (int j = 0; j < 8; ++j)
        if (out[row_of(j, 0)] > amax)
            amax = out[row_of(j, 0)];
    double s = 0.0;
    for (int j = 0; j < 8; ++j)
    {
        pi[j] = exp(out[row_of(j, 0)] - amax);
        s += pi[j];
    }
    for (int j = 0; j < 8; ++j)
    {
        pi[j] /= s;
        kappa[j] = softplus(out[row_of(j, 1)]);
        const double th = out[row_of(j, 2)], ph = out[row_of(j, 3)];
        mu[3 * j] = sin(th) * cos(ph);
        mu[3 * j + 1] = sin(th) * sin(ph);
        mu[3 * j + 2] = cos(th);
    }
}

double oracle_guide_pdf(const double out[32], const double w[3])
{
    double pi[8], kappa[8], mu[24];
    oracle_guide_decode(out, pi, kappa, mu);
    double q = 0.0;
    for (int j = 0; j < 8; ++j)
    {
        const double c = mu[3 * j] * w[0] + mu[3 * j + 1] * w[1] + mu[3 * j + 2] * w[2];
        q += pi[j] * exp(log_norm(kappa[j]) + kappa[j] * (c - 1.0));
    }
    return q;
}

void oracle_guide_sample(const double out[32], double u0, double u1, double u2, double w[3])
{
    double pi[8], kappa[8], mu[24];
    oracle_guide_decode(out, pi, kappa, mu);
    int j = 7;
    double c = 0.0;
    for (int k = 0; k < 8; ++k)
    {
        c += pi[k];
        if (u0 < c)
        {
            j = k;
            break;
        }
    }
    const double kap = kappa[j];
    /* w_z = 1 + t, t = log(u + (1 - u) e^{-2 kappa}) / kappa via log1p/expm1;
     * 1 - w_z^2 = -t (2 + t) */
    double t;
    if (kap < 1e-4)
        t = 2.0 * u1 - 2.0;
    else
        t = log1p((1.0 - u1) * expm1(-2.0 * kap)) / kap;
    if (t > 0.0)
        t = 0.0;
    if (t < -2.0)
        t = -2.0;
    const double wz = 1.0 + t;
    const double r = sqrt(fmax(0.0, -t * (2.0 + t)));
    const double phi = 2.0 * PI_D * u2;
    const double lx = r * cos(phi), ly = r * sin(phi);
    /* Duff et al. 2017 basis around mu (Common.h:102-116) */
    const double x = mu[3 * j], y = mu[3 * j + 1], z = mu[3 * j + 2];
    const double sign = copysign(1.0, z);
    const double a = -1.0 / (sign + z);
    const double b = x * y * a;
    const double T[3] = { 1.0 + sign * x * x * a, sign * b, -sign * x };
    const double B[3] = { b, sign + y * y * a, -y };
    for (int k = 0; k < 3; ++k)
        w[k] = lx * T[k] + ly * B[k] + wz * mu[3 * j + k];
}

void oracle_guide_pdf_batch(const double out[32], const double *dirs, uint32_t n, double *pdf)
{
    for (uint32_t i = 0; i < n; ++i)
        pdf[i] = oracle_guide_pdf(out, dirs + 3 * (size_t)i);
}

void oracle_guide_sample_batch(const double out[32], const double *u, uint32_t n, double *dirs)
{
    for (uint32_t i = 0; i < n; ++i)
        oracle_guide_sample(out, u[3 * (size_t)i], u[3 * (size_t)i + 1], u[3 * (size_t)i + 2],
                            dirs + 3 * (size_t)i);
}

/* Loss of one sample and d loss / d outputs (NOUT). */
static double sample_loss_grad(const double out[NOUT], const float *s, double g[NOUT])
{
    const double wgt = s[12];
    const double om[3] = { s[9], s[10], s[11] };
    double pi[8], kappa[8], mu[24], lv[8];
    oracle_guide_decode(out, pi, kappa, mu);
    double lmax = -1e300;
    for (int j = 0; j < 8; ++j)
    {
        const double c = mu[3 * j] * om[0] + mu[3 * j + 1] * om[1] + mu[3 * j + 2] * om[2];
        lv[j] = log(pi[j]) + log_norm(kappa[j]) + kappa[j] * (c - 1.0);
        if (lv[j] > lmax)
            lmax = lv[j];
    }
    double sum = 0.0;
    for (int j = 0; j < 8; ++j)
        sum += exp(lv[j] - lmax);
    const double logq = lmax + log(sum);
    for (int r = 0; r < NOUT; ++r)
        g[r] = 0.0;
    for (int j = 0; j < 8; ++j)
    {
        const double resp = exp(lv[j] - logq); /* pi_j v_j / q */
        const double th = out[row_of(j, 2)], ph = out[row_of(j, 3)];
        const double c = mu[3 * j] * om[0] + mu[3 * j + 1] * om[1] + mu[3 * j + 2] * om[2];
        const double dth[3] = { cos(th) * cos(ph), cos(th) * sin(ph), -sin(th) };
        const double dph[3] = { -sin(th) * sin(ph), sin(th) * cos(ph), 0.0 };
        const double wdth = om[0] * dth[0] + om[1] * dth[1] + om[2] * dth[2];
        const double wdph = om[0] * dph[0] + om[1] * dph[1] + om[2] * dph[2];
        g[row_of(j, 0)] = -wgt * (resp - pi[j]);
        g[row_of(j, 1)] = -wgt * resp * (dlog_norm(kappa[j]) + c - 1.0) *
                          sigmoid(out[row_of(j, 1)]);
        g[row_of(j, 2)] = -wgt * resp * kappa[j] * wdth;
        g[row_of(j, 3)] = -wgt * resp * kappa[j] * wdph;
    }
    return -wgt * logq;
}

double oracle_guide_loss_grad(const double *params, const float *samples, uint32_t n,
                              int emulate_bf16, double *grad)
{
    const int em = emulate_bf16;
    memset(grad, 0, sizeof(double) * ORACLE_GUIDE_NPARAMS);
    if (n == 0)
        return 0.0;
    double loss = 0.0;
    double act[NL][NIN], out[NOUT], dcur[NH], dnext[NIN];
    for (uint32_t s = 0; s < n; ++s)
    {
        const float *smp = samples + 16 * (size_t)s;
        forward_keep(params, smp, em, act, out);
        double g[NOUT];
        loss += sample_loss_grad(out, smp, g);
        for (int r = 0; r < NOUT; ++r)
            dcur[r] = bf(g[r], em);
        for (int l = NL - 1; l >= 0; --l)
        {
            const int no = layer_out(l);
            double *gW = grad + w_off(l), *gb = grad + b_off(l);
            const double *W = params + w_off(l);
            for (int o = 0; o < no; ++o)
            {
                gb[o] += dcur[o];
                for (int i = 0; i < NIN; ++i)
                    gW[(size_t)o * NIN + i] += dcur[o] * act[l][i];
            }
            if (l == 0)
                break;
            for (int i = 0; i < NIN; ++i)
            {
                double d = 0.0;
                for (int o = 0; o < no; ++o)
                    d += bf(W[(size_t)o * NIN + i], em) * dcur[o];
                /* ReLU of the previous layer: its output act[l][i] > 0 */
                dnext[i] = act[l][i] > 0.0 ? bf(d, em) : 0.0;
            }
            memcpy(dcur, dnext, sizeof dnext);
        }
    }
    const double inv = 1.0 / (double)n;
    for (size_t p = 0; p < ORACLE_GUIDE_NPARAMS; ++p)
        grad[p] *= inv;
    return loss * inv;
}

void oracle_guide_adam(double *params, double *m, double *v, const double *grad,
                       uint32_t step, double lr, double beta1, double beta2, double eps)
{
    const double c1 = 1.0 - pow(beta1, (double)step), c2 = 1.0 - pow(beta2, (double)step);
    for (size_t p = 0; p < ORACLE_GUIDE_NPARAMS; ++p)
    {
        m[p] = beta1 * m[p] + (1.0 - beta1) * grad[p];
        v[p] = beta2 * v[p] + (1.0 - beta2) * grad[p] * grad[p];
        params[p] -= lr * (m[p] / c1) / (sqrt(v[p] / c2) + eps);
    }
}
