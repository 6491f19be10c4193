// wl_guide_api.cpp — C ABI of the learned guiding network: creation and
// initialisation, parameter access, inference, gradients and training.  The
// reference has no counterpart; the path tracer attaches a guide through
// wl_scene_set_guide.
#include "guide_layout.h"
#include "wl_internal.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <new>
#include <vector>

using wl::set_error;

namespace
{

// SplitMix64 (Steele, Lea, Flood 2014) for the weight initialisation.
struct SplitMix
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// He-uniform hidden layers (bound sqrt(6 / fan_in)), zero hidden biases; the
// output layer starts 10x smaller with its lobe means spread over the sphere
// (Fibonacci points), concentration softplus(1).
void init_params(uint32_t seed, float *P)
{
    SplitMix rng{ 0x5EEDull ^ ((uint64_t)seed << 17) };
    const double bound = std::sqrt(6.0 / 64.0);
    for (int l = 0; l < 5; ++l)
    {
        const int out = l < 4 ? 64 : 32;
        const double scale = l < 4 ? 1.0 : 0.1;
        float *W = P + wl::guide_w_off(l), *b = P + wl::guide_b_off(l);
        for (int k = 0; k < out * 64; ++k)
            W[k] = (float)((2.0 * rng.uniform() - 1.0) * bound * scale);
        for (int k = 0; k < out; ++k)
            b[k] = 0.0f;
    }
    float *b4 = P + wl::guide_b_off(4);
    const double golden = 3.14159265358979323846 * (3.0 - std::sqrt(5.0));
    for (int j = 0; j < wl::kGuideLobes; ++j)
    {
        const double z = 1.0 - 2.0 * (j + 0.5) / wl::kGuideLobes;
        b4[wl::guide_row(j, 0)] = 0.0f;
        b4[wl::guide_row(j, 1)] = 1.0f;
        b4[wl::guide_row(j, 2)] = (float)std::acos(z);
        b4[wl::guide_row(j, 3)] = (float)std::fmod(golden * j, 2.0 * 3.14159265358979323846);
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

wl_status check_config(const wl_guide_config *c)
{
    if (c->batch_size == 0 || c->batch_size % 32 != 0 || c->batch_size > (1u << 22))
        return set_error(WL_ERROR_INVALID_VALUE,
                         "batch_size must be a positive multiple of 32, at most 2^22");
    if (c->max_samples == 0 || c->max_samples > (1u << 26))
        return set_error(WL_ERROR_INVALID_VALUE, "max_samples must be in [1, 2^26]");
    if (!(c->learning_rate > 0.f) || !(c->beta1 >= 0.f && c->beta1 < 1.f) ||
        !(c->beta2 >= 0.f && c->beta2 < 1.f) || !(c->epsilon > 0.f))
        return set_error(WL_ERROR_INVALID_VALUE, "bad Adam hyper-parameters");
    if (!(c->mix_alpha >= 0.f && c->mix_alpha < 1.f))
        return set_error(WL_ERROR_INVALID_VALUE, "mix_alpha must be in [0, 1)");
    if (!(c->max_weight > 0.f))
        return set_error(WL_ERROR_INVALID_VALUE, "max_weight must be positive");
    return WL_SUCCESS;
}

} // namespace

extern "C" {

wl_status wl_guide_default_config(wl_guide_config *c)
{
    if (!c)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    std::memset(c, 0, sizeof *c);
    c->seed = 1;
    c->batch_size = 16384;
    c->steps_per_batch = 8;
    c->max_samples = 1u << 20;
    c->learning_rate = 5e-3f;
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->epsilon = 1e-8f;
    c->mix_alpha = 0.5f;
    c->max_weight = 100.0f;
    c->guided_bounces = 1;
    return WL_SUCCESS;
}

wl_status wl_guide_create(wl_context ctx, const wl_guide_config *cfg, wl_guide *out)
{
    if (!ctx || !out)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    wl_guide_config c;
    if (cfg)
        c = *cfg;
    else
        wl_guide_default_config(&c);
    wl_status st = check_config(&c);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(hipSetDevice(ctx->device));
    wl_guide g = new (std::nothrow) wl_guide_s;
    if (!g)
        return set_error(WL_ERROR_OUT_OF_MEMORY, "guide allocation");
    g->ctx = ctx;
    g->cfg = c;
    const size_t P = wl::kGuideParams;
    const uint32_t cap = c.batch_size;
    const size_t sizes[] = {
        4 * P, 4 * P, 4 * P,                                 // params, m, v
        16 * (size_t)wl::kGuideFwdVecs, 16 * (size_t)wl::kGuideBwdVecs,
        4 * P * wl::kGuideSplits,                            // partials
        2 * (size_t)5 * 64 * cap, 2 * (size_t)(4 * 64 + 32) * cap, // staging
        4 * (size_t)cap,                                     // loss
        64,                                                  // counts
        64 * (size_t)c.max_samples,                          // training samples
        16 * (size_t)wl::kGuideFwdVecs, 4 * P                // published network
    };
    size_t total = 0;
    for (size_t b : sizes)
        total += align256(b);
    hipError_t e = hipMalloc(&g->d_alloc, total);
    if (e != hipSuccess)
    {
        delete g;
        return wl::hip_error(e, "guide buffers");
    }
    char *p = (char *)g->d_alloc;
    void *ptr[13];
    for (int k = 0; k < 13; ++k)
    {
        ptr[k] = p;
        p += align256(sizes[k]);
    }
    wl::GuideDev &D = g->dev;
    D.params = (float *)ptr[0];
    D.adam_m = (float *)ptr[1];
    D.adam_v = (float *)ptr[2];
    D.fwd = ptr[3];
    D.bwd = ptr[4];
    D.partials = (float *)ptr[5];
    D.xt = ptr[6];
    D.dt = ptr[7];
    D.loss = (float *)ptr[8];
    D.count = (uint32_t *)ptr[9];
    D.train = (float *)ptr[10];
    D.cap = cap;
    D.train_cap = c.max_samples;
    g->pub_fwd = ptr[11];
    g->pub_params = (float *)ptr[12];
    wl::ctx_retain(ctx);
    // default priority, like every stream of the library
    e = hipStreamCreateWithFlags(&g->tstream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&g->ev_pub, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&g->ev_trained, hipEventDisableTiming);
    if (e != hipSuccess)
    {
        wl_guide_destroy(g);
        return wl::hip_error(e, "guide training stream");
    }
    std::vector<float> h(P);
    init_params(c.seed, h.data());
    st = wl_guide_set_params(g, h.data(), (uint32_t)P);
    if (st == WL_SUCCESS)
    {
        e = hipMemset(D.count, 0, 64);
        if (e != hipSuccess)
            st = wl::hip_error(e, "guide counters");
    }
    if (st != WL_SUCCESS)
    {
        wl_guide_destroy(g);
        return st;
    }
    *out = g;
    return WL_SUCCESS;
}

wl_status wl_guide_destroy(wl_guide g)
{
    if (!g)
        return set_error(WL_ERROR_INVALID_VALUE, "guide is NULL");
    hipSetDevice(g->ctx->device);
    hipDeviceSynchronize();
    if (g->ev_trained)
        hipEventDestroy(g->ev_trained);
    if (g->ev_pub)
        hipEventDestroy(g->ev_pub);
    if (g->tstream)
        hipStreamDestroy(g->tstream);
    hipFree(g->d_alloc);
    wl_context ctx = g->ctx;
    delete g;
    wl::ctx_release(ctx);
    return WL_SUCCESS;
}

wl_status wl_guide_set_params(wl_guide g, const float *params, uint32_t n)
{
    if (!g || !params)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (n != (uint32_t)wl::kGuideParams)
        return set_error(WL_ERROR_INVALID_VALUE, "n must be WL_GUIDE_PARAMS");
    WL_HIP_CHECK(hipSetDevice(g->ctx->device));
    WL_HIP_CHECK(hipDeviceSynchronize());
    WL_HIP_CHECK(hipMemcpy(g->dev.params, params, 4 * (size_t)n, hipMemcpyHostToDevice));
    WL_HIP_CHECK(hipMemset(g->dev.adam_m, 0, 4 * (size_t)n));
    WL_HIP_CHECK(hipMemset(g->dev.adam_v, 0, 4 * (size_t)n));
    WL_HIP_CHECK(wl::guide_pack(g->dev, nullptr));
    WL_HIP_CHECK(hipDeviceSynchronize());
    g->pending = false; // the device is idle: any training has ended
    wl_status st = wl::guide_publish(g, nullptr);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(hipDeviceSynchronize());
    g->steps = 0;
    return WL_SUCCESS;
}

wl_status wl_guide_get_params(wl_guide g, float *params, uint32_t n)
{
    if (!g || !params)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (n != (uint32_t)wl::kGuideParams)
        return set_error(WL_ERROR_INVALID_VALUE, "n must be WL_GUIDE_PARAMS");
    WL_HIP_CHECK(hipSetDevice(g->ctx->device));
    wl_status st = wl::guide_settle(g, nullptr);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(hipDeviceSynchronize());
    WL_HIP_CHECK(hipMemcpy(params, g->dev.params, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return WL_SUCCESS;
}

wl_status wl_guide_set_mix(wl_guide g, float mix_alpha)
{
    if (!g)
        return set_error(WL_ERROR_INVALID_VALUE, "guide is NULL");
    if (!(mix_alpha >= 0.f && mix_alpha < 1.f))
        return set_error(WL_ERROR_INVALID_VALUE, "mix_alpha must be in [0, 1)");
    g->cfg.mix_alpha = mix_alpha;
    return WL_SUCCESS;
}

wl_status wl_guide_get_info(wl_guide g, wl_guide_info *info)
{
    if (!g || !info)
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(g->ctx->device));
    wl_status st = wl::guide_settle(g, nullptr);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(hipDeviceSynchronize());
    uint32_t c[4];
    WL_HIP_CHECK(hipMemcpy(c, g->dev.count, sizeof c, hipMemcpyDeviceToHost));
    std::memset(info, 0, sizeof *info);
    info->steps = g->steps;
    info->last_samples = c[2];
    return WL_SUCCESS;
}

wl_status wl_guide_infer(wl_guide g, const float *d_queries, uint32_t n, float *d_out,
                         void *stream)
{
    if (!g || (n && (!d_queries || !d_out)))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(g->ctx->device));
    wl_status st = wl::guide_settle(g, (hipStream_t)stream);
    if (st != WL_SUCCESS)
        return st;
    WL_HIP_CHECK(wl::guide_infer(g->dev, d_queries, n, d_out, (hipStream_t)stream));
    return WL_SUCCESS;
}

wl_status wl_guide_mixture_sample(wl_context ctx, const float *d_out, const float *d_u,
                                  uint32_t n, float *d_dirs, float *d_pdf, void *stream)
{
    if (!ctx || (n && (!d_out || !d_u || !d_dirs || !d_pdf)))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(ctx->device));
    WL_HIP_CHECK(wl::guide_mixture_sample(d_out, d_u, n, d_dirs, d_pdf, (hipStream_t)stream));
    return WL_SUCCESS;
}

wl_status wl_guide_mixture_pdf(wl_context ctx, const float *d_out, const float *d_dirs,
                               uint32_t n, float *d_pdf, void *stream)
{
    if (!ctx || (n && (!d_out || !d_dirs || !d_pdf)))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(ctx->device));
    WL_HIP_CHECK(wl::guide_mixture_pdf(d_out, d_dirs, n, d_pdf, (hipStream_t)stream));
    return WL_SUCCESS;
}

wl_status wl_guide_gradient(wl_guide g, const float *d_samples, uint32_t n, float *d_grad,
                            float *d_loss, void *stream)
{
    if (!g || !d_grad || (n && !d_samples))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    if (n > g->dev.cap)
        return set_error(WL_ERROR_LIMIT_EXCEEDED, "n exceeds the guide's batch_size");
    WL_HIP_CHECK(hipSetDevice(g->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    wl_status st = wl::guide_settle(g, s);
    if (st != WL_SUCCESS)
        return st;
    uint32_t *cnt = g->dev.count + 1;
    WL_HIP_CHECK(wl::guide_set_count(cnt, n, s));
    WL_HIP_CHECK(wl::guide_step(g->dev, d_samples, cnt, n, 0, std::max<uint32_t>(n, 1), nullptr,
                                d_grad, d_loss, g->ctx->num_cus, s));
    g->last_stream = s;
    return WL_SUCCESS;
}

} // extern "C"

namespace wl
{
// Adam steps over samples[0, min(*n_dev, n_cap)) (device-side count).
wl_status guide_train_device(wl_guide g, const float *samples, const uint32_t *n_dev,
                             uint32_t n_cap, uint32_t steps, hipStream_t s)
{
    for (uint32_t k = 0; k < steps; ++k)
    {
        const uint64_t t = ++g->steps;
        GuideAdam a;
        a.lr = g->cfg.learning_rate;
        a.beta1 = g->cfg.beta1;
        a.beta2 = g->cfg.beta2;
        a.eps = g->cfg.epsilon;
        a.c1 = (float)(1.0 - std::pow((double)a.beta1, (double)t));
        a.c2 = (float)(1.0 - std::pow((double)a.beta2, (double)t));
        WL_HIP_CHECK(guide_step(g->dev, samples, n_dev, n_cap, k, g->cfg.batch_size, &a,
                                nullptr, nullptr, g->ctx->num_cus, s));
    }
    g->last_stream = s;
    return WL_SUCCESS;
}

wl_status guide_publish(wl_guide g, hipStream_t s)
{
    WL_HIP_CHECK(hipMemcpyAsync(g->pub_fwd, g->dev.fwd, 16 * (size_t)kGuideFwdVecs,
                                hipMemcpyDeviceToDevice, s));
    WL_HIP_CHECK(hipMemcpyAsync(g->pub_params, g->dev.params, 4 * (size_t)kGuideParams,
                                hipMemcpyDeviceToDevice, s));
    return WL_SUCCESS;
}

wl_status guide_settle(wl_guide g, hipStream_t s)
{
    if (!g->pending)
        return WL_SUCCESS;
    WL_HIP_CHECK(hipStreamWaitEvent(s, g->ev_trained, 0));
    g->pending = false;
    return guide_publish(g, s);
}
} // namespace wl

extern "C" wl_status wl_guide_train(wl_guide g, const float *d_samples, uint32_t n,
                                    uint32_t steps, void *stream)
{
    if (!g || (n && !d_samples))
        return set_error(WL_ERROR_INVALID_VALUE, "NULL argument");
    WL_HIP_CHECK(hipSetDevice(g->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    wl_status st = wl::guide_settle(g, s);
    if (st != WL_SUCCESS)
        return st;
    uint32_t *cnt = g->dev.count + 1;
    WL_HIP_CHECK(wl::guide_set_count(cnt, n, s));
    st = wl::guide_train_device(g, d_samples, cnt, n, steps, s);
    if (st != WL_SUCCESS)
        return st;
    return wl::guide_publish(g, s);
}
