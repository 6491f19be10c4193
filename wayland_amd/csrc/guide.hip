// guide.hip — training and inference kernels of the guiding network (gfx950).
//
// Inference and training run the MLP on BF16 MFMA tiles with FP32
// accumulation (guide_device.h).  A training step is four kernels:
//
//   fwdbwd   per wave, 16 samples at a time: encode, forward (MFMA), loss and
//            its output gradient (cross-lane softmax / log-sum-exp over the
//            4 lane groups), backward deltas (MFMA with the transposed
//            weights), and the BF16 layer inputs and deltas written
//            feature-major to a staging buffer;
//   dw       weight gradients dW_l = sum_s delta_s x_s^T as MFMA tiles with
//            the samples as the K dimension, split over kGuideSplits slices
//            (one partial sum per slice, no atomics);
//   update   sums the slices, divides by the batch size, Adam, and writes
//            each new weight into its slots of the packed BF16 operands
//            (k_guide_pack does the whole packing when weights are set).
#include "guide_device.h"
#include "wl_guide.h"

#include <algorithm>

namespace wl
{

constexpr int kGuideBlock = 256;

// ---- pack ------------------------------------------------------------------
__global__ void k_guide_pack(const float *__restrict__ P, bf16x8 *__restrict__ fwd,
                             bf16x8 *__restrict__ bwd)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < kGuideFwdVecs)
    {
        const int l = v < 2048 ? v / 512 : 4;
        const int rem = v < 2048 ? v % 512 : v - 2048;
        const int t = rem / 128, kb = (rem % 128) / 64, lane = rem % 64;
        const float *W = P + guide_w_off(l);
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (__bf16)W[(16 * t + lane % 16) * 64 + guide_phi(kb, lane / 16, j)];
        fwd[v] = o;
    }
    else if (v < kGuideFwdVecs + kGuideBwdVecs)
    {
        const int u = v - kGuideFwdVecs;
        int l, t, kb, lane;
        if (u < 1536)
        {
            l = 1 + u / 512;
            const int rem = u % 512;
            t = rem / 128;
            kb = (rem % 128) / 64;
            lane = rem % 64;
        }
        else
        {
            l = 4;
            const int rem = u - 1536;
            t = rem / 64;
            kb = 0;
            lane = rem % 64;
        }
        const float *W = P + guide_w_off(l);
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (__bf16)W[guide_phi(kb, lane / 16, j) * 64 + 16 * t + lane % 16];
        bwd[u] = o;
    }
}

// ---- inference on query records -------------------------------------------------
__global__ __launch_bounds__(kGuideBlock)
void k_guide_infer(const bf16x8 *__restrict__ fwd,
                                                             const float *__restrict__ P,
                                                             const float *__restrict__ queries,
                                                             uint32_t n, float *__restrict__ out)
{
    __shared__ bf16x8 sA[kGuideFwdVecs];
    __shared__ float sb[kGuideBiases];
    __shared__ float sq[9 * kGuideBlock];
    float *q = sq + 9 * threadIdx.x;
    guide_stage<kGuideBlock, kGuideFwdVecs>(sA, fwd);
    guide_stage_biases<kGuideBlock>(sb, P);
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t nwaves = gridDim.x * (kGuideBlock / 64);
    for (uint32_t tile = blockIdx.x * (kGuideBlock / 64) + threadIdx.x / 64; tile < tiles;
         tile += nwaves)
    {
        const uint32_t s = tile * 16 + (lane & 15);
        const bool valid = s < n;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            q[k] = valid ? queries[9 * (size_t)s + k] : 0.f;
        bf16x8 x[2];
        guide_encode(q, g, x);
        floatx4 o[2];
        guide_forward<false>(sA, sb, lane, x, o, nullptr);
        if (valid)
        {
#pragma unroll
            for (int t = 0; t < 2; ++t)
                *reinterpret_cast<float4 *>(out + 32 * (size_t)s + 16 * t + 4 * g) =
                    make_float4(o[t][0], o[t][1], o[t][2], o[t][3]);
        }
    }
}

// ---- mixture sampling / density on given network outputs -----------------------
// The same decode chain as the path tracer's guided vertices: the outputs are
// placed in the lane layout guide_forward leaves them in (lane group g holds
// lobes 2 g and 2 g + 1 of 16 queries), decoded across the lane groups into
// the stored form (guide_decode_lanes, here into LDS instead of the path
// tracer's per-query buffer), then read back with guide_load and sampled
// with the stored lobe weights.
template <bool SAMPLE>
__global__ __launch_bounds__(kGuideBlock)
void k_guide_mix(const float *__restrict__ outs, const float *__restrict__ in, uint32_t n,
                 float *__restrict__ dirs, float *__restrict__ pdf)
{
    __shared__ float dec[(kGuideBlock / 64) * 16 * kGuideDecoded];
    const int lane = threadIdx.x & 63, g = lane >> 4, wave = threadIdx.x / 64;
    float *wdec = dec + wave * 16 * kGuideDecoded;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t per_block = kGuideBlock / 64;
    // block-uniform loop: every wave reaches the barriers equally often
    for (uint32_t t0 = blockIdx.x * per_block; t0 < tiles; t0 += gridDim.x * per_block)
    {
        const uint32_t i = (t0 + wave) * 16 + (lane & 15);
        const bool valid = i < n;
        floatx4 o[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                o[t][c] = valid ? outs[32 * (size_t)i + 16 * t + 4 * g + c] : 0.0f;
        guide_decode_lanes(o, g, wdec + kGuideDecoded * (lane & 15), valid);
        __syncthreads();
        if (valid && lane < 16)
        {
            const float *src = wdec + kGuideDecoded * lane;
            GuideMix m;
            guide_load(src, m);
            V3 w;
            if (SAMPLE)
            {
                w = guide_sample(m, src + 40, in[3 * (size_t)i], in[3 * (size_t)i + 1],
                                 in[3 * (size_t)i + 2]);
                dirs[3 * (size_t)i] = w.x;
                dirs[3 * (size_t)i + 1] = w.y;
                dirs[3 * (size_t)i + 2] = w.z;
            }
            else
            {
                w = mk(in[3 * (size_t)i], in[3 * (size_t)i + 1], in[3 * (size_t)i + 2]);
            }
            pdf[i] = guide_pdf(m, w);
        }
        __syncthreads();
    }
}

__global__ void k_guide_set_count(uint32_t *c, uint32_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        c[0] = n;
}

// c[2] = min(c[0], cap) (kept for reporting), c[0] = 0
__global__ void k_guide_roll_count(uint32_t *c, uint32_t cap)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        c[2] = min(c[0], cap);
        c[0] = 0;
    }
}

// Mini-batch of `step`: [offset, offset + size) of the first min(n, n_cap)
// samples.
__device__ __forceinline__ void batch_range(uint32_t n, uint32_t n_cap, uint32_t step,
                                            uint32_t batch, uint32_t &offset, uint32_t &size)
{
    n = min(n, n_cap);
    offset = n ? (uint32_t)(((uint64_t)step * batch) % n) : 0u;
    size = n ? min(batch, n - offset) : 0u;
}

// ---- training: forward, loss gradient, backward, staging ----------------------------

// d log-normaliser / d kappa = 1/kappa - 2 / (e^{2 kappa} - 1)
__device__ __forceinline__ float vmf_dlog_norm(float kappa)
{
    if (kappa < 1e-3f)
        return 1.0f - kappa * (1.0f / 3.0f);
    return 1.0f / kappa - 2.0f / expm1f(2.0f * kappa);
}

// Loss -w log q(om) of the lane's sample and its gradient with respect to
// the lane's outputs out[t][i] (lobe 2 g + t, parameter i).  All 64 lanes
// must call it (cross-lane reductions over the 4 lane groups).
__device__ __forceinline__ float guide_loss_grad(const floatx4 out[2], V3 om, float w,
                                                 float d[2][4])
{
    const float amax = wave_max4(fmaxf(out[0][0], out[1][0]));
    const float e0 = expf(out[0][0] - amax), e1 = expf(out[1][0] - amax);
    const float se = wave_sum4(e0 + e1);
    const float lse = amax + logf(se);
    float lv[2], kap[2], cth[2], pi[2], wdth[2], wdph[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
    {
        pi[t] = (t == 0 ? e0 : e1) / se;
        kap[t] = guide_softplus(out[t][1]);
        float st, ct, sp, cp;
        guide_sincos(out[t][2], &st, &ct);
        guide_sincos(out[t][3], &sp, &cp);
        const V3 mu = mk(st * cp, st * sp, ct);
        cth[t] = dot(mu, om);
        wdth[t] = om.x * (ct * cp) + om.y * (ct * sp) - om.z * st;
        wdph[t] = -om.x * (st * sp) + om.y * (st * cp);
        lv[t] = (out[t][0] - lse) + vmf_log_norm(kap[t]) + kap[t] * (cth[t] - 1.0f);
    }
    const float lmax = wave_max4(fmaxf(lv[0], lv[1]));
    const float sl = wave_sum4(expf(lv[0] - lmax) + expf(lv[1] - lmax));
    const float logq = lmax + logf(sl);
#pragma unroll
    for (int t = 0; t < 2; ++t)
    {
        const float resp = expf(lv[t] - logq);
        const float sig = 1.0f / (1.0f + expf(-out[t][1]));
        d[t][0] = -w * (resp - pi[t]);
        d[t][1] = -w * resp * (vmf_dlog_norm(kap[t]) + cth[t] - 1.0f) * sig;
        d[t][2] = -w * resp * kap[t] * wdth[t];
        d[t][3] = -w * resp * kap[t] * wdph[t];
    }
    return -w * logq;
}

__global__ __launch_bounds__(kGuideBlock) void k_guide_fwdbwd(GuideDev G,
                                                              const float *__restrict__ samples,
                                                              const uint32_t *__restrict__ n_dev, uint32_t n_cap,
                                                              uint32_t step, uint32_t batch)
{
    __shared__ bf16x8 sA[kGuideFwdVecs];
    __shared__ bf16x8 sB[kGuideBwdVecs];
    __shared__ float sb[kGuideBiases];
    __shared__ float sq[9 * kGuideBlock];
    float *q = sq + 9 * threadIdx.x;
    const bf16x8 *fwd = reinterpret_cast<const bf16x8 *>(G.fwd);
    const bf16x8 *bwd = reinterpret_cast<const bf16x8 *>(G.bwd);
    guide_stage<kGuideBlock, kGuideFwdVecs>(sA, fwd);
    guide_stage<kGuideBlock, kGuideBwdVecs>(sB, bwd);
    guide_stage_biases<kGuideBlock>(sb, G.params);
    __syncthreads();
    uint32_t offset, size;
    batch_range(*n_dev, n_cap, step, batch, offset, size);
    const uint32_t padded = (size + 31) & ~31u; // dw reads whole 32-sample steps
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t cap = G.cap;
    __bf16 *XT = reinterpret_cast<__bf16 *>(G.xt);
    __bf16 *DT = reinterpret_cast<__bf16 *>(G.dt);
    const uint32_t nwaves = gridDim.x * (kGuideBlock / 64);
    for (uint32_t tile = blockIdx.x * (kGuideBlock / 64) + threadIdx.x / 64; tile < padded / 16;
         tile += nwaves)
    {
        const uint32_t s = tile * 16 + (lane & 15);
        const bool valid = s < size;
        const float *rec = samples + 16 * ((size_t)offset + s);
#pragma unroll
        for (int k = 0; k < 9; ++k)
            q[k] = valid ? rec[k] : 0.f;
        const V3 om = valid ? mk(rec[9], rec[10], rec[11]) : mk(0, 0, 1);
        float w = valid ? rec[12] : 0.f;
        if (!(w >= 0.f && w < 3.0e38f))
            w = 0.f; // non-finite or negative weights do not train
        bf16x8 x[2];
        guide_encode(q, g, x);
        floatx4 out[2];
        bf16x8 keep[5][2];
        guide_forward<true>(sA, sb, lane, x, out, keep);
        float d[2][4];
        const float loss = guide_loss_grad(out, om, w, d);
        if (valid && g == 0)
            G.loss[s] = loss;
        // layer inputs, feature-major
#pragma unroll
        for (int l = 0; l < 5; ++l)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    XT[((size_t)l * 64 + guide_phi(kb, g, j)) * cap + s] =
                        valid ? keep[l][kb][j] : (__bf16)0.0f;
        // output delta: rows 16 t + 4 g + i = phi(0, g, 4 t + i)
        bf16x8 dB[2];
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            dB[0][j] = (__bf16)d[j / 4][j % 4];
            dB[1][j] = (__bf16)0.0f;
            DT[((size_t)256 + guide_phi(0, g, j)) * cap + s] = dB[0][j];
        }
        // back through layers 4..1: delta_in = (W_l^T delta_out) * relu'(x_l)
#pragma unroll
        for (int l = 4; l >= 1; --l)
        {
            floatx4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                acc[t] = floatx4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
                for (int kb = 0; kb < (l == 4 ? 1 : 2); ++kb)
                    acc[t] = mfma16(sB[guide_bwd_vec(l, t, kb) + lane], dB[kb], acc[t]);
            }
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                {
                    const float v = (float)keep[l][kb][j] > 0.f ? acc[2 * kb + j / 4][j % 4] : 0.f;
                    dB[kb][j] = (__bf16)v;
                    DT[((size_t)(l - 1) * 64 + guide_phi(kb, g, j)) * cap + s] = dB[kb][j];
                }
        }
    }
}

// ---- training: weight gradients --------------------------------------------------
// Work item = (tile, split); tiles 0..63: layers 0..3, 4 x 4 tiles of 16 x 16;
// tiles 64..71: layer 4, 2 x 4 tiles.  Tiles with im == 0 also sum the bias
// gradient of their 16 rows.
__global__ __launch_bounds__(kGuideBlock) void k_guide_dw(GuideDev G,
                                                          const uint32_t *__restrict__ n_dev, uint32_t n_cap,
                                                          uint32_t step, uint32_t batch)
{
    const uint32_t item = blockIdx.x * (kGuideBlock / 64) + threadIdx.x / 64;
    if (item >= 72u * kGuideSplits)
        return;
    const int tile = (int)(item % 72u), ks = (int)(item / 72u);
    int l, om, im;
    if (tile < 64)
    {
        l = tile / 16;
        om = (tile % 16) / 4;
        im = tile % 4;
    }
    else
    {
        l = 4;
        om = (tile - 64) / 4;
        im = (tile - 64) % 4;
    }
    uint32_t offset, size;
    batch_range(*n_dev, n_cap, step, batch, offset, size);
    const uint32_t steps = (size + 31) / 32;
    const uint32_t per = (steps + kGuideSplits - 1) / kGuideSplits;
    const uint32_t st0 = min(steps, (uint32_t)ks * per), st1 = min(steps, st0 + per);
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t cap = G.cap;
    const __bf16 *D = reinterpret_cast<const __bf16 *>(G.dt) +
                      (size_t)(l < 4 ? l * 64 : 256) * cap + (size_t)(om * 16 + (lane & 15)) * cap;
    const __bf16 *X = reinterpret_cast<const __bf16 *>(G.xt) + (size_t)l * 64 * cap +
                      (size_t)(im * 16 + (lane & 15)) * cap;
    floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
    float bs = 0.f;
    // four 32-sample steps per round, their eight loads issued before the
    // first MFMA waits on them
    constexpr uint32_t kRound = 4;
    uint32_t st = st0;
    for (; st + kRound <= st1; st += kRound)
    {
        bf16x8 a[kRound], b[kRound];
#pragma unroll
        for (uint32_t k = 0; k < kRound; ++k)
        {
            const size_t s0 = (size_t)(st + k) * 32 + 8 * g;
            a[k] = *reinterpret_cast<const bf16x8 *>(D + s0);
            b[k] = *reinterpret_cast<const bf16x8 *>(X + s0);
        }
#pragma unroll
        for (uint32_t k = 0; k < kRound; ++k)
        {
            acc = mfma16(a[k], b[k], acc);
            if (im == 0)
            {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    bs += (float)a[k][j];
            }
        }
    }
    for (; st < st1; ++st)
    {
        const size_t s0 = (size_t)st * 32 + 8 * g;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(D + s0);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(X + s0);
        acc = mfma16(a, b, acc);
        if (im == 0)
        {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                bs += (float)a[j];
        }
    }
    float *P = G.partials + (size_t)ks * kGuideParams;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        P[guide_w_off(l) + (om * 16 + 4 * g + i) * 64 + im * 16 + (lane & 15)] = acc[i];
    if (im == 0)
    {
        bs = wave_sum4(bs);
        if (g == 0)
            P[guide_b_off(l) + om * 16 + (lane & 15)] = bs;
    }
}

// ---- training: reduction, Adam ---------------------------------------------------
__global__ void k_guide_update(GuideDev G, const uint32_t *__restrict__ n_dev, uint32_t n_cap, uint32_t step,
                               uint32_t batch, int do_adam, GuideAdam a, float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGuideParams)
        return;
    uint32_t offset, size;
    batch_range(*n_dev, n_cap, step, batch, offset, size);
    if (size == 0)
    {
        if (grad)
            grad[p] = 0.f;
        return;
    }
    float s = 0.f;
#pragma unroll 8
    for (int k = 0; k < kGuideSplits; ++k)
        s += G.partials[(size_t)k * kGuideParams + p];
    const float gr = s / (float)size;
    if (grad)
        grad[p] = gr;
    if (do_adam)
    {
        const float m = a.beta1 * G.adam_m[p] + (1.0f - a.beta1) * gr;
        const float v = a.beta2 * G.adam_v[p] + (1.0f - a.beta2) * gr * gr;
        G.adam_m[p] = m;
        G.adam_v[p] = v;
        const float w = G.params[p] - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
        G.params[p] = w;
        // the weight's element in the packed operands (the inverse of k_guide_pack)
        int l = 0;
        while (l < 4 && p >= guide_w_off(l + 1))
            ++l;
        const int r = p - guide_w_off(l);
        if (r < (l < 4 ? 64 : 32) * 64)
        {
            const int out = r / 64, in = r % 64;
            const __bf16 wb = (__bf16)w;
            const int fkb = (in / 16) / 2, fj = 4 * ((in / 16) % 2) + in % 4;
            const int flane = (out % 16) + 16 * ((in % 16) / 4);
            reinterpret_cast<__bf16 *>(G.fwd)[8 * (guide_fwd_vec(l, out / 16, fkb) + flane) + fj] = wb;
            if (l >= 1)
            {
                const int bkb = (out / 16) / 2, bj = 4 * ((out / 16) % 2) + out % 4;
                const int blane = (in % 16) + 16 * ((out % 16) / 4);
                reinterpret_cast<__bf16 *>(G.bwd)[8 * (guide_bwd_vec(l, in / 16, bkb) + blane) + bj] =
                    wb;
            }
        }
    }
}

__global__ void k_guide_loss_mean(const float *__restrict__ loss, const uint32_t *__restrict__ n_dev, uint32_t n_cap,
                                  uint32_t step, uint32_t batch, float *__restrict__ out)
{
    __shared__ float red[256];
    uint32_t offset, size;
    batch_range(*n_dev, n_cap, step, batch, offset, size);
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < size; i += blockDim.x)
        s += loss[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = size ? red[0] / (float)size : 0.f;
}

// ---- host launchers -------------------------------------------------------------------
hipError_t guide_pack(const GuideDev &G, hipStream_t s)
{
    const int total = kGuideFwdVecs + kGuideBwdVecs;
    hipLaunchKernelGGL(k_guide_pack, dim3((total + 255) / 256), dim3(256), 0, s, G.params,
                       reinterpret_cast<bf16x8 *>(G.fwd), reinterpret_cast<bf16x8 *>(G.bwd));
    return hipGetLastError();
}

hipError_t guide_infer(const GuideDev &G, const float *queries, uint32_t n, float *out,
                       hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t grid = std::min<uint32_t>((tiles + 3) / 4, 1024);
    hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(kGuideBlock), 0, s,
                       reinterpret_cast<const bf16x8 *>(G.fwd), G.params, queries, n, out);
    return hipGetLastError();
}

hipError_t guide_mixture_sample(const float *outs, const float *u, uint32_t n, float *dirs,
                                float *pdf, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t tiles = (n + 15) / 16;
    hipLaunchKernelGGL(k_guide_mix<true>, dim3(std::min<uint32_t>((tiles + 3) / 4, 4096)),
                       dim3(kGuideBlock), 0, s, outs, u, n, dirs, pdf);
    return hipGetLastError();
}

hipError_t guide_mixture_pdf(const float *outs, const float *dirs, uint32_t n, float *pdf,
                             hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t tiles = (n + 15) / 16;
    hipLaunchKernelGGL(k_guide_mix<false>, dim3(std::min<uint32_t>((tiles + 3) / 4, 4096)),
                       dim3(kGuideBlock), 0, s, outs, dirs, n, nullptr, pdf);
    return hipGetLastError();
}

hipError_t guide_set_count(uint32_t *count, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_guide_set_count, dim3(1), dim3(64), 0, s, count, n);
    return hipGetLastError();
}

hipError_t guide_roll_count(uint32_t *count, uint32_t cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_guide_roll_count, dim3(1), dim3(64), 0, s, count, cap);
    return hipGetLastError();
}

hipError_t guide_step(const GuideDev &G, const float *samples, const uint32_t *n_dev,
                      uint32_t n_cap, uint32_t step, uint32_t batch, const GuideAdam *adam,
                      float *grad, float *loss, int num_cus, hipStream_t s)
{
    if (batch == 0 || batch > G.cap)
        return hipErrorInvalidValue;
    // one workgroup per 64 samples, at most two per CU (64 KiB LDS each)
    const uint32_t g_fb = std::max<uint32_t>(
        1, std::min<uint32_t>((batch + 63) / 64, 2u * (uint32_t)std::max(1, num_cus)));
    hipLaunchKernelGGL(k_guide_fwdbwd, dim3(g_fb), dim3(kGuideBlock), 0, s, G, samples, n_dev,
                       n_cap, step, batch);
    const uint32_t items = 72u * kGuideSplits;
    hipLaunchKernelGGL(k_guide_dw, dim3((items + 3) / 4), dim3(kGuideBlock), 0, s, G, n_dev,
                       n_cap, step, batch);
    GuideAdam a{};
    if (adam)
        a = *adam;
    hipLaunchKernelGGL(k_guide_update, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, G,
                       n_dev, n_cap, step, batch, adam ? 1 : 0, a, grad);
    if (loss)
        hipLaunchKernelGGL(k_guide_loss_mean, dim3(1), dim3(256), 0, s, G.loss, n_dev, n_cap,
                           step, batch, loss);
    return hipGetLastError();
}

} // namespace wl
