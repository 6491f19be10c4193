// guide_device.h — the learned guiding distribution on gfx950: input
// encoding, the 64-64-64-64-64-32 MLP on BF16 MFMA tiles, and the 8-lobe von
// Mises-Fisher mixture its outputs parameterise.  Shared by the guide
// kernels (guide.hip) and the path tracer's shade kernel (pathtrace.hip).
//
// Parameters (FP32 master copy, 18720 floats): layer l = 0..4 holds
// W_l[out][in] row-major then b_l[out], out = 64 (l < 4) or 32 (l = 4),
// in = 64.  Output row r is lobe 2 * ((r % 16) / 4) + r / 16, parameter r % 4:
// 0 logit, 1 raw concentration (kappa = softplus), 2 theta, 3 phi.
//
// MFMA tiling (v_mfma_f32_16x16x32_bf16, one wave, 16 queries per tile).
// Lane l, lane group g = l / 16, element j:
//   A (16 x 32):  row l % 16,    k = 8 g + j       (8 bf16 per lane)
//   B (32 x 16):  column l % 16, k = 8 g + j       (8 bf16 per lane)
//   D (16 x 16):  column l % 16, row 4 g + j       (4 floats per lane)
// Layers run transposed: H_out^T = W^T-tile x H_in^T, so a lane always works
// on query l % 16 and the result tile t of a layer leaves lane group g with
// features 16 t + 4 g + i.  The next layer's B operand takes, in k-block kb,
// element j, the feature phi(kb, g, j) = 16 (2 kb + j / 4) + 4 g + j % 4 —
// exactly the values the lane already holds — and the packed weights use the
// same k order, so activations never cross lanes between layers.
#pragma once
#include "guide_layout.h"
#include "shading_device.h"

namespace wl
{

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ floatx4 mfma16(bf16x8 a, bf16x8 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---- input encoding ------------------------------------------------------------
// Query q[9]: position in [0,1]^3, normal, outgoing direction.  Feature e,
// with block B = e / 16, lane group g = (e % 16) / 4 and pair p = 2 B +
// (e % 4) / 2 (the features a lane of group g holds, in its register order):
//   g < 3, p < 6:  sin (e even) / cos (e odd) of 2^p * pi * q[g]
//   g < 3, p = 6:  normal[g] (e even), outgoing direction[g] (e odd)
//   otherwise 0.
// So a lane encodes one axis: one sincospi, then five angle doublings
// (sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a)), whose
// error (about 2^5 FP32 ulps at f = 5) stays far below the BF16 operand's.
// q lives in LDS: the axis a lane reads depends on its lane group, and a
// private array indexed per lane would go to scratch memory.
__device__ __forceinline__ void guide_encode(const float *q, int g, bf16x8 x[2])
{
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k)
        v[k] = 0.f;
    if (g < 3)
    {
        float s, c;
        sincospif(q[g], &s, &c); // exact argument reduction
#pragma unroll
        for (int f = 0; f < 6; ++f)
        {
            v[2 * f] = s;
            v[2 * f + 1] = c;
            const float s2 = 2.0f * (s * c), c2 = (c - s) * (c + s);
            s = s2;
            c = c2;
        }
        v[12] = q[3 + g];
        v[13] = q[6 + g];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        x[0][j] = (__bf16)v[j];
        x[1][j] = (__bf16)v[8 + j];
    }
}

// ---- staging of the packed operands and biases into LDS -------------------------
// Every load of a thread is issued before its first LDS store: a loop of
// load, wait, store pays the memory latency once per iteration (23 times for
// the training kernel's operands), which set these kernels' time.
template <int BLOCK, int N, class T>
__device__ __forceinline__ void guide_stage(T *__restrict__ dst, const T *__restrict__ src)
{
    constexpr int kFull = N / BLOCK, kRest = N % BLOCK;
    T v[kFull + 1];
#pragma unroll
    for (int k = 0; k < kFull; ++k)
        v[k] = src[threadIdx.x + k * BLOCK];
    if (kRest != 0 && (int)threadIdx.x < kRest)
        v[kFull] = src[threadIdx.x + kFull * BLOCK];
#pragma unroll
    for (int k = 0; k < kFull; ++k)
        dst[threadIdx.x + k * BLOCK] = v[k];
    if (kRest != 0 && (int)threadIdx.x < kRest)
        dst[threadIdx.x + kFull * BLOCK] = v[kFull];
}

// Biases in the order guide_forward reads them: 64 per hidden layer, then
// the 32 of the output layer (BLOCK >= 256 threads).
template <int BLOCK>
__device__ __forceinline__ void guide_stage_biases(float *__restrict__ dst,
                                                   const float *__restrict__ params)
{
    static_assert(BLOCK >= 256, "one hidden-layer bias per thread");
    const int i = threadIdx.x;
    float b0 = 0.f, b1 = 0.f;
    if (i < 256)
        b0 = params[guide_b_off(i / 64) + i % 64];
    if (i < kGuideBiases - 256)
        b1 = params[guide_b_off(4) + i];
    if (i < 256)
        dst[i] = b0;
    if (i < kGuideBiases - 256)
        dst[256 + i] = b1;
}

// ---- forward pass of one 16-query tile -------------------------------------
// A: packed forward operands (LDS or global), bias: 288 FP32 biases in layer
// order.  x: layer-0 input on entry.  keep (optional): the B operand that
// entered each layer.  out: output tiles t = 0, 1 (rows 16 t + 4 g + i).
template <bool KEEP>
__device__ __forceinline__ void guide_forward(const bf16x8 *A, const float *bias, int lane,
                                              bf16x8 x[2], floatx4 out[2], bf16x8 keep[5][2])
{
    const int g = lane >> 4;
    // inference keeps one layer's operands live at a time; training unrolls
    // the layers so the stored inputs stay in registers
#pragma unroll(KEEP ? 4 : 2)
    for (int l = 0; l < 4; ++l)
    {
        floatx4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            const float4 b = *reinterpret_cast<const float4 *>(bias + 64 * l + 16 * t + 4 * g);
            acc[t] = floatx4{ b.x, b.y, b.z, b.w };
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                acc[t] = mfma16(A[guide_fwd_vec(l, t, kb) + lane], x[kb], acc[t]);
        }
        if (KEEP)
        {
            keep[l][0] = x[0];
            keep[l][1] = x[1];
        }
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                x[kb][j] = (__bf16)fmaxf(acc[2 * kb + j / 4][j % 4], 0.0f);
    }
    if (KEEP)
    {
        keep[4][0] = x[0];
        keep[4][1] = x[1];
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
    {
        const float4 b = *reinterpret_cast<const float4 *>(bias + 256 + 16 * t + 4 * g);
        out[t] = floatx4{ b.x, b.y, b.z, b.w };
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
            out[t] = mfma16(A[guide_fwd_vec(4, t, kb) + lane], x[kb], out[t]);
    }
}

// ---- cross-lane reductions over the 4 lane groups (lanes l, l^16, l^32, l^48)
__device__ __forceinline__ float wave_max4(float v)
{
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ float wave_sum4(float v)
{
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

// ---- von Mises-Fisher mixture ------------------------------------------------
__device__ __forceinline__ float guide_softplus(float k) { return k > 15.f ? k : log1pf(expf(k)); }

// sin and cos of a lobe angle (a network output, |x| of order 10): the
// argument is reduced to [-pi, pi] by k = rint(x / 2 pi) and a two-part 2 pi
// (Cody & Waite: the float nearest 2 pi, then the remainder), and the
// hardware sine and cosine take it from there.  The general sincosf carries
// a large-argument reduction whose selects cost more than the network's
// MFMAs here.  Error below 1e-6 absolute for |x| < 1e4; the decoded mixture
// stays a consistent density whatever the rounding, as sampling and pdf read
// the same decoded values.
__device__ __forceinline__ void guide_sincos(float x, float *s, float *c)
{
    const float k = rintf(x * 0.159154943091895336f);
    float r = fmaf(-k, 6.28318548202514648f, x); // float(2 pi)
    r = fmaf(k, 1.74845553e-7f, r);              // 2 pi - float(2 pi) = -1.74845553e-7
    *s = __sinf(r);
    *c = __cosf(r);
}

// log of kappa / (2 pi (1 - e^{-2 kappa}))
__device__ __forceinline__ float vmf_log_norm(float kappa)
{
    if (kappa < 1e-4f)
        return logf((1.0f + kappa) * 0.0795774715f);
    return logf(kappa) - 1.83787706641f - logf(-expm1f(-2.0f * kappa));
}

// Decoded mixture: per lobe the log of (weight x vMF normaliser), the
// concentration and the mean direction (5 floats, the registers a shade
// kernel keeps across its BSDF and light samples).
struct GuideMix
{
    float lw[kGuideLobes];
    float kappa[kGuideLobes];
    V3 mu[kGuideLobes];
};

// Decoded mixture as the path tracer stores it per query (48 floats):
// float4 (lw, kappa, mu.x, mu.y) for lobes 0..7, then mu.z[8], then the lobe
// weights pi[8].
constexpr int kGuideDecoded = 48;

// Decode of the outputs a lane holds after guide_forward (lobes 2 g and
// 2 g + 1) into the stored form; every lane of the wave must call it.
__device__ __forceinline__ void guide_decode_lanes(const floatx4 out[2], int g, float *dst,
                                                   bool write)
{
    const float amax = wave_max4(fmaxf(out[0][0], out[1][0]));
    const float se = wave_sum4(expf(out[0][0] - amax) + expf(out[1][0] - amax));
    const float lse = amax + logf(se);
#pragma unroll
    for (int t = 0; t < 2; ++t)
    {
        const int j = 2 * g + t;
        const float kap = guide_softplus(out[t][1]);
        const float lpi = out[t][0] - lse;
        float st, ct, sp, cp;
        guide_sincos(out[t][2], &st, &ct);
        guide_sincos(out[t][3], &sp, &cp);
        if (write)
        {
            reinterpret_cast<float4 *>(dst)[j] =
                make_float4(lpi + vmf_log_norm(kap), kap, st * cp, st * sp);
            dst[32 + j] = ct;
            dst[40 + j] = expf(lpi);
        }
    }
}

__device__ __forceinline__ void guide_load(const float *src, GuideMix &m)
{
    const float4 *q = reinterpret_cast<const float4 *>(src);
    const float4 z0 = q[8], z1 = q[9];
    const float mz[8] = { z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w };
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j)
    {
        const float4 v = q[j];
        m.lw[j] = v.x;
        m.kappa[j] = v.y;
        m.mu[j] = mk(v.z, v.w, mz[j]);
    }
}

__device__ __forceinline__ float guide_pdf(const GuideMix &m, V3 w)
{
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j)
        q += expf(m.lw[j] + m.kappa[j] * (dot(m.mu[j], w) - 1.0f));
    return q;
}

// u0 picks the lobe by the cumulative weights pi (read from memory, or
// recomputed when pi is NULL), (u1, u2) the direction (Jakob 2012 inversion).
__device__ __forceinline__ V3 guide_sample(const GuideMix &m, const float *pi, float u0,
                                           float u1, float u2)
{
    float c = 0.f, kap = m.kappa[kGuideLobes - 1];
    V3 mu = m.mu[kGuideLobes - 1];
    bool found = false;
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j)
    {
        c += pi ? pi[j] : expf(m.lw[j] - vmf_log_norm(m.kappa[j])); // lobe weight
        if (!found && u0 < c)
        {
            found = true;
            kap = m.kappa[j];
            mu = m.mu[j];
        }
    }
    // w_z = 1 + t, t = log(u + (1 - u) e^{-2 kappa}) / kappa, written with
    // log1p / expm1 so small kappa does not cancel; 1 - w_z^2 = -t (2 + t)
    float t = kap < 1e-4f ? 2.0f * u1 - 2.0f
                          : log1pf((1.0f - u1) * expm1f(-2.0f * kap)) / kap;
    t = clampf(t, -2.0f, 0.0f);
    const float wz = 1.0f + t;
    const float r = sqrtf(fmaxf(0.0f, -t * (2.0f + t)));
    float sp, cp;
    sincospif(2.0f * u2, &sp, &cp);
    V3 T, Bv;
    onb(mu, T, Bv);
    return T * (r * cp) + Bv * (r * sp) + mu * wz;
}

} // namespace wl
