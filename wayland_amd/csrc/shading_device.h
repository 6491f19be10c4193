// shading_device.h — device-side vector math, Disney BSDF, area-light
// sampling and the TEA/LCG random numbers of the EasyRender path tracer,
// written for gfx950.
//
// The expressions follow the reference's source order so that results agree
// with a host compile of it to the last bit wherever only + - * / sqrt are
// involved (this file is compiled with -ffp-contract=off; division and sqrt
// are correctly rounded).  sinf/cosf/logf are the device library's and may
// differ from a host libm by an ulp.  glm semantics: dot = (x + y) + z,
// normalize(v) = v * (1 / sqrt(dot(v, v))), length = sqrt(dot(v, v)),
// vec / scalar divides each component.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wayland_hip.h"

namespace wl
{
#define WL_DEV __device__ __forceinline__

struct V3
{
    float x, y, z;
};

WL_DEV V3 mk(float x, float y, float z) { return V3{ x, y, z }; }
WL_DEV V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
WL_DEV V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
WL_DEV V3 operator*(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
WL_DEV V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
WL_DEV V3 operator*(float s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
WL_DEV V3 operator/(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
WL_DEV V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
WL_DEV V3 add_scalar(V3 a, float s) { return mk(a.x + s, a.y + s, a.z + s); }
WL_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
WL_DEV V3 cross(V3 a, V3 b)
{
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
WL_DEV float length(V3 a) { return sqrtf(dot(a, a)); }
WL_DEV V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
WL_DEV float clampf(float a, float lo, float hi) { return a < lo ? lo : (a > hi ? hi : a); }
WL_DEV V3 clamp3(V3 v, float lo, float hi)
{
    return mk(clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi));
}
WL_DEV float lerpf(float a, float b, float w) { return a * (1 - w) + b * w; }
WL_DEV V3 lerp3(V3 a, V3 b, float w)
{
    return mk(lerpf(a.x, b.x, w), lerpf(a.y, b.y, w), lerpf(a.z, b.z, w));
}
WL_DEV V3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = kPi * 2.0f;

// ---- RNG: TEA (4 rounds) seeding + 32-bit LCG, 24-bit floats ----------------
WL_DEV uint32_t tea4(uint32_t v0, uint32_t v1)
{
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; n++)
    {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
WL_DEV float rnd(uint32_t &s)
{
    s = 1664525u * s + 1013904223u;
    return (float)(s & 0x00FFFFFFu) / (float)0x01000000;
}

// ---- frame + material math -------------------------------------------------
WL_DEV void onb(V3 n, V3 &u, V3 &v)
{
    float x = n.x, y = n.y, z = n.z;
    float sign = copysignf(1.0f, z);
    float a = -1.0f / (sign + z);
    float b = x * y * a;
    u = mk(1.0f + sign * x * x * a, sign * b, -sign * x);
    v = mk(b, sign + y * y * a, -y);
}

WL_DEV float schlick(float u)
{
    float m = clampf(1.0f - u, 0.0f, 1.0f);
    float m2 = m * m;
    return m2 * m2 * m;
}
WL_DEV float fresnel(float ci, float ct, float eta)
{
    const float rs = (ci - ct * eta) / (ci + eta * ct);
    const float rp = (ci * eta - ct) / (ci * eta + ct);
    return 0.5f * (rs * rs + rp * rp);
}
WL_DEV float gtr1(float NDotH, float a)
{
    if (a >= 1.0f)
        return (1.0f / kPi);
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return (a2 - 1.0f) / (kPi * logf(a2) * t);
}
WL_DEV float gtr2(float NDotH, float a)
{
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return a2 / (kPi * t * t);
}
WL_DEV float smithg(float NDotv, float alphaG)
{
    float a = alphaG * alphaG;
    float b = NDotv * NDotv;
    return 1.0f / (NDotv + sqrtf(a + b - a * b));
}
// Beckmann rational approximation with a fixed alpha of 0.5
WL_DEV float lambda_beckmann(V3 w, V3 n)
{
    float wDotn = dot(w, n);
    float absTanTheta = fabsf(length(cross(w, n)) / wDotn);
    if (__builtin_isinf(absTanTheta))
        return 0.0f;
    float a = 1.0f / (0.5f * absTanTheta);
    if (a >= 1.6f)
        return 0.0f;
    return (1.0f - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
}

struct Mat
{
    V3 color;
    float metallic, roughness, specular, specularTint, eta, trans, subsurface,
        anisotropic, sheen, sheenTint, clearcoat, clearcoatgloss;
};

WL_DEV Mat load_mat(const wl_disney_material *__restrict__ p)
{
    const float4 *q = reinterpret_cast<const float4 *>(p);
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    Mat m;
    m.color = mk(a.x, a.y, a.z);
    m.metallic = a.w;
    m.roughness = b.x;
    m.specular = b.y;
    m.specularTint = b.z;
    m.eta = b.w;
    m.trans = c.x;
    m.subsurface = c.y;
    m.anisotropic = c.z;
    m.sheen = c.w;
    m.sheenTint = d.x;
    m.clearcoat = d.y;
    m.clearcoatgloss = d.z;
    return m;
}

// Microfacet transmission lobe (reference Evaluate.h:10-76).
WL_DEV V3 eval_transmit(const Mat &mat, V3 Ns, V3 V, V3 L, V3 tex)
{
    V3 N = Ns;
    V3 combined = tex * mat.color;
    float NDotL = dot(N, L);
    float NDotV = dot(N, V);
    float eta = 1 / mat.eta;
    if (NDotL > 0)
    {
        eta = 1 / eta;
        N = -N;
    }
    if (NDotL == 0 || NDotV == 0)
        return mk(0, 0, 0);
    float refract = ((1 - NDotV * NDotV) * eta * eta >= 1) ? 1.0f : 0.0f;
    V3 wh = normalize(L + V * eta);
    if (dot(wh, N) < 0)
        wh = -wh;
    if (dot(L, wh) * dot(V, wh) > 0)
        return mk(0, 0, 0);
    float sqrtDenom = dot(L, wh) + eta * dot(V, wh);
    float factor = 1 / eta;
    V3 T = mat.trans * combined;
    float Gs = 1 / (1 + lambda_beckmann(V, N) + lambda_beckmann(L, N));
    float a = fmaxf(0.001f, mat.roughness);
    float Ds = gtr2(dot(N, wh), a);
    float F = fresnel(fabsf(dot(V, wh)), fabsf(dot(L, wh)), eta);
    float num = Ds * Gs * eta * eta * fabsf(dot(L, wh)) * fabsf(dot(V, wh)) * factor * factor;
    float den = NDotL * NDotL * sqrtDenom * sqrtDenom;
    return ((1 - refract) * (1.f - F)) * T * fabsf(num / den);
}

// Disney BSDF value (reference Evaluate.h:78-191).
WL_DEV V3 eval_disney(const Mat &mat, V3 Ns, V3 Ng, V3 V, V3 L, V3 tex)
{
    V3 combined = tex * mat.color;
    float NDotL = dot(Ns, L);
    float gNDotL = dot(Ng, L);
    float NDotV = dot(Ns, V);
    float gNDotV = dot(Ng, V);
    float eta = 1 / mat.eta;
    if (mat.trans < 0.1f && (gNDotL * gNDotV <= 0.0f || NDotL * NDotV <= 0))
        return mk(0, 0, 0);
    if (mat.trans > 0.9f && (gNDotL * gNDotV * NDotL * NDotV <= 0))
        return mk(0, 0, 0);
    if (NDotL * NDotV <= 0)
        return eval_transmit(mat, Ns, V, L, tex);
    if (mat.metallic + combined.x + combined.y + combined.z <= 0)
        return mk(0, 0, 0);
    if (NDotL < 0.0f && NDotV < 0.0f)
    {
        Ns = -Ns;
        eta = 1 / eta;
        NDotL *= -1;
        NDotV *= -1;
    }
    V3 H = normalize(L + V);
    float NDotH = dot(Ns, H);
    float LDotH = dot(L, H);
    float refract = ((1 - NDotV * NDotV) * eta * eta >= 1) ? 1.0f : 0.0f;
    V3 Cdlin = combined;
    float Cdlum = 0.3f * Cdlin.x + 0.6f * Cdlin.y + 0.1f * Cdlin.z;
    V3 Ctint = Cdlum > 0.0f ? Cdlin / Cdlum : mk(1, 1, 1);
    V3 Cspec0 = lerp3((mat.specular * 0.08f) * lerp3(mk(1, 1, 1), Ctint, mat.specularTint),
                      Cdlin, mat.metallic);
    V3 Csheen = lerp3(mk(1, 1, 1), Ctint, mat.sheenTint);
    float FL = schlick(NDotL), FV = schlick(NDotV);
    float Fd90 = 0.5f + 2.0f * LDotH * LDotH * mat.roughness;
    float Fd = lerpf(1.0f, Fd90, FL) * lerpf(1.0f, Fd90, FV);
    float Fss90 = LDotH * LDotH * mat.roughness;
    float Fss = lerpf(1.0f, Fss90, FL) * lerpf(1.0f, Fss90, FV);
    float ss = 1.25f * (Fss * (1.0f / (NDotL + NDotV) - 0.5f) + 0.5f);
    float a = fmaxf(0.001f, mat.roughness);
    float Ds = gtr2(NDotH, a);
    float FH = schlick(LDotH);
    V3 Fs = lerp3(Cspec0, mk(1, 1, 1), FH);
    float roughg = sqrtf(mat.roughness * 0.5f + 0.5f);
    float Gs = smithg(NDotL, roughg) * smithg(NDotV, roughg);
    V3 Fsheen = (FH * mat.sheen) * Csheen;
    float trans = mat.trans;
    float cosThetaI = fabsf(dot(Ns, V));
    float sin2ThetaI = 1 - cosThetaI * cosThetaI;
    float sin2ThetaT = eta * eta * sin2ThetaI;
    float cosThetaT = 1;
    if (sin2ThetaT <= 1)
        cosThetaT = sqrtf(1 - sin2ThetaT);
    float F = fresnel(cosThetaI, cosThetaT, eta);
    V3 out = ((((1.0f / kPi) * lerpf(Fd, ss, mat.subsurface)) * Cdlin + Fsheen) *
              (1.0f - mat.metallic)) *
             (1 - trans * (1 - F) * (1 - refract));
    if (trans > 0)
        out = add_scalar(out, Gs * Ds * (1 - trans * (1 - refract) * (1 - F)));
    else
        out = out + (Gs * Ds) * Fs;
    return out;
}

// Density of sample_disney's directions, as the reference computes it
// (reference Pdf.h:42-118).
WL_DEV float pdf_disney(const Mat &mat, V3 Ns, V3 V, V3 L)
{
    float transRatio = mat.trans;
    float eta = mat.eta;
    if (dot(Ns, V) < 0)
    {
        eta = 1 / eta;
        Ns = -Ns;
    }
    float NdotV = fabsf(dot(V, Ns));
    V3 wh = -normalize(V + L * eta);
    float clearcoatAlpha = lerpf(0.1f, 0.001f, mat.clearcoatgloss);
    float diffuseRatio = 0.5f * (1.f - mat.metallic);
    float specularRatio = 1.f - diffuseRatio;
    V3 half = normalize(L + V);
    float cosTheta = fabsf(dot(half, Ns));
    float pdfGTR2 = gtr2(cosTheta, mat.roughness) * cosTheta;
    float pdfGTR1 = gtr1(cosTheta, clearcoatAlpha) * cosTheta;
    float ratio = 1.0f / (1.0f + mat.clearcoat);
    float pdfSpec = lerpf(pdfGTR1, pdfGTR2, ratio) / (4.0f * fabsf(dot(L, half)));
    float pdfDiff = fabsf(dot(L, Ns)) * (1.f / kPi);
    float cosThetaI = fabsf(dot(wh, V));
    float sin2ThetaI = 1 - cosThetaI * cosThetaI;
    float sin2ThetaT = 1 / (eta * eta) * sin2ThetaI;
    float refractRatio = 0;
    float temp = (1 - NdotV * NdotV) / (eta * eta);
    if (temp < 1)
        refractRatio = 1 - fresnel(NdotV, sqrtf(1 - temp), eta);
    float pdf = (diffuseRatio * pdfDiff + specularRatio * pdfSpec) * (1 - transRatio * refractRatio);
    if (sin2ThetaT <= 1 && dot(L, wh) * dot(V, wh) < 0)
    {
        float sqrtDenom = eta * dot(L, wh) + dot(V, wh);
        float dwh_dwi = fabsf((eta * eta * dot(L, wh)) / (sqrtDenom * sqrtDenom));
        float a = fmaxf(0.001f, mat.roughness);
        float Ds = gtr2(fabsf(dot(wh, Ns)), a);
        float pdfTrans = Ds * fabsf(dot(Ns, wh)) * dwh_dwi;
        pdf += transRatio * pdfTrans * refractRatio;
    }
    cosThetaI = fabsf(dot(half, V));
    sin2ThetaI = 1 - cosThetaI * cosThetaI;
    sin2ThetaT = 1 / (eta * eta) * sin2ThetaI;
    if (sin2ThetaT > 1)
        pdf += (diffuseRatio * pdfDiff + specularRatio * pdfSpec) * transRatio * refractRatio;
    return pdf;
}

// Direction sampling, 7 draws (+2 on the diffuse lobe) in the reference's
// order (reference Sample.h:134-231, 24-28, 43-53).
WL_DEV V3 sample_disney(const Mat &mat, V3 Ns, V3 V, uint32_t &seed)
{
    float r1 = rnd(seed);
    float r2 = rnd(seed);
    float r3 = rnd(seed);
    float r4 = rnd(seed);
    float eta = 1 / mat.eta;
    if (dot(Ns, V) < 0)
    {
        eta = 1 / eta;
        Ns = -Ns;
    }
    float NdotV = fabsf(dot(Ns, V));
    float transprob = rnd(seed);
    float refractprob = rnd(seed);
    float probability = rnd(seed);
    float diffuseRatio = 0.5f * (1.0f - mat.metallic);
    if (transprob < mat.trans)
    {
        float refractRatio = 0;
        float temp = (1 - NdotV * NdotV) * (eta * eta);
        if (temp < 1)
            refractRatio = 1 - fresnel(NdotV, sqrtf(1 - temp), eta);
        if (refractprob < refractRatio)
        {
            V3 Ox, Oy;
            onb(Ns, Ox, Oy);
            float a = mat.roughness;
            float phi = r3 * kTwoPi;
            float cosTheta = sqrtf((1.0f - r4) / (1.0f + (a * a - 1.0f) * r4));
            float sinTheta = sqrtf(1.0f - (cosTheta * cosTheta));
            float sinPhi = sinf(phi);
            float cosPhi = cosf(phi);
            V3 half = mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
            half = half.x * Ox + half.y * Oy + half.z * Ns;
            if (dot(V, Ns) == 0)
                return -V;
            float cosThetaI = dot(half, V);
            float sin2ThetaI = 1 - cosThetaI * cosThetaI;
            float sin2ThetaT = eta * eta * sin2ThetaI;
            if (sin2ThetaT <= 1)
            {
                float cosThetaT = sqrtf(1 - sin2ThetaT);
                V3 L = normalize(eta * -V + (eta * cosThetaI - cosThetaT) * half);
                if (cosThetaI < 0)
                {
                    float LdotH = dot(L, half);
                    L = L - (2 * LdotH) * half;
                }
                return L;
            }
            return half * 2.f * dot(half, V) - V;
        }
    }
    V3 Ox, Oy;
    onb(Ns, Ox, Oy);
    if (probability < diffuseRatio)
    {
        float r = sqrtf(rnd(seed)), theta = rnd(seed) * kTwoPi;
        float dx = r * cosf(theta), dy = r * sinf(theta);
        float z = sqrtf(1 - dx * dx - dy * dy);
        return Ox * dx + Oy * dy + Ns * z;
    }
    float a = mat.roughness;
    float phi = r1 * kTwoPi;
    float cosTheta = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
    float sinTheta = sqrtf(1.0f - (cosTheta * cosTheta));
    float sinPhi = sinf(phi);
    float cosPhi = cosf(phi);
    V3 half = mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
    half = half.z * Ns + half.x * Ox + half.y * Oy;
    return (2.0f * dot(V, half)) * half - V;
}

WL_DEV float tri_area(V3 a, V3 b, V3 c) { return 0.5f * length(cross(b - a, c - a)); }

// ACES filmic fit (reference Tonemap.h:34-42).
WL_DEV float aces1(float r)
{
    r = r * 0.6f;
    float v = (r * (2.51f * r + 0.03f)) / (r * (2.43f * r + 0.59f) + 0.14f);
    return clampf(v, 0.0f, 1.0f);
}

// Texture fetch with the reference's texture object settings
// (DeviceManager.cpp:205-220): normalized coordinates, wrap addressing,
// bilinear filtering, texels read as u8 / 255.  The filter weights keep 8
// fractional bits (the fixed-point precision the CUDA programming guide states
// for hardware linear filtering), rounded to nearest.  Texels are RGBA8 words,
// rows top to bottom; a non-finite coordinate reads as 0.
WL_DEV float tex_weight_q8(float f) { return floorf(f * 256.0f + 0.5f) * (1.0f / 256.0f); }

WL_DEV V3 texel_rgb(uint32_t c)
{
    return mk((float)(c & 255u) / 255.0f, (float)((c >> 8) & 255u) / 255.0f,
              (float)((c >> 16) & 255u) / 255.0f);
}

WL_DEV V3 tex_bilinear(const uint32_t *texels, uint32_t w, uint32_t h, float u, float v)
{
    float fu = u - floorf(u), fv = v - floorf(v);
    if (!(fu >= 0.0f && fu <= 1.0f))
        fu = 0.0f;
    if (!(fv >= 0.0f && fv <= 1.0f))
        fv = 0.0f;
    const float x = fu * (float)w - 0.5f, y = fv * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = tex_weight_q8(x - fx), b = tex_weight_q8(y - fy);
    // x in [-0.5, w - 0.5] so the lower texel is in [-1, w - 1]
    int i0 = (int)fx, j0 = (int)fy;
    i0 = i0 < 0 ? i0 + (int)w : i0;
    j0 = j0 < 0 ? j0 + (int)h : j0;
    const int i1 = i0 + 1 == (int)w ? 0 : i0 + 1;
    const int j1 = j0 + 1 == (int)h ? 0 : j0 + 1;
    const V3 t00 = texel_rgb(texels[(size_t)j0 * w + i0]);
    const V3 t10 = texel_rgb(texels[(size_t)j0 * w + i1]);
    const V3 t01 = texel_rgb(texels[(size_t)j1 * w + i0]);
    const V3 t11 = texel_rgb(texels[(size_t)j1 * w + i1]);
    return ((t00 * ((1.0f - a) * (1.0f - b)) + t10 * (a * (1.0f - b))) + t01 * ((1.0f - a) * b)) +
           t11 * (a * b);
}

} // namespace wl
