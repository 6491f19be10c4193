// example_pt.hip — the examples/PathTracer program (PathTracer.cu:49-144)
// as a gfx950 kernel, and FromFloatToChar (Conversion.cu:6-19).
//
// The reference recurses through optixTrace from its closest-hit program and
// multiplies the returned colour on the way back:
//   C = (((C_child * color) * cosWeight) * 2) * RR_factor.
// Here one lane runs the whole path iteratively, keeps the per-bounce factors
// in a small private array and applies them innermost first, so the result
// rounds exactly as the recursion does.  The seed is threaded through the
// path and across samples the way the payload registers carry it.
#include "shading_device.h"
#include "traverse_device.h"
#include "wl_program.h"

namespace wl
{

constexpr int kExBlock = 256;
constexpr int kExMaxLevels = 32;

struct ExFactor
{
    float r, g, b, cw;
};

__device__ __forceinline__ V3 ex_sample_dir(V3 n, uint32_t &seed)
{
    V3 x, y;
    onb(n, x, y);
    const float angle = rnd(seed) * 2 * kPi, h = rnd(seed);
    return (sinf(angle) * x + cosf(angle) * y) * sqrtf(1 - h * h) + h * n;
}

__global__ __launch_bounds__(kExBlock) void k_example_pt(ExampleLaunch L)
{
    __shared__ uint32_t stack[WL_BVH8_STACK_DEPTH * kExBlock];
    const uint32_t W = L.width, H = L.height;
    const uint32_t npix = W * H;
    const wl_example_params &P = L.params;
    const V3 gaze = ld3(P.camera.gaze), up = ld3(P.camera.up), pos = ld3(P.camera.position);
    const V3 hvec = cross(gaze, up) * P.camera.horizontalRatio;
    const V3 vvec = up * P.camera.verticalRatio;
    const float csx = 1.f / (float)W, csy = 1.f / (float)H;
    for (uint32_t base = blockIdx.x * kExBlock; base < npix; base += gridDim.x * kExBlock)
    {
        const uint32_t i = base + threadIdx.x;
        if (i >= npix)
            continue;
        // launch index .x is the row (PathTracer.cu:51): i = W * row + col
        const uint32_t row = i / H, col = i % H;
        const float vpos = 1 - 2.f * (float)row * csx;
        const float hpos = 2.f * (float)col * csy - 1;
        const V3 rayDir = normalize(gaze + hpos * hvec + vpos * vvec);
        const uint32_t idx = W * row + col;
        uint32_t seed = tea4(idx, 10086);
        V3 result = mk(0, 0, 0);
        const float weight = 1.f / (float)P.sampleNum;
        ExFactor fac[kExMaxLevels];
        uint8_t rrbit[kExMaxLevels];
        for (uint32_t smp = 0; smp < P.sampleNum; ++smp)
        {
            const float r1 = rnd(seed);
            V3 dir = rayDir + ((r1 * 2) * csy) * hvec;
            const float r2 = rnd(seed);
            dir = dir + ((r2 * 2) * csx) * vvec;
            V3 o = pos;
            float tmin = 0.f;
            uint32_t depth = P.maxDepth;
            int n = 0;
            V3 leaf = mk(0, 0, 0);
            for (;;)
            {
                TravRay r;
                trav_setup(r, o.x, o.y, o.z, dir.x, dir.y, dir.z, tmin, 30.f);
                uint32_t nn = 0, nt = 0;
                const TravResult h = trav_run<false, false, kExBlock, true>(
                    L.nodes, L.tris, r, true, stack + threadIdx.x, nn, nt);
                if (h.prim == WL_INVALID)
                {
                    leaf = mk(0, 0, 0); // __miss__Empty
                    break;
                }
                const wl_example_hit_data *hd = reinterpret_cast<const wl_example_hit_data *>(
                    L.hit_records + (size_t)h.input * L.hit_stride + WL_SBT_RECORD_HEADER_SIZE);
                const V3 emission = ld3(hd->emission);
                if (depth == 0 || !(emission.x == 0 && emission.y == 0 && emission.z == 0))
                {
                    leaf = emission;
                    break;
                }
                bool rr = false;
                if (depth <= 5)
                {
                    if (rnd(seed) > 0.99f)
                    {
                        leaf = emission;
                        break;
                    }
                    rr = true;
                }
                const int32_t *ix = hd->indices + 3 * (size_t)h.local;
                const float *nrm = hd->normals;
                const V3 n0 = ld3(nrm + 3 * (size_t)ix[0]), n1 = ld3(nrm + 3 * (size_t)ix[1]),
                         n2 = ld3(nrm + 3 * (size_t)ix[2]);
                const V3 normal = normalize(n0 * (1 - h.b1 - h.b2) + n1 * h.b1 + n2 * h.b2);
                const V3 hit = o + dir * h.t;
                const V3 nd = ex_sample_dir(normal, seed);
                const float cw = fmaxf(0.f, dot(nd, normal));
                if (n < kExMaxLevels)
                {
                    fac[n] = ExFactor{ hd->color[0], hd->color[1], hd->color[2], cw };
                    rrbit[n] = rr ? 1 : 0;
                }
                ++n;
                o = hit;
                dir = nd;
                tmin = 1e-5f;
                depth -= 1;
            }
            V3 c = leaf;
            for (int k = (n < kExMaxLevels ? n : kExMaxLevels) - 1; k >= 0; --k)
            {
                c = c * mk(fac[k].r, fac[k].g, fac[k].b);
                c = c * fac[k].cw;
                c = c * 2.0f;
                c = c * (rrbit[k] ? 1.0f / 0.99f : 1.0f);
            }
            result = result + c * weight;
        }
        float4 *out = reinterpret_cast<float4 *>(P.colorBuffer);
        out[idx] = make_float4(result.x, result.y, result.z, 1.0f);
    }
}

hipError_t launch_example_pt(const ExampleLaunch &L, int num_cus, hipStream_t st)
{
    const uint32_t npix = L.width * L.height;
    const uint32_t g = std::min<uint32_t>((npix + kExBlock - 1) / kExBlock, (uint32_t)num_cus * 5);
    hipLaunchKernelGGL(k_example_pt, dim3(g), dim3(kExBlock), 0, st, L);
    return hipGetLastError();
}

__global__ void k_float_to_u8(const float *src, size_t n, uint8_t *dst)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        float v = src[i];
        if (v > 1.0f)
            v = 1.0f;
        dst[i] = (uint8_t)(v * 255);
    }
}

hipError_t launch_float_to_u8(const float *src, size_t n, uint8_t *dst, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    const size_t g = std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_float_to_u8, dim3((uint32_t)g), dim3(256), 0, st, src, n, dst);
    return hipGetLastError();
}

} // namespace wl
