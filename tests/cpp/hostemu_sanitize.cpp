// Drives the CPU emulation library (wayland_amd/csrc/hostemu.cpp: the BVH8
// builder, the traversal code of the gfx950 kernels and the serving
// schedule's index arithmetic) under AddressSanitizer and
// UndefinedBehaviorSanitizer (built and run by tests/test_sanitize.py).
// Exits non-zero if the emulation reports an out-of-range index; the
// sanitizers abort on any invalid access or undefined behaviour.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <cmath>
#include <vector>

extern "C" {
int wlh_build(const float *tris, uint32_t ntris, void **out, char *err, int errlen);
void wlh_free(void *p);
void wlh_info(void *p, uint32_t *out6);
uint64_t wlh_trace(void *p, const float *rays, uint32_t n, float *hits, int any,
                   uint32_t *counts, int nthreads);
uint64_t wlh_serve_check(const uint32_t *req4, uint32_t n, uint32_t capacity, uint64_t *out4);
}

static uint32_t g_state = 2024u;
static float frand()
{
    g_state = g_state * 1664525u + 1013904223u;
    return (float)(g_state >> 8) * (1.0f / 16777216.0f);
}

int main()
{
    uint64_t faults = 0;
    // soups of 1, 7, 300 and 5000 triangles, including degenerate ones
    for (uint32_t nt : { 1u, 7u, 300u, 5000u })
    {
        std::vector<float> tris(9 * (size_t)nt);
        for (auto &v : tris)
            v = frand();
        for (uint32_t t = 0; t < nt; t += 13) // a zero-area triangle
            for (int k = 0; k < 3; ++k)
                tris[9 * (size_t)t + 6 + k] = tris[9 * (size_t)t + 3 + k];
        void *h = nullptr;
        char err[256];
        if (wlh_build(tris.data(), nt, &h, err, sizeof err) != 0)
        {
            fprintf(stderr, "build failed: %s\n", err);
            return 2;
        }
        uint32_t info[6];
        wlh_info(h, info);
        const uint32_t nr = 4000;
        std::vector<float> rays(8 * (size_t)nr), hits(4 * (size_t)nr);
        std::vector<uint32_t> counts(2 * (size_t)nr);
        for (uint32_t i = 0; i < nr; ++i)
        {
            float *r = &rays[8 * (size_t)i];
            r[0] = frand() * 1.6f - 0.3f;
            r[1] = frand() * 1.6f - 0.3f;
            r[2] = frand() * 1.6f - 0.3f;
            r[3] = i % 9 ? 0.0f : 0.2f;
            const float z = 1 - 2 * frand(), ph = 6.2831853f * frand();
            const float s = std::sqrt(std::fmax(0.f, 1 - z * z));
            r[4] = s * std::cos(ph);
            r[5] = s * std::sin(ph);
            r[6] = i % 17 ? z : 0.0f; // some axis-parallel directions
            r[7] = i % 5 ? INFINITY : 0.4f;
        }
        faults += wlh_trace(h, rays.data(), nr, hits.data(), 0, counts.data(), 3);
        faults += wlh_trace(h, rays.data(), nr, hits.data(), 1, nullptr, 2);
        printf("%u tris: %u nodes, depth %u\n", nt, info[0], info[2]);
        wlh_free(h);
    }
    // serving schedules
    for (uint32_t trial = 0; trial < 40; ++trial)
    {
        const uint32_t cap = 1u << (8 + trial % 6);
        const uint32_t n = 1 + trial % 23;
        std::vector<uint32_t> req(4 * (size_t)n);
        for (uint32_t k = 0; k < n; ++k)
        {
            uint32_t w = 1 + (uint32_t)(frand() * 40), hh = 1 + (uint32_t)(frand() * 40);
            while (w * hh > cap)
                hh = hh > 1 ? hh / 2 : 1, w = w > 1 && w * hh > cap ? w / 2 : w;
            req[4 * k] = w;
            req[4 * k + 1] = hh;
            req[4 * k + 2] = (uint32_t)(frand() * 7);
            req[4 * k + 3] = (uint32_t)(frand() * 100000);
        }
        uint64_t out[4];
        faults += wlh_serve_check(req.data(), n, cap, out);
    }
    printf(faults ? "HOSTEMU_SANITIZE_FAILED %llu\n" : "HOSTEMU_SANITIZE_OK\n",
           (unsigned long long)faults);
    return faults ? 1 : 0;
}
