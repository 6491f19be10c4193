// wl_program.h — program discovery, pipelines, SBTs and launches: the
// built-in integrators behind the OptiX program interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wayland_hip.h"
#include "bvh8.h"

namespace wl
{

// Everything k_example_pt reads.
struct ExampleLaunch
{
    wl_example_params params;
    const WlBvh8Node *nodes;
    const WlTriRecord *tris;
    const uint8_t *hit_records; // device copy of the SBT hit records
    size_t hit_stride;
    uint32_t width, height;
};

hipError_t launch_example_pt(const ExampleLaunch &L, int num_cus, hipStream_t st);
hipError_t launch_float_to_u8(const float *src, size_t n, uint8_t *dst, hipStream_t st);

} // namespace wl
