// wl_guide.h — device buffers of a guiding network and the host functions
// that enqueue its kernels (guide.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wl
{

// Samples per weight-gradient split: the split-K reduction of dW runs over
// kGuideSplits slices of the mini-batch.
constexpr int kGuideSplits = 32;

struct GuideDev
{
    float *params;   // 18720 FP32 master weights
    float *adam_m, *adam_v;
    void *fwd;       // packed BF16 forward operands (2304 x 16 B)
    void *bwd;       // packed BF16 backward operands (1792 x 16 B)
    float *partials; // kGuideSplits x 18720 gradient partial sums
    void *xt;        // BF16 layer inputs, feature-major [5][64][cap]
    void *dt;        // BF16 deltas, feature-major [4 * 64 + 32][cap]
    float *loss;     // per-sample loss [cap]
    uint32_t *count; // [0] samples in the training buffer, [1] scratch
    float *train;    // training samples, 16 floats each [train_cap]
    uint32_t cap;        // mini-batch capacity (multiple of 32)
    uint32_t train_cap;  // training buffer capacity
};

struct GuideAdam
{
    float lr, beta1, beta2, eps;
    float c1, c2; // bias corrections 1 - beta^t
};

hipError_t guide_pack(const GuideDev &G, hipStream_t s);
hipError_t guide_infer(const GuideDev &G, const float *queries, uint32_t n, float *out,
                       hipStream_t s);
hipError_t guide_mixture_sample(const float *outs, const float *u, uint32_t n, float *dirs,
                                float *pdf, hipStream_t s);
hipError_t guide_mixture_pdf(const float *outs, const float *dirs, uint32_t n, float *pdf,
                             hipStream_t s);
hipError_t guide_set_count(uint32_t *count, uint32_t n, hipStream_t s);
// count[2] = min(count[0], cap), count[0] = 0
hipError_t guide_roll_count(uint32_t *count, uint32_t cap, hipStream_t s);
// One gradient evaluation over mini-batch `step` of samples[0, n) with
// n = min(*n_dev, n_cap): offset = (step * batch) mod n, size
// min(batch, n - offset).  With adam set the parameters take one Adam step
// and are re-packed; with grad set the mean gradient is written there; with
// loss set the mean loss.  An empty batch changes nothing.
hipError_t guide_step(const GuideDev &G, const float *samples, const uint32_t *n_dev,
                      uint32_t n_cap, uint32_t step, uint32_t batch, const GuideAdam *adam,
                      float *grad, float *loss, int num_cus, hipStream_t s);

} // namespace wl
