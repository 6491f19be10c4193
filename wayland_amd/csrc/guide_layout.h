// guide_layout.h — sizes and index maps of the guiding network's parameters
// and packed MFMA operands (host and device).  The tiling they describe is
// documented in guide_device.h.
#pragma once

#if defined(__HIPCC__)
#define WL_GUIDE_HD __host__ __device__
#else
#define WL_GUIDE_HD
#endif

namespace wl
{

constexpr int kGuideParams = 18720;
constexpr int kGuideOut = 32;
constexpr int kGuideLobes = 8;
// packed BF16 operands, in 16-byte units (8 bf16 per lane and tile)
constexpr int kGuideFwdVecs = 2304; // layers 0-3: 4 tiles x 2 k-blocks x 64 lanes; layer 4: 2 x 2 x 64
constexpr int kGuideBwdVecs = 1792; // layers 1-3: 4 x 2 x 64; layer 4: 4 x 1 x 64
constexpr int kGuideBiases = 4 * 64 + 32;

WL_GUIDE_HD constexpr int guide_w_off(int l) { return l * (64 * 64 + 64); }
WL_GUIDE_HD constexpr int guide_b_off(int l) { return guide_w_off(l) + (l < 4 ? 64 : 32) * 64; }
// output row of a lobe's parameter (0 logit, 1 raw kappa, 2 theta, 3 phi)
WL_GUIDE_HD constexpr int guide_row(int lobe, int param)
{
    return 16 * (lobe % 2) + 4 * (lobe / 2) + param;
}
// feature held by lane group g in k-block kb, element j of a B operand
WL_GUIDE_HD constexpr int guide_phi(int kb, int g, int j)
{
    return 16 * (2 * kb + j / 4) + 4 * g + (j % 4);
}
// forward operand of layer l, output tile t, k-block kb (16-byte units)
WL_GUIDE_HD constexpr int guide_fwd_vec(int l, int t, int kb)
{
    return (l < 4 ? l * 512 : 2048) + (t * 2 + kb) * 64;
}
// backward operand (W_l^T) of layer l = 1..4, input tile t, k-block kb
WL_GUIDE_HD constexpr int guide_bwd_vec(int l, int t, int kb)
{
    return (l < 4 ? (l - 1) * 512 + (t * 2 + kb) * 64 : 1536 + t * 64);
}

} // namespace wl
