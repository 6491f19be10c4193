"""GPU renders against the converged FP64 CPU oracle (65,536 samples per
pixel, committed under tests/golden/ with each pixel's one-sample variance):

* the Cornell box workload (256 x 256, 16 spp, depth 8, NEE + MIS): its
  relMSE against the oracle matches the value the fixture's variance
  predicts for 16 samples, and the image mean and every 16 x 16 block mean
  sit within their standard errors of the oracle (no bias);
* the glossy caustic at 1024 x 1024 with learned guiding (training on): the
  caustic window's block means over independent guided runs sit within
  their standard errors of the oracle, so the one-sample MIS of guide and
  BSDF leaves the estimate unbiased; every pixel of the image is finite.

The estimator is PathTracing.cu:40-159 (the oracle's restatement).
"""
import os
import sys

import numpy as np
import pytest

from wayland_amd import scenes

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_fp64_references as fx  # noqa: E402

pytestmark = pytest.mark.gpu


def _relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


def _block(a, b):
    h, w = a.shape[:2]
    return a.reshape(h // b, b, w // b, b, -1).sum(axis=(1, 3))


def test_cornell_converges_to_fp64_oracle(gpu_ctx):
    import wayland_amd as wl
    fix = fx.load("cornell_256_fp64")
    assert fix is not None, "tests/golden/cornell_256_fp64.* missing"
    ref, var, meta = fix
    res, spp = meta["width"], 16
    print(f"case: cornell {res}^2 x {spp} spp vs FP64 oracle ({meta['spp']} spp)", flush=True)
    sc = scenes.cornell_box(res, res)
    pt = wl.PathTracer(gpu_ctx, wl.Accel(gpu_ctx, sc), sc)
    cfg = wl.pt_config(res, res, spp=spp, max_depth=meta["max_depth"],
                       rr_depth=meta["rr_depth"], rr_prob=meta["rr_prob"])
    film, _ = pt.render(cfg, stats=False)
    img = pt.resolve(film, spp).cpu().numpy().astype(np.float64).reshape(res, res, 3)
    e = _relmse(img, ref)
    exp = fx.expected_relmse(ref, var, spp, meta["spp"])
    # the realised relMSE scatters around the expected value by a few per
    # cent (five independent 16-spp oracle renders: -6 % .. +9 %)
    print(f"relMSE {e:.6f}, expected {exp:.6f}, ratio {e / exp:.4f}")
    assert abs(e / exp - 1.0) < 0.15
    # bias: image mean and 16 x 16 block means against their standard errors
    sv = var / spp + var / meta["spp"]
    z_all = (img - ref).sum() / np.sqrt(sv.sum())
    zb = _block(img - ref, 16) / np.sqrt(_block(sv, 16))
    print(f"image-mean z {z_all:.2f}, block |z| max {np.abs(zb).max():.2f}")
    assert abs(z_all) < 4.0
    assert np.abs(zb).max() < 5.0
    pt.close()


def test_guided_caustic_1024_unbiased(gpu_ctx):
    import torch

    import wayland_amd as wl
    fix = fx.load("caustic_crop128_fp64")
    assert fix is not None, "tests/golden/caustic_crop128_fp64.* missing"
    ref, var, meta = fix
    res = meta["width"]
    x0, y0, ww, wh = meta["window"]
    runs, spp = 16, 16
    print(f"case: guided glossy caustic {res}^2, {runs} runs x {spp} spp vs FP64 oracle",
          flush=True)
    sc = scenes.glossy_caustic(res, res)
    pt = wl.PathTracer(gpu_ctx, wl.Accel(gpu_ctx, sc), sc)
    crops = []
    for r in range(runs):
        guide = wl.Guide(gpu_ctx, seed=r + 1)
        pt.set_guide(guide, wl.GUIDE_SAMPLE | wl.GUIDE_TRAIN)
        cfg = wl.pt_config(res, res, spp=1, max_depth=meta["max_depth"],
                           rr_depth=meta["rr_depth"], rr_prob=meta["rr_prob"],
                           max_batch_paths=res * res)
        film = torch.zeros((res * res, 6), dtype=torch.float32, device="cuda")
        for k in range(spp):
            cfg.frame_offset = (r << 20) + k
            pt.render(cfg, film, stats=False)
        steps = guide.info()["steps"]
        pt.set_guide(None)
        guide.close()
        img = pt.resolve(film, spp).cpu().numpy().astype(np.float64).reshape(res, res, 3)
        assert np.isfinite(img).all()
        assert steps == spp * guide.cfg.steps_per_batch
        crops.append(img[y0:y0 + wh, x0:x0 + ww])
    crops = np.stack(crops)
    # 32 x 32 blocks of the window (channel means): mean over the runs vs the
    # oracle, standard error from the spread of the independent runs (a t
    # statistic with runs - 1 degrees of freedom; |t15| > 5 has probability
    # 1.6e-4 per block)
    bm = np.stack([_block(c.mean(2, keepdims=True), 32) for c in crops]) / 1024.0
    rb = _block(ref.mean(2, keepdims=True), 32) / 1024.0
    se = bm.std(0, ddof=1) / np.sqrt(runs)
    z = (bm.mean(0) - rb) / np.maximum(se, 1e-12)
    ratio = crops.mean() / ref.mean()
    print(f"block |t| max {np.abs(z).max():.2f}, window mean ratio {ratio:.4f}")
    assert np.abs(z).max() < 5.0
    assert abs(ratio - 1.0) < 0.03
    pt.close()
