"""The headless frame loop on the GPU: the PNG it writes after N progressive
frames of tests/golden/cbox.pbrt is the oracle's film update of the same
frames (running mean + ACES u8, PathTracing.cu:140-159), flipped vertically
as Renderer::WriteOutput writes it."""
import glob
import os

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_frame_limit_run_writes_the_oracle_image(gpu_ctx, tmp_path):
    from PIL import Image

    from wayland_amd import render
    print("case: headless cbox.pbrt 40x32, 4 frames", flush=True)
    W, H, N = 40, 32, 4
    s = render.RendererSetting.from_args(["-s", os.path.join(GOLDEN, "cbox.pbrt"), "-fr", str(N),
                                          "-o", str(tmp_path), "-r", str(W), str(H)])
    r = render.Renderer(s, ctx=gpu_ctx)
    frames, seconds = r.run()
    r.close()
    assert frames == N and seconds > 0
    (png,) = glob.glob(str(tmp_path / "cbox-PathTracing-*.png"))
    (txt,) = glob.glob(str(tmp_path / "cbox-PathTracing-*.txt"))
    assert f"\tFrameCount: {N}\n" in open(txt).read()
    img = np.asarray(Image.open(png))
    assert img.shape == (H, W, 4)
    opt = ob.OraclePT(r.loaded.scene)
    oacc = np.zeros((W * H, 4), np.float32)
    ocol = np.zeros((W * H, 4), np.uint8)
    for f in range(N):
        o, _ = opt.render_frame(opt.config(W, H), f)
        ob.film_update(oacc, ocol, o, f)
    want = ocol.reshape(H, W, 4)[::-1]
    close = (np.abs(img.astype(int) - want.astype(int)) <= 1).all(-1).mean()
    print(f"pixels within 1 LSB: {close:.5f}")
    assert close >= 0.999
    assert (img[..., 3] == 255).all()


def test_time_limit_run(gpu_ctx):
    from wayland_amd import render
    print("case: headless cbox.pbrt 64x64, 0.2 s limit", flush=True)
    s = render.RendererSetting.from_args(["-s", os.path.join(GOLDEN, "cbox.pbrt"), "-t", "0.2",
                                          "-r", "64", "64"])
    r = render.Renderer(s, ctx=gpu_ctx)
    frames, seconds = r.run()
    r.close()
    print(f"{frames} frames in {seconds:.3f} s")
    assert frames >= 1 and seconds > 0.2
    assert r.frame_id == frames
