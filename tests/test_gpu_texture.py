"""Textured shading on the GPU against the oracle, per sample with identical
TEA seeds (PathTracing.cu:261-268: the filtered texel multiplies the base
colour for the bounce and for the NEE at that vertex).

The Cornell box gets three bindings: a 37x23 image on the back wall, a 16x16
image on the floor with texcoords spanning [-1.5, 2.5] (wrap addressing), and
a textured short box without texcoords (reads (0, 0)).  A pbrt file with an
imagemap texture, and a serving request on the textured scene, take the same
path."""
import os

import numpy as np
import pytest

import oracle_bind as ob
from wayland_amd import scenes

pytestmark = pytest.mark.gpu

REL = 1e-5


def _frac_within(a, b, rel=REL, floor=1e-6):
    err = np.abs(a - b) / np.maximum(np.abs(b), floor)
    return float((err.reshape(-1, 3).max(1) <= rel).mean())


def test_textured_cornell_per_sample_parity(gpu_ctx):
    import wayland_amd as wl
    print("case: textured cornell 64x64 frames 0-2", flush=True)
    sc = scenes.textured_cornell_box(64, 64)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    opt = ob.OraclePT(sc)
    plain = scenes.cornell_box(64, 64)
    opt_plain = ob.OraclePT(plain)
    for frame in range(3):
        g = pt.render_samples(wl.pt_config(64, 64), frame).cpu().numpy()
        o, _ = opt.render_frame(opt.config(64, 64), frame)
        frac = _frac_within(g, o)
        print(f"frame {frame}: within 1e-5 {frac:.5f}", flush=True)
        assert frac >= 0.999
        # the textures change the image
        p, _ = opt_plain.render_frame(opt_plain.config(64, 64), frame)
        assert _frac_within(g, p) < 0.9
    # removing the textures gives back the untextured render
    pt.set_textures([], [])
    g = pt.render_samples(wl.pt_config(64, 64), 0).cpu().numpy()
    p, _ = opt_plain.render_frame(opt_plain.config(64, 64), 0)
    assert _frac_within(g, p) >= 0.999
    pt.close()
    accel.close()


def test_pbrt_imagemap_per_sample_parity(gpu_ctx, tmp_path):
    from PIL import Image

    import wayland_amd as wl
    from wayland_amd import pbrt
    print("case: cbox.pbrt + imagemap back wall 48x48 frames 0-1", flush=True)
    img = np.random.default_rng(3).integers(0, 256, (20, 30, 4), dtype=np.uint8)
    img[..., 3] = 255
    Image.fromarray(img, "RGBA").save(tmp_path / "wall.png")
    src = open(os.path.join(os.path.dirname(__file__), "golden", "cbox.pbrt")).read()
    src = src.replace('\nNamedMaterial "BackWall"\n',
                      '\nTexture "w" "color" "imagemap" "string filename" [ "wall.png" ]\n'
                      'MakeNamedMaterial "TexWall" "string type" [ "matte" ] "texture Kd" [ "w" ]\n'
                      'NamedMaterial "TexWall"\n', 1)
    (tmp_path / "t.pbrt").write_text(src)
    s = pbrt.load_pbrt(str(tmp_path / "t.pbrt"), 48, 48)
    assert len(s.scene.textures) == 1
    assert sum(m.texture is not None for m in s.scene.meshes) == 1
    accel = wl.Accel(gpu_ctx, s.scene)
    pt = wl.PathTracer(gpu_ctx, accel, s.scene)
    opt = ob.OraclePT(s.scene)
    for frame in range(2):
        g = pt.render_samples(wl.pt_config(48, 48), frame).cpu().numpy()
        o, _ = opt.render_frame(opt.config(48, 48), frame)
        frac = _frac_within(g, o)
        print(f"frame {frame}: within 1e-5 {frac:.5f}", flush=True)
        assert frac >= 0.999
    pt.close()
    accel.close()


def test_textured_scene_serves_bit_identically(gpu_ctx):
    import wayland_amd as wl
    print("case: serve textured cornell, 2 requests", flush=True)
    sc = scenes.textured_cornell_box(32, 32)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    reqs = [{"pt": pt, "width": 32, "height": 32, "spp": 3},
            {"pt": pt, "width": 16, "height": 24, "spp": 2, "frame_offset": 5}]
    server = wl.Server(gpu_ctx, batch_paths=1 << 12)
    films, _, _ = server.render(reqs)
    for k, r in enumerate(reqs):
        cfg = wl.pt_config(r["width"], r["height"], spp=r["spp"],
                           frame_offset=r.get("frame_offset", 0))
        ref, _ = pt.render(cfg, stats=False)
        assert np.array_equal(films[k].cpu().numpy().view(np.uint32),
                              ref.cpu().numpy().view(np.uint32)), f"request {k}"
    server.close()
    pt.close()
    accel.close()
