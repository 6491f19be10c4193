"""Headless frame loop (wayland_amd/render.py) on the CPU: command line, output
names and statistics in the reference's format (Renderer.cpp:60-106,
157-209), PNG encoding and the vertical flip."""
import io
import os
import re

import numpy as np
import pytest

from wayland_amd import render


def test_command_line_mirrors_renderer_setting():
    s = render.RendererSetting.from_args([])
    assert (s.frame_limit, s.time_limit, s.program, s.resolution, s.write_output) == \
        (-1, -1.0, "PathTracing", (1000, 1000), False)
    s = render.RendererSetting.from_args(
        ["-fr", "12", "-t", "2.5", "-s", "a/b/cbox.pbrt", "-o", "out", "-p", "BDPT",
         "-r", "640", "360"])
    assert (s.frame_limit, s.time_limit, s.scene_path, s.output_path, s.program,
            s.resolution, s.write_output) == (12, 2.5, "a/b/cbox.pbrt", "out", "BDPT",
                                              (640, 360), True)
    assert s.scene_name == "cbox"
    # an unknown program name leaves the default (programMap lookup fails)
    assert render.RendererSetting.from_args(["-p", "Nope"]).program == "PathTracing"


@pytest.mark.parametrize("t,label", [(1.99, "1.9s"), (12.04, "12.0s"), (0.5, "0.5s"),
                                     (123.456, "123.4s")])
def test_render_time_label_truncates(t, label):
    assert render.render_time_label(t) == label


def test_time_stamp_replaces_two_colons():
    s = render.time_stamp(0.0)
    assert re.fullmatch(r"\[\w{3} \w{3} [ \d]\d \d\d_\d\d_\d\d \d{4}\]", s), s


def test_stats_text_format():
    txt = render.stats_text("cbox", "PathTracing", (64, 48), 2.5, 10)
    assert txt == ("{\n\tScene: cbox\n\tProgram: PathTracing\n\tResolution: 64 48\n"
                   "\tRenderTime: 2.5\n\tFrameCount: 10\n\tFrameTime: 0.25\n"
                   "\tFramesPerSecond: 4\n}\n")
    # six significant digits, as an ostream prints a float
    assert "\tFrameTime: 0.333333\n" in render.stats_text("x", "PathTracing", (1, 1), 1.0, 3)


def test_png_round_trip_and_flip():
    from PIL import Image
    rng = np.random.default_rng(0)
    fb = rng.integers(0, 256, (5 * 7, 4), dtype=np.uint8)
    img = render.flip_frame_buffer(fb, 7, 5)
    # the reference reverses the whole buffer, then each row: a vertical flip
    rev = fb[::-1].reshape(5, 7, 4)[:, ::-1]
    assert np.array_equal(img, rev)
    dec = np.asarray(Image.open(io.BytesIO(render.png_bytes(img))))
    assert dec.shape == (5, 7, 4) and np.array_equal(dec, img)


def test_write_output_names_and_files(tmp_path):
    from PIL import Image
    r = render.Renderer.__new__(render.Renderer)
    r.setting = render.RendererSetting.from_args(["-s", "x/cbox.pbrt", "-o", str(tmp_path),
                                                  "-r", "4", "3"])
    r.width, r.height = 4, 3
    r.frame_buffer = np.arange(48, dtype=np.uint8).reshape(12, 4)
    base = r.write_output(1.25, 5, now=0.0)
    name = os.path.basename(base)
    assert name.startswith("cbox-PathTracing-1.2s-[")
    assert open(base + ".txt").read() == render.stats_text("cbox", "PathTracing", (4, 3),
                                                            1.25, 5)
    dec = np.asarray(Image.open(base + ".png"))
    assert np.array_equal(dec[2, 0], [0, 1, 2, 3])      # buffer row 0 -> image bottom


def test_headless_run_needs_a_limit():
    r = render.Renderer.__new__(render.Renderer)
    r.setting = render.RendererSetting.from_args(["-s", "x.pbrt"])
    with pytest.raises(ValueError):
        r.run()


def test_other_programs_raise():
    with pytest.raises(NotImplementedError):
        render.Renderer(render.RendererSetting.from_args(["-s", "x.pbrt", "-p", "AO"]))


def test_save_image_cpu_jpeg(tmp_path):
    from PIL import Image
    yy, xx = np.mgrid[0:32, 0:48]
    rgba = np.zeros((32, 48, 4), np.uint8)
    rgba[..., 0], rgba[..., 1], rgba[..., 2], rgba[..., 3] = 5 * xx, 7 * yy, 128, 255
    path = str(tmp_path / "img.jpg")
    render.save_image_cpu(path, 48, 32, rgba.reshape(-1, 4))
    with Image.open(path) as im:
        assert im.format == "JPEG" and im.size == (48, 32)
        dec = np.asarray(im.convert("RGB")).astype(int)
    # a smooth image survives quality-90 coding to a few units
    assert np.abs(dec - rgba[..., :3]).mean() < 2.0
