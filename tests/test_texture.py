"""The texture fetch on the CPU: the oracle's bilinear filter
(oracle/oracle_impl.inc tex_bilinear_rgba) pinned by properties that fix
it, and the pbrt loader's imagemap binding.

The filter restates the reference's texture object (DeviceManager.cpp:205-220:
normalized coordinates, wrap addressing, linear filtering, u8 read as
normalized float).  Its weights keep 8 fractional bits, the precision the
CUDA programming guide gives for hardware linear filtering; the rounding of
the weight (to nearest here) is not specified there, so results between two
1/256 steps are parity unpinned against real texture hardware."""
import math
import os

import numpy as np
import pytest

import oracle_bind as ob
from wayland_amd import pbrt


def _img(h, w, seed=0):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4), dtype=np.uint8)


def test_constant_image_is_constant_everywhere():
    img = np.zeros((5, 7, 4), np.uint8)
    img[..., :3] = (200, 17, 90)
    for u, v in [(0.3, 0.8), (-2.7, 5.1), (0.0, 0.0), (1.0, 1.0), (float("nan"), 0.5),
                 (float("inf"), -float("inf")), (1e7, -1e-9)]:
        got = ob.tex_bilinear(img, u, v)
        # the four weights sum to 1 exactly; each product rounds once
        assert np.allclose(got, np.float32([200, 17, 90]) / np.float32(255), rtol=3e-7,
                           atol=0), (u, v)


def test_texel_centres_return_the_texel():
    img = _img(6, 9, 1)
    for j in range(6):
        for i in range(9):
            got = ob.tex_bilinear(img, (i + 0.5) / 9, (j + 0.5) / 6)
            assert np.array_equal(got, img[j, i, :3].astype(np.float32) / np.float32(255))


def test_rows_run_top_to_bottom():
    # row 0 of the array is v near 0 (stbi_load order)
    img = np.zeros((4, 4, 4), np.uint8)
    img[0, :, 0] = 255
    assert ob.tex_bilinear(img, 0.5, 0.125)[0] == 1.0
    assert ob.tex_bilinear(img, 0.5, 0.875)[0] == 0.0


def test_linear_ramp_interpolates_with_quantised_weight():
    # R grows by 16 per column: between two texel centres the result is the
    # linear blend with the fraction rounded to a multiple of 1/256
    w = 16
    img = np.zeros((1, w, 4), np.uint8)
    img[0, :, 0] = 16 * np.arange(w)
    rng = np.random.default_rng(2)
    for u in rng.uniform(0.5 / w, (w - 0.5) / w, 200):
        x = u * w - 0.5
        i = math.floor(x)
        r = float(ob.tex_bilinear(img, u, 0.5)[0]) * 255 / 16 - i
        assert abs(r * 256 - round(r * 256)) < 1e-3          # weight is k / 256
        assert abs(r - (x - i)) <= 0.5 / 256 + 1e-6           # nearest step


def test_wrap_addressing():
    img = _img(3, 5, 3)
    # the seam at u = 0 blends the last and first columns half and half
    got = ob.tex_bilinear(img, 0.0, 0.5)
    want = (img[1, 4, :3].astype(np.float64) + img[1, 0, :3]) / 2 / 255
    assert np.allclose(got, want, atol=1e-6)
    # integer shifts of dyadic coordinates change nothing
    for u, v in [(0.375, 0.625), (0.0625, 0.8125), (0.96875, 0.03125)]:
        base = ob.tex_bilinear(img, u, v)
        for du, dv in [(1, 0), (-3, 2), (7, -5)]:
            assert np.array_equal(ob.tex_bilinear(img, u + du, v + dv), base)


def test_transpose_symmetry():
    # swapping the image axes and the coordinates gives the same texel blend
    img = _img(7, 4, 4)
    imgT = np.ascontiguousarray(img.transpose(1, 0, 2))
    rng = np.random.default_rng(5)
    for u, v in rng.uniform(-1, 2, (100, 2)):
        a = ob.tex_bilinear(img, u, v)
        b = ob.tex_bilinear(imgT, v, u)
        assert np.allclose(a, b, atol=2e-7)


def test_pbrt_imagemap_binding(tmp_path):
    from PIL import Image
    img = _img(8, 16, 6)
    Image.fromarray(img, "RGBA").save(tmp_path / "t.png")
    (tmp_path / "s.pbrt").write_text("""Camera "perspective"
WorldBegin
Texture "tex" "color" "imagemap" "string filename" [ "t.png" ]
Texture "sc" "color" "scale" "texture tex1" "tex" "rgb tex2" [0.5 0.5 0.5]
MakeNamedMaterial "A" "string type" [ "matte" ] "texture Kd" [ "tex" ]
MakeNamedMaterial "B" "string type" [ "disney" ] "texture color" [ "tex" ] "float metallic" [0.2]
MakeNamedMaterial "C" "string type" [ "matte" ] "rgb Kd" [0.1 0.2 0.3]
NamedMaterial "A"
Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 1 1 0] "float uv" [0 0 2 0 2 1] "integer indices" [0 1 2]
NamedMaterial "B"
Shape "trianglemesh" "point3 P" [0 0 1 1 0 1 1 1 1] "integer indices" [0 1 2]
NamedMaterial "C"
Shape "trianglemesh" "point3 P" [0 0 2 1 0 2 1 1 2] "integer indices" [0 1 2]
WorldEnd
""")
    s = pbrt.load_pbrt(str(tmp_path / "s.pbrt"))
    sc = s.scene
    assert len(sc.textures) == 1 and np.array_equal(sc.textures[0], img)
    a, b, c = sc.meshes
    assert a.texture == 0 and np.allclose(a.uv, [[0, 0], [2, 0], [2, 1]])
    assert np.allclose(a.material[:3], 1.0)             # textured Kd keeps the default 1
    assert b.texture == 0 and b.uv is None and b.material[3] == np.float32(0.2)
    assert c.texture is None
    flat = sc.flatten()
    assert flat["mesh_texture"].tolist() == [0, 0, 0xFFFFFFFF]
    assert flat["texture_size"].tolist() == [[16, 8]]
    assert np.array_equal(flat["texels"], img.reshape(-1, 4))
    # a mesh without uv reads texcoords (0, 0)
    assert np.array_equal(flat["texcoords"][3:6], np.zeros((3, 2), np.float32))
    assert any("scale" in n for n in s.notes)
