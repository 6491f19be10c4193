"""pbrt-v3 ingestion (wayland_amd/pbrt.py) on the CPU.

The camera frame, the transform directives and the PLY reader are checked
against closed forms and against files written here.  The material mapping is
checked against the values that Material.cpp:33-107 assigns.  The Cornell box
fixture is tests/golden/cbox.pbrt, a CC0 file from scene/cornell-box."""
import math
import os
import struct

import numpy as np
import pytest

from wayland_amd import pbrt, scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CBOX = os.path.join(GOLDEN, "cbox.pbrt")
WATER = os.path.join(GOLDEN, "water-caustic", "water.pbrt")


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


HEADER = """Film "image" "integer xresolution" [ 200 ] "integer yresolution" [ 100 ]
{xf}
Camera "perspective" "float fov" [ 40 ]
WorldBegin
"""


def test_cbox_structure():
    s = pbrt.load_pbrt(CBOX)
    sc = s.scene
    assert (s.width, s.height, s.fov) == (1024, 1024, 19.5)
    assert len(sc.meshes) == 8 and sc.num_triangles == 36
    lights = [m for m in sc.meshes if m.light_L is not None]
    assert len(lights) == 1 and lights[0].light_L == (50.0, 35.0, 15.0)
    assert lights[0].two_sided is False
    assert np.allclose(lights[0].vertices[:, 1], 1.98)
    # the light's own material: matte Kd 0 -> black Disney base
    assert np.array_equal(lights[0].material[:3], [0, 0, 0])
    # left wall red, right wall green (MakeNamedMaterial Kd)
    cols = np.stack([m.material[:3] for m in sc.meshes])
    for c in ((0.63, 0.065, 0.05), (0.14, 0.45, 0.091)):
        assert np.isclose(cols, c).all(1).sum() == 1
    # every wall vertex inside the [-1,1]x[0,2]x[-1,1] room
    v = np.concatenate([m.vertices for m in sc.meshes])
    assert v.min() >= -1.0 - 1e-6 and v.max() <= 2.0 + 1e-6
    # flatten packs the normals each mesh carries
    flat = sc.flatten()
    assert flat["mesh_has_normal"].tolist() == [1] * 8
    assert flat["mesh_light"].tolist()[-1] == 0


def test_cbox_camera_closed_form():
    # Transform [1 0 0 0  0 1 0 0  0 0 -1 0  0 -1 6.8 1] is world->camera;
    # its inverse puts the eye at (0, 1, 6.8) looking down -z with +y up.
    s = pbrt.load_pbrt(CBOX, 300, 200)
    cam = s.scene.camera
    t = math.tan(math.radians(19.5 / 2))
    assert np.allclose(cam.pos, [0, 1, 6.8], atol=1e-6)
    assert np.allclose(cam.lookat, [0, 0, -1], atol=1e-7)
    assert np.allclose(cam.up, [0, t, 0], rtol=1e-6, atol=1e-7)
    # right = cross(lookAt, up) * W/H (Camera.cpp:27, PathTracingProgramManager.cpp:25-28)
    assert np.allclose(cam.right, [t * 1.5, 0, 0], rtol=1e-6, atol=1e-7)
    # the same frame as the synthetic Cornell box's camera
    ref = scenes.cornell_box(300, 200).camera
    for a, b in ((cam.pos, ref.pos), (cam.lookat, ref.lookat), (cam.up, ref.up),
                 (cam.right, ref.right)):
        assert np.allclose(a, b, atol=1e-6)


def test_cbox_walls_match_the_synthetic_room():
    # the synthetic Cornell box keeps the file's five walls and the light
    # position; only the two boxes differ (axis-aligned there)
    s = pbrt.load_pbrt(CBOX).scene
    ref = scenes.cornell_box(64, 64)
    room = np.concatenate([m.vertices for m in s.meshes[:5]])
    room_ref = np.concatenate([m.vertices for m in ref.meshes])
    for p in np.round(room, 5):
        assert np.isclose(room_ref, p, atol=1e-5).all(1).any(), p


@pytest.mark.parametrize("xf,eye,look,up", [
    ('LookAt 1 2 3  1 2 -7  0 1 0', (1, 2, 3), (0, 0, -1), (0, 1, 0)),
    ('LookAt 0 0 0  5 0 0  0 0 1', (0, 0, 0), (1, 0, 0), (0, 0, 1)),
    # a camera placed by Translate then Rotate: world->camera = T * R,
    # camera->world = R^-1 * T^-1; eye = -(translation) rotated back
    ('Translate 0 0 4\nRotate 90 0 1 0', (4, 0, 0), (-1, 0, 0), (0, 1, 0)),
    ('Scale -1 1 1\nLookAt 0 1 5  0 1 0  0 1 0', (0, 1, 5), (0, 0, -1), (0, 1, 0)),
])
def test_camera_directives(tmp_path, xf, eye, look, up):
    f = _write(tmp_path, "c.pbrt", HEADER.format(xf=xf) + 'Material "matte"\n'
               'Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]\n'
               "WorldEnd\n")
    s = pbrt.load_pbrt(f)
    cam = s.scene.camera
    t = math.tan(math.radians(20))
    assert np.allclose(cam.pos, eye, atol=1e-5)
    assert np.allclose(cam.lookat, look, atol=1e-6)
    assert np.allclose(cam.up, np.asarray(up) * t, atol=1e-6)
    right = np.cross(look, np.asarray(up) * t) * 2.0
    assert np.allclose(cam.right, right, atol=1e-6)


def test_shape_transforms_and_normal_quirk(tmp_path):
    body = """AttributeBegin
  Translate 0 1 0
  Scale 2 2 2
  Material "matte" "rgb Kd" [0.1 0.2 0.3]
  Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 0 1] "normal N" [0 1 0 0 1 0 0 1 0]
      "integer indices" [0 1 2]
AttributeEnd
TransformBegin
  Rotate 90 0 0 1
  Shape "trianglemesh" "point3 P" [1 0 0 0 0 0 0 0 1] "integer indices" [0 1 2]
TransformEnd
Shape "trianglemesh" "point3 P" [1 0 0 0 0 0 0 0 1] "integer indices" [0 1 2]
WorldEnd
"""
    f = _write(tmp_path, "t.pbrt", HEADER.format(xf="") + 'Material "matte"\n' + body)
    s = pbrt.load_pbrt(f).scene
    m0, m1, m2 = s.meshes
    assert np.allclose(m0.vertices, [[0, 1, 0], [2, 1, 0], [0, 1, 2]])
    # normals go through the full matrix with w = 1 (Shape.cpp:39-40)
    assert np.allclose(m0.normals, [[0, 3, 0]] * 3)
    assert np.allclose(m0.material[:3], [0.1, 0.2, 0.3])
    # Rotate 90 about z maps x to y; the material outside the block is the
    # one set before it; the transform is restored after TransformEnd
    assert np.allclose(m1.vertices, [[0, 1, 0], [0, 0, 0], [0, 0, 1]], atol=1e-6)
    assert np.allclose(m1.material[:3], [1, 1, 1])
    assert np.allclose(m2.vertices, [[1, 0, 0], [0, 0, 0], [0, 0, 1]])


def test_material_mappings(tmp_path):
    text = HEADER.format(xf="") + """
MakeNamedMaterial "p" "string type" [ "plastic" ] "rgb Kd" [0.2 0.3 0.4] "rgb Ks" [0.3 0.4 0] "float roughness" [0.25]
MakeNamedMaterial "pd" "string type" [ "plastic" ]
MakeNamedMaterial "g" "string type" [ "glass" ] "rgb Kr" [1 0.5 0.25] "rgb Kt" [0 0.5 0.75] "float eta" [1.33]
MakeNamedMaterial "d" "string type" [ "disney" ] "rgb color" [0.9 0.8 0.7] "float metallic" [0.3]
    "float roughness" [0.2] "float eta" [1.8] "float spectrans" [1] "float anisotropic" [0.1]
    "float sheen" [0.4] "float sheentint" [0.6] "float clearcoat" [0.7] "float clearcoatgloss" [0.8]
    "float speculartint" [0.9]
MakeNamedMaterial "dd" "string type" [ "disney" ]
MakeNamedMaterial "m" "string type" [ "matte" ]
"""
    for name in ("p", "pd", "g", "d", "dd", "m"):
        text += (f'NamedMaterial "{name}"\nShape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0] '
                 '"integer indices" [0 1 2]\n')
    s = pbrt.load_pbrt(_write(tmp_path, "m.pbrt", text + "WorldEnd\n")).scene
    F = {k: i for i, k in enumerate(scenes.MATERIAL_FIELDS)}
    p, pd, g, d, dd, m = (mm.material for mm in s.meshes)
    # plastic: color = Kd + Ks, metallic = clamp(|Ks|, 0, 1) (Material.cpp:70-87)
    assert np.allclose(p[:3], [0.5, 0.7, 0.4])
    assert p[F["metallic"]] == pytest.approx(0.5)
    assert p[F["roughness"]] == pytest.approx(0.25)
    assert (p[F["trans"]], p[F["eta"]], p[F["specular"]]) == (0.0, 1.5, 0.0)
    assert np.allclose(pd[:3], [0.5, 0.5, 0.5]) and pd[F["roughness"]] == pytest.approx(0.1)
    assert pd[F["metallic"]] == pytest.approx(math.sqrt(3) * 0.25)
    # glass: color = (Kr + Kt) / 2, roughness 0.01, specular 0.5, trans 1 (Material.cpp:89-107)
    assert np.allclose(g[:3], [0.5, 0.5, 0.5])
    assert (g[F["roughness"]], g[F["specular"]], g[F["trans"]]) == (np.float32(0.01), 0.5, 1.0)
    assert g[F["eta"]] == pytest.approx(1.33)
    # disney: fields from the file, specular and subsurface forced to 0 (Material.cpp:33-49)
    want = dict(color_r=0.9, color_g=0.8, color_b=0.7, metallic=0.3, roughness=0.2, specular=0,
                specularTint=0.9, eta=1.8, trans=1, subsurface=0, anisotropic=0.1, sheen=0.4,
                sheenTint=0.6, clearcoat=0.7, clearcoatgloss=0.8)
    for k, v in want.items():
        assert d[F[k]] == pytest.approx(v), k
    # the parser's defaults (minipbrt.h:675-687, 750-751): color / Kd 1
    assert np.allclose(dd[:3], 1) and dd[F["roughness"]] == pytest.approx(0.5)
    assert dd[F["eta"]] == pytest.approx(1.5) and dd[F["sheenTint"]] == pytest.approx(0.5)
    assert dd[F["clearcoatgloss"]] == pytest.approx(1.0)
    assert np.array_equal(m, scenes.matte_material((1, 1, 1)))


def test_area_light_scoping(tmp_path):
    text = HEADER.format(xf="") + """Material "matte"
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [1 2 3] "bool twosided" "true"
  Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]
  Shape "trianglemesh" "point3 P" [0 0 1 1 0 1 0 1 1] "integer indices" [0 1 2]
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse"
  Shape "trianglemesh" "point3 P" [0 0 2 1 0 2 0 1 2] "integer indices" [0 1 2]
AttributeEnd
Shape "trianglemesh" "point3 P" [0 0 3 1 0 3 0 1 3] "integer indices" [0 1 2]
WorldEnd
"""
    s = pbrt.load_pbrt(_write(tmp_path, "l.pbrt", text)).scene
    assert [m.light_L for m in s.meshes] == [(1.0, 2.0, 3.0), (1.0, 2.0, 3.0), (1.0, 1.0, 1.0), None]
    assert [m.two_sided for m in s.meshes] == [True, True, False, False]
    flat = s.flatten()
    assert flat["light_mesh"].tolist() == [0, 1, 2]


def _ply(path, fmt, verts, normals, faces, extra_face_prop=False):
    endian = {"binary_little_endian": "<", "binary_big_endian": ">"}.get(fmt)
    hdr = ["ply", f"format {fmt} 1.0", "comment written by the test",
           f"element vertex {len(verts)}", "property float x", "property float y",
           "property float z"]
    if normals is not None:
        hdr += ["property float nx", "property float ny", "property float nz"]
    hdr += ["property float u", "property float v", f"element face {len(faces)}",
            "property list uchar int vertex_indices"]
    if extra_face_prop:
        hdr += ["property int flags"]
    hdr += ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        for k, v in enumerate(verts):
            row = list(v) + (list(normals[k]) if normals is not None else []) + [0.5, 0.25]
            if endian:
                f.write(struct.pack(endian + "f" * len(row), *row))
            else:
                f.write((" ".join(repr(float(x)) for x in row) + "\n").encode())
        for face in faces:
            vals = [len(face)] + list(face) + ([7] if extra_face_prop else [])
            if endian:
                f.write(struct.pack(endian + "B" + "i" * len(face) + ("i" if extra_face_prop else ""),
                                    *vals))
            else:
                f.write((" ".join(str(x) for x in vals) + "\n").encode())


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("quads,extra", [(False, False), (True, False), (True, True)])
def test_ply_reader_round_trip(tmp_path, fmt, quads, extra):
    rng = np.random.default_rng(3)
    verts = rng.normal(size=(12, 3)).astype(np.float32)
    normals = rng.normal(size=(12, 3)).astype(np.float32)
    if quads:
        faces = [(0, 1, 2, 3), (4, 5, 6), (7, 8, 9, 10, 11)]
        want = [(0, 1, 2), (0, 2, 3), (4, 5, 6), (7, 8, 9), (7, 9, 10), (7, 10, 11)]
    else:
        faces = [tuple(rng.permutation(12)[:3]) for _ in range(9)]
        want = faces
    path = str(tmp_path / "m.ply")
    _ply(path, fmt, verts, normals, faces, extra)
    v, n, f = pbrt.read_ply(path)
    assert np.array_equal(v, verts) and np.array_equal(n, normals)
    assert f.dtype == np.uint32 and f.tolist() == [list(w) for w in want]
    _ply(path, fmt, verts, None, faces, extra)
    v, n, f = pbrt.read_ply(path)
    assert n is None and f.tolist() == [list(w) for w in want]


def test_plymesh_shape_and_include(tmp_path):
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    os.makedirs(tmp_path / "models")
    _ply(str(tmp_path / "models" / "q.ply"), "binary_little_endian", verts, None, [(0, 1, 2, 3)])
    _write(tmp_path, "inc.pbrt", 'Shape "plymesh" "string filename" [ "models/q.ply" ]\n')
    text = HEADER.format(xf="") + """Material "glass"
AttributeBegin
  Transform [ 1 0 0 0 0 1.5 0 0 0 0 1 0 0 0 0 1 ]
  Include "inc.pbrt"
AttributeEnd
WorldEnd
"""
    s = pbrt.load_pbrt(_write(tmp_path, "w.pbrt", text))
    (m,) = s.scene.meshes
    assert m.indices.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert np.allclose(m.vertices, verts * [1, 1.5, 1])
    assert m.material[8] == 1.0 and len(s.notes) == 1


@pytest.mark.parametrize("snippet,exc", [
    ('Material "matte"\nShape "sphere" "float radius" [1]', NotImplementedError),
    ('Material "metal"', NotImplementedError),
    ('Texture "t" "color" "checkerboard"', NotImplementedError),
    ('Material "matte" "texture Kd" "t"\n'
     'Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]', ValueError),
    ('Texture "s" "color" "scale" "rgb tex2" [1 1 1]\nMaterial "matte" "texture Kd" "s"\n'
     'Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]',
     NotImplementedError),
    ('LightSource "distant" "rgb L" [1 1 1]', NotImplementedError),
    ('Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]', ValueError),
    ('NamedMaterial "nope"', ValueError),
])
def test_unsupported_input_raises(tmp_path, snippet, exc):
    f = _write(tmp_path, "u.pbrt", HEADER.format(xf="") + snippet + "\nWorldEnd\n")
    with pytest.raises(exc):
        pbrt.load_pbrt(f)


def test_water_caustic_scene_loads():
    # scene/water-caustic/water.pbrt: two plymesh water surfaces scaled 1.5x
    # in y, a tiny very bright light; the reference's loader stops at plymesh
    s = pbrt.load_pbrt(WATER)
    sc = s.scene
    assert len(sc.meshes) == 10 and sc.num_triangles == 103487
    water = sc.meshes[-2:]
    for m in water:
        assert m.material[8] == 1.0 and m.material[7] == pytest.approx(1.8)
        assert m.vertices[:, 1].max() <= 1.5 * 1.12
    assert sc.meshes[7].light_L == (34000.0, 23800.0, 8000.0)
