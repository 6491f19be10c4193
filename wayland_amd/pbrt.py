"""pbrt-v3 scene ingestion: the subset the reference's SceneManager
translates (src/Frontend/Core/SceneManager.cpp:14-246 over the minipbrt
parser), to this package's Scene.

Handled as the reference handles it:
  * the current transform (Transform, ConcatTransform, Translate, Scale,
    Rotate, LookAt, Identity), AttributeBegin/End and TransformBegin/End;
  * Camera "perspective": camera-to-world = inverse of the transform at the
    Camera statement; position, up, lookAt as PinholeCamera (Camera.cpp:15-28):
    up scaled by tan(fov / 2) with fov on the vertical axis, right =
    cross(lookAt, up) scaled by the film's aspect ratio
    (PathTracingProgramManager.cpp:22-28);
  * Film "image" x/yresolution;
  * materials matte, plastic, glass and disney, mapped to the Disney BSDF as
    Material.cpp:33-107 maps them (MakeNamedMaterial / NamedMaterial /
    Material);
  * AreaLightSource "diffuse" ("rgb L", "bool twosided") on the shapes that
    follow it in its attribute block (Light.cpp:8-13);
  * Shape "trianglemesh" (P, N, uv, indices): positions through the shape's
    transform; normals too, with w = 1 as TriangleMesh does (Shape.cpp:7-49:
    TransformVec3(t, N, 1), so a translation moves them).
Defaults for parameters a file leaves out are the vendored parser's
(3rdParty/minipbrt/minipbrt.h:256-258, 289-295, 353-355, 675-687, 705-708,
750-751, 795-798), which the reference reads: matte Kd 1, disney color 1,
plastic Kd = Ks = 0.25 and roughness 0.1, glass Kr = Kt = 1 and eta 1.5, area
light L 1 one-sided, perspective fov 90, film 640 x 480.
Textures: Texture "imagemap" (decoded to RGBA8, top row first, as stbi_load
does in Texture.cpp:13-16) bound through "texture Kd" (matte, plastic),
"texture Kt" (glass) or "texture color" (disney), with the mesh's "uv"/"st"
texcoords (zero when absent, as the reference's zero-filled buffer gives);
"scale" textures are skipped as TransformTexture skips them
(SceneManager.cpp:157-159).
Beyond the reference: Shape "plymesh" (ascii or binary PLY with x y z, optional
nx ny nz, and polygon faces fanned into triangles; SceneManager.cpp:207-221 has
this path commented out) and Include.  Other textures, shapes, materials and
lights raise, as the reference does; a shape with no material raises (the
reference would index past its material list); ReverseOrientation is accepted
and, as in the reference, has no effect.
"""
from __future__ import annotations

import math
import os
import re
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .scenes import (Camera, Mesh, Scene, disney_material, glass_material, matte_material,
                     pinhole_camera)

_TOKEN = re.compile(r'"[^"]*"|\[|\]|[^\s\[\]"#]+|#[^\n]*')


def _tokens(text: str):
    for m in _TOKEN.finditer(text):
        t = m.group(0)
        if not t.startswith("#"):
            yield t


def _num(t: str) -> float:
    return float(t)


@dataclass
class _Param:
    type: str
    values: list


def _parse_params(toks: List[str], i: int):
    """Parameters after a directive: "type name" value or [values]."""
    params: Dict[str, _Param] = {}
    while i < len(toks) and toks[i].startswith('"'):
        decl = toks[i][1:-1].split()
        if len(decl) != 2:
            break
        ptype, name = decl
        i += 1
        vals = []
        if toks[i] == "[":
            i += 1
            while toks[i] != "]":
                vals.append(toks[i])
                i += 1
            i += 1
        else:
            vals.append(toks[i])
            i += 1
        if ptype in ("string", "texture", "bool"):
            vals = [v.strip('"') for v in vals]
        else:
            vals = [_num(v) for v in vals]
        params[name] = _Param(ptype, vals)
    return params, i


def _get(params, name, default):
    p = params.get(name)
    if p is None:
        return default
    if p.type == "texture":
        raise NotImplementedError(f"texture-valued parameter '{name}' is not supported")
    return p.values


def _rgb(params, name, default):
    p = params.get(name)
    if p is None:
        return np.asarray(default, np.float32)
    if p.type in ("rgb", "color"):
        return np.asarray(p.values[:3], np.float32)
    if p.type == "float" and len(p.values) == 1:
        return np.full(3, p.values[0], np.float32)
    if p.type == "texture":
        raise NotImplementedError(f"textured '{name}' is not supported")
    raise NotImplementedError(f"'{p.type} {name}' (only rgb) is not supported")


def _color(params, name, default):
    """A texturable colour: (rgb, texture name or None).  A textured
    parameter keeps the default value, which the texel then multiplies
    (minipbrt's ColorTex holds both; Material.cpp:33-107 takes the value and
    the texture id)."""
    p = params.get(name)
    if p is not None and p.type == "texture":
        return np.asarray(default, np.float32), str(p.values[0])
    return _rgb(params, name, default), None


def _float(params, name, default):
    return float(_get(params, name, [default])[0])


def _material(mtype: str, params):
    """pbrt material -> (the reference's Disney parameters, texture name or
    None) as Material.cpp maps them; the texture is Kd's for matte and
    plastic, Kt's for glass and color's for disney."""
    if mtype == "matte":
        kd, tex = _color(params, "Kd", (1.0, 1.0, 1.0))
        return matte_material(kd), tex
    if mtype == "plastic":
        kd, tex = _color(params, "Kd", (0.25, 0.25, 0.25))
        ks = _rgb(params, "Ks", (0.25, 0.25, 0.25))
        ks_len = np.float32(np.sqrt(np.float32(np.dot(ks, ks))))
        return disney_material(color=kd + ks, metallic=float(np.clip(ks_len, 0, 1)),
                               roughness=_float(params, "roughness", 0.1), sheenTint=0.5,
                               specular=0.0, specularTint=0.0, eta=1.5, trans=0.0,
                               clearcoat=0.0, clearcoatgloss=1.0), tex
    if mtype == "glass":
        kt, tex = _color(params, "Kt", (1, 1, 1))
        return glass_material(_rgb(params, "Kr", (1, 1, 1)), kt,
                              eta=_float(params, "eta", 1.5)), tex
    if mtype == "disney":
        color, tex = _color(params, "color", (1.0, 1.0, 1.0))
        return disney_material(color=color,
                               metallic=_float(params, "metallic", 0.0),
                               roughness=_float(params, "roughness", 0.5),
                               specular=0.0,
                               specularTint=_float(params, "speculartint", 0.0),
                               eta=_float(params, "eta", 1.5),
                               trans=_float(params, "spectrans", 0.0),
                               subsurface=0.0,
                               anisotropic=_float(params, "anisotropic", 0.0),
                               sheen=_float(params, "sheen", 0.0),
                               sheenTint=_float(params, "sheentint", 0.5),
                               clearcoat=_float(params, "clearcoat", 0.0),
                               clearcoatgloss=_float(params, "clearcoatgloss", 1.0)), tex
    raise NotImplementedError(f"material type '{mtype}' is not supported")


def read_ply(path: str, with_uv: bool = False):
    """PLY mesh (ascii or binary_little_endian / binary_big_endian):
    returns positions (N, 3), normals (N, 3) or None, triangle indices (M, 3)
    and, with with_uv, texcoords (N, 2) or None (properties u v, s t or
    texture_u texture_v); polygons are fanned into triangles."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header")
    end = data.index(b"\n", end) + 1
    header = data[:end].decode("ascii").splitlines()
    fmt = None
    elements = []
    for line in header:
        w = line.split()
        if not w:
            continue
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elements.append([w[1], int(w[2]), []])
        elif w[0] == "property":
            elements[-1][2].append(w[1:])
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"{path}: unsupported PLY format {fmt}")
    sizes = {"char": "b", "int8": "b", "uchar": "B", "uint8": "B", "short": "h", "int16": "h",
             "ushort": "H", "uint16": "H", "int": "i", "int32": "i", "uint": "I",
             "uint32": "I", "float": "f", "float32": "f", "double": "d", "float64": "d"}
    endian = ">" if fmt == "binary_big_endian" else "<"
    body = data[end:]
    pos = 0
    words = body.split() if fmt == "ascii" else None
    wi = 0
    verts = normals = faces = uv = None
    for name, count, props in elements:
        if name == "vertex":
            names = [p[-1] for p in props]
            if any(p[0] == "list" for p in props):
                raise ValueError(f"{path}: list property in vertex element")
            if fmt == "ascii":
                arr = np.array(words[wi:wi + count * len(props)], np.float64).reshape(count, -1)
                wi += count * len(props)
            else:
                dt = np.dtype([(n, endian + sizes[p[0]]) for n, p in zip(names, props)])
                rec = np.frombuffer(body, dt, count, pos)
                pos += dt.itemsize * count
                arr = np.stack([rec[n].astype(np.float64) for n in names], 1)
            col = {n: k for k, n in enumerate(names)}
            verts = arr[:, [col["x"], col["y"], col["z"]]].astype(np.float32)
            if all(n in col for n in ("nx", "ny", "nz")):
                normals = arr[:, [col["nx"], col["ny"], col["nz"]]].astype(np.float32)
            for un, vn in (("u", "v"), ("s", "t"), ("texture_u", "texture_v")):
                if un in col and vn in col:
                    uv = arr[:, [col[un], col[vn]]].astype(np.float32)
                    break
        elif name == "face":
            tris = []
            lp = next(p for p in props if p[0] == "list")
            cnt_t, idx_t = sizes[lp[1]], sizes[lp[2]]
            other = [p for p in props if p is not lp]
            if fmt == "ascii":
                for _ in range(count):
                    n = int(words[wi])
                    idx = [int(x) for x in words[wi + 1:wi + 1 + n]]
                    wi += 1 + n + len(other)
                    for k in range(1, n - 1):
                        tris.append((idx[0], idx[k], idx[k + 1]))
            else:
                csz, isz = struct.calcsize(cnt_t), struct.calcsize(idx_t)
                osz = sum(struct.calcsize(sizes[p[0]]) for p in other)
                # all-triangle faces with only the index list: one strided view
                stride = csz + 3 * isz
                if not other and len(body) - pos >= count * stride:
                    dt = np.dtype([("n", endian + cnt_t), ("i", endian + idx_t, (3,))])
                    rec = np.frombuffer(body, dt, count, pos)
                    if (rec["n"] == 3).all():
                        tris = rec["i"].astype(np.uint32)
                        pos += count * stride
                        count = 0
                for _ in range(count):
                    (n,) = struct.unpack_from(endian + cnt_t, body, pos)
                    pos += csz
                    idx = struct.unpack_from(endian + idx_t * n, body, pos)
                    pos += isz * n + osz
                    for k in range(1, n - 1):
                        tris.append((idx[0], idx[k], idx[k + 1]))
            faces = (tris if isinstance(tris, np.ndarray)
                     else np.asarray(tris, np.uint32).reshape(-1, 3))
        else:
            # skip an unknown element (fixed-size properties only)
            if fmt == "ascii":
                wi += count * len(props)
            else:
                pos += count * sum(struct.calcsize(sizes[p[0]]) for p in props)
    if verts is None or faces is None:
        raise ValueError(f"{path}: no vertex or face element")
    if with_uv:
        return verts, normals, faces, uv
    return verts, normals, faces


def _rotate(angle_deg, axis):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s, c = math.sin(math.radians(angle_deg)), math.cos(math.radians(angle_deg))
    x, y, z = a
    m = np.eye(4)
    m[:3, :3] = [[x * x + (1 - x * x) * c, x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                 [x * y * (1 - c) + z * s, y * y + (1 - y * y) * c, y * z * (1 - c) - x * s],
                 [x * z * (1 - c) - y * s, y * z * (1 - c) + x * s, z * z + (1 - z * z) * c]]
    return m


def _look_at(eye, look, up):
    eye, look, up = (np.asarray(v, np.float64) for v in (eye, look, up))
    d = look - eye
    d /= np.linalg.norm(d)
    right = np.cross(up / np.linalg.norm(up), d)
    right /= np.linalg.norm(right)
    new_up = np.cross(d, right)
    cam_to_world = np.eye(4)
    cam_to_world[:3, 0], cam_to_world[:3, 1], cam_to_world[:3, 2] = right, new_up, d
    cam_to_world[:3, 3] = eye
    return np.linalg.inv(cam_to_world)


@dataclass
class PbrtScene:
    scene: Scene
    width: int
    height: int
    fov: float
    camera_to_world: np.ndarray
    notes: List[str] = field(default_factory=list)


def _transform_points(t, p, w):
    """m * vec4(p, w) in float32, truncated to xyz (TransformVec3, Common.h:22-31),
    summed column by column as glm does."""
    m = t.astype(np.float32)
    p = np.asarray(p, np.float32).reshape(-1, 3)
    w = np.float32(w)
    out = np.empty_like(p)
    for r in range(3):
        out[:, r] = ((m[r, 0] * p[:, 0] + m[r, 1] * p[:, 1]) + m[r, 2] * p[:, 2]) + m[r, 3] * w
    return out


def load_pbrt(path: str, width: Optional[int] = None, height: Optional[int] = None) -> PbrtScene:
    """Parse a pbrt-v3 file into a Scene (camera for width x height, default
    the Film's resolution)."""
    base = os.path.dirname(os.path.abspath(path))
    text = open(path).read()
    toks = list(_tokens(text))
    # expand Include (relative to the including file's directory)
    out = []
    i = 0
    while i < len(toks):
        if toks[i] in ("Include", "Import"):
            inc = os.path.join(base, toks[i + 1].strip('"'))
            out.extend(_tokens(open(inc).read()))
            i += 2
        else:
            out.append(toks[i])
            i += 1
    toks = out

    ctm = np.eye(4)
    stack = []
    named: Dict[str, tuple] = {}
    textures: Dict[str, Optional[int]] = {}   # name -> image index (None: skipped)
    images: List[np.ndarray] = []
    material = None  # (Disney parameters, texture name or None)
    area = None  # (L, two_sided)
    fov = 90.0
    cam_to_world = None
    film_w, film_h = 640, 480
    meshes: List[Mesh] = []
    notes: List[str] = []
    i = 0
    while i < len(toks):
        d = toks[i]
        i += 1
        if d == "Transform":
            assert toks[i] == "["
            vals = [_num(v) for v in toks[i + 1:i + 17]]
            i += 18
            ctm = np.asarray(vals, np.float64).reshape(4, 4).T  # column-major in the file
        elif d == "ConcatTransform":
            vals = [_num(v) for v in toks[i + 1:i + 17]]
            i += 18
            ctm = ctm @ np.asarray(vals, np.float64).reshape(4, 4).T
        elif d == "Identity":
            ctm = np.eye(4)
        elif d == "Translate":
            t = np.eye(4)
            t[:3, 3] = [_num(v) for v in toks[i:i + 3]]
            i += 3
            ctm = ctm @ t
        elif d == "Scale":
            ctm = ctm @ np.diag([_num(v) for v in toks[i:i + 3]] + [1.0])
            i += 3
        elif d == "Rotate":
            ctm = ctm @ _rotate(_num(toks[i]), [_num(v) for v in toks[i + 1:i + 4]])
            i += 4
        elif d == "LookAt":
            v = [_num(x) for x in toks[i:i + 9]]
            i += 9
            ctm = ctm @ _look_at(v[0:3], v[3:6], v[6:9])
        elif d in ("AttributeBegin", "TransformBegin"):
            stack.append((ctm.copy(), material, area) if d == "AttributeBegin" else (ctm.copy(),))
        elif d in ("AttributeEnd", "TransformEnd"):
            top = stack.pop()
            ctm = top[0]
            if len(top) == 3:
                material, area = top[1], top[2]
        elif d in ("WorldBegin",):
            ctm = np.eye(4)
        elif d in ("WorldEnd",):
            pass
        elif d == "Camera":
            ctype = toks[i].strip('"')
            params, i = _parse_params(toks, i + 1)
            if ctype != "perspective":
                raise NotImplementedError(f"camera type '{ctype}' is not supported")
            fov = _float(params, "fov", 90.0)
            cam_to_world = np.linalg.inv(ctm)
        elif d == "Film":
            params, i = _parse_params(toks, i + 1)
            film_w = int(_float(params, "xresolution", 640))
            film_h = int(_float(params, "yresolution", 480))
        elif d in ("Integrator", "Sampler", "PixelFilter", "Accelerator", "ColorSpace"):
            params, i = _parse_params(toks, i + 1)
        elif d == "MakeNamedMaterial":
            name = toks[i].strip('"')
            params, i = _parse_params(toks, i + 1)
            named[name] = _material(params["type"].values[0], params)
        elif d == "NamedMaterial":
            name = toks[i].strip('"')
            i += 1
            if name not in named:
                raise ValueError(f"unknown named material '{name}'")
            material = named[name]
        elif d == "Material":
            mtype = toks[i].strip('"')
            params, i = _parse_params(toks, i + 1)
            material = _material(mtype, params)
        elif d == "AreaLightSource":
            ltype = toks[i].strip('"')
            params, i = _parse_params(toks, i + 1)
            if ltype != "diffuse":
                raise NotImplementedError(f"area light '{ltype}' is not supported")
            two = str(_get(params, "twosided", ["false"])[0]).lower() == "true"
            area = (_rgb(params, "L", (1, 1, 1)), two)
        elif d == "Shape":
            stype = toks[i].strip('"')
            params, i = _parse_params(toks, i + 1)
            if stype == "trianglemesh":
                P = np.asarray(_get(params, "P", None), np.float32).reshape(-1, 3)
                idx = np.asarray(_get(params, "indices", None), np.float64).astype(np.uint32)
                N = _get(params, "N", None)
                N = None if N is None else np.asarray(N, np.float32).reshape(-1, 3)
                UV = _get(params, "uv", None) or _get(params, "st", None)
                UV = None if UV is None else np.asarray(UV, np.float32).reshape(-1, 2)
            elif stype == "plymesh":
                fn = os.path.join(base, str(_get(params, "filename", [""])[0]))
                P, N, idx, UV = read_ply(fn, with_uv=True)
                notes.append(f"plymesh {os.path.basename(fn)}: loaded (the reference's "
                             "SceneManager has this path disabled)")
            else:
                raise NotImplementedError(f"shape '{stype}' is not supported")
            if material is None:
                raise ValueError(f"shape {len(meshes)} has no material")
            verts = _transform_points(ctm, P, 1.0)
            normals = None if N is None else _transform_points(ctm, N, 1.0)
            tex = None
            if material[1] is not None:
                if material[1] not in textures:
                    raise ValueError(f"unknown texture '{material[1]}'")
                tex = textures[material[1]]
                if tex is None:
                    raise NotImplementedError(f"material uses the skipped texture '{material[1]}'")
            meshes.append(Mesh(verts, idx.reshape(-1, 3).astype(np.uint32), normals, material[0],
                               light_L=None if area is None else tuple(float(x) for x in area[0]),
                               two_sided=False if area is None else area[1], uv=UV, texture=tex))
        elif d == "Texture":
            tname, tclass = toks[i].strip('"'), toks[i + 2].strip('"')
            params, i = _parse_params(toks, i + 3)
            if tclass == "imagemap":
                fn = os.path.join(base, str(_get(params, "filename", [""])[0]))
                textures[tname] = len(images)
                images.append(load_image_rgba(fn))
            elif tclass == "scale":
                # TransformTexture leaves scale textures out (SceneManager.cpp:157-159)
                textures[tname] = None
                notes.append(f"texture '{tname}': scale textures are skipped, as the reference "
                             "skips them")
            else:
                raise NotImplementedError(f"texture class '{tclass}' is not supported")
        elif d == "ReverseOrientation":
            pass
        elif d in ("MakeNamedMedium", "MediumInterface", "LightSource",
                   "ObjectBegin", "ObjectInstance",
                   "CoordinateSystem", "CoordSysTransform", "ActiveTransform",
                   "TransformTimes"):
            raise NotImplementedError(f"'{d}' is not supported")
        else:
            raise ValueError(f"unexpected token '{d}'")
    if cam_to_world is None:
        raise ValueError("the scene has no camera")
    W = width or film_w
    H = height or film_h
    camera = camera_from_pbrt(cam_to_world, fov, W, H)
    name = os.path.splitext(os.path.basename(path))[0]
    return PbrtScene(Scene(meshes, camera, (0.0, 0.0, 0.0), name, images), W, H, fov,
                     cam_to_world, notes)


def load_image_rgba(path: str) -> np.ndarray:
    """An image file as (H, W, 4) uint8, top row first, grey and RGB widened to
    RGBA: what stbi_load(path, ..., 4) gives Texture::TexLoadImage
    (Texture.cpp:13-16).  Decoded with Pillow; a JPEG decoder other than stb's
    may differ from it by a unit in a few texels."""
    from PIL import Image
    with Image.open(path) as im:
        return np.ascontiguousarray(np.asarray(im.convert("RGBA"), np.uint8))


def camera_from_pbrt(cam_to_world: np.ndarray, fov: float, width: int, height: int) -> Camera:
    """PinholeCamera(fov, cameraToWorld) (Camera.cpp:15-28): position, up and
    lookAt from the camera-to-world matrix, then the frame of pinhole_camera
    (up scaled by tan(fov / 2), right = cross(lookAt, up) times the aspect)."""
    t = cam_to_world.astype(np.float32)
    pos = _transform_points(t, np.zeros((1, 3), np.float32), 1.0)[0]
    up = _transform_points(t, np.array([[0, 1, 0]], np.float32), 0.0)[0]
    look = _transform_points(t, np.array([[0, 0, 1]], np.float32), 0.0)[0]
    return pinhole_camera(pos, look, up, fov, width, height)
