"""Synthetic scenes and ray sets for the path-tracing workloads.

Every workload is generated from a seed, in numpy, so the GPU path, the CPU
oracle and the benchmark all see bit-identical float32 inputs:

* ``cornell_box``        — 36-triangle Cornell box (geometry frame of
  ``scene/cornell-box/cbox.pbrt``; the two boxes axis-aligned), one quad
  area light, Matte->Disney materials (``src/Frontend/Core/Material.cpp:51-67``).
* ``icosphere_soup``     — N displaced icospheres in the unit cube
  (traversal workload; triangle soup, no shared vertices).
* ``traversal_rays``     — incoherent rays: origins uniform in the unit cube,
  directions uniform on the sphere.
* ``wavefront_scene``    — the icosphere soup with per-object Disney
  parameters, two quad lights and a constant environment.
* ``glossy_caustic``     — glass icosphere over a rough-metal floor grid with a
  small quad light behind it.
* ``textured_cornell_box`` — the Cornell box with two RGBA8 textures.

A scene is a list of meshes plus camera/background; ``Scene.flatten()`` packs it
into the flat arrays the C ABI and the oracle take (mesh-local indices,
per-mesh vertex/triangle offsets), the layout of
``src/Frontend/Core/DeviceManager.cpp:111-181``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

INVALID = 0xFFFFFFFF

# Field order of the 64-byte Disney material (src/Frontend/Device/Material.h:7-21)
MATERIAL_FIELDS = ("color_r", "color_g", "color_b", "metallic", "roughness",
                   "specular", "specularTint", "eta", "trans", "subsurface",
                   "anisotropic", "sheen", "sheenTint", "clearcoat",
                   "clearcoatgloss", "pad")


def disney_material(color=(0.5, 0.5, 0.5), metallic=0.0, roughness=0.5,
                    specular=0.0, specularTint=0.0, eta=1.5, trans=0.0,
                    subsurface=0.0, anisotropic=0.0, sheen=0.0, sheenTint=0.5,
                    clearcoat=0.0, clearcoatgloss=1.0) -> np.ndarray:
    """A Disney material as 16 float32 (defaults of Device/Material.h:9-20)."""
    return np.array([color[0], color[1], color[2], metallic, roughness,
                     specular, specularTint, eta, trans, subsurface,
                     anisotropic, sheen, sheenTint, clearcoat, clearcoatgloss,
                     0.0], dtype=np.float32)


def matte_material(kd) -> np.ndarray:
    """pbrt ``matte`` -> Disney, as ``Material.cpp:51-67`` maps it."""
    return disney_material(color=kd, metallic=0.0, roughness=0.5,
                           sheenTint=0.5, specular=0.0, specularTint=0.0,
                           eta=1.5, trans=0.0, clearcoat=0.0,
                           clearcoatgloss=1.0)


def glass_material(kr=(1, 1, 1), kt=(1, 1, 1), eta=1.5) -> np.ndarray:
    """pbrt ``glass`` -> Disney, as ``Material.cpp:89-107`` maps it."""
    c = (np.asarray(kr, np.float32) + np.asarray(kt, np.float32)) / np.float32(2)
    return disney_material(color=c, metallic=0.0, roughness=0.01,
                           sheenTint=0.5, specular=0.5, specularTint=0.0,
                           eta=eta, trans=1.0)


@dataclass
class Mesh:
    vertices: np.ndarray                 # (V, 3) float32
    indices: np.ndarray                  # (T, 3) uint32, mesh-local
    normals: Optional[np.ndarray] = None  # (V, 3) float32 or None
    material: np.ndarray = field(default_factory=lambda: disney_material())
    light_L: Optional[tuple] = None      # emitted radiance if an area light
    two_sided: bool = False
    uv: Optional[np.ndarray] = None      # (V, 2) float32 texcoords or None
    texture: Optional[int] = None        # index into Scene.textures or None


@dataclass
class Camera:
    """Pinhole frame of ``Device/Camera.h:9-15`` (right already aspect-scaled)."""
    pos: np.ndarray
    lookat: np.ndarray
    up: np.ndarray
    right: np.ndarray


def pinhole_camera(pos, look_dir, up_dir, fov_deg, width, height) -> Camera:
    """Camera frame as ``Core/Camera.cpp:15-28`` + ``PathTracingProgramManager.cpp:22-28``:
    up scaled by tan(fov/2) (fov on the vertical axis), right = cross(lookAt, up)
    scaled by the aspect ratio; float32 arithmetic throughout."""
    f32 = np.float32
    look = np.asarray(look_dir, f32)
    look = (look * (f32(1) / np.sqrt(np.dot(look, look), dtype=f32))).astype(f32)
    up = np.asarray(up_dir, f32)
    up = (up * (f32(1) / np.sqrt(np.dot(up, up), dtype=f32))).astype(f32)
    half = f32(f32(fov_deg) / f32(2))
    ang = f32(float(half) * math.pi / 180.0)
    up = (up * f32(math.tan(float(ang)))).astype(f32)
    right = np.array([look[1] * up[2] - up[1] * look[2],
                      look[2] * up[0] - up[2] * look[0],
                      look[0] * up[1] - up[0] * look[1]], f32)
    right = (right * f32(f32(width) / f32(height))).astype(f32)
    return Camera(np.asarray(pos, f32), look, up, right)


@dataclass
class Scene:
    meshes: List[Mesh]
    camera: Optional[Camera] = None
    bg: tuple = (0.0, 0.0, 0.0)
    name: str = "scene"
    textures: List[np.ndarray] = field(default_factory=list)  # (H, W, 4) uint8, top row first

    @property
    def num_triangles(self) -> int:
        return int(sum(m.indices.shape[0] for m in self.meshes))

    def flatten(self) -> dict:
        """Pack into flat arrays (see module docstring)."""
        vo, to, tc, hn, ml = [], [], [], [], []
        verts, norms, idx, mats = [], [], [], []
        lights_mesh, lights_L, lights_2s = [], [], []
        nv = nt = 0
        for mi, m in enumerate(self.meshes):
            v = np.ascontiguousarray(m.vertices, np.float32).reshape(-1, 3)
            i = np.ascontiguousarray(m.indices, np.uint32).reshape(-1, 3)
            vo.append(nv)
            to.append(nt)
            tc.append(i.shape[0])
            hn.append(1 if m.normals is not None else 0)
            verts.append(v)
            norms.append(np.ascontiguousarray(m.normals, np.float32).reshape(-1, 3)
                         if m.normals is not None else np.zeros_like(v))
            idx.append(i)
            mats.append(np.asarray(m.material, np.float32).reshape(16))
            if m.light_L is not None:
                ml.append(len(lights_mesh))
                lights_mesh.append(mi)
                lights_L.append(np.asarray(m.light_L, np.float32))
                lights_2s.append(1 if m.two_sided else 0)
            else:
                ml.append(INVALID)
            nv += v.shape[0]
            nt += i.shape[0]
        u32 = np.uint32
        out = dict(
            vertices=np.concatenate(verts) if verts else np.zeros((0, 3), np.float32),
            normals=np.concatenate(norms) if norms else np.zeros((0, 3), np.float32),
            indices=np.concatenate(idx) if idx else np.zeros((0, 3), u32),
            mesh_vertex_offset=np.array(vo, u32),
            mesh_tri_offset=np.array(to, u32),
            mesh_tri_count=np.array(tc, u32),
            mesh_has_normal=np.array(hn, u32),
            mesh_light=np.array(ml, u32),
            mesh_material=np.stack(mats) if mats else np.zeros((0, 16), np.float32),
            light_mesh=np.array(lights_mesh, u32),
            light_L=np.stack(lights_L) if lights_L else np.zeros((0, 3), np.float32),
            light_two_sided=np.array(lights_2s, u32),
            bg=np.asarray(self.bg, np.float32),
        )
        # textures: texcoords indexed as vertices (zeros for a mesh without),
        # images back to back as RGBA8 words
        tex_off, tex_size, texels = [], [], []
        n = 0
        for img in self.textures:
            img = np.ascontiguousarray(img, np.uint8)
            h, w = img.shape[:2]
            tex_off.append(n)
            tex_size.append((w, h))
            texels.append(img.reshape(-1, 4))
            n += w * h
        uvs = [np.ascontiguousarray(m.uv, np.float32).reshape(-1, 2) if m.uv is not None
               else np.zeros((np.asarray(m.vertices).reshape(-1, 3).shape[0], 2), np.float32)
               for m in self.meshes]
        out.update(
            texcoords=np.concatenate(uvs) if uvs else np.zeros((0, 2), np.float32),
            mesh_texture=np.array([INVALID if (m.texture is None or not self.textures)
                                   else m.texture for m in self.meshes], u32),
            texture_offset=np.array(tex_off, u32),
            texture_size=np.array(tex_size, u32).reshape(-1, 2),
            texels=np.concatenate(texels) if texels else np.zeros((0, 4), np.uint8),
        )
        return out

    def triangles(self) -> np.ndarray:
        """World-space triangles in global primitive order, (T, 9) float32."""
        out = []
        for m in self.meshes:
            v = np.asarray(m.vertices, np.float32).reshape(-1, 3)
            i = np.asarray(m.indices, np.int64).reshape(-1, 3)
            out.append(v[i].reshape(-1, 9))
        if not out:
            return np.zeros((0, 9), np.float32)
        return np.ascontiguousarray(np.concatenate(out), np.float32)


# --------------------------------------------------------------------------
# Cornell box (36 triangles)
# --------------------------------------------------------------------------
def _quad(p0, p1, p2, p3, n):
    v = np.array([p0, p1, p2, p3], np.float32)
    nn = np.tile(np.asarray(n, np.float32), (4, 1))
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), nn


def _box(lo, hi):
    """Axis-aligned box, 12 triangles, outward winding, flat normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [  # (4 corners counter-clockwise seen from outside, normal)
        ([x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1], [0, 0, 1]),
        ([x1, y0, z0], [x0, y0, z0], [x0, y1, z0], [x1, y1, z0], [0, 0, -1]),
        ([x1, y0, z1], [x1, y0, z0], [x1, y1, z0], [x1, y1, z1], [1, 0, 0]),
        ([x0, y0, z0], [x0, y0, z1], [x0, y1, z1], [x0, y1, z0], [-1, 0, 0]),
        ([x0, y1, z1], [x1, y1, z1], [x1, y1, z0], [x0, y1, z0], [0, 1, 0]),
        ([x0, y0, z0], [x1, y0, z0], [x1, y0, z1], [x0, y0, z1], [0, -1, 0]),
    ]
    vs, ids, ns = [], [], []
    for k, (a, b, c, d, n) in enumerate(faces):
        vs += [a, b, c, d]
        ns += [n] * 4
        ids += [[4 * k, 4 * k + 1, 4 * k + 2], [4 * k, 4 * k + 2, 4 * k + 3]]
    return (np.array(vs, np.float32), np.array(ids, np.uint32),
            np.array(ns, np.float32))


def cornell_box(width=256, height=256, seed=0, light_L=17.0) -> Scene:
    """Plain-PT CPU workload: 5 walls x 2 tris, two axis-aligned boxes x 12
    tris, one quad light (L = light_L) at y = 1.98; Lambertian (Matte) Disney
    materials with base colour ~ U[0.1, 0.9]^3 from ``seed``."""
    rng = np.random.default_rng(seed)
    walls = [
        _quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1], [0, 1, 0]),     # floor
        _quad([1, 2, 1], [-1, 2, 1], [-1, 2, -1], [1, 2, -1], [0, -1, 0]),    # ceiling
        _quad([-1, 0, -1], [-1, 2, -1], [1, 2, -1], [1, 0, -1], [0, 0, 1]),   # back
        _quad([1, 0, -1], [1, 2, -1], [1, 2, 1], [1, 0, 1], [-1, 0, 0]),      # right
        _quad([-1, 0, 1], [-1, 2, 1], [-1, 2, -1], [-1, 0, -1], [1, 0, 0]),   # left
    ]
    boxes = [_box([0.05, 0.0, 0.05], [0.65, 0.6, 0.65]),
             _box([-0.7, 0.0, -0.65], [-0.1, 1.2, -0.05])]
    meshes = []
    for v, i, n in walls + boxes:
        col = rng.uniform(0.1, 0.9, 3).astype(np.float32)
        meshes.append(Mesh(v, i, n, matte_material(col)))
    lv, li, ln = _quad([-0.24, 1.98, -0.22], [0.23, 1.98, -0.22],
                       [0.23, 1.98, 0.16], [-0.24, 1.98, 0.16], [0, -1, 0])
    meshes.append(Mesh(lv, li, ln, matte_material((0, 0, 0)),
                       light_L=(light_L, light_L, light_L), two_sided=False))
    cam = pinhole_camera([0, 1, 6.8], [0, 0, -1], [0, 1, 0], 19.5, width, height)
    return Scene(meshes, cam, (0.0, 0.0, 0.0), "cornell")


# --------------------------------------------------------------------------
# Icospheres
# --------------------------------------------------------------------------
def textured_cornell_box(width=256, height=256) -> Scene:
    """The Cornell box with textures: a 37 x 23 random RGBA8 image on the back
    wall (uv = the wall's x, y), a 16 x 16 checker-and-gradient image on the
    floor with texcoords spanning [-1.5, 2.5] (wrap addressing), and the
    first box bound to the wall image without texcoords (reads (0, 0))."""
    sc = cornell_box(width, height)
    rng = np.random.default_rng(9)
    back = rng.integers(0, 256, (23, 37, 4), dtype=np.uint8)
    yy, xx = np.mgrid[0:16, 0:16]
    floor = np.zeros((16, 16, 4), np.uint8)
    floor[..., 0] = np.where((xx // 4 + yy // 4) % 2 == 0, 240, 30)
    floor[..., 1] = 8 * xx
    floor[..., 2] = 8 * yy
    floor[..., 3] = 255
    sc.textures = [back, floor]
    sc.name = "cornell_textured"
    for m in sc.meshes:
        if m.light_L is not None:
            continue
        v = np.asarray(m.vertices).reshape(-1, 3)
        if v.shape[0] == 4 and np.allclose(v[:, 2], -1):      # back wall
            m.uv = np.stack([(v[:, 0] + 1) / 2, 1 - v[:, 1] / 2], 1).astype(np.float32)
            m.texture = 0
        elif v.shape[0] == 4 and np.allclose(v[:, 1], 0):     # floor
            m.uv = np.stack([v[:, 0] * 2 + 0.5, v[:, 2] * 2 + 0.5], 1).astype(np.float32)
            m.texture = 1
    boxes = [m for m in sc.meshes if np.asarray(m.vertices).reshape(-1, 3).shape[0] == 24]
    boxes[0].texture = 0
    return sc


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def unit_icosphere_soup(subdiv: int) -> np.ndarray:
    """Unit icosphere as a triangle soup, (20 * 4**subdiv, 3, 3) float64,
    outward winding.  Midpoints are computed symmetrically, so shared edges
    get bit-identical vertices (the soup is watertight)."""
    v, f = _icosahedron()
    tri = v[f]  # (20, 3, 3)
    for _ in range(subdiv):
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        ab = a + b
        bc = b + c
        ca = c + a
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        tri = np.stack([
            np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
            np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    return tri


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """Hash-lattice value noise in [-1, 1] with smoothstep-trilinear blending."""
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    fr = q - i0
    w = fr * fr * (3.0 - 2.0 * fr)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        x = x ^ (x >> 16)
        return (x & 0xFFFF).astype(np.float64) / 32767.5 - 1.0

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wx = w[..., 0] if dx else 1.0 - w[..., 0]
                wy = w[..., 1] if dy else 1.0 - w[..., 1]
                wz = w[..., 2] if dz else 1.0 - w[..., 2]
                out = out + wx * wy * wz * h(i0[..., 0] + dx, i0[..., 1] + dy,
                                             i0[..., 2] + dz)
    return out


def icosphere_soup(count=64, subdiv=7, seed=1, amp=0.05, r_range=(0.03, 0.12)):
    """``count`` displaced icospheres (radius ~ U[r_range], centre ~ U[r, 1-r]^3,
    vertex radius r * (1 + amp * n(p))) as a list of (T, 3, 3) float32 soups."""
    rng = np.random.default_rng(seed)
    unit = unit_icosphere_soup(subdiv)
    flat = unit.reshape(-1, 3)
    spheres = []
    for s in range(count):
        r = rng.uniform(*r_range)
        c = rng.uniform(r, 1.0 - r, 3)
        n = value_noise(flat * 1.0 + s * 17.0, seed * 1000 + s)
        pos = c + (r * (1.0 + amp * n))[:, None] * flat
        spheres.append(pos.reshape(-1, 3, 3).astype(np.float32))
    return spheres


def soup_mesh(tris: np.ndarray, material=None, normals=False) -> Mesh:
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 3, 3)
    v = t.reshape(-1, 3)
    idx = np.arange(v.shape[0], dtype=np.uint32).reshape(-1, 3)
    return Mesh(v, idx, None, material if material is not None else disney_material())


def icosphere_scene(count=25, subdiv=6, seed=1, amp=0.05) -> Scene:
    """Traversal workload scene: 25 x subdiv 6 = 2,048,000 triangles by
    default (the 2M-triangle point); 64 x subdiv 7 = 20,971,520 is the other."""
    meshes = [soup_mesh(t) for t in icosphere_soup(count, subdiv, seed, amp)]
    return Scene(meshes, None, (0.0, 0.0, 0.0), f"icospheres-{count}x{subdiv}")


def traversal_rays(n: int, seed=2, tmin=0.0, tmax=np.inf) -> np.ndarray:
    """Incoherent rays, (n, 8) float32 rows (ox, oy, oz, tmin, dx, dy, dz,
    tmax): origins ~ U[0,1]^3, directions uniform on S^2 (z = 1-2u, phi = 2 pi v)."""
    rng = np.random.default_rng(seed)
    o = rng.random((n, 3), dtype=np.float64)
    u = rng.random(n)
    v = rng.random(n)
    z = 1.0 - 2.0 * u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.empty((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = tmin
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    return rays


def wavefront_scene(count=25, subdiv=6, seed=4, width=1920, height=1080,
                    env=0.1, light_L=10.0) -> Scene:
    """Wavefront-PT workload: the icosphere soup with per-object Disney
    parameters (metallic ~ U[0,1], roughness ~ U[0.05,1], base colour ~
    U[0.1,0.9]^3, 25% of objects specTrans = 1 with IOR 1.5), two 0.3 x 0.3
    quad lights at y = 1.2 (L = light_L) and a constant environment."""
    rng = np.random.default_rng(seed)
    meshes = []
    for t in icosphere_soup(count, subdiv, seed=1):
        col = rng.uniform(0.1, 0.9, 3)
        met = rng.uniform(0, 1)
        rough = rng.uniform(0.05, 1.0)
        trans = 1.0 if rng.uniform() < 0.25 else 0.0
        mat = disney_material(color=col, metallic=met, roughness=rough,
                              trans=trans, eta=1.5)
        meshes.append(soup_mesh(t, mat))
    for cx, cz in ((0.3, 0.35), (0.7, 0.65)):
        h = 0.15
        v, i, n = _quad([cx - h, 1.2, cz - h], [cx + h, 1.2, cz - h],
                        [cx + h, 1.2, cz + h], [cx - h, 1.2, cz + h], [0, -1, 0])
        meshes.append(Mesh(v, i, n, matte_material((0, 0, 0)),
                           light_L=(light_L,) * 3, two_sided=False))
    cam = pinhole_camera([0.5, 0.5, 2.2], [0, 0, -1], [0, 1, 0], 40.0, width, height)
    return Scene(meshes, cam, (env, env, env), f"wavefront-{count}x{subdiv}")


def glossy_caustic(width=1024, height=1024, sphere_subdiv=6, floor_n=243) -> Scene:
    """Glossy-caustic workload: glass icosphere (IOR 1.5, Glass mapping) over a
    rough-metal floor grid (metallic 1, roughness 0.2) with a 0.1 x 0.1 quad
    light behind the sphere; the camera looks over the lit floor in front of
    the sphere."""
    meshes = []
    sph = unit_icosphere_soup(sphere_subdiv) * 0.3 + np.array([0.0, 0.32, 0.0])
    meshes.append(soup_mesh(sph.astype(np.float32), glass_material(eta=1.5)))
    # floor grid in [-2,2]^2 at y = 0, 2 * floor_n^2 triangles
    g = np.linspace(-2.0, 2.0, floor_n + 1, dtype=np.float64)
    X, Z = np.meshgrid(g, g, indexing="ij")
    P = np.stack([X, np.zeros_like(X), Z], -1).astype(np.float32).reshape(-1, 3)
    ii, jj = np.meshgrid(np.arange(floor_n), np.arange(floor_n), indexing="ij")
    a = (ii * (floor_n + 1) + jj).reshape(-1)
    b = a + (floor_n + 1)
    idx = np.concatenate([np.stack([a, a + 1, b + 1], 1),
                          np.stack([a, b + 1, b], 1)]).astype(np.uint32)
    floor_mat = disney_material(color=(0.9, 0.8, 0.7), metallic=1.0, roughness=0.2)
    meshes.append(Mesh(P, idx, np.tile(np.float32([0, 1, 0]), (P.shape[0], 1)), floor_mat))
    h = 0.05
    lv, li, ln = _quad([-h, 0.35, -0.55 - h], [-h, 0.35 + 2 * h, -0.55 - h],
                       [h, 0.35 + 2 * h, -0.55 - h], [h, 0.35, -0.55 - h], [0, 0, 1])
    meshes.append(Mesh(lv, li, ln, matte_material((0, 0, 0)), light_L=(400.0,) * 3,
                       two_sided=False))
    # looking over the floor in front of the sphere: the light's cone on the
    # metal floor with the sphere's shadow and the caustic focused into it
    cam = pinhole_camera([0.0, 0.8, 3.5], [0.0, -0.8, -2.0], [0.0, 1.0, 0.0], 35.0,
                         width, height)
    return Scene(meshes, cam, (0.0, 0.0, 0.0), "glossy-caustic")


def serving_scenes():
    """The scenes of the batched-serving workload: four Cornell-box variants
    (base colours from seeds 0..3) and four icosphere variants (8 spheres at
    subdivision 4, materials and lights of the wavefront workload, seeds
    10..13), each with a square 256 x 256 camera."""
    out = [cornell_box(256, 256, seed=k) for k in range(4)]
    out += [wavefront_scene(8, 4, seed=10 + k, width=256, height=256) for k in range(4)]
    return out


def serving_requests(rng, n, res_lo, res_hi, spp_hi, seed_shift, num_scenes=8):
    """n render requests of the serving workload drawn from rng: scene index
    uniform over the serving scenes, square side uniform in [res_lo, res_hi],
    spp uniform in [1, spp_hi], depth uniform in [8, 16], frame offset
    (k + seed_shift) * 2^16 for request k.  The 64-request mix is
    serving_requests(np.random.default_rng(5), 64, 128, 1024, 64, 0)."""
    reqs = []
    for k in range(n):
        res = int(rng.integers(res_lo, res_hi + 1))
        reqs.append({"scene": int(rng.integers(0, num_scenes)), "width": res, "height": res,
                     "spp": int(rng.integers(1, spp_hi + 1)),
                     "max_depth": int(rng.integers(8, 17)),
                     "frame_offset": (k + seed_shift) << 16})
    return reqs
