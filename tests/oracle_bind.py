"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests and
the benchmark's CPU-baseline leg.  The oracle is what the GPU results are
compared against: nothing in the shipped package imports this module."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_meshes", C.c_uint32),
        ("vertices", f32p), ("normals", f32p), ("indices", u32p),
        ("mesh_vertex_offset", u32p), ("mesh_tri_offset", u32p),
        ("mesh_tri_count", u32p), ("mesh_has_normal", u32p),
        ("mesh_light", u32p), ("mesh_material", C.c_void_p),
        ("num_lights", C.c_uint32),
        ("light_mesh", u32p), ("light_L", f32p), ("light_two_sided", u32p),
        ("cam_pos", C.c_float * 3), ("cam_lookat", C.c_float * 3),
        ("cam_up", C.c_float * 3), ("cam_right", C.c_float * 3),
        ("bg", C.c_float * 3),
        ("num_textures", C.c_uint32),
        ("texcoords", f32p), ("mesh_texture", u32p), ("texture_offset", u32p),
        ("texture_size", u32p), ("texels", C.c_void_p),
    ]


class PTConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("rr_prob", C.c_float), ("epsilon", C.c_float),
                ("strategy", C.c_uint32)]


class ExampleConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("sample_num", C.c_uint32), ("max_depth", C.c_uint32),
                ("cam_pos", C.c_float * 3), ("cam_gaze", C.c_float * 3),
                ("cam_up", C.c_float * 3),
                ("horizontal_ratio", C.c_float), ("vertical_ratio", C.c_float)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_tea.restype = C.c_uint32
        L.oracle_tea.argtypes = [C.c_uint32, C.c_uint32, u32p, C.c_uint32, u32p]
        L.oracle_tea4.restype = C.c_uint32
        L.oracle_tea4.argtypes = [C.c_uint32, C.c_uint32]
        L.oracle_lcg.restype = C.c_uint32
        L.oracle_lcg.argtypes = [u32p]
        L.oracle_rnd.restype = C.c_float
        L.oracle_rnd.argtypes = [u32p]
        L.oracle_intersect_f32.restype = C.c_int
        L.oracle_intersect_f32.argtypes = [C.c_void_p, f32p, f32p, f32p, f32p]
        L.oracle_trace_closest_brute.argtypes = [f32p, C.c_uint32, C.c_void_p,
                                                 C.c_uint32, C.c_void_p, C.c_int]
        L.oracle_trace_any_brute.argtypes = [f32p, C.c_uint32, C.c_void_p,
                                             C.c_uint32, u8p, C.c_int]
        L.oracle_bvh_build.restype = C.c_void_p
        L.oracle_bvh_build.argtypes = [f32p, C.c_uint32]
        L.oracle_bvh_free.argtypes = [C.c_void_p]
        L.oracle_bvh_info.argtypes = [C.c_void_p, u32p, u32p]
        L.oracle_bvh_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32,
                                               C.c_void_p, C.c_int]
        L.oracle_bvh_trace_any.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32,
                                           u8p, C.c_int]
        for sfx, fp in (("f32", f32p), ("f64", f64p)):
            getattr(L, f"oracle_disney_eval_{sfx}").argtypes = [C.c_void_p, fp, fp, fp, fp, fp]
            getattr(L, f"oracle_disney_pdf_{sfx}").argtypes = [C.c_void_p, fp, fp, fp, fp]
            getattr(L, f"oracle_disney_pdf_{sfx}").restype = C.c_float if sfx == "f32" else C.c_double
            getattr(L, f"oracle_disney_sample_{sfx}").argtypes = [C.c_void_p, fp, fp, fp, u32p, fp]
            getattr(L, f"oracle_disney_sample_batch_{sfx}").argtypes = [
                C.c_void_p, fp, fp, fp, u32p, C.c_uint32, fp]
            getattr(L, f"oracle_disney_pdf_batch_{sfx}").argtypes = [
                C.c_void_p, fp, fp, fp, fp, C.c_uint32, fp]
            getattr(L, f"oracle_disney_eval_batch_{sfx}").argtypes = [
                C.c_void_p, fp, fp, fp, fp, C.c_uint32, fp]
            getattr(L, f"oracle_material_math_{sfx}").argtypes = [C.c_int, fp]
            getattr(L, f"oracle_material_math_{sfx}").restype = (
                C.c_float if sfx == "f32" else C.c_double)
        L.oracle_onb_f32.argtypes = [f32p, f32p, f32p]
        L.oracle_aces_f32.argtypes = [f32p, f32p]
        L.oracle_tex_bilinear_f32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float,
                                              C.c_float, f32p]
        L.oracle_pt_create.restype = C.c_void_p
        L.oracle_pt_create.argtypes = [C.POINTER(SceneDesc)]
        L.oracle_pt_free.argtypes = [C.c_void_p]
        L.oracle_pt_sample_f32.restype = C.c_uint32
        L.oracle_pt_sample_f32.argtypes = [C.c_void_p, C.POINTER(PTConfig), C.c_uint32,
                                           C.c_uint32, C.c_uint32, f32p, u32p]
        L.oracle_pt_trace_f32.restype = C.c_uint32
        L.oracle_pt_trace_f32.argtypes = [C.c_void_p, C.POINTER(PTConfig), C.c_uint32,
                                          C.c_uint32, C.c_uint32, f32p, f32p, C.c_uint32]
        L.oracle_pt_render_frame_f32.argtypes = [C.c_void_p, C.POINTER(PTConfig),
                                                 C.c_uint32, f32p, C.c_int,
                                                 C.POINTER(C.c_uint64)]
        L.oracle_pt_render_f64.argtypes = [C.c_void_p, C.POINTER(PTConfig), C.c_uint32,
                                           C.c_uint32, f64p, C.c_int]
        L.oracle_pt_render_f64_window.argtypes = [
            C.c_void_p, C.POINTER(PTConfig), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
            C.c_uint32, C.c_uint32, C.c_uint32, f64p, f64p, C.c_int]
        L.oracle_film_update.argtypes = [f32p, u8p, f32p, C.c_uint32, C.c_uint32]
        L.oracle_example_render.argtypes = [C.c_void_p, f32p, f32p, C.POINTER(ExampleConfig),
                                            C.c_uint32, C.c_uint32, f32p, C.c_int]
        L.oracle_bf16_round.restype = C.c_float
        L.oracle_bf16_round.argtypes = [C.c_float]
        L.oracle_guide_encode.argtypes = [f32p, f64p]
        L.oracle_guide_forward.argtypes = [f64p, f32p, C.c_int, f64p]
        L.oracle_guide_decode.argtypes = [f64p, f64p, f64p, f64p]
        L.oracle_guide_pdf.restype = C.c_double
        L.oracle_guide_pdf.argtypes = [f64p, f64p]
        L.oracle_guide_sample.argtypes = [f64p, C.c_double, C.c_double, C.c_double, f64p]
        L.oracle_guide_pdf_batch.argtypes = [f64p, f64p, C.c_uint32, f64p]
        L.oracle_guide_sample_batch.argtypes = [f64p, f64p, C.c_uint32, f64p]
        L.oracle_guide_loss_grad.restype = C.c_double
        L.oracle_guide_loss_grad.argtypes = [f64p, f32p, C.c_uint32, C.c_int, f64p]
        L.oracle_guide_adam.argtypes = [f64p, f64p, f64p, f64p, C.c_uint32, C.c_double,
                                        C.c_double, C.c_double, C.c_double]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


# ---- RNG --------------------------------------------------------------------
def tea(v0, v1, key, rounds):
    k = np.asarray(key, np.uint32)
    out1 = C.c_uint32(0)
    v = lib().oracle_tea(v0, v1, _p(k, u32p), rounds, C.byref(out1))
    return int(v), int(out1.value)


def tea4(v0, v1):
    return int(lib().oracle_tea4(v0, v1))


def rnd_sequence(seed, n):
    s = C.c_uint32(seed)
    return np.array([lib().oracle_rnd(C.byref(s)) for _ in range(n)], np.float32)


def lcg_sequence(seed, n):
    s = C.c_uint32(seed)
    out = []
    for _ in range(n):
        lib().oracle_lcg(C.byref(s))
        out.append(s.value)
    return out


# ---- traversal --------------------------------------------------------------
HIT_DTYPE = np.dtype([("t", np.float32), ("prim", np.uint32),
                      ("b1", np.float32), ("b2", np.float32)])


def intersect(ray, tri):
    r = np.ascontiguousarray(ray, np.float32)
    v = np.ascontiguousarray(tri, np.float32)
    t, b1, b2 = C.c_float(), C.c_float(), C.c_float()
    ok = lib().oracle_intersect_f32(r.ctypes.data, _p(v, f32p), C.byref(t),
                                    C.byref(b1), C.byref(b2))
    return bool(ok), t.value, b1.value, b2.value


def closest_brute(tris, rays, nthreads=None):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    lib().oracle_trace_closest_brute(_p(tris, f32p), tris.shape[0], rays.ctypes.data,
                                     rays.shape[0], hits.ctypes.data,
                                     nthreads or nthreads_default())
    return hits


def any_brute(tris, rays, nthreads=None):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    occ = np.zeros(rays.shape[0], np.uint8)
    lib().oracle_trace_any_brute(_p(tris, f32p), tris.shape[0], rays.ctypes.data,
                                 rays.shape[0], _p(occ, u8p), nthreads or nthreads_default())
    return occ


class OracleBVH:
    def __init__(self, tris):
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.ntris = tris.shape[0]
        self.h = lib().oracle_bvh_build(_p(tris, f32p), tris.shape[0])

    def info(self):
        n, d = C.c_uint32(), C.c_uint32()
        lib().oracle_bvh_info(self.h, C.byref(n), C.byref(d))
        return n.value, d.value

    def closest(self, rays, nthreads=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(rays.shape[0], HIT_DTYPE)
        lib().oracle_bvh_trace_closest(self.h, rays.ctypes.data, rays.shape[0],
                                       hits.ctypes.data, nthreads or nthreads_default())
        return hits

    def any(self, rays, nthreads=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], np.uint8)
        lib().oracle_bvh_trace_any(self.h, rays.ctypes.data, rays.shape[0],
                                   _p(occ, u8p), nthreads or nthreads_default())
        return occ

    def __del__(self):
        try:
            if self.h:
                lib().oracle_bvh_free(self.h)
                self.h = None
        except Exception:
            pass


# ---- BSDF -------------------------------------------------------------------
def _v(x, dt):
    return np.ascontiguousarray(x, dt).reshape(3)


def disney_eval(mat, Ns, Ng, V, L, f64=False):
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    out = np.zeros(3, dt)
    a = [_v(x, dt) for x in (Ns, Ng, V, L)]
    getattr(lib(), f"oracle_disney_eval_{sfx}")(m.ctypes.data, *[_p(x, fp) for x in a], _p(out, fp))
    return out


def disney_pdf(mat, Ns, Ng, V, L, f64=False):
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    a = [_v(x, dt) for x in (Ns, Ng, V, L)]
    return getattr(lib(), f"oracle_disney_pdf_{sfx}")(m.ctypes.data, *[_p(x, fp) for x in a])


def disney_sample(mat, Ns, Ng, V, seed, f64=False):
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    a = [_v(x, dt) for x in (Ns, Ng, V)]
    s = C.c_uint32(seed)
    out = np.zeros(3, dt)
    getattr(lib(), f"oracle_disney_sample_{sfx}")(m.ctypes.data, *[_p(x, fp) for x in a],
                                                  C.byref(s), _p(out, fp))
    return out, s.value


def disney_sample_batch(mat, Ns, Ng, V, seeds, f64=True):
    """Directions (n, 3) sampled from the seeds (n,), one per seed."""
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    a = [_v(x, dt) for x in (Ns, Ng, V)]
    sd = np.ascontiguousarray(seeds, np.uint32).reshape(-1)
    out = np.zeros((sd.size, 3), dt)
    getattr(lib(), f"oracle_disney_sample_batch_{sfx}")(
        m.ctypes.data, *[_p(x, fp) for x in a], _p(sd, u32p), sd.size, _p(out, fp))
    return out


def disney_pdf_batch(mat, Ns, Ng, V, L, f64=True):
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    a = [_v(x, dt) for x in (Ns, Ng, V)]
    Ls = np.ascontiguousarray(L, dt).reshape(-1, 3)
    out = np.zeros(Ls.shape[0], dt)
    getattr(lib(), f"oracle_disney_pdf_batch_{sfx}")(
        m.ctypes.data, *[_p(x, fp) for x in a], _p(Ls, fp), Ls.shape[0], _p(out, fp))
    return out


def disney_eval_batch(mat, Ns, Ng, V, L, f64=True):
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    a = [_v(x, dt) for x in (Ns, Ng, V)]
    Ls = np.ascontiguousarray(L, dt).reshape(-1, 3)
    out = np.zeros((Ls.shape[0], 3), dt)
    getattr(lib(), f"oracle_disney_eval_batch_{sfx}")(
        m.ctypes.data, *[_p(x, fp) for x in a], _p(Ls, fp), Ls.shape[0], _p(out, fp))
    return out


MM_SCHLICK, MM_FRESNEL, MM_GTR1, MM_GTR2, MM_SMITHG, MM_LAMBDA = range(6)


def material_math(which, *args, f64=True):
    """One MaterialMath function (Common.h:119-181) of the oracle."""
    dt, fp, sfx = (np.float64, f64p, "f64") if f64 else (np.float32, f32p, "f32")
    a = np.zeros(8, dt)
    flat = np.concatenate([np.atleast_1d(np.asarray(x, dt)).reshape(-1) for x in args])
    a[:flat.size] = flat
    return getattr(lib(), f"oracle_material_math_{sfx}")(which, _p(a, fp))


def onb(n):
    n = _v(n, np.float32)
    u = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    lib().oracle_onb_f32(_p(n, f32p), _p(u, f32p), _p(v, f32p))
    return u, v


def aces(rad):
    r = _v(rad, np.float32)
    out = np.zeros(3, np.float32)
    lib().oracle_aces_f32(_p(r, f32p), _p(out, f32p))
    return out


def tex_bilinear(rgba, u, v):
    """Filtered texel of an (H, W, 4) uint8 image at normalized (u, v)."""
    img = np.ascontiguousarray(rgba, np.uint8)
    out = np.zeros(3, np.float32)
    lib().oracle_tex_bilinear_f32(img.ctypes.data, img.shape[1], img.shape[0], u, v,
                                  _p(out, f32p))
    return out


# ---- path tracer ------------------------------------------------------------
class OraclePT:
    """The oracle's scene + integrator (PathTracing.cu restated)."""

    def __init__(self, scene):
        self.scene = scene
        f = scene.flatten()
        self._keep = f
        d = SceneDesc()
        d.num_meshes = len(scene.meshes)
        d.vertices = _p(f["vertices"], f32p)
        d.normals = _p(f["normals"], f32p)
        d.indices = _p(f["indices"], u32p)
        for k in ("mesh_vertex_offset", "mesh_tri_offset", "mesh_tri_count",
                  "mesh_has_normal", "mesh_light", "light_mesh", "light_two_sided"):
            setattr(d, k, _p(f[k], u32p))
        d.mesh_material = f["mesh_material"].ctypes.data
        d.num_lights = f["light_mesh"].shape[0]
        d.light_L = _p(f["light_L"], f32p)
        cam = scene.camera
        if cam is not None:
            d.cam_pos[:] = list(cam.pos)
            d.cam_lookat[:] = list(cam.lookat)
            d.cam_up[:] = list(cam.up)
            d.cam_right[:] = list(cam.right)
        d.bg[:] = list(f["bg"])
        d.num_textures = f["texture_offset"].shape[0]
        if d.num_textures:
            d.texcoords = _p(f["texcoords"], f32p)
            d.mesh_texture = _p(f["mesh_texture"], u32p)
            d.texture_offset = _p(f["texture_offset"], u32p)
            d.texture_size = _p(f["texture_size"], u32p)
            d.texels = f["texels"].ctypes.data
        self.h = lib().oracle_pt_create(C.byref(d))

    @staticmethod
    def config(width, height, max_depth=25, rr_depth=5, rr_prob=0.9,
               epsilon=1e-3, strategy=2):
        return PTConfig(width, height, max_depth, rr_depth, rr_prob, epsilon, strategy)

    def sample(self, cfg, x, y, frame):
        rgb = np.zeros(3, np.float32)
        sh = C.c_uint32()
        n = lib().oracle_pt_sample_f32(self.h, C.byref(cfg), x, y, frame,
                                       _p(rgb, f32p), C.byref(sh))
        return rgb, n, sh.value

    def trace(self, cfg, x, y, frame, max_vertices=64):
        """One sample with its per-vertex record (see oracle_pt_trace_f32)."""
        rgb = np.zeros(3, np.float32)
        tr = np.zeros((max_vertices, 8), np.float32)
        n = lib().oracle_pt_trace_f32(self.h, C.byref(cfg), x, y, frame, _p(rgb, f32p),
                                      _p(tr, f32p), max_vertices)
        return rgb, tr[:n]

    def render_frame(self, cfg, frame, nthreads=None):
        out = np.zeros((cfg.height, cfg.width, 3), np.float32)
        rays = C.c_uint64()
        lib().oracle_pt_render_frame_f32(self.h, C.byref(cfg), frame, _p(out, f32p),
                                         nthreads or nthreads_default(), C.byref(rays))
        return out, rays.value

    def render_f64(self, cfg, frame0, nframes, nthreads=None):
        out = np.zeros((cfg.height, cfg.width, 3), np.float64)
        lib().oracle_pt_render_f64(self.h, C.byref(cfg), frame0, nframes, _p(out, f64p),
                                   nthreads or nthreads_default())
        return out

    def render_f64_window(self, cfg, x0, y0, ww, wh, frame0, nframes, nthreads=None,
                          stride=1):
        """FP64 mean and per-pixel one-sample variance over the ww x wh
        pixels (x0 + stride i, y0 + stride j) of the cfg.width x cfg.height
        image (pixels keep their whole-image seeds; stride 1: a window)."""
        assert x0 + stride * (ww - 1) < cfg.width and y0 + stride * (wh - 1) < cfg.height
        mean = np.zeros((wh, ww, 3), np.float64)
        var = np.zeros((wh, ww, 3), np.float64)
        lib().oracle_pt_render_f64_window(self.h, C.byref(cfg), x0, y0, ww, wh, stride, frame0,
                                          nframes, _p(mean, f64p), _p(var, f64p),
                                          nthreads or nthreads_default())
        return mean, var

    def example_render(self, mesh_color, mesh_emission, ecfg, row0, row1, nthreads=None):
        mc = np.ascontiguousarray(mesh_color, np.float32)
        me = np.ascontiguousarray(mesh_emission, np.float32)
        out = np.zeros(((row1 - row0) * ecfg.height, 4), np.float32)
        lib().oracle_example_render(self.h, _p(mc, f32p), _p(me, f32p), C.byref(ecfg),
                                    row0, row1, _p(out, f32p), nthreads or nthreads_default())
        return out

    def __del__(self):
        try:
            if self.h:
                lib().oracle_pt_free(self.h)
                self.h = None
        except Exception:
            pass


def film_update(acc_rgba, color_rgba, this_frame_rgb, frame):
    npix = this_frame_rgb.reshape(-1, 3).shape[0]
    tf = np.ascontiguousarray(this_frame_rgb, np.float32)
    lib().oracle_film_update(_p(acc_rgba, f32p), _p(color_rgba, u8p), _p(tf, f32p),
                             npix, frame)


# ---- learned guiding ----------------------------------------------------------
GUIDE_PARAMS = 18720


def bf16_round(x: float) -> float:
    return float(lib().oracle_bf16_round(float(x)))


def guide_encode(q):
    q = np.ascontiguousarray(q, np.float32)
    x = np.zeros(64, np.float64)
    lib().oracle_guide_encode(_p(q, f32p), _p(x, f64p))
    return x


def guide_forward(params, queries, emulate_bf16=True):
    P = np.ascontiguousarray(params, np.float64)
    Q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
    out = np.zeros((Q.shape[0], 32), np.float64)
    for i in range(Q.shape[0]):
        lib().oracle_guide_forward(_p(P, f64p), _p(Q[i], f32p), int(emulate_bf16),
                                   _p(out[i], f64p))
    return out


def guide_decode(out):
    o = np.ascontiguousarray(out, np.float64)
    pi, kap, mu = np.zeros(8), np.zeros(8), np.zeros(24)
    lib().oracle_guide_decode(_p(o, f64p), _p(pi, f64p), _p(kap, f64p), _p(mu, f64p))
    return pi, kap, mu.reshape(8, 3)


def guide_pdf(out, dirs):
    o = np.ascontiguousarray(out, np.float64)
    D = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    pdf = np.zeros(D.shape[0])
    lib().oracle_guide_pdf_batch(_p(o, f64p), _p(D, f64p), D.shape[0], _p(pdf, f64p))
    return pdf


def guide_sample(out, u):
    o = np.ascontiguousarray(out, np.float64)
    U = np.ascontiguousarray(u, np.float64).reshape(-1, 3)
    w = np.zeros((U.shape[0], 3))
    lib().oracle_guide_sample_batch(_p(o, f64p), _p(U, f64p), U.shape[0], _p(w, f64p))
    return w


def guide_loss_grad(params, samples, emulate_bf16=False):
    P = np.ascontiguousarray(params, np.float64)
    S = np.ascontiguousarray(samples, np.float32).reshape(-1, 16)
    g = np.zeros(GUIDE_PARAMS, np.float64)
    loss = lib().oracle_guide_loss_grad(_p(P, f64p), _p(S, f32p), S.shape[0],
                                        int(emulate_bf16), _p(g, f64p))
    return loss, g


def guide_adam(params, m, v, grad, step, lr, beta1, beta2, eps):
    """In place on float64 arrays."""
    lib().oracle_guide_adam(_p(params, f64p), _p(m, f64p), _p(v, f64p),
                            _p(np.ascontiguousarray(grad, np.float64), f64p),
                            step, lr, beta1, beta2, eps)
