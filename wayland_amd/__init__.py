"""wayland_amd — MI355X (gfx950) rendering backend for EasyRender.

Python host-side mirror of the C ABI in include/wayland_hip.h, used by the
tests and the benchmark.  Device buffers are torch tensors on a ROCm device
(plumbing only); all work runs in the HIP kernels of libwayland_hip.so, on
torch's current stream so that it orders with surrounding torch work.

    ctx   = Context(0)                      # ContextManager
    accel = Accel(ctx, scene)               # StaticAccelStructure
    hits  = accel.trace_closest(rays)       # optixTrace, closest hit
    occ   = accel.trace_any(rays)           # optixTrace, terminate on first hit
    pt    = PathTracer(ctx, accel, scene)   # PathTracing program
    film  = pt.render(cfg)                  # wavefront PT into a compensated film
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import WL_INVALID_INDEX, WlError  # noqa: F401

__all__ = ["Context", "Accel", "PathTracer", "WlError", "WL_INVALID_INDEX",
           "HIT_DTYPE", "pt_config"]

HIT_DTYPE = np.dtype([("t", np.float32), ("prim", np.uint32),
                      ("b1", np.float32), ("b2", np.float32)])

STRATEGY_UPT, STRATEGY_NEE, STRATEGY_MIS = 0, 1, 2
PT_SORT_BY_MATERIAL = 1
# paths per batch of PathTracer.render_noalloc (WL_PT_NOALLOC_MAX_BATCH_PATHS)
NOALLOC_MAX_BATCH_PATHS = 1 << 18


def _torch():
    import torch
    return torch


def _stream_ptr(device=None):
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def device_count() -> int:
    n = C.c_int(0)
    _lib.load().wl_device_count(C.byref(n))
    return n.value


class Context:
    """One device + one stream (reference: ContextManager)."""

    def __init__(self, device: int = 0):
        self.device = device
        h = C.c_void_p()
        _lib.call("wl_context_create", device, C.byref(h))
        self.h = h

    def synchronize(self):
        _lib.call("wl_context_synchronize", self.h)

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accel:
    """Triangle acceleration structure over a scenes.Scene (one build input
    per mesh), reference StaticAccelStructure + TriangleBuildInputArray."""

    def __init__(self, ctx: Context, scene=None, meshes=None):
        self.ctx = ctx
        meshes = meshes if meshes is not None else scene.meshes
        keep = []
        arr = (_lib.TriInput * len(meshes))()
        for i, m in enumerate(meshes):
            v = np.ascontiguousarray(m.vertices, np.float32).reshape(-1, 3)
            idx = np.ascontiguousarray(m.indices, np.uint32).reshape(-1, 3)
            keep += [v, idx]
            arr[i].vertices = v.ctypes.data
            arr[i].num_vertices = v.shape[0]
            arr[i].indices = idx.ctypes.data
            arr[i].num_triangles = idx.shape[0]
            arr[i].geometry_flags = 1
            arr[i].num_sbt_records = 1
        h = C.c_void_p()
        _lib.call("wl_accel_build_triangles", ctx.h, arr, len(meshes), 0, C.byref(h))
        self.h = h
        self.num_inputs = len(meshes)

    def info(self) -> dict:
        inf = _lib.AccelInfo()
        _lib.call("wl_accel_get_info", self.h, C.byref(inf))
        return {k: getattr(inf, k) for k, _ in inf._fields_ if k != "pad_"}

    def input_offsets(self) -> np.ndarray:
        out = np.zeros(self.num_inputs, np.uint32)
        _lib.call("wl_accel_input_offsets", self.h, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out

    # rays: torch float32 tensor (n, 8) on the device
    def trace_closest(self, rays, hits=None, counts=None):
        torch = _torch()
        n = rays.shape[0]
        assert rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[1] == 8
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        assert hits.shape == (n, 4) and hits.is_contiguous()
        if counts is not None:
            assert counts.shape == (n, 2) and counts.dtype == torch.int32
            _lib.call("wl_trace_closest_counted", self.h, C.c_void_p(rays.data_ptr()), n,
                      C.c_void_p(hits.data_ptr()), C.c_void_p(counts.data_ptr()),
                      _stream_ptr(rays.device))
        else:
            _lib.call("wl_trace_closest", self.h, C.c_void_p(rays.data_ptr()), n,
                      C.c_void_p(hits.data_ptr()), _stream_ptr(rays.device))
        return hits

    def trace_any(self, rays, bits=None):
        torch = _torch()
        n = rays.shape[0]
        assert rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[1] == 8
        words = (n + 31) // 32
        if bits is None:
            bits = torch.empty(max(words, 1), dtype=torch.int32, device=rays.device)
        assert bits.numel() >= words
        _lib.call("wl_trace_any", self.h, C.c_void_p(rays.data_ptr()), n,
                  C.c_void_p(bits.data_ptr()), _stream_ptr(rays.device))
        return bits

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_accel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hits_to_numpy(hits) -> np.ndarray:
    """(n, 4) float32 device tensor -> structured array with HIT_DTYPE."""
    a = hits.detach().cpu().numpy()
    return np.ascontiguousarray(a).view(HIT_DTYPE).reshape(-1)


def unpack_bits(bits, n) -> np.ndarray:
    w = bits.detach().cpu().numpy().view(np.uint32)
    b = np.unpackbits(w.view(np.uint8), bitorder="little")
    return b[:n].astype(np.uint8)


def pt_config(width, height, spp=1, frame_offset=0, max_depth=25, rr_depth=5,
              rr_prob=0.9, epsilon=1e-3, strategy=STRATEGY_MIS, flags=0,
              max_batch_paths=0):
    return _lib.PTConfig(width, height, spp, frame_offset, max_depth, rr_depth,
                         rr_prob, epsilon, strategy, flags, max_batch_paths, 0)


def identify_programs(source: str):
    """Module::IdentifyPrograms on source text -> list of (kind, entries)."""
    data = source.encode()
    cap = 64
    arr = (_lib.ProgramGroup * cap)()
    n = C.c_uint32()
    _lib.call("wl_identify_programs", data, len(data), arr, cap, C.byref(n))
    return [(arr[i].kind, tuple(arr[i].entry[k].value.decode() for k in range(3)))
            for i in range(n.value)], arr, n.value


class Pipeline:
    """Program groups -> a built-in gfx950 integrator (reference Pipeline)."""

    def __init__(self, ctx: Context, source: str, max_trace_depth: int = 1):
        self.ctx = ctx
        self.groups, arr, n = identify_programs(source)
        h = C.c_void_p()
        b = C.c_uint32()
        _lib.call("wl_pipeline_create", ctx.h, arr, n, max_trace_depth, C.byref(h), C.byref(b))
        self.h, self.builtin = h, b.value

    def group_index(self, entry: str) -> int:
        for i, (_, names) in enumerate(self.groups):
            if entry in names:
                return i
        raise KeyError(entry)

    def record(self, group_index: int, data=None) -> bytes:
        """One SBT record: packed header + data, padded to 16 bytes."""
        size = 0 if data is None else C.sizeof(data)
        buf = (C.c_uint8 * _lib.record_stride(size))()
        _lib.call("wl_sbt_pack_header", self.h, group_index, buf)
        if data is not None:
            C.memmove(C.addressof(buf) + _lib.SBT_HEADER, C.addressof(data), size)
        return bytes(buf)

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_pipeline_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SBT:
    """Shader binding table (reference ShaderBindingTable)."""

    def __init__(self, pipeline: Pipeline, raygen: bytes, miss: list, hit: list):
        self.pipeline = pipeline
        ms = len(miss[0]) if miss else 0
        hs = len(hit[0]) if hit else 0
        mb = b"".join(miss)
        hb = b"".join(hit)
        h = C.c_void_p()
        _lib.call("wl_sbt_create", pipeline.h, raygen, len(raygen), mb or None, len(miss), ms,
                  hb or None, len(hit), hs, C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_sbt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Texture:
    """A texture object for PathTracing hit records (cudaCreateTextureObject,
    DeviceManager.cpp:192-227); `handle` goes into the record's texture field."""

    def __init__(self, ctx: Context, image):
        img = np.ascontiguousarray(image, np.uint8)
        if img.ndim != 3 or img.shape[2] != 4:
            raise ValueError("texture images must be (H, W, 4) uint8")
        desc = _lib.TextureImage(img.ctypes.data, img.shape[1], img.shape[0])
        h = C.c_uint64()
        _lib.call("wl_texture_create", ctx.h, C.byref(desc), C.byref(h))
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().wl_texture_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def traversable(accel: Accel) -> int:
    v = C.c_uint64()
    _lib.call("wl_accel_traversable", accel.h, C.byref(v))
    return v.value


def launch(pipeline: Pipeline, params, sbt: SBT, width: int, height: int, depth: int = 1):
    """optixLaunch: params is a ctypes struct in the program's LaunchParams layout."""
    _lib.call("wl_launch", pipeline.h, _stream_ptr(), C.byref(params), C.sizeof(params),
              sbt.h, width, height, depth)


def float_to_u8(src):
    torch = _torch()
    dst = torch.empty(src.numel(), dtype=torch.uint8, device=src.device)
    _lib.call("wl_float_to_u8", C.c_void_p(src.data_ptr()), src.numel(),
              C.c_void_p(dst.data_ptr()), _stream_ptr(src.device))
    return dst


GUIDE_PARAMS = 18720
GUIDE_SAMPLE = 1
GUIDE_TRAIN = 2


class Guide:
    """Learned guiding network: 64 -> 4 x 64 -> 32 MLP (BF16 MFMA, FP32
    master weights) giving an 8-lobe von Mises-Fisher mixture.  No
    counterpart in the reference."""

    def __init__(self, ctx: Context, **cfg):
        c = _lib.GuideConfig()
        _lib.call("wl_guide_default_config", C.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown guide option {k}")
            setattr(c, k, v)
        self.ctx, self.cfg = ctx, c
        h = C.c_void_p()
        _lib.call("wl_guide_create", ctx.h, C.byref(c), C.byref(h))
        self.h = h

    def get_params(self) -> np.ndarray:
        out = np.empty(GUIDE_PARAMS, np.float32)
        _lib.call("wl_guide_get_params", self.h,
                  out.ctypes.data_as(C.POINTER(C.c_float)), GUIDE_PARAMS)
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != GUIDE_PARAMS:
            raise ValueError("wrong parameter count")
        _lib.call("wl_guide_set_params", self.h,
                  p.ctypes.data_as(C.POINTER(C.c_float)), GUIDE_PARAMS)

    def set_mix(self, alpha: float):
        _lib.call("wl_guide_set_mix", self.h, float(alpha))

    def info(self) -> dict:
        i = _lib.GuideInfo()
        _lib.call("wl_guide_get_info", self.h, C.byref(i))
        return {"steps": i.steps, "last_samples": i.last_samples}

    def infer(self, queries, out=None):
        torch = _torch()
        q = queries.contiguous()
        n = q.shape[0]
        if out is None:
            out = torch.empty((n, 32), dtype=torch.float32, device=q.device)
        _lib.call("wl_guide_infer", self.h, C.c_void_p(q.data_ptr()), n,
                  C.c_void_p(out.data_ptr()), _stream_ptr(q.device))
        return out

    def gradient(self, samples):
        torch = _torch()
        smp = samples.contiguous()
        grad = torch.empty(GUIDE_PARAMS, dtype=torch.float32, device=smp.device)
        loss = torch.empty(1, dtype=torch.float32, device=smp.device)
        _lib.call("wl_guide_gradient", self.h, C.c_void_p(smp.data_ptr()), smp.shape[0],
                  C.c_void_p(grad.data_ptr()), C.c_void_p(loss.data_ptr()),
                  _stream_ptr(smp.device))
        return grad, loss

    def train(self, samples, steps: int):
        smp = samples.contiguous()
        _lib.call("wl_guide_train", self.h, C.c_void_p(smp.data_ptr()), smp.shape[0],
                  int(steps), _stream_ptr(smp.device))

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mixture_sample(ctx: Context, outs, u):
    """Directions and mixture densities for given network outputs and
    uniforms (3 per query)."""
    torch = _torch()
    o, uu = outs.contiguous(), u.contiguous()
    n = o.shape[0]
    dirs = torch.empty((n, 3), dtype=torch.float32, device=o.device)
    pdf = torch.empty(n, dtype=torch.float32, device=o.device)
    _lib.call("wl_guide_mixture_sample", ctx.h, C.c_void_p(o.data_ptr()),
              C.c_void_p(uu.data_ptr()), n, C.c_void_p(dirs.data_ptr()),
              C.c_void_p(pdf.data_ptr()), _stream_ptr(o.device))
    return dirs, pdf


def mixture_pdf(ctx: Context, outs, dirs):
    torch = _torch()
    o, d = outs.contiguous(), dirs.contiguous()
    n = o.shape[0]
    pdf = torch.empty(n, dtype=torch.float32, device=o.device)
    _lib.call("wl_guide_mixture_pdf", ctx.h, C.c_void_p(o.data_ptr()),
              C.c_void_p(d.data_ptr()), n, C.c_void_p(pdf.data_ptr()), _stream_ptr(o.device))
    return pdf


class Server:
    """Batched serving of many render requests (megabatches replayed from one
    captured hipGraph).  A request is a dict: pt (PathTracer), width, height,
    spp, and optionally frame_offset, max_depth, rr_depth, rr_prob, epsilon,
    strategy, film (a zeroed (W*H, 6) float32 CUDA tensor)."""

    def __init__(self, ctx: Context, **cfg):
        c = _lib.ServeConfig()
        _lib.call("wl_serve_default_config", C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        h = C.c_void_p()
        _lib.call("wl_server_create", ctx.h, C.byref(c), C.byref(h))
        self.ctx, self.h = ctx, h

    @staticmethod
    def _requests(reqs):
        torch = _torch()
        arr = (_lib.RenderRequest * len(reqs))()
        films = []
        for k, r in enumerate(reqs):
            W, H = r["width"], r["height"]
            film = r.get("film")
            if film is None:
                film = torch.zeros((W * H, 6), dtype=torch.float32, device="cuda")
            films.append(film)
            a = arr[k]
            a.scene = r["pt"].h
            a.film = film.data_ptr()
            a.width, a.height, a.spp = W, H, r["spp"]
            a.frame_offset = r.get("frame_offset", 0)
            a.max_depth = r.get("max_depth", 25)
            a.rr_depth = r.get("rr_depth", 5)
            a.rr_prob = r.get("rr_prob", 0.9)
            a.epsilon = r.get("epsilon", 1e-3)
            a.strategy = r.get("strategy", STRATEGY_MIS)
        return arr, films

    def config(self) -> dict:
        c = _lib.ServeConfig()
        _lib.call("wl_server_get_config", self.h, C.byref(c))
        return {"batch_paths": c.batch_paths, "shade_blocks_per_cu": c.shade_blocks_per_cu,
                "use_graph": c.use_graph, "lanes": c.lanes}

    def set_config(self, **cfg):
        c = _lib.ServeConfig()
        _lib.call("wl_server_get_config", self.h, C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        _lib.call("wl_server_set_config", self.h, C.byref(c))

    def render(self, reqs):
        """Returns (films, stats dict, per-request completion seconds)."""
        arr, films = self._requests(reqs)
        st = _lib.ServeStats()
        done = (C.c_double * len(reqs))()
        _lib.call("wl_server_render", self.h, arr, len(reqs), C.byref(st), done,
                  _stream_ptr())
        stats = {k: getattr(st, k) for k, _ in st._fields_ if k != "pad_"}
        return films, stats, list(done)

    def autotune(self, reqs, max_spp=4) -> dict:
        arr, _ = self._requests([dict(r, film=None) for r in reqs])
        c = _lib.ServeConfig()
        _lib.call("wl_server_autotune", self.h, arr, len(reqs), int(max_spp), C.byref(c))
        return {"batch_paths": c.batch_paths, "shade_blocks_per_cu": c.shade_blocks_per_cu,
                "use_graph": c.use_graph, "lanes": c.lanes}

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PathTracer:
    """The PathTracing program on a scene (reference:
    src/Frontend/Programs/PathTracing/*)."""

    def __init__(self, ctx: Context, accel: Accel, scene):
        self.ctx, self.accel, self.scene = ctx, accel, scene
        n = len(scene.meshes)
        meshes = (_lib.MeshShading * n)()
        keep = []
        light_of = {}
        lights = []
        for i, m in enumerate(scene.meshes):
            if m.light_L is not None:
                light_of[i] = len(lights)
                lights.append((i, m.light_L, m.two_sided))
        for i, m in enumerate(scene.meshes):
            if m.normals is not None:
                nrm = np.ascontiguousarray(m.normals, np.float32).reshape(-1, 3)
                keep.append(nrm)
                meshes[i].normals = nrm.ctypes.data
            else:
                meshes[i].normals = None
            meshes[i].material.values[:] = [float(x) for x in np.asarray(m.material, np.float32)]
            meshes[i].area_light = light_of.get(i, WL_INVALID_INDEX)
        larr = (_lib.AreaLight * max(1, len(lights)))()
        for k, (mi, L, two) in enumerate(lights):
            larr[k].mesh = mi
            larr[k].L[:] = [float(x) for x in L]
            larr[k].two_sided = 1 if two else 0
        cam = _lib.PinholeCamera()
        if scene.camera is not None:
            cam.pos[:] = [float(x) for x in scene.camera.pos]
            cam.lookat[:] = [float(x) for x in scene.camera.lookat]
            cam.up[:] = [float(x) for x in scene.camera.up]
            cam.right[:] = [float(x) for x in scene.camera.right]
        bg = (C.c_float * 3)(*[float(x) for x in scene.bg])
        h = C.c_void_p()
        _lib.call("wl_scene_create", ctx.h, accel.h, meshes, n, larr, len(lights),
                  C.byref(cam), bg, C.byref(h))
        self.h = h
        if getattr(scene, "textures", None):
            self.set_textures(scene.textures, [(m.uv, m.texture) for m in scene.meshes])

    def set_textures(self, images, bindings):
        """images: (H, W, 4) uint8 arrays, top row first; bindings: one
        (uv (V, 2) or None, image index or None) per mesh.  No images removes
        the scene's textures."""
        imgs = [np.ascontiguousarray(im, np.uint8) for im in images]
        for im in imgs:
            if im.ndim != 3 or im.shape[2] != 4:
                raise ValueError("texture images must be (H, W, 4) uint8")
        iarr = (_lib.TextureImage * max(1, len(imgs)))()
        for k, im in enumerate(imgs):
            iarr[k].rgba = im.ctypes.data
            iarr[k].height, iarr[k].width = im.shape[0], im.shape[1]
        n = len(bindings)
        marr = (_lib.MeshTexture * max(1, n))()
        keep = []
        for k, (uv, tex) in enumerate(bindings):
            if uv is not None:
                u = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
                keep.append(u)
                marr[k].texcoords = u.ctypes.data
            else:
                marr[k].texcoords = None
            marr[k].texture = WL_INVALID_INDEX if tex is None else int(tex)
        _lib.call("wl_scene_set_textures", self.h, iarr, len(imgs), marr, n)

    def render(self, cfg, film=None, stats=True):
        torch = _torch()
        npix = cfg.width * cfg.height
        if film is None:
            film = torch.zeros((npix, 6), dtype=torch.float32, device="cuda")
        st = _lib.PTStats()
        _lib.call("wl_pt_render", self.h, C.byref(cfg), C.c_void_p(film.data_ptr()),
                  C.byref(st) if stats else None, _stream_ptr(film.device))
        return film, ({k: getattr(st, k) for k, _ in st._fields_ if k != "pad_"}
                      if stats else None)

    def render_noalloc(self, cfg, film):
        """wl_pt_render_noalloc: the same film as render(), from the library's
        fixed scratch array (no allocation; batches of at most
        NOALLOC_MAX_BATCH_PATHS paths)."""
        _lib.call("wl_pt_render_noalloc", self.h, C.byref(cfg), C.c_void_p(film.data_ptr()),
                  _stream_ptr(film.device))
        return film

    def trace_samples(self, cfg, frame, max_vertices=32):
        """Per-sample radiance of one frame plus each path's per-vertex record
        (H, W, max_vertices, 8): depth, mesh, pdf, mean f|cos|/pdf, mean
        throughput reaching the vertex, 0, 0, trans (zeros past its end)."""
        torch = _torch()
        out = torch.empty((cfg.height, cfg.width, 3), dtype=torch.float32, device="cuda")
        tr = torch.zeros((cfg.height, cfg.width, max_vertices, 8), dtype=torch.float32,
                         device="cuda")
        _lib.call("wl_pt_trace_samples", self.h, C.byref(cfg), frame, C.c_void_p(out.data_ptr()),
                  C.c_void_p(tr.data_ptr()), max_vertices, _stream_ptr(out.device))
        return out, tr

    def render_samples(self, cfg, frame, out=None):
        torch = _torch()
        if out is None:
            out = torch.empty((cfg.height, cfg.width, 3), dtype=torch.float32, device="cuda")
        _lib.call("wl_pt_render_samples", self.h, C.byref(cfg), frame,
                  C.c_void_p(out.data_ptr()), _stream_ptr(out.device))
        return out

    def launch_frame(self, cfg, frame, radiance_rgba, color_rgba):
        _lib.call("wl_pt_launch_frame", self.h, C.byref(cfg), frame,
                  C.c_void_p(radiance_rgba.data_ptr()), C.c_void_p(color_rgba.data_ptr()),
                  _stream_ptr(radiance_rgba.device))

    def resolve(self, film, nsamples, out=None):
        torch = _torch()
        npix = film.shape[0]
        if out is None:
            out = torch.empty((npix, 3), dtype=torch.float32, device=film.device)
        _lib.call("wl_film_resolve", self.ctx.h, C.c_void_p(film.data_ptr()), npix,
                  nsamples, C.c_void_p(out.data_ptr()), _stream_ptr(film.device))
        return out

    def guide_infer_hits(self, rays, hits):
        """The path tracer's guide inference on (n, 8) rays and their (n, 4)
        closest hits: (n, 48) decoded mixtures as a guided vertex stores
        them (see wl_guide_infer_hits)."""
        torch = _torch()
        n = rays.shape[0]
        out = torch.zeros((n, 48), dtype=torch.float32, device=rays.device)
        _lib.call("wl_guide_infer_hits", self.h, C.c_void_p(rays.data_ptr()),
                  C.c_void_p(hits.data_ptr()), n, C.c_void_p(out.data_ptr()),
                  _stream_ptr(rays.device))
        return out

    def set_guide(self, guide, mode=GUIDE_SAMPLE | GUIDE_TRAIN):
        """Attach a Guide (mode GUIDE_SAMPLE, optionally | GUIDE_TRAIN) or
        detach (guide None).  The guide must outlive the attachment."""
        _lib.call("wl_scene_set_guide", self.h, guide.h if guide is not None else None,
                  int(mode) if guide is not None else 0)
        self.guide = guide

    def close(self):
        if getattr(self, "h", None):
            _lib.load().wl_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
