"""ctypes binding of libwayland_hip.so (the C ABI in include/wayland_hip.h).

Loading fails loudly if the library has not been built: there is no CPU
fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# WL_LIBRARY_PATH selects another build of the same ABI (e.g. a tuning build)
LIB_PATH = os.environ.get("WL_LIBRARY_PATH") or os.path.join(PKG, "libwayland_hip.so")

WL_INVALID_INDEX = 0xFFFFFFFF


class WlError(RuntimeError):
    def __init__(self, status, fn, message):
        self.status = status
        super().__init__(f"{fn} failed: {status_name(status)}: {message}")


class TriInput(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint32),
                ("indices", C.c_void_p), ("num_triangles", C.c_uint32),
                ("geometry_flags", C.c_uint32), ("num_sbt_records", C.c_uint32)]


class AccelInfo(C.Structure):
    _fields_ = [("num_triangles", C.c_uint64), ("num_inputs", C.c_uint32),
                ("num_nodes", C.c_uint32), ("inner_nodes", C.c_uint32),
                ("leaves", C.c_uint32), ("depth", C.c_uint32), ("pad_", C.c_uint32),
                ("device_bytes", C.c_uint64), ("sah_cost", C.c_double),
                ("build_seconds", C.c_double)]


class DisneyMaterial(C.Structure):
    _fields_ = [("values", C.c_float * 16)]


class PinholeCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("lookat", C.c_float * 3),
                ("up", C.c_float * 3), ("right", C.c_float * 3)]


class MeshShading(C.Structure):
    _fields_ = [("normals", C.c_void_p), ("material", DisneyMaterial),
                ("area_light", C.c_uint32), ("pad_", C.c_uint32 * 3)]


class TextureImage(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class MeshTexture(C.Structure):
    _fields_ = [("texcoords", C.c_void_p), ("texture", C.c_uint32), ("pad_", C.c_uint32)]


class AreaLight(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("L", C.c_float * 3), ("two_sided", C.c_uint32)]


class PTConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("frame_offset", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("rr_prob", C.c_float), ("epsilon", C.c_float),
                ("strategy", C.c_uint32), ("flags", C.c_uint32),
                ("max_batch_paths", C.c_uint32), ("pad_", C.c_uint32)]


class PTStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("max_bounce", C.c_uint32),
                ("pad_", C.c_uint32), ("seconds", C.c_double)]


class GuideConfig(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("batch_size", C.c_uint32),
                ("steps_per_batch", C.c_uint32), ("max_samples", C.c_uint32),
                ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("epsilon", C.c_float), ("mix_alpha", C.c_float), ("max_weight", C.c_float),
                ("guided_bounces", C.c_uint32), ("guided_vertices", C.c_uint32)]


class GuideInfo(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("last_samples", C.c_uint32), ("pad_", C.c_uint32)]


class RenderRequest(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("film", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("spp", C.c_uint32), ("frame_offset", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("rr_prob", C.c_float),
                ("epsilon", C.c_float), ("strategy", C.c_uint32), ("pad_", C.c_uint32)]


class ServeConfig(C.Structure):
    _fields_ = [("batch_paths", C.c_uint32), ("shade_blocks_per_cu", C.c_uint32),
                ("use_graph", C.c_uint32), ("lanes", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("paths", C.c_uint64),
                ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("batches", C.c_uint32), ("launches", C.c_uint32)]


class ProgramGroup(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("entry", (C.c_char * 64) * 3)]


class ExampleCamera(C.Structure):
    _fields_ = [("horizontalRatio", C.c_float), ("verticalRatio", C.c_float),
                ("position", C.c_float * 3), ("gaze", C.c_float * 3), ("up", C.c_float * 3)]


class ExampleParams(C.Structure):
    _fields_ = [("sampleNum", C.c_uint32), ("maxDepth", C.c_uint32),
                ("camera", ExampleCamera), ("pad_", C.c_uint32),
                ("colorBuffer", C.c_void_p), ("traversable", C.c_uint64)]


class ExampleHitData(C.Structure):
    _fields_ = [("color", C.c_float * 3), ("emission", C.c_float * 3),
                ("indices", C.c_void_p), ("normals", C.c_void_p)]


class PathTracingParams(C.Structure):
    _fields_ = [("frameID", C.c_uint32), ("fbSize", C.c_int32 * 2), ("pad0_", C.c_uint32),
                ("radianceBuffer", C.c_void_p), ("colorBuffer", C.c_void_p),
                ("camera", PinholeCamera), ("areaLightCount", C.c_uint32),
                ("pad1_", C.c_uint32), ("areaLights", C.c_void_p), ("traversable", C.c_uint64)]


class DeviceAreaLight(C.Structure):
    _fields_ = [("twoSided", C.c_uint8), ("hasNormal", C.c_uint8), ("pad0_", C.c_uint8 * 2),
                ("triangleNum", C.c_uint32), ("L", C.c_float * 3), ("pad1_", C.c_uint32),
                ("vertices", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p)]


class PathTracingHitData(C.Structure):
    _fields_ = [("hasTexture", C.c_uint8), ("hasNormal", C.c_uint8), ("pad0_", C.c_uint8 * 6),
                ("texture", C.c_uint64), ("meshID", C.c_uint32), ("areaLightID", C.c_uint32),
                ("materialID", C.c_uint32), ("disneyMat", C.c_float * 15),
                ("indices", C.c_void_p), ("vertices", C.c_void_p), ("normals", C.c_void_p),
                ("texcoords", C.c_void_p)]


SBT_HEADER = 32


def record_stride(data_size):
    return (SBT_HEADER + data_size + 15) // 16 * 16


# name -> (restype, argtypes); every function declared in include/wayland_hip.h
_VP = C.c_void_p
_U32 = C.c_uint32
SIGNATURES = {
    "wl_status_string": (C.c_char_p, [C.c_int]),
    "wl_last_error_message": (C.c_char_p, []),
    "wl_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "wl_context_create": (C.c_int, [C.c_int, C.POINTER(_VP)]),
    "wl_context_destroy": (C.c_int, [_VP]),
    "wl_context_get_stream": (C.c_int, [_VP, C.POINTER(_VP)]),
    "wl_context_synchronize": (C.c_int, [_VP]),
    "wl_set_cache_path": (C.c_int, [_VP, C.c_char_p]),
    "wl_accel_build_triangles": (C.c_int, [_VP, C.POINTER(TriInput), _U32, _U32, C.POINTER(_VP)]),
    "wl_accel_destroy": (C.c_int, [_VP]),
    "wl_accel_get_info": (C.c_int, [_VP, C.POINTER(AccelInfo)]),
    "wl_accel_depth": (C.c_int, [_VP, C.POINTER(_U32)]),
    "wl_accel_input_offsets": (C.c_int, [_VP, C.POINTER(_U32)]),
    "wl_trace_closest": (C.c_int, [_VP, _VP, _U32, _VP, _VP]),
    "wl_trace_any": (C.c_int, [_VP, _VP, _U32, _VP, _VP]),
    "wl_trace_closest_counted": (C.c_int, [_VP, _VP, _U32, _VP, _VP, _VP]),
    "wl_scene_create": (C.c_int, [_VP, _VP, C.POINTER(MeshShading), _U32,
                                  C.POINTER(AreaLight), _U32, C.POINTER(PinholeCamera),
                                  C.POINTER(C.c_float), C.POINTER(_VP)]),
    "wl_scene_destroy": (C.c_int, [_VP]),
    "wl_scene_set_camera": (C.c_int, [_VP, C.POINTER(PinholeCamera)]),
    "wl_scene_set_textures": (C.c_int, [_VP, C.POINTER(TextureImage), _U32,
                                        C.POINTER(MeshTexture), _U32]),
    "wl_pt_render": (C.c_int, [_VP, C.POINTER(PTConfig), _VP, C.POINTER(PTStats), _VP]),
    "wl_pt_render_noalloc": (C.c_int, [_VP, C.POINTER(PTConfig), _VP, _VP]),
    "wl_guide_infer_hits": (C.c_int, [_VP, _VP, _VP, C.c_uint32, _VP, _VP]),
    "wl_pt_trace_samples": (C.c_int, [_VP, C.POINTER(PTConfig), C.c_uint32, _VP, _VP,
                                      C.c_uint32, _VP]),
    "wl_pt_render_samples": (C.c_int, [_VP, C.POINTER(PTConfig), _U32, _VP, _VP]),
    "wl_pt_launch_frame": (C.c_int, [_VP, C.POINTER(PTConfig), _U32, _VP, _VP, _VP]),
    "wl_film_resolve": (C.c_int, [_VP, _VP, _U32, _U32, _VP, _VP]),
    "wl_guide_default_config": (C.c_int, [C.POINTER(GuideConfig)]),
    "wl_guide_create": (C.c_int, [_VP, C.POINTER(GuideConfig), C.POINTER(_VP)]),
    "wl_guide_destroy": (C.c_int, [_VP]),
    "wl_guide_set_params": (C.c_int, [_VP, C.POINTER(C.c_float), _U32]),
    "wl_guide_get_params": (C.c_int, [_VP, C.POINTER(C.c_float), _U32]),
    "wl_guide_set_mix": (C.c_int, [_VP, C.c_float]),
    "wl_guide_get_info": (C.c_int, [_VP, C.POINTER(GuideInfo)]),
    "wl_guide_infer": (C.c_int, [_VP, _VP, _U32, _VP, _VP]),
    "wl_guide_mixture_sample": (C.c_int, [_VP, _VP, _VP, _U32, _VP, _VP, _VP]),
    "wl_guide_mixture_pdf": (C.c_int, [_VP, _VP, _VP, _U32, _VP, _VP]),
    "wl_guide_gradient": (C.c_int, [_VP, _VP, _U32, _VP, _VP, _VP]),
    "wl_guide_train": (C.c_int, [_VP, _VP, _U32, _U32, _VP]),
    "wl_scene_set_guide": (C.c_int, [_VP, _VP, _U32]),
    "wl_serve_default_config": (C.c_int, [C.POINTER(ServeConfig)]),
    "wl_server_create": (C.c_int, [_VP, C.POINTER(ServeConfig), C.POINTER(_VP)]),
    "wl_server_destroy": (C.c_int, [_VP]),
    "wl_server_get_config": (C.c_int, [_VP, C.POINTER(ServeConfig)]),
    "wl_server_set_config": (C.c_int, [_VP, C.POINTER(ServeConfig)]),
    "wl_server_render": (C.c_int, [_VP, C.POINTER(RenderRequest), _U32, C.POINTER(ServeStats),
                                   C.POINTER(C.c_double), _VP]),
    "wl_server_autotune": (C.c_int, [_VP, C.POINTER(RenderRequest), _U32, _U32,
                                     C.POINTER(ServeConfig)]),
    "wl_identify_programs": (C.c_int, [C.c_char_p, C.c_size_t, _VP, _U32, C.POINTER(_U32)]),
    "wl_pipeline_create": (C.c_int, [_VP, _VP, _U32, _U32, C.POINTER(_VP), C.POINTER(_U32)]),
    "wl_pipeline_destroy": (C.c_int, [_VP]),
    "wl_sbt_pack_header": (C.c_int, [_VP, _U32, _VP]),
    "wl_sbt_create": (C.c_int, [_VP, _VP, C.c_size_t, _VP, _U32, C.c_size_t, _VP, _U32,
                                C.c_size_t, C.POINTER(_VP)]),
    "wl_sbt_destroy": (C.c_int, [_VP]),
    "wl_texture_create": (C.c_int, [_VP, C.POINTER(TextureImage), C.POINTER(C.c_uint64)]),
    "wl_texture_destroy": (C.c_int, [C.c_uint64]),
    "wl_accel_traversable": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "wl_launch": (C.c_int, [_VP, _VP, _VP, C.c_size_t, _VP, _U32, _U32, _U32]),
    "wl_float_to_u8": (C.c_int, [_VP, C.c_size_t, _VP, _VP]),
}

_lib = None


def load():
    """Load the library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "or python wayland_amd/build.py")
    # torch ships its own HIP runtime under the same soname.  Loading torch
    # first makes the library bind to that one runtime; loaded the other way
    # round, the process holds two runtimes and torch then sees no device
    # (and tensor pointers would belong to the other runtime).
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def status_name(s):
    try:
        return load().wl_status_string(s).decode()
    except Exception:
        return str(s)


def check(status, fn):
    if status != 0:
        msg = load().wl_last_error_message()
        raise WlError(status, fn, msg.decode() if msg else "")
    return status


def call(name, *args):
    fn = getattr(load(), name)
    return check(fn(*args), name)
