"""ctypes binding of wayland_amd/libwl_hostemu.so: the library's own BVH8
builder and traversal code compiled for the CPU with index checks (tests
only)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "wayland_amd", "libwl_hostemu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import sys
            sys.path.insert(0, ROOT)
            from wayland_amd import build as b
            b.build_hostemu()
        L = C.CDLL(LIB)
        L.wlh_build.restype = C.c_int
        L.wlh_build.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        L.wlh_free.argtypes = [C.c_void_p]
        L.wlh_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
        L.wlh_trace.restype = C.c_uint64
        L.wlh_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int,
                                C.c_void_p, C.c_int]
        L.wlh_serve_check.restype = C.c_uint64
        L.wlh_serve_check.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.wlh_dyn_chunks.restype = C.c_uint32
        L.wlh_dyn_chunks.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        _lib = L
    return _lib


class HostBVH8:
    def __init__(self, tris):
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        h = C.c_void_p()
        err = C.create_string_buffer(256)
        rc = lib().wlh_build(tris.ctypes.data, tris.shape[0], C.byref(h), err, 256)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        self.h = h

    def info(self):
        out = (C.c_uint32 * 6)()
        lib().wlh_info(self.h, out)
        return dict(zip(["nodes", "tris", "depth", "inner", "leaves", "sah"], list(out)))

    def trace(self, rays, any_hit=False, nthreads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        counts = np.zeros((n, 2), np.uint32)
        faults = lib().wlh_trace(self.h, rays.ctypes.data, n, hits.ctypes.data,
                                 1 if any_hit else 0, counts.ctypes.data, nthreads)
        return hits, counts, int(faults)

    def __del__(self):
        try:
            lib().wlh_free(self.h)
        except Exception:
            pass


def serve_check(reqs, capacity):
    """Serving schedule of reqs ((width, height, spp, frame_offset) each) at
    `capacity` paths per megabatch, walked with the multi-request kernels'
    index arithmetic on the CPU: (violations, megabatches, items, paths,
    largest megabatch)."""
    a = np.ascontiguousarray(np.asarray(reqs, np.uint32).reshape(-1, 4))
    out = (C.c_uint64 * 4)()
    bad = lib().wlh_serve_check(a.ctypes.data, a.shape[0], capacity, out)
    return (int(bad),) + tuple(int(x) for x in out)
