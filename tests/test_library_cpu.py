"""CPU checks of the shipped library: it loads, exports every entry point the
C ABI header declares, and reports errors through the ABI (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
           if f.endswith(".h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(wl_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("wl_context_create", "wl_accel_build_triangles", "wl_trace_closest",
                 "wl_trace_any", "wl_pt_render", "wl_status_string"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from wayland_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_abi():
    from wayland_amd import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES), \
        set(declared_symbols()) - set(_lib.SIGNATURES)


def test_status_strings_and_null_arguments():
    from wayland_amd import _lib
    lib = _lib.load()
    assert lib.wl_status_string(0) == b"WL_SUCCESS"
    assert lib.wl_status_string(5) == b"WL_ERROR_LIMIT_EXCEEDED"
    # NULL handles are rejected with INVALID_VALUE and a message, not a crash
    assert lib.wl_context_destroy(None) == 1
    assert lib.wl_last_error_message()
    assert lib.wl_accel_destroy(None) == 1
    assert lib.wl_trace_closest(None, None, 0, None, None) == 1
    assert lib.wl_scene_destroy(None) == 1
    out = C.c_void_p()
    assert lib.wl_context_create(0, None) == 1


def test_no_device_reported_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from wayland_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    st = lib.wl_context_create(0, C.byref(h))
    assert st in (3, 6)  # WL_ERROR_DEVICE or WL_ERROR_NO_DEVICE
    assert not h.value
