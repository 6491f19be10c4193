"""The C++ face of the ABI (include/wayland_hip.hpp) on the CPU: the header
compiles on its own under C++20, and the examples/PathTracer driver built on
it exists and rejects a bad command line without touching a device."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "example_pathtracer")
LIB = os.path.join(ROOT, "tests", "cpp", "libexample_pathtracer.so")


def test_header_compiles_alone():
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-x", "c++",
                        os.path.join(ROOT, "include", "wayland_hip.hpp"),
                        "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
                        "-D__HIP_PLATFORM_AMD__"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ example not built")
def test_example_usage_error():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.skipif(not os.path.exists(LIB), reason="C++ example not built")
def test_example_library_exports_entry():
    import torch  # noqa: F401  (one HIP runtime in the process)
    lib = ctypes.CDLL(LIB)
    assert hasattr(lib, "example_pathtracer_main")
