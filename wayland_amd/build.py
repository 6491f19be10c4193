"""Builds wayland_amd/libwayland_hip.so for gfx950 with hipcc (in tree).

Every HIP source is compiled with
  -ffp-contract=off   no fused multiply-add unless written as fmaf(), so the
                      watertight triangle test and the BSDF arithmetic round
                      exactly like the host restatement;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libwayland_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WL_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "bvh_builder.cpp",
    "wl_api.cpp",
    "wl_pt_api.cpp",
    "wl_program.cpp",
    "wl_guide_api.cpp",
    "wl_serve.cpp",
    "traverse.hip",
    "guide.hip",
    "pathtrace.hip",
    "example_pt.hip",
    "wl_pt_noalloc.hip",
]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-c", path, "-o", obj] + COMMON + HIP_FLAGS
    else:
        cmd = [HIPCC, "-x", "c++", "-c", path, "-o", obj] + COMMON + [
            "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_hostemu()
    build_cpp_examples()
    if verbose:
        print("built", LIB, "and", HOSTEMU)
    return LIB


CPP_EXAMPLE = os.path.join(ROOT, "tests", "cpp", "example_pathtracer")
CPP_EXAMPLE_LIB = os.path.join(ROOT, "tests", "cpp", "libexample_pathtracer.so")


def build_cpp_examples() -> str:
    """The examples/PathTracer driver over include/wayland_hip.hpp (tests only):
    host C++20, linked against the library with an rpath to it."""
    src = CPP_EXAMPLE + ".cpp"
    if not os.path.exists(src):
        return ""
    common = ["-x", "c++", "-std=c++20", "-O2", src, "-D__HIP_PLATFORM_AMD__",
              "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
              "-L" + PKG, "-lwayland_hip", "-L/opt/rocm/lib", "-lamdhip64",
              "-Wl,-rpath,$ORIGIN/../../wayland_amd"]
    for out, extra in ((CPP_EXAMPLE, ["-DWL_EXAMPLE_EXECUTABLE"]),
                       (CPP_EXAMPLE_LIB, ["-shared", "-fPIC"])):
        tmp = out + ".tmp"
        cmd = [HIPCC, "-o", tmp] + extra + common
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, out)
    return CPP_EXAMPLE


HOSTEMU = os.path.join(PKG, "libwl_hostemu.so")


def build_hostemu() -> str:
    """CPU emulation of the builder + traversal (tests only)."""
    tmp = HOSTEMU + ".tmp"
    cmd = [HIPCC, "-x", "c++", "-shared", "-o", tmp,
           os.path.join(CSRC, "hostemu.cpp"), os.path.join(CSRC, "bvh_builder.cpp"),
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-lpthread"] + COMMON
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hostemu build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, HOSTEMU)
    return HOSTEMU


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
