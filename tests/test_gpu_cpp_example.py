"""examples/PathTracer/main.cpp's flow in C++ over include/wayland_hip.hpp
(tests/cpp/example_pathtracer.cpp), run in-process through its exported
entry point on the reference's cornell.obj: the colour buffer matches the
oracle's restatement of PathTracer.cu as the Python launch test does, and
the bytes equal FromFloatToChar of it.  (The example's main also calls the
denoiser, which throws "not supported"; it prints that and goes on.)"""
import ctypes
import os
import struct

import numpy as np
import pytest

import oracle_bind as ob
from wayland_amd import obj, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libexample_pathtracer.so")
SOURCE = r'''
extern "C" __constant__ LaunchParam param;
extern "C" __global__ void __raygen__RenderFrame() { }
extern "C" __global__ void __miss__Empty() { }
extern "C" __global__ void __closesthit__PT() { }
'''


def _dump(path, meshes, mats):
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(meshes)))
        for m in meshes:
            f.write(struct.pack("<II", m.vertices.shape[0], m.indices.shape[0]))
            f.write(np.asarray(mats[m.material].kd, "<f4").tobytes())
            f.write(np.asarray(mats[m.material].ke, "<f4").tobytes())
            f.write(np.ascontiguousarray(m.vertices, "<f4").tobytes())
            f.write(np.ascontiguousarray(m.normals, "<f4").tobytes())
            f.write(np.ascontiguousarray(m.indices, "<u4").tobytes())


def test_cpp_example_matches_oracle(gpu_ctx, tmp_path):
    print("case: C++ examples/PathTracer driver, cornell.obj 48x48, 8 spp, maxDepth 9",
          flush=True)
    meshes, mats = obj.load_obj(os.path.join(ROOT, "tests", "golden", "cornell.obj"))
    _dump(str(tmp_path / "meshes.bin"), meshes, mats)
    (tmp_path / "PathTracer.cu").write_text(SOURCE)
    W = H = 48
    spp, depth = 8, 9
    out = tmp_path / "out.bin"
    lib = ctypes.CDLL(LIB)
    args = [b"example", str(tmp_path / "meshes.bin").encode(), str(tmp_path / "PathTracer.cu").encode(),
            str(W).encode(), str(H).encode(), str(spp).encode(), str(depth).encode(),
            str(out).encode()]
    argv = (ctypes.c_char_p * len(args))(*args)
    rc = lib.example_pathtracer_main(len(args), argv)
    assert rc == 0
    raw = out.read_bytes()
    g = np.frombuffer(raw[:16 * W * H], np.float32).reshape(W * H, 4)
    u8 = np.frombuffer(raw[16 * W * H:], np.uint8)
    sc = scenes.Scene([scenes.Mesh(m.vertices, m.indices, m.normals) for m in meshes])
    opt = ob.OraclePT(sc)
    v = float(np.tan(np.float32(22.5)))
    ec = ob.ExampleConfig(W, H, spp, depth)
    ec.cam_pos[:] = [-0.23, 2.58, 5.0]
    ec.cam_gaze[:] = [0.0, 0.0, -1.0]
    ec.cam_up[:] = [0.0, 1.0, 0.0]
    ec.horizontal_ratio, ec.vertical_ratio = v, v
    mc = np.stack([mats[m.material].kd for m in meshes])
    me = np.stack([mats[m.material].ke for m in meshes])
    o = opt.example_render(mc, me, ec, 0, W)
    err = np.abs(g - o) / np.maximum(np.abs(o), 1e-4)
    frac = float((err[:, :3].max(1) <= 1e-5).mean())
    exact = float((g.view(np.uint32) == o.view(np.uint32)).all(-1).mean())
    print(f"pixels within 1e-5: {frac:.5f}, bit-exact {exact:.5f}")
    assert frac >= 0.99 and exact >= 0.97
    assert np.array_equal(u8, (np.minimum(g.reshape(-1), 1.0) * 255).astype(np.uint8))
