"""Scenes read from pbrt-v3 files render on the GPU as the oracle renders them.

The Cornell box comes from the fixture tests/golden/cbox.pbrt.  The water
scene is built like scene/water-caustic/water.pbrt: the Cornell room plus a
transmissive Disney surface (eta 1.8, spectrans 1, roughness 0.001).  That
surface is a binary PLY written by the test and placed through a Transform
that scales y by 1.5.  Parity is per sample, on identical TEA seeds.  The
bounds on the fraction of samples within 1e-5 are those of the GPU path
tracer tests."""
import os
import struct

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL = 1e-5


def _frac_within(a, b, rel=REL, floor=1e-6):
    err = np.abs(a - b) / np.maximum(np.abs(b), floor)
    return float((err.reshape(-1, 3).max(1) <= rel).mean())


def _water_ply(path, n=48):
    """Wavy height field over [-1, 1]^2 at y about 0.75, with vertex normals."""
    xs = np.linspace(-1, 1, n, dtype=np.float64)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = 0.75 + 0.03 * np.sin(7 * X + 1) * np.cos(5 * Z)
    dYdx = 0.03 * 7 * np.cos(7 * X + 1) * np.cos(5 * Z)
    dYdz = -0.03 * 5 * np.sin(7 * X + 1) * np.sin(5 * Z)
    N = np.stack([-dYdx, np.ones_like(X), -dYdz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    V = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    N = N.reshape(-1, 3).astype(np.float32)
    faces = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, i * n + j + 1, (i + 1) * n + j + 1, (i + 1) * n + j
            faces.append((a, b, c, d))
    hdr = ("ply\nformat binary_little_endian 1.0\n"
           f"element vertex {len(V)}\nproperty float x\nproperty float y\nproperty float z\n"
           "property float nx\nproperty float ny\nproperty float nz\n"
           f"element face {len(faces)}\nproperty list uchar int vertex_indices\nend_header\n")
    with open(path, "wb") as f:
        f.write(hdr.encode())
        f.write(np.concatenate([V, N], 1).astype("<f4").tobytes())
        for q in faces:
            f.write(struct.pack("<B4i", 4, *q))


def _water_scene(tmp_path):
    src = open(os.path.join(GOLDEN, "cbox.pbrt")).read()
    world_end = src.rindex("WorldEnd")
    os.makedirs(tmp_path / "models", exist_ok=True)
    _water_ply(str(tmp_path / "models" / "water.ply"))
    water = """
MakeNamedMaterial "Water" "string type" [ "disney" ] "float eta" [ 1.8 ]
    "float spectrans" [ 1.0 ] "float roughness" [ 0.001 ]
NamedMaterial "Water"
AttributeBegin
    Transform [ 1 0 0 0 0 1.5 0 0 0 0 1 0 0 0 0 1 ]
    Shape "plymesh" "string filename" [ "models/water.ply" ]
AttributeEnd
"""
    path = tmp_path / "water.pbrt"
    path.write_text(src[:world_end] + water + src[world_end:])
    return str(path)


def _parity(ctx, sc, w, h, frames, max_depth=25, rr_depth=5):
    import wayland_amd as wl
    accel = wl.Accel(ctx, sc)
    pt = wl.PathTracer(ctx, accel, sc)
    opt = ob.OraclePT(sc)
    out = []
    for frame in frames:
        cfg = wl.pt_config(w, h, max_depth=max_depth, rr_depth=rr_depth)
        g = pt.render_samples(cfg, frame).cpu().numpy()
        o, _ = opt.render_frame(opt.config(w, h, max_depth=max_depth, rr_depth=rr_depth), frame)
        assert np.isfinite(g).all()
        out.append(_frac_within(g, o))
        print(f"frame {frame}: within 1e-5 {out[-1]:.5f}", flush=True)
    pt.close()
    accel.close()
    return out


def test_cbox_file_per_sample_parity(gpu_ctx):
    from wayland_amd import pbrt
    print("case: cbox.pbrt 64x64 frames 0-2", flush=True)
    s = pbrt.load_pbrt(os.path.join(GOLDEN, "cbox.pbrt"), 64, 64)
    fr = _parity(gpu_ctx, s.scene, 64, 64, range(3))
    assert min(fr) >= 0.999


def test_plymesh_water_per_sample_parity(gpu_ctx, tmp_path):
    from wayland_amd import pbrt
    print("case: cbox + plymesh water (disney eta 1.8 spectrans 1) 64x64 frames 0-2",
          flush=True)
    s = pbrt.load_pbrt(_water_scene(tmp_path), 64, 64)
    assert s.scene.num_triangles == 36 + 2 * 47 * 47
    fr = _parity(gpu_ctx, s.scene, 64, 64, range(3), max_depth=16, rr_depth=4)
    assert min(fr) >= 0.995
