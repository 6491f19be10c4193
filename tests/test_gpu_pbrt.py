"""Scenes read from pbrt-v3 files render on the GPU as the oracle renders them.

The Cornell box comes from the fixture tests/golden/cbox.pbrt.  The water
scene is built like scene/water-caustic/water.pbrt: the Cornell room plus a
transmissive Disney surface (eta 1.8, spectrans 1, roughness 0.001).  That
surface is a binary PLY written by the test and placed through a Transform
that scales y by 1.5.  Parity is per sample, on identical TEA seeds."""
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
    """The near-specular eta 1.8 surface makes long paths ill-conditioned: the
    FP32 oracle itself agrees with its FP64 build on only about 92 % of the
    samples at depth 16, and an ulp of difference between the GPU's and the
    host's sqrt/sin/log moves a refracted path onto another triangle.  So the
    bar is per sample where paths are short (depth 3) or never meet the water,
    and for the rest the GPU-minus-oracle differences must have zero mean
    (paired z-score over 16 frames)."""
    import wayland_amd as wl
    from wayland_amd import pbrt
    print("case: cbox + plymesh water (disney eta 1.8 spectrans 1) 64x64", flush=True)
    s = pbrt.load_pbrt(_water_scene(tmp_path), 64, 64)
    sc = s.scene
    assert sc.num_triangles == 36 + 2 * 47 * 47
    water = len(sc.meshes) - 1
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    opt = ob.OraclePT(sc)

    def both(depth, frame):
        g = pt.render_samples(wl.pt_config(64, 64, max_depth=depth, rr_depth=4), frame)
        o, _ = opt.render_frame(opt.config(64, 64, max_depth=depth, rr_depth=4), frame)
        return g.cpu().numpy(), o

    for frame in range(3):
        frac = _frac_within(*both(3, frame))
        print(f"depth 3 frame {frame}: within 1e-5 {frac:.5f}", flush=True)
        assert frac >= 0.999
    cfg = opt.config(64, 64, max_depth=16, rr_depth=4)
    diffs = []
    for frame in range(16):
        g, o = both(16, frame)
        assert np.isfinite(g).all()
        err = (np.abs(g - o) / np.maximum(np.abs(o), 1e-6)).reshape(-1, 3).max(1)
        frac = float((err <= REL).mean())
        diffs.append((g.astype(np.float64) - o).reshape(-1, 3))
        if frame < 3:
            dry = np.array([not (opt.trace(cfg, i % 64, i // 64, frame)[1][:, 1] == water).any()
                            for i in range(64 * 64)])
            dry_frac = float((err[dry] <= REL).mean())
            print(f"depth 16 frame {frame}: within 1e-5 {frac:.5f}, on the {dry.sum()} "
                  f"samples that never meet the water {dry_frac:.5f}", flush=True)
            assert dry_frac >= 0.999
        assert frac >= 0.97
    d = np.concatenate(diffs)
    z = d.mean(0) / np.maximum(d.std(0) / np.sqrt(len(d)), 1e-30)
    print(f"paired z of GPU - oracle over 16 frames: {np.round(z, 2)}")
    assert np.abs(z).max() < 4.0
    pt.close()
    accel.close()


def test_water_caustic_file_per_sample_parity(gpu_ctx):
    """The reference's scene file scene/water-caustic/water.pbrt (CC0, a
    fixture under tests/golden/water-caustic/, with its two binary PLY water
    surfaces: 10 meshes, 103,487 triangles) rendered per sample against the
    oracle at 64 x 64.  At depth 3 the bar is per sample.  At depth 16 the
    eta 1.8, roughness 0.001 water (which covers most of the view) makes
    long paths ill-conditioned, so every sample outside 1e-5 is replayed
    vertex by vertex against the oracle's trace (tests/pt_replay.py): each
    must be a branch flip after an agreeing prefix (a different surface,
    lobe, path end or shadow-ray outcome) or rounding below 1e-3.  A drift
    of the throughput along the same branches is allowed only after the
    path has met the water, whose near-specular refraction amplifies
    ulp-level direction differences (a slightly different point of the same
    wavy mesh); on every other path it fails.  The GPU-minus-oracle
    differences must have zero mean (paired z)."""
    import wayland_amd as wl
    from wayland_amd import pbrt
    import pt_replay as rp
    print("case: water.pbrt (103,487 tris, plymesh water) 64x64", flush=True)
    s = pbrt.load_pbrt(os.path.join(GOLDEN, "water-caustic", "water.pbrt"), 64, 64)
    sc = s.scene
    assert len(sc.meshes) == 10 and sc.num_triangles == 103487
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    opt = ob.OraclePT(sc)
    for frame in range(3):
        g = pt.render_samples(wl.pt_config(64, 64, max_depth=3, rr_depth=4), frame)
        o, _ = opt.render_frame(opt.config(64, 64, max_depth=3, rr_depth=4), frame)
        frac = _frac_within(g.cpu().numpy(), o)
        print(f"depth 3 frame {frame}: within 1e-5 {frac:.5f}", flush=True)
        assert frac >= 0.999
    cfg = wl.pt_config(64, 64, max_depth=16, rr_depth=4)
    ocfg = opt.config(64, 64, max_depth=16, rr_depth=4)
    kinds, total, d = rp.replay_frames(pt, opt, cfg, ocfg, range(16), 24)
    n_out = sum(len(v) for v in kinds.values())
    print(f"depth 16: outliers {n_out} of {total} samples ({n_out / total:.4f}): " +
          ", ".join(f"{k} {len(v)}" for k, v in sorted(kinds.items())))
    water = {len(sc.meshes) - 2, len(sc.meshes) - 1}
    dry_drift = [e for e in kinds.get("drift", []) if not water & set(e[4])]
    print(f"drifts after the water {len(kinds.get('drift', [])) - len(dry_drift)}, "
          f"elsewhere {len(dry_drift)}")
    assert not dry_drift, dry_drift[:10]
    assert n_out / total < 0.08
    z = rp.paired_z(d)
    print(f"paired z of GPU - oracle over 16 frames: {np.round(z, 2)}")
    assert np.abs(z).max() < 4.0
    pt.close()
    accel.close()
