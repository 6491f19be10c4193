"""Writes the converged FP64 reference images under tests/golden/ with the
CPU oracle (oracle/liboracle.so, the FP64 build of the PathTracing.cu
restatement): 65,536 samples per pixel, per-pixel mean and the variance of
one sample.

  cornell_256_fp64      Cornell box workload: 256 x 256, depth 8, NEE + MIS,
                        RR after bounce 5 (q 0.9), whole image
  caustic_crop128_fp64  glossy caustic (1024 x 1024 view, depth 16): the
                        128 x 128 window at x 448..575, y 416..543 (y counted
                        from the bottom row, as the film stores it), centred
                        on the caustic that the glass sphere focuses onto the
                        metal floor
  caustic_crop128_depth8_fp64
                        the same window at depth 8
  caustic_grid256_fp64  the whole glossy-caustic view at depth 16, sampled
                        on a 256 x 256 grid: pixels (1 + 4 i, 1 + 4 j)

Samples use frame ids 2^30 .. 2^30 + 65535, disjoint from every frame the GPU
tests and the benchmark render, so the reference is independent of them.

Each fixture is <name>.f32 (little-endian float32: mean[h][w][3] then
var[h][w][3]) with <name>.json (scene, view, window, configuration, sample
count, SHA-256 of the .f32).  Usage: python tests/golden/make_fp64_references.py
[name ...]; takes about 25 min (Cornell) and 5 min (caustic) on 8 cores.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

SPP = 65536
FRAME0 = 1 << 30

FIXTURES = {
    "cornell_256_fp64": dict(scene="cornell_box", width=256, height=256, window=(0, 0, 256, 256),
                             max_depth=8, rr_depth=5, rr_prob=0.9),
    "caustic_crop128_fp64": dict(scene="glossy_caustic", width=1024, height=1024,
                                 window=(448, 416, 128, 128), max_depth=16, rr_depth=5,
                                 rr_prob=0.9),
    "caustic_crop128_depth8_fp64": dict(scene="glossy_caustic", width=1024, height=1024,
                                        window=(448, 416, 128, 128), max_depth=8, rr_depth=5,
                                        rr_prob=0.9),
    "caustic_grid256_fp64": dict(scene="glossy_caustic", width=1024, height=1024,
                                 window=(1, 1, 256, 256), stride=4, max_depth=16, rr_depth=5,
                                 rr_prob=0.9),
}


def load(name: str):
    """(mean, var, meta) of a fixture: mean and var as float64 [h][w][3];
    None if the fixture files are absent."""
    base = os.path.join(HERE, name)
    if not (os.path.exists(base + ".f32") and os.path.exists(base + ".json")):
        return None
    with open(base + ".json") as f:
        meta = json.load(f)
    blob = open(base + ".f32", "rb").read()
    if hashlib.sha256(blob).hexdigest() != meta["sha256"]:
        raise ValueError(f"{name}.f32 does not match its recorded SHA-256")
    _, _, ww, wh = meta["window"]
    a = np.frombuffer(blob, dtype="<f4").astype(np.float64).reshape(2, wh, ww, 3)
    return a[0], a[1], meta


def expected_relmse(mean, var, spp: int, ref_spp: int = SPP) -> float:
    """E[relMSE] of an unbiased spp-sample estimate scored against the
    fixture: mean over pixel-channels of (var/spp + var/ref_spp)/(R^2 + 0.01)
    (the estimate's variance plus the fixture's own)."""
    return float(np.mean((var / spp + var / ref_spp) / (mean ** 2 + 0.01)))


def make(name: str, nthreads: int | None = None) -> dict:
    import oracle_bind as ob
    from wayland_amd import scenes
    spec = FIXTURES[name]
    sc = getattr(scenes, spec["scene"])(spec["width"], spec["height"])
    opt = ob.OraclePT(sc)
    cfg = opt.config(spec["width"], spec["height"], max_depth=spec["max_depth"],
                     rr_depth=spec["rr_depth"], rr_prob=spec["rr_prob"])
    x0, y0, ww, wh = spec["window"]
    stride = spec.get("stride", 1)
    t = time.time()
    mean, var = opt.render_f64_window(cfg, x0, y0, ww, wh, FRAME0, SPP,
                                      nthreads or os.cpu_count(), stride=stride)
    secs = time.time() - t
    blob = np.concatenate([mean.astype("<f4").ravel(), var.astype("<f4").ravel()]).tobytes()
    with open(os.path.join(HERE, name + ".f32"), "wb") as f:
        f.write(blob)
    meta = dict(spec, name=name, spp=SPP, frame0=FRAME0, strategy="MIS", epsilon=1e-3,
                layout="float32 LE mean[h][w][3] then var[h][w][3], row 0 = bottom row y0",
                sha256=hashlib.sha256(blob).hexdigest(), seconds=round(secs, 1),
                threads=nthreads or os.cpu_count())
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    return meta


if __name__ == "__main__":
    for n in sys.argv[1:] or list(FIXTURES):
        print(make(n), flush=True)
