"""CPU checks of the library's own BVH8 builder and traversal code (compiled
for the host from the same traverse_device.h the gfx950 kernels use, with
every node / triangle / stack index checked against its allocation)."""
import numpy as np
import pytest

import hostemu_bind as hb
import oracle_bind as ob
from wayland_amd import scenes


def _check(tris, rays):
    h = hb.HostBVH8(tris)
    hits, counts, faults = h.trace(rays)
    assert faults == 0
    ref = ob.closest_brute(tris, rays) if tris.shape[0] * rays.shape[0] < 4e8 \
        else ob.OracleBVH(tris).closest(rays)
    assert np.array_equal(hits[:, 1].view(np.uint32), ref["prim"])
    assert np.array_equal(hits[:, 0].view(np.uint32), ref["t"].view(np.uint32))
    hit = ref["prim"] != 0xFFFFFFFF
    assert np.array_equal(hits[hit, 2].view(np.uint32), ref["b1"][hit].view(np.uint32))
    assert np.array_equal(hits[hit, 3].view(np.uint32), ref["b2"][hit].view(np.uint32))
    ha, _, fa = h.trace(rays, any_hit=True)
    assert fa == 0
    occ = ha[:, 1].view(np.uint32) != 0xFFFFFFFF
    assert np.array_equal(occ, hit)
    return h, counts


def test_cornell():
    sc = scenes.cornell_box()
    rays = scenes.traversal_rays(20000, seed=3)
    rays[:, 0:3] = rays[:, 0:3] * 2 - 1 + np.array([0, 1, 0], np.float32)
    _check(sc.triangles(), rays)


def test_icospheres():
    sc = scenes.icosphere_scene(10, 4, seed=4)
    _check(sc.triangles(), scenes.traversal_rays(20000, seed=4))


def test_single_triangle_and_tiny_scenes():
    for n in (1, 2, 3, 4, 5, 9, 33):
        tris = scenes.icosphere_scene(1, 2, seed=n).triangles()[:n]
        rays = scenes.traversal_rays(2000, seed=n)
        _check(tris, rays)


def test_empty_scene():
    tris = np.zeros((0, 9), np.float32)
    h = hb.HostBVH8(tris)
    hits, _, faults = h.trace(scenes.traversal_rays(100, seed=1))
    assert faults == 0 and (hits[:, 1].view(np.uint32) == 0xFFFFFFFF).all()


def test_degenerate_and_coincident_triangles():
    rng = np.random.default_rng(5)
    base = rng.uniform(0, 1, (200, 9)).astype(np.float32)
    dup = np.repeat(base[:20], 40, axis=0)          # 800 coincident copies
    line = base[:50].copy()
    line[:, 6:9] = line[:, 0:3]                     # zero-area triangles
    flat = base[:100].copy()
    flat[:, 1::3] = 0.5                             # all in the plane y = 0.5
    tris = np.concatenate([base, dup, line, flat])
    rays = scenes.traversal_rays(5000, seed=6)
    _check(tris, rays)


def test_large_scene_depth_and_subsample():
    sc = scenes.icosphere_scene(25, 6)
    tris = sc.triangles()
    rays = scenes.traversal_rays(1 << 16, seed=7)
    h, counts = _check(tris, rays)
    info = h.info()
    assert info["depth"] <= 32
    assert counts[:, 0].mean() > 1
