"""GPU parity of the BVH8 traversal kernels (wl_trace_closest / wl_trace_any)
against the oracle's watertight traversal: closest-hit primitive id, t and
barycentrics must be bit-identical; any-hit occlusion must be identical."""
import numpy as np
import pytest

import oracle_bind as ob
from wayland_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu_hits(accel, rays_np):
    import torch
    import wayland_amd as wl
    rays = torch.from_numpy(np.ascontiguousarray(rays_np)).cuda()
    hits = accel.trace_closest(rays)
    bits = accel.trace_any(rays)
    torch.cuda.synchronize()
    return wl.hits_to_numpy(hits), wl.unpack_bits(bits, rays_np.shape[0])


def _assert_bit_exact(h, ref):
    assert np.array_equal(h["prim"], ref["prim"]), \
        f"prim mismatch at {np.nonzero(h['prim'] != ref['prim'])[0][:10]}"
    assert np.array_equal(h["t"].view(np.uint32), ref["t"].view(np.uint32))
    hit = ref["prim"] != 0xFFFFFFFF
    assert np.array_equal(h["b1"][hit].view(np.uint32), ref["b1"][hit].view(np.uint32))
    assert np.array_equal(h["b2"][hit].view(np.uint32), ref["b2"][hit].view(np.uint32))


def test_cornell_closest_and_any_bit_exact(gpu_ctx):
    import wayland_amd as wl
    print("case: cornell 36 tris, 65536 rays", flush=True)
    sc = scenes.cornell_box()
    accel = wl.Accel(gpu_ctx, sc)
    rays = scenes.traversal_rays(65536, seed=11)
    rays[:, 0:3] = rays[:, 0:3] * 2 - 1 + np.array([0, 1, 0], np.float32)
    h, occ = _gpu_hits(accel, rays)
    tris = sc.triangles()
    _assert_bit_exact(h, ob.closest_brute(tris, rays))
    assert np.array_equal(occ, ob.any_brute(tris, rays))


def test_icospheres_small_vs_bruteforce(gpu_ctx):
    import wayland_amd as wl
    print("case: 16 icospheres subdiv 3, 20000 rays, brute force", flush=True)
    sc = scenes.icosphere_scene(16, 3, seed=5)
    accel = wl.Accel(gpu_ctx, sc)
    rays = scenes.traversal_rays(20000, seed=12)
    h, occ = _gpu_hits(accel, rays)
    tris = sc.triangles()
    _assert_bit_exact(h, ob.closest_brute(tris, rays))
    assert np.array_equal(occ, ob.any_brute(tris, rays))


def test_edge_and_vertex_rays_watertight(gpu_ctx):
    """Rays from each sphere centre towards vertices and edge midpoints of its
    own (closed) mesh must all hit: a watertight test never lets a ray slip
    between triangles that share an edge or a vertex."""
    import wayland_amd as wl
    print("case: watertight vertex/edge rays", flush=True)
    soup = scenes.unit_icosphere_soup(3).astype(np.float32)  # closed, 1280 tris
    sc = scenes.Scene([scenes.soup_mesh(soup)])
    accel = wl.Accel(gpu_ctx, sc)
    tgt = np.concatenate([soup.reshape(-1, 3),
                          0.5 * (soup[:, 0] + soup[:, 1]), 0.5 * (soup[:, 1] + soup[:, 2])])
    d = tgt / np.linalg.norm(tgt, axis=1, keepdims=True)
    rays = np.zeros((d.shape[0], 8), np.float32)
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    h, occ = _gpu_hits(accel, rays)
    assert (h["prim"] != 0xFFFFFFFF).all()
    assert occ.all()
    _assert_bit_exact(h, ob.closest_brute(sc.triangles(), rays))


def test_tmin_tmax_and_empty(gpu_ctx):
    import torch
    import wayland_amd as wl
    print("case: tmin/tmax window + n=0", flush=True)
    sc = scenes.cornell_box()
    accel = wl.Accel(gpu_ctx, sc)
    rays = scenes.traversal_rays(4096, seed=13)
    rays[:, 0:3] = rays[:, 0:3] * 2 - 1 + np.array([0, 1, 0], np.float32)
    rays[:, 3] = 0.2
    rays[:, 7] = np.random.default_rng(0).uniform(0.3, 2.0, 4096).astype(np.float32)
    h, occ = _gpu_hits(accel, rays)
    tris = sc.triangles()
    _assert_bit_exact(h, ob.closest_brute(tris, rays))
    assert np.array_equal(occ, ob.any_brute(tris, rays))
    empty = torch.empty((0, 8), dtype=torch.float32, device="cuda")
    accel.trace_closest(empty)
    torch.cuda.synchronize()


def test_2m_scene_subsample_bit_exact(gpu_ctx):
    """Full-size traversal workload (2,048,000 triangles, 2^23 incoherent
    rays): every 64th ray is checked bit for bit against the oracle's BVH2
    traversal; the rest by size-independent properties."""
    import torch
    import wayland_amd as wl
    print("case: 25 icospheres subdiv 6 (2,048,000 tris), 8388608 rays", flush=True)
    sc = scenes.icosphere_scene(25, 6)
    accel = wl.Accel(gpu_ctx, sc)
    rays = scenes.traversal_rays(1 << 23, seed=2)
    h, occ = _gpu_hits(accel, rays)
    sub = slice(0, None, 64)
    obvh = ob.OracleBVH(sc.triangles())
    _assert_bit_exact(h[sub], obvh.closest(rays[sub]))
    assert np.array_equal(occ[sub], obvh.any(rays[sub]))
    # any-hit agrees with closest-hit everywhere
    assert np.array_equal(occ.astype(bool), h["prim"] != 0xFFFFFFFF)
    # hit distances are in range and hit ids are valid triangles
    hit = h["prim"] != 0xFFFFFFFF
    assert (h["prim"][hit] < sc.num_triangles).all()
    assert (h["t"][hit] >= 0).all() and np.isinf(h["t"][~hit]).all()
    del rays
    torch.cuda.empty_cache()


def test_64x7_scene_subsample_bit_exact(gpu_ctx):
    """The other point of the traversal workload: 64 icospheres at
    subdivision 7 (20,971,520 triangles), 2^23 incoherent rays; every 256th
    ray bit for bit against the oracle's BVH2, the rest by size-independent
    properties."""
    import torch
    import wayland_amd as wl
    print("case: 64 icospheres subdiv 7 (20,971,520 tris), 8388608 rays", flush=True)
    sc = scenes.icosphere_scene(64, 7)
    assert sc.num_triangles == 20_971_520
    accel = wl.Accel(gpu_ctx, sc)
    print(f"accel: {accel.info()}", flush=True)
    rays = scenes.traversal_rays(1 << 23, seed=2)
    h, occ = _gpu_hits(accel, rays)
    sub = slice(0, None, 256)
    obvh = ob.OracleBVH(sc.triangles())
    _assert_bit_exact(h[sub], obvh.closest(rays[sub]))
    assert np.array_equal(occ[sub], obvh.any(rays[sub]))
    assert np.array_equal(occ.astype(bool), h["prim"] != 0xFFFFFFFF)
    hit = h["prim"] != 0xFFFFFFFF
    assert (h["prim"][hit] < sc.num_triangles).all()
    assert (h["t"][hit] >= 0).all() and np.isinf(h["t"][~hit]).all()
    del rays, obvh
    accel.close()
    torch.cuda.empty_cache()


def test_concurrent_streams_and_sizes(gpu_ctx):
    """The raw trace kernels hand out chunks of rays at run time from counters
    kept per stream: traces of one accel launched on two streams at once, and
    launches of many sizes one after another on one stream (each advancing
    the counters by its own chunk count), give the hits of a lone trace,
    which are checked against the oracle on a subsample."""
    import torch
    import wayland_amd as wl
    print("case: 8 icospheres subdiv 4, two streams, sizes 1 .. 2^20", flush=True)
    sc = scenes.icosphere_scene(8, 4)
    accel = wl.Accel(gpu_ctx, sc)
    ra_np = scenes.traversal_rays(1 << 20, seed=5)
    rb_np = scenes.traversal_rays((3 << 18) + 77, seed=6)
    ha, oa = _gpu_hits(accel, ra_np)
    hb, ob_ = _gpu_hits(accel, rb_np)
    obvh = ob.OracleBVH(sc.triangles())
    _assert_bit_exact(ha[::97], obvh.closest(ra_np[::97]))
    ra = torch.from_numpy(ra_np).cuda()
    rb = torch.from_numpy(rb_np).cuda()
    main = torch.cuda.current_stream()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    s1.wait_stream(main)
    s2.wait_stream(main)
    outs = []
    for _ in range(4):
        with torch.cuda.stream(s1):
            outs.append(("a", accel.trace_closest(ra), accel.trace_any(ra)))
        with torch.cuda.stream(s2):
            outs.append(("b", accel.trace_closest(rb), accel.trace_any(rb)))
    torch.cuda.synchronize()
    for which, h, bits in outs:
        ref_h, ref_o, n = (ha, oa, ra.shape[0]) if which == "a" else (hb, ob_, rb.shape[0])
        _assert_bit_exact(wl.hits_to_numpy(h), ref_h)
        assert np.array_equal(wl.unpack_bits(bits, n), ref_o)
    # sizes below, at and across the chunk and XCD boundaries, back to back
    for n in (1, 7, 8, 255, 256, 257, 2047, 2048, 2049, 65537, 1 << 20):
        h = wl.hits_to_numpy(accel.trace_closest(ra[:n].contiguous()))
        o = wl.unpack_bits(accel.trace_any(ra[:n].contiguous()), n)
        _assert_bit_exact(h, ha[:n])
        assert np.array_equal(o, oa[:n])
    accel.close()
