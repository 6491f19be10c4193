"""Batched serving on the GPU: requests of different scenes, sizes, sample
counts and depths share megabatches (replayed from one captured hipGraph),
and each request's film must equal its standalone wl_pt_render bit for bit
(sum and Kahan compensation)."""
import numpy as np
import pytest

from wayland_amd import scenes

pytestmark = pytest.mark.gpu


def _tracers(ctx, res):
    import wayland_amd as wl
    out = []
    for sc in (scenes.cornell_box(res, res, seed=0), scenes.cornell_box(res, res, seed=3),
               scenes.wavefront_scene(4, 2, width=res, height=res)):
        accel = wl.Accel(ctx, sc)
        out.append((wl.PathTracer(ctx, accel, sc), accel))
    return out


def _requests(tracers, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(9):
        res = int(rng.choice([16, 32, 48, 64]))
        reqs.append({"pt": tracers[k % 3][0], "width": res, "height": res,
                     "spp": int(rng.integers(1, 10)), "max_depth": int(rng.integers(4, 17)),
                     "frame_offset": int(rng.integers(0, 1000)),
                     "rr_depth": int(rng.integers(2, 6))})
    return reqs


def _standalone(r):
    import wayland_amd as wl
    cfg = wl.pt_config(r["width"], r["height"], spp=r["spp"], frame_offset=r["frame_offset"],
                       max_depth=r["max_depth"], rr_depth=r["rr_depth"])
    film, _ = r["pt"].render(cfg, stats=False)
    return film.cpu().numpy()


@pytest.mark.parametrize("use_graph,lanes", [(1, 2), (0, 2), (1, 1)])
def test_serve_matches_standalone_bitwise(gpu_ctx, use_graph, lanes):
    # lanes = 2: the requests are split into two groups whose megabatches run
    # concurrently on two streams
    import wayland_amd as wl
    print(f"case: serve 9 requests, batch 2^13 paths, graph {use_graph}, lanes {lanes}",
          flush=True)
    tracers = _tracers(gpu_ctx, 64)
    reqs = _requests(tracers, 7)
    server = wl.Server(gpu_ctx, batch_paths=1 << 13, use_graph=use_graph, lanes=lanes)
    films, stats, done = server.render(reqs)
    print(f"batches {stats['batches']}, paths {stats['paths']}, seconds {stats['seconds']:.4f}")
    assert stats["paths"] == sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    assert stats["batches"] > 1
    assert all(d > 0 for d in done)
    for k, r in enumerate(reqs):
        ref = _standalone(r)
        got = films[k].cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"request {k}"
    server.close()


def test_serve_graph_replay_twice(gpu_ctx):
    # the second call replays the captured graph (same key) with new films
    import wayland_amd as wl
    print("case: serve twice with one graph", flush=True)
    tracers = _tracers(gpu_ctx, 32)
    reqs = _requests(tracers, 11)
    server = wl.Server(gpu_ctx, batch_paths=1 << 12)
    f1, s1, _ = server.render(reqs)
    f2, s2, _ = server.render(reqs)
    for a, b in zip(f1, f2):
        assert np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))
    assert s1["extension_rays"] == s2["extension_rays"]
    server.close()


def test_serve_arena_refilled_per_call(gpu_ctx):
    # the BVH arena is rebuilt on every call: a second call with another set
    # and order of scenes (other roots at the same arena addresses, replayed
    # graph) must still match the standalone renders bit for bit
    import wayland_amd as wl
    print("case: serve, second call with other scenes in the arena", flush=True)
    tracers = _tracers(gpu_ctx, 32)
    reqs = _requests(tracers, 17)
    server = wl.Server(gpu_ctx, batch_paths=1 << 12)
    server.render(reqs)
    second = [r for r in reqs if r["pt"] is not tracers[0][0]][::-1]
    films, _, _ = server.render(second)
    for k, r in enumerate(second):
        ref = _standalone(r)
        assert np.array_equal(films[k].cpu().numpy().view(np.uint32), ref.view(np.uint32)), \
            f"request {k}"
    server.close()


def test_serve_autotune_then_render(gpu_ctx):
    import wayland_amd as wl
    print("case: serve autotune", flush=True)
    tracers = _tracers(gpu_ctx, 64)
    reqs = _requests(tracers, 13)
    server = wl.Server(gpu_ctx)
    cfg = server.autotune(reqs, max_spp=2)
    print(f"chosen {cfg}")
    assert cfg["batch_paths"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)
    assert cfg["shade_blocks_per_cu"] in (4, 8, 16)
    assert cfg["lanes"] in (1, 2)
    films, _, _ = server.render(reqs[:3])
    for k in range(3):
        ref = _standalone(reqs[k])
        assert np.array_equal(films[k].cpu().numpy().view(np.uint32), ref.view(np.uint32))
    server.close()


def test_serve_rejects_oversized_request(gpu_ctx):
    import wayland_amd as wl
    tracers = _tracers(gpu_ctx, 64)
    server = wl.Server(gpu_ctx, batch_paths=1 << 10)
    with pytest.raises(wl.WlError):
        server.render([{"pt": tracers[0][0], "width": 64, "height": 64, "spp": 1}])
    server.close()


def test_serve_rejects_overlapping_films(gpu_ctx):
    """Two requests whose films share pixels would race (their megabatches
    update both films concurrently): rejected before any work is enqueued,
    with every film left untouched."""
    import torch

    import wayland_amd as wl
    tracers = _tracers(gpu_ctx, 32)
    server = wl.Server(gpu_ctx, batch_paths=1 << 14)
    base = torch.zeros((2 * 32 * 32, 6), dtype=torch.float32, device="cuda")
    same = [{"pt": tracers[0][0], "width": 32, "height": 32, "spp": 1, "film": base[:1024]},
            {"pt": tracers[1][0], "width": 32, "height": 32, "spp": 1, "film": base[:1024]}]
    part = [{"pt": tracers[0][0], "width": 32, "height": 32, "spp": 1, "film": base[:1024]},
            {"pt": tracers[2][0], "width": 32, "height": 32, "spp": 1, "film": base[512:1536]}]
    for reqs in (same, part):
        with pytest.raises(wl.WlError) as e:
            server.render(reqs)
        assert e.value.status == 1  # WL_ERROR_INVALID_VALUE
    assert not bool(base.any())
    # adjacent, disjoint films are fine
    ok = [{"pt": tracers[0][0], "width": 32, "height": 32, "spp": 1, "film": base[:1024]},
          {"pt": tracers[1][0], "width": 32, "height": 32, "spp": 1, "film": base[1024:]}]
    server.render(ok)
    assert bool(base[:1024, :3].any()) and bool(base[1024:, :3].any())
    server.close()
