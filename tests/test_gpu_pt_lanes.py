"""Large unguided renders run each batch as two concurrent pixel halves (the
second on the scene's second stream).  The film must equal the one-lane
result bit for bit: the serving path renders the same request through its
own (single-lane) megabatch kernels, which are tested equal to small
standalone renders."""
import numpy as np
import pytest

from wayland_amd import scenes

pytestmark = pytest.mark.gpu


def test_two_lane_render_matches_single_lane(gpu_ctx):
    import wayland_amd as wl
    W, H, spp = 1024, 512, 8  # 2^22 paths in one batch: the two-lane path
    print(f"case: two-lane render {W}x{H} x {spp} spp vs the serving path", flush=True)
    sc = scenes.wavefront_scene(4, 3, width=W, height=H)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    cfg = wl.pt_config(W, H, spp=spp, max_depth=12, rr_depth=4, frame_offset=77)
    film, st = pt.render(cfg, stats=True)
    req = {"pt": pt, "width": W, "height": H, "spp": spp, "max_depth": 12, "rr_depth": 4,
           "frame_offset": 77}
    server = wl.Server(gpu_ctx, batch_paths=1 << 22)
    films, sst, _ = server.render([req])
    server.close()
    a = film.cpu().numpy()
    b = films[0].cpu().numpy()
    assert st["paths"] == W * H * spp
    assert st["extension_rays"] == sst["extension_rays"]
    assert st["shadow_rays"] == sst["shadow_rays"]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # a second call reuses the lanes' buffers and gives the same film
    film2, _ = pt.render(cfg, stats=False)
    assert np.array_equal(film2.cpu().numpy().view(np.uint32), a.view(np.uint32))
    pt.close()
    accel.close()


def test_two_lane_render_over_tiles_and_sample_ranges(gpu_ctx):
    """Batches of 2^22 paths on a 2304 x 2048 image at 2 spp: two pixel tiles
    times two sample ranges, each batch split into two concurrent halves,
    and a second render on another stream; every film equals the one-lane
    render (batches of 2^21 paths, below the two-lane threshold) bit for
    bit."""
    import torch

    import wayland_amd as wl
    W, H, spp = 2304, 2048, 2
    print(f"case: two-lane render {W}x{H} x {spp} spp, batches 2^22 vs 2^21", flush=True)
    sc = scenes.wavefront_scene(4, 3, width=W, height=H)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    two = wl.pt_config(W, H, spp=spp, max_depth=10, rr_depth=4, frame_offset=3,
                       max_batch_paths=1 << 22)
    one = wl.pt_config(W, H, spp=spp, max_depth=10, rr_depth=4, frame_offset=3,
                       max_batch_paths=1 << 21)
    film_two, st = pt.render(two, stats=True)
    film_one, st1 = pt.render(one, stats=True)
    assert st["paths"] == W * H * spp
    assert st["extension_rays"] == st1["extension_rays"]
    a = film_two.cpu().numpy().view(np.uint32)
    assert np.array_equal(a, film_one.cpu().numpy().view(np.uint32))
    # again on a stream of its own
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        film_s = torch.zeros((W * H, 6), dtype=torch.float32, device="cuda")
        pt.render(two, film_s, stats=False)
    s.synchronize()
    assert np.array_equal(film_s.cpu().numpy().view(np.uint32), a)
    pt.close()
    accel.close()
