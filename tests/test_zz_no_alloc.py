"""wl_pt_render_noalloc (the render that allocates nothing: its batch buffers
live in the library's fixed scratch array) against wl_pt_render, through the
C ABI.  Two calls in a row, the second on a smaller input, each film equal
to wl_pt_render's for the same configuration bit for bit; then the host-side
rejections (batch beyond the array, guide attached)."""
import numpy as np
import pytest

from wayland_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(film):
    return film.cpu().numpy().view(np.uint32)


def test_noalloc_render_matches_render(gpu_ctx):
    import torch

    import wayland_amd as wl
    cap = wl.NOALLOC_MAX_BATCH_PATHS
    # call 1: 768 x 512 x 2 spp = 3 x cap paths, so two pixel tiles (cap and
    # cap / 2 pixels) times two sample passes share the scratch array
    W, H, spp = 768, 512, 2
    print(f"case: noalloc {W}x{H} x {spp} spp, icosphere scene", flush=True)
    sc = scenes.wavefront_scene(4, 3, width=W, height=H)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    cfg = wl.pt_config(W, H, spp=spp, max_depth=12, rr_depth=4, frame_offset=5,
                       max_batch_paths=cap)
    ref, _ = pt.render(cfg, stats=False)
    film = torch.zeros((W * H, 6), dtype=torch.float32, device="cuda")
    cfg0 = wl.pt_config(W, H, spp=spp, max_depth=12, rr_depth=4, frame_offset=5)
    pt.render_noalloc(cfg0, film)  # max_batch_paths 0 = the array's capacity
    torch.cuda.synchronize()
    assert np.array_equal(_bits(film), _bits(ref))
    assert float(film[:, :3].sum()) > 0
    pt.close()
    accel.close()

    # call 2: a smaller input (one batch of 64 x 48 x 5 paths), Cornell box
    print("case: noalloc 64x48 x 5 spp, Cornell box", flush=True)
    sc2 = scenes.cornell_box(64, 48)
    accel2 = wl.Accel(gpu_ctx, sc2)
    pt2 = wl.PathTracer(gpu_ctx, accel2, sc2)
    cfg2 = wl.pt_config(64, 48, spp=5, max_depth=8, frame_offset=11, max_batch_paths=cap)
    ref2, _ = pt2.render(cfg2, stats=False)
    film2 = torch.zeros((64 * 48, 6), dtype=torch.float32, device="cuda")
    pt2.render_noalloc(cfg2, film2)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(film2), _bits(ref2))

    # rejections before any work is enqueued
    big = wl.pt_config(64, 48, spp=1, max_batch_paths=cap + 1)
    with pytest.raises(wl.WlError) as e:
        pt2.render_noalloc(big, film2)
    assert e.value.status == 5  # WL_ERROR_LIMIT_EXCEEDED
    guide = wl.Guide(gpu_ctx)
    pt2.set_guide(guide, wl.GUIDE_SAMPLE)
    with pytest.raises(wl.WlError) as e:
        pt2.render_noalloc(cfg2, film2)
    assert e.value.status == 4  # WL_ERROR_NOT_SUPPORTED
    pt2.set_guide(None)
    guide.close()
    # the film is untouched by the rejected calls
    assert np.array_equal(_bits(film2), _bits(ref2))
    pt2.close()
    accel2.close()
