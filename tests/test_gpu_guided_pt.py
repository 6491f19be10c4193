"""Guided path tracing on the GPU: the learned proposal is one-sample-MIS'd
with the reference BSDF sampler, so the guided estimator must converge to the
plain one (unbiasedness, checked block by block with the spread of
independent runs), and on the glossy-caustic workload the trained network
must lower the error against a converged plain-PT reference."""
import time

import numpy as np
import pytest

from wayland_amd import scenes

pytestmark = pytest.mark.gpu


def _block_means(pt, res, spp, runs, frame0, block=8):
    import wayland_amd as wl
    out = []
    for r in range(runs):
        cfg = wl.pt_config(res, res, spp=spp, frame_offset=frame0 + r * spp,
                           max_batch_paths=res * res)
        film, _ = pt.render(cfg, stats=False)
        img = pt.resolve(film, spp).cpu().numpy().reshape(res, res, 3)
        out.append(img.reshape(res // block, block, res // block, block, 3).mean(axis=(1, 3)))
    return np.stack(out)


def _zscores(a, b):
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    return (ma - mb) / np.maximum(se, 1e-12), ma, mb


@pytest.mark.parametrize("scene,bounces,vertices", [
    ("cornell", 1, 0), ("cornell", 0, 1), ("cornell", 0, 0), ("disney", 0, 2),
    ("disney", 3, 1)])
def test_guided_unbiased(gpu_ctx, scene, bounces, vertices):
    """Block means of guided (training on) and plain renders agree within
    statistical error.  Cornell box: every material guidable; Disney scene:
    transmissive spheres (never guided) in front of guidable ones, so a
    path's first guided vertex often lies on a later bounce.  bounces /
    vertices: the guide's guided_bounces / guided_vertices limits."""
    import wayland_amd as wl
    print(f"case: guided vs plain, {scene} 64x64, 16 x 16 spp each, "
          f"guided_bounces {bounces}, guided_vertices {vertices}", flush=True)
    res = 64
    sc = (scenes.cornell_box(res, res) if scene == "cornell"
          else scenes.wavefront_scene(12, 3, width=res, height=res))
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    plain = _block_means(pt, res, 16, 16, 0)
    guide = wl.Guide(gpu_ctx, batch_size=4096, steps_per_batch=2,
                     guided_bounces=bounces, guided_vertices=vertices)
    pt.set_guide(guide, wl.GUIDE_SAMPLE | wl.GUIDE_TRAIN)
    guided = _block_means(pt, res, 16, 16, 1 << 20)
    info = guide.info()
    pt.set_guide(None)
    z, mg, mp = _zscores(guided, plain)
    # image means: z over the independent runs
    zi, _, _ = _zscores(guided.mean(axis=(1, 2, 3))[:, None], plain.mean(axis=(1, 2, 3))[:, None])
    print(f"steps {info['steps']}, last samples {info['last_samples']}, "
          f"max |z| {np.abs(z).max():.2f}, image-mean z {zi[0]:.2f}, "
          f"mean ratio {mg.mean() / mp.mean():.4f}")
    assert info["steps"] == 16 * 16 * 2
    assert info["last_samples"] > 0
    assert np.isfinite(guided).all()
    assert np.abs(z).max() < 5.0
    assert abs(zi[0]) < 4.0
    assert abs(mg.mean() / mp.mean() - 1) < (0.01 if scene == "cornell" else 0.03)
    guide.close()


def _relmse(img, ref, trim=0.0):
    """relMSE over pixels (mean over channels per pixel); trim drops that
    fraction of the largest per-pixel errors (fireflies)."""
    e = np.sort(((img - ref) ** 2 / (ref ** 2 + 0.01)).reshape(-1, 3).mean(1))
    return float(e[: e.size - int(e.size * trim)].mean())


def test_guided_glossy_caustic_gain(gpu_ctx):
    """Glossy caustic at 256 x 256: guided PT (training included) has a
    lower relMSE than plain PT at the same sample count against a converged
    plain reference.  Three guides trained from different seeds (training
    order also varies with the order of the collected records): over 20 such
    runs the guided / plain ratio lay in 0.23 to 0.41, but one training in
    about 20 ends markedly worse (0.82 seen once), so the median of the three
    is tested."""
    import torch
    import wayland_amd as wl
    print("case: glossy caustic 256x256, guided (3 seeds) vs plain", flush=True)
    res, spp = 256, 64
    sc = scenes.glossy_caustic(res, res)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    cfg_ref = wl.pt_config(res, res, spp=8192, frame_offset=1 << 22, max_depth=16)
    film, _ = pt.render(cfg_ref, stats=False)
    ref = pt.resolve(film, 8192).cpu().numpy()

    def run():
        cfg = wl.pt_config(res, res, spp=1, max_depth=16, max_batch_paths=res * res)
        film = torch.zeros((res * res, 6), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(spp):
            cfg.frame_offset = s
            pt.render(cfg, film=film, stats=False)
        torch.cuda.synchronize()
        return pt.resolve(film, spp).cpu().numpy(), time.perf_counter() - t0

    img_p, t_p = run()
    e_p = _relmse(img_p, ref, 0.001)
    ratios = []
    for seed in (1, 2, 3):
        guide = wl.Guide(gpu_ctx, seed=seed)
        pt.set_guide(guide, wl.GUIDE_SAMPLE | wl.GUIDE_TRAIN)
        img_g, t_g = run()
        pt.set_guide(None)
        guide.close()
        assert np.isfinite(img_g).all()
        e_g = _relmse(img_g, ref, 0.001)
        ratios.append(e_g / e_p)
        print(f"seed {seed}: relMSE (trim 0.001) plain {e_p:.4e} ({t_p:.3f} s), guided "
              f"{e_g:.4e} ({t_g:.3f} s), ratio {e_g / e_p:.3f}, untrimmed ratio "
              f"{_relmse(img_g, ref) / _relmse(img_p, ref):.3f}")
    # the firefly-robust error: guided well below plain at the same sample count
    assert float(np.median(ratios)) < 0.6, ratios


def test_guided_vertex_with_infinite_bsdf_pdf_stays_finite(gpu_ctx):
    """Frame (19 << 20) + 62 of the glossy caustic at 1024 x 1024 draws, at
    pixel 895139 (x 163, y 874), a GGX reflection on the metal floor that
    folds back onto -V: the reference pdf there is +inf (1 / |L.H|) and the
    direction lies below the surface.  At a guided vertex the miss weight of
    that ray takes its limit 1 / (1 - alpha) instead of inf / inf, so the
    pixel stays finite (the guide's parameters play no part: the BSDF branch
    and its direction come from the path's own random numbers)."""
    import torch

    import wayland_amd as wl
    res = 1024
    print("case: glossy caustic 1024^2, guided, frame (19 << 20) + 62", flush=True)
    sc = scenes.glossy_caustic(res, res)
    pt = wl.PathTracer(gpu_ctx, wl.Accel(gpu_ctx, sc), sc)
    guide = wl.Guide(gpu_ctx, seed=20)
    pt.set_guide(guide, wl.GUIDE_SAMPLE)
    cfg = wl.pt_config(res, res, spp=1, max_depth=16, frame_offset=(19 << 20) + 62,
                       max_batch_paths=res * res)
    film = torch.zeros((res * res, 6), dtype=torch.float32, device="cuda")
    pt.render(cfg, film, stats=False)
    pt.set_guide(None)
    guide.close()
    img = pt.resolve(film, 1).cpu().numpy().reshape(res * res, 3)
    print(f"pixel 895139: {img[895139]}")
    assert np.isfinite(img[895139]).all()
    assert np.isfinite(img).all()
    pt.close()


def test_trained_network_is_published(gpu_ctx):
    """A training render trains on the guide's own stream while its next
    batch renders, and the tracer samples from a published copy of the
    network.  After a training render, a render that does not train samples
    from exactly the trained network: its film equals, bit for bit, the film
    of a fresh guide given the same parameters, both when the sampling render
    follows the training render directly and when get_params comes first;
    inference sees the trained network too."""
    import torch
    import wayland_amd as wl
    print("case: cornell 64x64, training render then sampling render", flush=True)
    res = 64
    sc = scenes.cornell_box(res, res)
    pt = wl.PathTracer(gpu_ctx, wl.Accel(gpu_ctx, sc), sc)
    cfg = wl.pt_config(res, res, spp=4, frame_offset=77, max_batch_paths=res * res)

    def sample_film(guide):
        pt.set_guide(guide, wl.GUIDE_SAMPLE)
        film, _ = pt.render(cfg, stats=False)
        return film.cpu().numpy().view(np.uint32)

    def train(guide, frame0):
        pt.set_guide(guide, wl.GUIDE_SAMPLE | wl.GUIDE_TRAIN)
        tcfg = wl.pt_config(res, res, spp=4, frame_offset=frame0, max_batch_paths=res * res)
        pt.render(tcfg, stats=False)  # 4 batches; the last one's training is pending

    g = wl.Guide(gpu_ctx, seed=5, batch_size=4096, steps_per_batch=2)
    fresh = wl.Guide(gpu_ctx, seed=6, batch_size=4096, steps_per_batch=2)
    p0 = g.get_params()
    train(g, 1 << 20)
    f_direct = sample_film(g)  # straight after the training render
    p1 = g.get_params()
    assert g.info()["steps"] == 8
    assert not np.array_equal(p0, p1)
    fresh.set_params(p1)
    assert np.array_equal(f_direct, sample_film(fresh))
    train(g, 2 << 20)
    p2 = g.get_params()  # waits for the pending training
    f_after = sample_film(g)
    fresh.set_params(p2)
    assert np.array_equal(f_after, sample_film(fresh))
    assert not np.array_equal(f_direct, f_after)
    # inference reads the trained network as well
    train(g, 3 << 20)
    q = torch.rand((4096, 9), device="cuda") * 2 - 1
    out_g = g.infer(q).cpu().numpy()
    fresh.set_params(g.get_params())
    assert np.array_equal(out_g.view(np.uint32), fresh.infer(q).cpu().numpy().view(np.uint32))
    pt.set_guide(None)
    g.close()
    fresh.close()
    pt.close()
