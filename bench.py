"""Benchmark of the traversal hot path on one MI355X.

Workload (one "step" = one launch of the closest-hit kernel over the whole
ray batch): 25 seeded displaced icospheres, subdivision 6 = 2,048,000
triangles in the unit cube (the 2M-triangle point of the traversal workload;
``--scene 64x7`` selects 64 x subdivision 7 = 20,971,520 triangles), and
2^23 = 8,388,608 incoherent rays (origins uniform in the cube, directions
uniform on the sphere), FP32, rays and BVH resident in HBM before timing.

Prints one JSON line: value = closest-hit Grays/s; any-hit Grays/s, the
roofline of the closest-hit kernel (algorithmic bytes per launch from the
instrumented node/triangle fetch counts on a 1/64 ray subsample / measured
kernel time, vs 8 TB/s) and the CPU baseline (the oracle's BVH2 traversal of
the same rays on the host cores) ride along.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + Grays/s; relMSE at equal time vs FP64 oracle, guided vs plain PT"
HBM_PEAK_GBS = 8000.0
RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES = 32, 16, 128, 48
PT_RAY_HIT_BYTES, PT_SHADE_BYTES, PT_SHADOW_BYTES = 48, 380, 52


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    # 100 launches of about 1 ms after 20 untimed ones: the clocks settle
    # within the warm-up (3 warm-up launches and 20 timed ones read 4-5 %
    # lower)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--scene", default="25x6", help="COUNTxSUBDIV icospheres")
    p.add_argument("--rays", type=int, default=1 << 23)
    p.add_argument("--scene-b", default="64x7",
                   help="second traversal point, COUNTxSUBDIV ('' = skip)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a plain 1 GiB device copy (PMC scale factor)")
    p.add_argument("--pt-spp", type=int, default=64,
                   help="wavefront PT 1920x1080 at this spp on the same geometry "
                        "(0 = skip)")
    p.add_argument("--guide-spp", type=int, default=64,
                   help="guided glossy-caustic run length in spp (0 = skip)")
    p.add_argument("--guide-res", type=int, default=1024)
    p.add_argument("--guide-runs", type=int, default=60)
    p.add_argument("--guide-steps", type=int, default=0,
                   help="Adam steps per pass (0: the library default)")
    p.add_argument("--serve-requests", type=int, default=64,
                   help="batched-serving requests (0 = skip)")
    p.add_argument("--scene-file-frames", type=int, default=64,
                   help="headless frames of cbox.pbrt at 1024^2 (0 = skip the scene-file leg)")
    p.add_argument("--cornell-ref-spp", type=int, default=256,
                   help="FP64 oracle samples per pixel for the Cornell relMSE "
                        "(0 = skip the Cornell leg)")
    return p.parse_args()


def bench_traversal_point(args, ctx, wl, torch, timed):
    """The traversal workload's other point (default 64 icospheres at
    subdivision 7 = 20,971,520 triangles) with the same 2^23 rays:
    closest-hit and any-hit Grays/s and the algorithmic-byte roofline."""
    from wayland_amd import scenes
    count, subdiv = (int(x) for x in args.scene_b.split("x"))
    t0 = time.perf_counter()
    sc = scenes.icosphere_scene(count, subdiv)
    t_gen = time.perf_counter() - t0
    accel = wl.Accel(ctx, sc)
    info = accel.info()
    rays = torch.from_numpy(scenes.traversal_rays(args.rays, seed=2)).cuda()
    n = rays.shape[0]
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    bits = torch.empty(((n + 31) // 32,), dtype=torch.int32, device="cuda")
    t_c = timed(lambda: accel.trace_closest(rays, hits), args.steps, args.warmup)
    t_a = timed(lambda: accel.trace_any(rays, bits), args.steps, args.warmup)
    sub = rays[::64].contiguous()
    cnt = torch.empty((sub.shape[0], 2), dtype=torch.int32, device="cuda")
    accel.trace_closest(sub, None, cnt)
    torch.cuda.synchronize()
    c = cnt.double().mean(0).cpu().numpy()
    bpr = RAY_BYTES + HIT_BYTES + c[0] * NODE_BYTES + c[1] * TRI_BYTES
    out = {"workload": f"{count} displaced icospheres subdiv {subdiv} "
                       f"({sc.num_triangles} tris), {n} rays",
           "triangles": sc.num_triangles, "closest_grays_per_s": n / t_c / 1e9,
           "anyhit_grays_per_s": n / t_a / 1e9, "ms_closest": t_c * 1e3,
           "bvh": {"nodes": info["inner_nodes"], "depth": info["depth"],
                   "build_s": info["build_seconds"], "device_bytes": info["device_bytes"]},
           "scene_gen_s": t_gen,
           "roofline": {"bound": "hbm", "achieved": bpr * n / t_c / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": bpr * n / t_c / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_ray": bpr, "nodes_per_ray": c[0], "tris_per_ray": c[1]}}
    log(f"traversal point b: {out}")
    accel.close()
    del rays, hits, bits
    torch.cuda.empty_cache()
    return out


def bench_pt(args, ctx, accel, sc, wl, torch):
    """Wavefront PT workload: 1920x1080, depth 16, RR after bounce 4, on the
    same icosphere geometry with per-object Disney materials, two quad
    lights and a constant environment 0.1."""
    from wayland_amd import scenes
    count, subdiv = (int(x) for x in args.scene.split("x"))
    W, H = 1920, 1080
    wsc = scenes.wavefront_scene(count, subdiv, width=W, height=H)
    waccel = wl.Accel(ctx, wsc)
    pt = wl.PathTracer(ctx, waccel, wsc)
    film = torch.zeros((W * H, 6), dtype=torch.float32, device="cuda")
    # 2^27 paths per batch (32 GB of path state): 1080p x 64 spp in one batch
    cfg = wl.pt_config(W, H, spp=args.pt_spp, max_depth=16, rr_depth=4,
                       max_batch_paths=1 << 27)
    pt.render(cfg, film, stats=False)  # warm: allocates the batch buffers
    torch.cuda.synchronize()
    film.zero_()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    _, st = pt.render(cfg, film, stats=False)
    end.record()
    torch.cuda.synchronize()
    dt = start.elapsed_time(end) / 1e3
    _, st = pt.render(wl.pt_config(W, H, spp=1, max_depth=16, rr_depth=4,
                                   frame_offset=10_000), None, stats=True)
    rays_per_path = (st["extension_rays"] + st["shadow_rays"]) / st["paths"]
    paths = W * H * args.pt_spp
    # the same render with the material-sorted shade queue
    film.zero_()
    cfg_s = wl.pt_config(W, H, spp=args.pt_spp, max_depth=16, rr_depth=4,
                         max_batch_paths=1 << 27, flags=wl.PT_SORT_BY_MATERIAL)
    start.record()
    pt.render(cfg_s, film, stats=False)
    end.record()
    torch.cuda.synchronize()
    dt_sorted = start.elapsed_time(end) / 1e3
    pt.close()
    waccel.close()
    # path-state and queue bytes of the render (BVH node and triangle
    # fetches excluded): per extension ray the ray and hit (48 B) plus the
    # shade step (380 B: hit, path id, state read and write, shading record,
    # material, next ray, shadow ray and contribution); per shadow ray its ray,
    # path id and contribution (52 B)
    ext = st["extension_rays"] / st["paths"] * paths
    shd = st["shadow_rays"] / st["paths"] * paths
    state_bytes = ext * (PT_RAY_HIT_BYTES + PT_SHADE_BYTES) + shd * PT_SHADOW_BYTES
    traffic = traffic_from_profiles(f"pt_wavefront_{W}x{H}", "traffic_bytes_per_path")
    out = {"workload": f"wavefront PT {W}x{H} {args.pt_spp} spp depth 16 RR>4, "
                       f"{wsc.num_triangles} tris, 2 quad lights + env 0.1",
           "mpaths_per_s": paths / dt / 1e6,
           "grays_per_s": paths * rays_per_path / dt / 1e9,
           "rays_per_path": rays_per_path, "seconds": dt,
           "mpaths_per_s_material_sorted": paths / dt_sorted / 1e6,
           "roofline": {"bound": "hbm", "achieved": state_bytes / dt / 1e9,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": state_bytes / dt / 1e9 / HBM_PEAK_GBS,
                        "algorithmic_bytes": state_bytes,
                        "bytes": "path state + queues only (48 B ray/hit + 380 B shade per "
                                 "extension ray, 52 B per shadow ray); BVH fetches excluded",
                        "traffic_per_path": traffic}}
    log(f"pt: {out}")
    return out


def relmse(img, ref, trim=0.0, mask=None):
    """mean over pixels and channels of (I - R)^2 / (R^2 + 0.01); trim > 0
    drops that fraction of the largest per-pixel errors (fireflies); mask
    keeps only the pixels where it is true."""
    e = ((img - ref) ** 2 / (ref ** 2 + 0.01)).reshape(-1, 3).mean(1)
    if mask is not None:
        e = e[np.asarray(mask).reshape(-1)]
    if trim > 0:
        e = np.sort(e)[: e.size - int(e.size * trim)]
    return float(e.mean())


def bench_guide(args, ctx, wl, torch):
    """Learned guiding on the glossy-caustic workload (1024 x 1024, depth 16):
    guided PT (training included in its time) against plain PT given the same
    wall time, both scored by relMSE against the committed 65,536-spp FP64
    oracle images: the whole view on a 256 x 256 pixel grid (every 4th pixel;
    the headline), its floor pixels, its pixels off and on the glass sphere,
    and the 128 x 128 caustic window.  Plain PT runs twice per run index: with
    its own frame offsets (independent of the guided run) and with the guided
    run's (common random numbers: paths that meet no guided vertex are then
    the same samples on both sides; `common_seeds`, compared run by run)."""
    from wayland_amd import scenes
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_fp64_references as fx
    res = args.guide_res
    sc = scenes.glossy_caustic(res, res)
    accel = wl.Accel(ctx, sc)
    pt = wl.PathTracer(ctx, accel, sc)
    depth = 16
    grid = fx.load("caustic_grid256_fp64")
    crop = fx.load("caustic_crop128_fp64")
    if grid is None or crop is None or grid[2]["width"] != res:
        log("guide: FP64 references missing or for another resolution; skipped")
        return None
    gx0, gy0, gw, gh = grid[2]["window"]
    gs = grid[2]["stride"]
    cx0, cy0, cw, ch = crop[2]["window"]

    def progressive(spp, frame0):
        cfg = wl.pt_config(res, res, spp=1, max_depth=depth, max_batch_paths=res * res)
        f = torch.zeros((res * res, 6), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(spp):
            cfg.frame_offset = frame0 + k
            pt.render(cfg, f, stats=False)
        torch.cuda.synchronize()
        return pt.resolve(f, spp).cpu().numpy().reshape(res, res, 3), time.perf_counter() - t

    # grid pixels whose camera ray (through the pixel centre) first meets the
    # metal floor, where the guided vertex is; the glass sphere's pixels carry
    # glass-only paths that no guided vertex lies on
    cam = sc.camera
    xs = (gx0 + gs * np.arange(gw) + 0.5) / res * 2 - 1
    ys = (gy0 + gs * np.arange(gh) + 0.5) / res * 2 - 1
    U, V = np.meshgrid(xs, ys)  # row = y
    d = (np.asarray(cam.lookat, np.float64)[None, None] + U[..., None] * np.asarray(cam.right)
         + V[..., None] * np.asarray(cam.up))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    prays = np.zeros((gw * gh, 8), np.float32)
    prays[:, 0:3] = cam.pos
    prays[:, 3] = 1e-3
    prays[:, 4:7] = d.reshape(-1, 3)
    prays[:, 7] = 1e30
    prim = wl.hits_to_numpy(accel.trace_closest(torch.from_numpy(prays).cuda()))["prim"]
    flat = sc.flatten()
    f0, fn = int(flat["mesh_tri_offset"][1]), int(flat["mesh_tri_count"][1])
    floor = ((prim >= f0) & (prim < f0 + fn)).reshape(gh, gw)
    s0, sn = int(flat["mesh_tri_offset"][0]), int(flat["mesh_tri_count"][0])
    sphere = ((prim >= s0) & (prim < s0 + sn)).reshape(gh, gw)

    def scores(img):
        gi = img[gy0::gs, gx0::gs][:gh, :gw]
        ci = img[cy0:cy0 + ch, cx0:cx0 + cw]
        return (relmse(gi, grid[0]), relmse(gi, grid[0], mask=floor),
                relmse(ci, crop[0]), relmse(gi, grid[0], mask=~sphere),
                relmse(gi, grid[0], mask=sphere))

    # allocate the per-path guide buffers outside the timed runs
    gkw = {"steps_per_batch": args.guide_steps} if args.guide_steps else {}
    warm_guide = wl.Guide(ctx, **gkw)
    pt.set_guide(warm_guide, wl.GUIDE_SAMPLE | wl.GUIDE_TRAIN)
    progressive(1, 1 << 28)
    pt.set_guide(None)
    warm_guide.close()
    # independent runs: fresh network each time, training inside the timing
    runs = args.guide_runs
    g_sc, g_t, steps = [], [], 0
    for r in range(runs):
        guide = wl.Guide(ctx, seed=r + 1, **gkw)
        pt.set_guide(guide, wl.GUIDE_SAMPLE | wl.GUIDE_TRAIN)
        img_g, t_g = progressive(args.guide_spp, r << 20)
        steps = guide.info()["steps"]
        pt.set_guide(None)
        guide.close()
        g_sc.append(scores(img_g))
        g_t.append(t_g)
    # plain PT per-sample time (after a warm pass), then the same wall time as
    # the guided runs
    progressive(2, (1 << 26) + 64)
    _, t_probe = progressive(16, 1 << 26)
    spp_p = max(1, int(round(float(np.mean(g_t)) / (t_probe / 16))))
    for attempt in range(3):
        p_sc, p_t = [], []
        for r in range(runs):
            img_p, t_p = progressive(spp_p, (1 << 27) + (r << 20))
            p_sc.append(scores(img_p))
            p_t.append(t_p)
        # equal time within 3 %, else rescale the sample count and rerun
        ratio = float(np.mean(g_t)) / float(np.mean(p_t))
        if abs(ratio - 1.0) <= 0.03:
            break
        spp_p = max(1, int(round(spp_p * ratio)))
    # plain runs with the guided runs' frame offsets (common random numbers):
    # paths that meet no guided vertex, such as the glass-only paths of the
    # sphere's pixels, draw the same numbers on both sides, so their
    # fireflies are shared by run i of each side instead of being independent
    c_sc = []
    for r in range(runs):
        img_c, _ = progressive(spp_p, r << 20)
        c_sc.append(scores(img_c))
    g_sc, p_sc, c_sc = np.array(g_sc), np.array(p_sc), np.array(c_sc)

    def side(k):
        g, p, c = g_sc[:, k].tolist(), p_sc[:, k].tolist(), c_sc[:, k].tolist()
        per_run = np.asarray(g) / np.maximum(np.asarray(c), 1e-300)
        return {"relmse_guided": g, "relmse_plain": p,
                "relmse_ratio_equal_time": float(np.mean(g) / np.mean(p)),
                "relmse_ratio_ci95": bootstrap_ratio_ci(g, p),
                "relmse_median_ratio": float(np.median(g) / np.median(p)),
                "common_seeds": {
                    "relmse_plain": c,
                    "relmse_ratio_equal_time": float(np.mean(g) / np.mean(c)),
                    "relmse_ratio_ci95": bootstrap_paired_ratio_ci(g, c),
                    "median_of_run_ratios": float(np.median(per_run)),
                    "runs_guided_below_plain": int((per_run < 1.0).sum())}}

    out = {"workload": f"glossy caustic {res}x{res}, {sc.num_triangles} tris, depth {depth}",
           "runs": runs,
           "guided": {"spp": args.guide_spp, "seconds": float(np.mean(g_t)), "adam_steps": steps},
           "plain": {"spp": spp_p, "seconds": float(np.mean(p_t))},
           "common_seeds": "plain run i repeated with guided run i's frame offsets "
                           "(frames r << 20 onward): the first 64 samples per pixel share "
                           "their random numbers with the guided run; paired bootstrap",
           "reference": f"FP64 CPU oracle, {grid[2]['spp']} spp",
           "full_view": dict(side(0), pixels=f"{gw}x{gh} grid, every {gs}th pixel of the view"),
           "floor": dict(side(1), pixels=f"{int(floor.sum())} grid pixels on the metal floor"),
           "caustic_window": dict(side(2), pixels=f"x {cx0}..{cx0 + cw - 1}, "
                                                  f"y {cy0}..{cy0 + ch - 1}"),
           "off_sphere": dict(side(3), pixels=f"{int((~sphere).sum())} grid pixels whose "
                                              "camera ray misses the glass sphere"),
           "sphere": dict(side(4), pixels=f"{int(sphere.sum())} grid pixels whose camera "
                                          "ray meets the glass sphere first")}
    # share of the whole-view error that the sphere's pixels carry (mean over
    # runs of sum over sphere pixels / sum over all pixels)
    frac = float(sphere.mean())
    out["sphere_share_of_error"] = {
        "guided": float(np.mean(g_sc[:, 4] * frac / g_sc[:, 0])),
        "plain": float(np.mean(p_sc[:, 4] * frac / p_sc[:, 0]))}
    # headline: the whole view against the converged FP64 oracle, paired by
    # common random numbers (the precise comparison); the independent-seed
    # ratio beside it
    crn = out["full_view"]["common_seeds"]
    out["relmse_ratio_equal_time"] = crn["relmse_ratio_equal_time"]
    out["relmse_ratio_ci95"] = crn["relmse_ratio_ci95"]
    out["median_of_run_ratios"] = crn["median_of_run_ratios"]
    out["relmse_ratio_equal_time_independent_seeds"] = \
        out["full_view"]["relmse_ratio_equal_time"]
    log(f"guide: {out}")
    return out


def bootstrap_ratio_ci(a, b, n=10000, seed=0):
    """95 % percentile bootstrap interval of mean(a) / mean(b), resampling the
    independent runs of each side with replacement."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.size < 2 or b.size < 2:
        return None
    rng = np.random.default_rng(seed)
    ra = a[rng.integers(0, a.size, (n, a.size))].mean(1)
    rb = b[rng.integers(0, b.size, (n, b.size))].mean(1)
    q = np.percentile(ra / rb, [2.5, 97.5])
    return [float(q[0]), float(q[1])]


def bootstrap_paired_ratio_ci(a, b, n=10000, seed=0):
    """95 % percentile bootstrap interval of mean(a) / mean(b) for paired runs
    (run i of each side shares its random numbers): run indices are resampled
    jointly."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.size < 2 or a.size != b.size:
        return None
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, a.size, (n, a.size))
    q = np.percentile(a[idx].mean(1) / b[idx].mean(1), [2.5, 97.5])
    return [float(q[0]), float(q[1])]


def bench_serve(args, ctx, wl, torch):
    """Batched serving: 64 requests (seed 5) of Cornell and icosphere variants,
    square resolution uniform in 128..1024, spp uniform in 1..64, depth
    uniform in 8..16, through one server (megabatches replayed from a captured
    hipGraph, capacity and shade grid autotuned on the workload).  Against the
    same requests rendered one after another with wl_pt_render."""
    from wayland_amd import scenes
    rng = np.random.default_rng(5)
    tracers = [wl.PathTracer(ctx, wl.Accel(ctx, sc), sc) for sc in scenes.serving_scenes()]

    def workload(n, res_lo, res_hi, spp_hi, seed_shift):
        reqs = scenes.serving_requests(rng, n, res_lo, res_hi, spp_hi, seed_shift)
        for r in reqs:
            r["pt"] = tracers[r.pop("scene")]
        return reqs

    def run(reqs, label):
        paths = sum(r["width"] * r["height"] * r["spp"] for r in reqs)
        server = wl.Server(ctx)
        t0 = time.perf_counter()
        tuned = server.autotune(reqs, max_spp=2)
        t_tune = time.perf_counter() - t0
        server.render(reqs)  # warm (graph capture)
        films, st, done = server.render(reqs)
        server.close()
        # the same requests one by one (launches queued back to back)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for r in reqs:
            cfg = wl.pt_config(r["width"], r["height"], spp=r["spp"], max_depth=r["max_depth"],
                               frame_offset=r["frame_offset"])
            r["pt"].render(cfg, stats=False)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t
        done = np.array(done)
        out = {"workload": label, "paths": paths, "megabatches": st["batches"],
               "graph_launches": st["launches"],
               "autotuned": tuned, "autotune_seconds": t_tune,
               "mpaths_per_s": paths / st["seconds"] / 1e6, "seconds": st["seconds"],
               "grays_per_s": (st["extension_rays"] + st["shadow_rays"]) / st["seconds"] / 1e9,
               "latency_p50_s": float(np.percentile(done, 50)),
               "latency_p99_s": float(np.percentile(done, 99)),
               "sequential_mpaths_per_s": paths / t_seq / 1e6, "sequential_seconds": t_seq}
        log(f"serve: {out}")
        return out

    out = run(workload(args.serve_requests, 128, 1024, 64, 0),
              f"{args.serve_requests} requests, Cornell + icosphere variants, res 128..1024, "
              "spp 1..64, depth 8..16")
    out["small_requests"] = run(workload(256, 64, 256, 4, 1 << 10),
                                "256 requests, res 64..256, spp 1..4, depth 8..16")
    return out


def bench_cornell(args, ctx, wl, torch):
    """Plain-PT CPU workload on both sides: Cornell box 256 x 256, 16 spp,
    depth 8, NEE + MIS.  GPU render vs the oracle's host restatement of
    PathTracing.cu on the host cores (same TEA seeds), both scored by relMSE
    against the oracle's FP64 estimator at args.cornell_ref_spp samples."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from wayland_amd import scenes
    res, spp, depth = 256, 16, 8
    sc = scenes.cornell_box(res, res)
    accel = wl.Accel(ctx, sc)
    pt = wl.PathTracer(ctx, accel, sc)
    cfg = wl.pt_config(res, res, spp=spp, max_depth=depth)
    pt.render(cfg, stats=False)  # warm
    times = []
    for _ in range(5):
        film = torch.zeros((res * res, 6), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        pt.render(cfg, film, stats=False)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    t_gpu = float(np.median(times))
    img_gpu = pt.resolve(film, spp).cpu().numpy().reshape(res, res, 3)
    model, ncpu, avail = host_cpu()
    nthreads = max(1, min(HOST_SHARE_THREADS, avail))
    opt = ob.OraclePT(sc)
    ocfg = opt.config(res, res, max_depth=depth)
    opt.render_frame(ocfg, 0, nthreads)  # warm
    t = time.perf_counter()
    acc = np.zeros((res, res, 3), np.float64)
    for f in range(spp):
        img, _ = opt.render_frame(ocfg, f, nthreads)
        acc += img
    t_cpu = time.perf_counter() - t
    img_cpu = acc / spp
    # the converged reference: the committed 65,536-spp FP64 oracle image of
    # this workload; without it, an FP64 oracle render of cornell_ref_spp
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_fp64_references as fx
    fixture = fx.load("cornell_256_fp64")
    expected = None
    if fixture is not None:
        ref, ref_var, meta = fixture
        ref_spp, t_ref = meta["spp"], meta["seconds"]
        ref_kind = f"FP64 CPU oracle, {ref_spp} spp (committed fixture, {meta['threads']} threads)"
        expected = fx.expected_relmse(ref, ref_var, spp, ref_spp)
    else:
        t = time.perf_counter()
        ref_spp = args.cornell_ref_spp
        ref = opt.render_f64(ocfg, 1 << 20, ref_spp, nthreads)
        t_ref = time.perf_counter() - t
        ref_kind = f"FP64 CPU oracle, {ref_spp} spp"
    paths = res * res * spp
    out = {"workload": f"cornell box {res}x{res}, {spp} spp, depth {depth}, NEE+MIS, FP32, "
                       "TEA/LCG seeds (pixel, frame) as PathTracing.cu",
           "gpu_mpaths_per_s": paths / t_gpu / 1e6, "gpu_seconds": t_gpu,
           "cpu_mpaths_per_s": paths / t_cpu / 1e6, "cpu_seconds": t_cpu,
           "cpu_threads": nthreads, "cpu_kind": "port", "cpu_model": model,
           "host_cpus": ncpu,
           "relmse_gpu_vs_fp64_oracle": relmse(img_gpu, ref),
           "relmse_cpu_vs_fp64_oracle": relmse(img_cpu, ref),
           "relmse_expected_at_16spp": expected,
           "max_abs_gpu_minus_cpu": float(np.abs(img_gpu - img_cpu).max()),
           "fp64_oracle": ref_kind, "fp64_oracle_spp": ref_spp,
           "fp64_oracle_seconds": t_ref}
    log(f"cornell: {out}")
    return out


def bench_scene_files(args, ctx, wl, torch):
    """Reference-format inputs: (1) the headless frame loop on the Cornell box
    read from its pbrt file (tests/golden/cbox.pbrt, 1024 x 1024), one sample
    per frame through the progressive film with the frame buffer downloaded
    every frame, as Renderer::Run does; (2) the textured Cornell box against
    the same box untextured, 1024 x 1024, 16 spp, depth 25."""
    from wayland_amd import render, scenes
    res, frames = 1024, args.scene_file_frames
    setting = render.RendererSetting.from_args(
        ["-s", os.path.join(ROOT, "tests", "golden", "cbox.pbrt"), "-fr", "4",
         "-r", str(res), str(res)])
    r = render.Renderer(setting, ctx=ctx)
    r.run()  # warm
    setting.frame_limit = frames
    n, seconds = r.run()
    r.close()
    headless = {"workload": f"cbox.pbrt {res}x{res}, 1 spp per frame, progressive film + "
                            "u8 frame buffer downloaded per frame", "frames": n,
                "seconds": seconds, "frames_per_s": n / seconds,
                "mpaths_per_s": n * res * res / seconds / 1e6}
    out = {"headless_pbrt": headless}
    spp = 16
    rates = {}
    for name, sc in (("plain", scenes.cornell_box(res, res)),
                     ("textured", scenes.textured_cornell_box(res, res))):
        accel = wl.Accel(ctx, sc)
        pt = wl.PathTracer(ctx, accel, sc)
        cfg = wl.pt_config(res, res, spp=spp)
        pt.render(cfg, stats=False)
        times = []
        for _ in range(3):
            film = torch.zeros((res * res, 6), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            t = time.perf_counter()
            pt.render(cfg, film, stats=False)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        rates[name] = res * res * spp / float(np.median(times)) / 1e6
        pt.close()
        accel.close()
    out["textures"] = {"workload": f"cornell box {res}x{res}, {spp} spp, depth 25; back wall "
                                   "37x23 and floor 16x16 RGBA8 textures, wrap + bilinear",
                       "mpaths_per_s_textured": rates["textured"],
                       "mpaths_per_s_untextured": rates["plain"]}
    log(f"scene files: {out}")
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(tag, key="traffic_bytes_per_launch"):
    """HBM bytes measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated), recorded as `key: value` in profiles/<tag>_pmc.md."""
    path = os.path.join(ROOT, "profiles", f"{tag}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(key + r":\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_cpu():
    """(model string, logical CPUs of the host, CPUs this process may run on)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


# The host share of one GPU on the measurement boxes is 16 CPUs: the CPU
# baseline uses at most that many threads, whatever the host reports.
HOST_SHARE_THREADS = 16


def cpu_baseline(tris, rays_np):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    model, ncpu, avail = host_cpu()
    nthreads = max(1, min(HOST_SHARE_THREADS, avail))
    bvh = ob.OracleBVH(tris)
    sample = rays_np
    bvh.closest(sample[:4096], nthreads)  # warm
    # repeat the whole ray set until about 10 s of CPU time (threads x wall)
    # is spent; the rate is the median over the repetitions
    times = []
    while not times or (sum(times) * nthreads < 10.0 and len(times) < 50):
        t0 = time.perf_counter()
        bvh.closest(sample, nthreads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return {"value": sample.shape[0] / dt / 1e9, "unit": "Grays/s", "cores": nthreads,
            "kind": "port", "cpu_model": model, "host_cpus": ncpu, "cpus_available": avail,
            "sample": f"oracle BVH2 watertight closest-hit, all {sample.shape[0]} rays "
                      f"of the workload, {nthreads} threads (the host share of one GPU), "
                      f"median of {len(times)} passes of {dt:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the path would "
                                   "shard by rays (traversal) and by sample ranges "
                                   "(path tracer), which is not implemented here",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import wayland_amd as wl
    from wayland_amd import scenes

    count, subdiv = (int(x) for x in args.scene.split("x"))
    t0 = time.perf_counter()
    sc = scenes.icosphere_scene(count, subdiv)
    log(f"scene {count}x{subdiv}: {sc.num_triangles} tris, gen {time.perf_counter() - t0:.1f}s")
    ctx = wl.Context(0)
    t0 = time.perf_counter()
    accel = wl.Accel(ctx, sc)
    info = accel.info()
    log(f"accel build {time.perf_counter() - t0:.1f}s: {info}")
    rays_np = scenes.traversal_rays(args.rays, seed=2)
    rays = torch.from_numpy(rays_np).cuda()
    n = rays.shape[0]
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    bits = torch.empty(((n + 31) // 32,), dtype=torch.int32, device="cuda")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            fn()
        end.record()
        torch.cuda.synchronize()
        return start.elapsed_time(end) / 1e3 / steps

    t_closest = timed(lambda: accel.trace_closest(rays, hits), args.steps, args.warmup)
    log(f"closest: {t_closest * 1e3:.3f} ms/launch, {n / t_closest / 1e9:.3f} Grays/s")
    t_any = timed(lambda: accel.trace_any(rays, bits), args.steps, args.warmup)
    log(f"any: {t_any * 1e3:.3f} ms/launch, {n / t_any / 1e9:.3f} Grays/s")
    hit_rate = float((hits[:, 1].view(torch.int32) != -1).float().mean().item())

    # algorithmic bytes per ray from the instrumented kernel on a subsample
    sub = rays[::64].contiguous()
    cnt = torch.empty((sub.shape[0], 2), dtype=torch.int32, device="cuda")
    accel.trace_closest(sub, None, cnt)
    torch.cuda.synchronize()
    c = cnt.double().mean(0).cpu().numpy()
    bytes_per_ray = RAY_BYTES + HIT_BYTES + c[0] * NODE_BYTES + c[1] * TRI_BYTES
    achieved = bytes_per_ray * n / t_closest / 1e9
    tag = f"trace_closest_{count}x{subdiv}"
    traffic = traffic_from_profiles(tag)

    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()

    point_b = None
    if args.scene_b:
        point_b = bench_traversal_point(args, ctx, wl, torch, timed)

    pt_out = None
    if args.pt_spp > 0:
        pt_out = bench_pt(args, ctx, accel, sc, wl, torch)

    serve_out = None
    if args.serve_requests > 0:
        serve_out = bench_serve(args, ctx, wl, torch)

    files_out = None
    if args.scene_file_frames > 0:
        files_out = bench_scene_files(args, ctx, wl, torch)

    cornell_out = None
    if args.cornell_ref_spp > 0:
        cornell_out = bench_cornell(args, ctx, wl, torch)

    guide_out = None
    if args.guide_spp > 0:
        guide_out = bench_guide(args, ctx, wl, torch)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc.triangles(), rays_np)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": n / t_closest / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_closest * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"incoherent closest-hit traversal, {count} displaced "
                               f"icospheres subdiv {subdiv} ({sc.num_triangles} tris), "
                               f"{n} rays, origins U[0,1]^3, dirs uniform on S^2",
                   "triangles": sc.num_triangles, "rays": n},
        "anyhit_grays_per_s": n / t_any / 1e9,
        "hit_rate": hit_rate,
        "bvh": {"nodes": info["inner_nodes"], "depth": info["depth"],
                "build_s": info["build_seconds"], "device_bytes": info["device_bytes"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     # measured beyond-L2 bytes per launch over the same time:
                     # the share of HBM bandwidth the launch actually draws
                     "traffic_frac": (traffic / t_closest / 1e9 / HBM_PEAK_GBS
                                      if traffic else None),
                     "bytes_per_ray": bytes_per_ray, "nodes_per_ray": c[0],
                     "tris_per_ray": c[1],
                     "note": "achieved counts every node and triangle fetch (ray 32 B + "
                             "hit 16 B + nodes x 128 B + triangles x 48 B per ray), also "
                             "the top three BVH levels read from LDS and fetches that hit "
                             "in L2 or the Infinity Cache, so frac may exceed 1; "
                             "traffic_frac is the measured HBM-side share"},
        "cpu_baseline": cpu,
        "traversal_point_b": point_b,
        "pathtracer": pt_out,
        "guiding": guide_out,
        "cornell_cpu_workload": cornell_out,
        "serving": serve_out,
        "scene_files": files_out,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
