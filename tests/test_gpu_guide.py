"""GPU parity of the guiding network's kernels against the oracle
(oracle/oracle_guide.c, pinned by test_guide_oracle.py).

The oracle's BF16-emulating mode rounds weights, layer inputs and deltas to
bfloat16 where the MFMA kernels do and accumulates in FP64; the GPU
accumulates in FP32 in MFMA order, so a few values round to the other
neighbouring bfloat16.  Tolerances: outputs 2e-3 absolute + relative for
99.9 % of values; gradients 1e-2 relative (Frobenius, per layer)."""
import numpy as np
import pytest

import oracle_bind as ob
from guide_util import NP, b_off, random_outputs, random_params, random_samples, row, w_off

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _guide(ctx, params=None, **cfg):
    import wayland_amd as wl
    g = wl.Guide(ctx, **cfg)
    if params is not None:
        g.set_params(params.astype(np.float32))
    return g


@pytest.mark.parametrize("n", [1, 37, 1000])
def test_infer_matches_oracle(gpu_ctx, n):
    import torch
    print(f"case: guide infer n={n}", flush=True)
    P = random_params(11)
    g = _guide(gpu_ctx, P)
    S = random_samples(12, n)
    out = g.infer(_t(S[:, :9])).cpu().numpy()
    torch.cuda.synchronize()
    ref = ob.guide_forward(P.astype(np.float32).astype(np.float64), S[:, :9], emulate_bf16=True)
    err = np.abs(out - ref) / (1.0 + np.abs(ref))
    print(f"max err {err.max():.3e}, 99.9% {np.quantile(err, 0.999):.3e}")
    assert np.quantile(err, 0.999) < 2e-3
    assert err.max() < 3e-2
    g.close()


def test_default_init_matches_oracle(gpu_ctx):
    # parameters read back from the library's own initialisation
    print("case: guide default init forward", flush=True)
    g = _guide(gpu_ctx)
    P = g.get_params().astype(np.float64)
    S = random_samples(13, 256)
    out = g.infer(_t(S[:, :9])).cpu().numpy()
    ref = ob.guide_forward(P, S[:, :9], emulate_bf16=True)
    assert np.quantile(np.abs(out - ref) / (1 + np.abs(ref)), 0.999) < 2e-3
    # initial lobes: concentration softplus(1), Fibonacci means (|mu| = 1)
    pi, kap, mu = ob.guide_decode(ref[0])
    assert np.allclose(kap, np.log1p(np.e), atol=0.2)
    g.close()


@pytest.mark.parametrize("n", [256, 100])
def test_gradient_matches_oracle(gpu_ctx, n):
    print(f"case: guide gradient n={n}", flush=True)
    P = random_params(21)
    g = _guide(gpu_ctx, P)
    S = random_samples(22, n)
    grad, loss = g.gradient(_t(S))
    grad, loss = grad.cpu().numpy(), float(loss.cpu().numpy()[0])
    l_ref, g_ref = ob.guide_loss_grad(P.astype(np.float32).astype(np.float64), S,
                                      emulate_bf16=True)
    print(f"loss gpu {loss:.6f} oracle {l_ref:.6f}")
    assert loss == pytest.approx(l_ref, rel=1e-3)
    for l in range(5):
        n_out = 64 if l < 4 else 32
        for name, sl in (("W", slice(w_off(l), b_off(l))), ("b", slice(b_off(l), b_off(l) + n_out))):
            e = np.linalg.norm(grad[sl] - g_ref[sl]) / np.linalg.norm(g_ref[sl])
            print(f"layer {l} {name}: rel err {e:.2e}")
            assert e < 1e-2
    g.close()


def test_gradient_zero_weight_samples(gpu_ctx):
    # samples of weight 0 carry no gradient: a batch padded with them gives
    # the gradient of the weighted ones scaled by n_weighted / n
    print("case: guide gradient zero-weight padding", flush=True)
    P = random_params(23)
    g = _guide(gpu_ctx, P)
    S = random_samples(24, 64)
    S2 = np.concatenate([S, random_samples(25, 64)])
    S2[64:, 12] = 0.0
    g1, _ = g.gradient(_t(S))
    g2, _ = g.gradient(_t(S2))
    g1, g2 = g1.cpu().numpy(), g2.cpu().numpy()
    assert np.linalg.norm(2 * g2 - g1) / np.linalg.norm(g1) < 1e-5
    g.close()


def test_mixture_sample_and_pdf_match_oracle(gpu_ctx):
    import wayland_amd as wl
    print("case: guide mixture sample/pdf", flush=True)
    rng = np.random.default_rng(31)
    outs = np.stack([random_outputs(s, kappa_raw=(-8.0, 6.0)) for s in range(64)])
    outs = np.repeat(outs, 64, axis=0)
    u = rng.uniform(size=(outs.shape[0], 3))
    dirs, pdf = wl.mixture_sample(gpu_ctx, _t(outs), _t(u))
    dirs, pdf = dirs.cpu().numpy(), pdf.cpu().numpy()
    o32 = outs.astype(np.float32).astype(np.float64)
    u32 = u.astype(np.float32).astype(np.float64)
    ref_dirs = np.concatenate([ob.guide_sample(o32[i], u32[i:i + 1]) for i in range(len(u))])
    ref_pdf = np.array([ob.guide_pdf(o32[i], ref_dirs[i])[0] for i in range(len(u))])
    derr = np.linalg.norm(dirs - ref_dirs, axis=1)
    perr = np.abs(pdf - ref_pdf) / ref_pdf
    # a uniform within float rounding of a lobe boundary may pick the neighbour
    ok = derr < 1e-4
    print(f"dirs within 1e-4: {ok.mean():.5f}, pdf 99.9% rel {np.quantile(perr[ok], 0.999):.2e}")
    assert ok.mean() > 0.999
    assert np.quantile(perr[ok], 0.999) < 1e-3
    # density of given directions
    p2 = wl.mixture_pdf(gpu_ctx, _t(outs), _t(ref_dirs)).cpu().numpy()
    assert np.quantile(np.abs(p2 - ref_pdf) / ref_pdf, 0.999) < 1e-3


def test_train_step_matches_oracle_adam(gpu_ctx):
    print("case: guide one Adam step", flush=True)
    P = random_params(41)
    lr = 1e-3
    g = _guide(gpu_ctx, P, learning_rate=lr, batch_size=512)
    S = random_samples(42, 300)
    g.train(_t(S), 1)
    P1 = g.get_params().astype(np.float64)
    P0 = P.astype(np.float32).astype(np.float64)
    _, gr = ob.guide_loss_grad(P0, S, emulate_bf16=True)
    Pr, m, v = P0.copy(), np.zeros(NP), np.zeros(NP)
    ob.guide_adam(Pr, m, v, gr, 1, lr, 0.9, 0.99, 1e-8)
    d = np.abs((P1 - P0) - (Pr - P0))
    # first step ~ lr * sign(g): only gradients within rounding of zero differ
    print(f"params off by > 0.1 lr: {(d > 0.1 * lr).mean():.4f}")
    assert (d > 0.1 * lr).mean() < 0.01
    assert info_steps(g) == 1
    # the update kernel also rewrote the packed BF16 operands: inference now
    # follows the new master weights
    Q = random_samples(43, 200)[:, :9]
    out = g.infer(_t(Q)).cpu().numpy()
    ref = ob.guide_forward(P1, Q, emulate_bf16=True)
    assert np.quantile(np.abs(out - ref) / (1.0 + np.abs(ref)), 0.999) < 2e-3
    g.close()


def info_steps(g):
    return g.info()["steps"]


def test_training_learns_a_direction_field(gpu_ctx):
    """Samples drawn from a known position-dependent lobe: training lowers the
    loss and the network's density at the true direction rises far above the
    uniform 1/(4 pi)."""
    import torch
    print("case: guide training on a synthetic direction field", flush=True)
    rng = np.random.default_rng(51)
    g = _guide(gpu_ctx, None, learning_rate=5e-3, batch_size=4096)
    n = 65536
    S = np.zeros((n, 16), np.float32)
    pos = rng.uniform(0, 1, (n, 3))
    S[:, 0:3] = pos
    S[:, 3:6] = [0, 1, 0]
    S[:, 6:9] = [0, 0, 1]
    # true direction: rotates with x; samples within ~0.1 rad of it
    ang = np.pi * pos[:, 0]
    true = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], 1)
    jitter = rng.normal(scale=0.1, size=(n, 3))
    om = true + jitter
    S[:, 9:12] = om / np.linalg.norm(om, axis=1, keepdims=True)
    S[:, 12] = 1.0
    St = _t(S)
    _, loss0 = g.gradient(St[:4096])
    g.train(St, 400)
    _, loss1 = g.gradient(St[:4096])
    loss0, loss1 = float(loss0.cpu()[0]), float(loss1.cpu()[0])
    out = g.infer(St[:1024, :9]).cpu().numpy()
    q_true = np.array([ob.guide_pdf(out[i].astype(np.float64), true[i])[0] for i in range(1024)])
    print(f"loss {loss0:.3f} -> {loss1:.3f}, median density at truth {np.median(q_true):.2f}")
    assert loss1 < loss0 - 2.0
    assert np.median(q_true) > 20 * (1 / (4 * np.pi))
    g.close()


def test_path_tracer_inference_matches_oracle(gpu_ctx):
    """The guide inference the path tracer runs at a guided bounce
    (k_guide_infer_pt: the query built from a ray and its hit, encoding,
    network, decode across lane groups into the stored per-vertex form)
    against the oracle's query, forward pass and decode (oracle_guide.c) on
    camera and bounce rays of the glossy-caustic scene."""
    import torch

    import wayland_amd as wl
    from wayland_amd import scenes
    print("case: k_guide_infer_pt on 3000 rays of the glossy caustic vs oracle", flush=True)
    sc = scenes.glossy_caustic(64, 64, sphere_subdiv=3, floor_n=20)
    accel = wl.Accel(gpu_ctx, sc)
    pt = wl.PathTracer(gpu_ctx, accel, sc)
    P = random_params(31, out_scale=1.0)
    g = _guide(gpu_ctx, P)
    pt.set_guide(g, wl.GUIDE_SAMPLE)
    rng = np.random.default_rng(5)
    n = 3000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform([-1.5, 0.05, -1.5], [1.5, 1.0, 1.5], (n, 3))
    d = rng.normal(size=(n, 3))
    d[:, 1] = -np.abs(d[:, 1])  # mostly towards the floor and the sphere
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = 1e-3, 1e30
    tr = torch.from_numpy(rays).cuda()
    hits = accel.trace_closest(tr)
    dec = pt.guide_infer_hits(tr, hits).cpu().numpy()
    h = wl.hits_to_numpy(hits)
    tris = sc.triangles().reshape(-1, 3, 3)
    v = sc.flatten()["vertices"].reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    inv = (1.0 / (hi - lo)).astype(np.float32)
    ok = h["prim"] != 0xFFFFFFFF
    assert ok.sum() > n // 2
    errs = []
    for i in np.nonzero(ok)[0]:
        t = tris[h["prim"][i]]
        ng = np.cross(t[1] - t[0], t[2] - t[0]).astype(np.float64)
        ng /= np.linalg.norm(ng)
        if ng @ rays[i, 4:7] > 0:
            ng = -ng
        pos = rays[i, 0:3].astype(np.float32) + rays[i, 4:7] * h["t"][i]
        q = np.concatenate([(pos - lo) * inv, ng, -rays[i, 4:7]]).astype(np.float32)
        out = ob.guide_forward(P.astype(np.float32).astype(np.float64), q[None], True)[0]
        pi, kap, mu = ob.guide_decode(out)
        got = dec[i]
        gpi, gkap = got[40:48], got[1:32:4]
        gmu = np.stack([got[2:32:4], got[3:32:4], got[32:40]], 1)
        errs.append(max(np.abs(gpi - pi).max(), (np.abs(gkap - kap) / (1 + kap)).max(),
                        np.abs(gmu - mu).max()))
    errs = np.array(errs)
    print(f"{len(errs)} hits: max err {errs.max():.3e}, 99.9% {np.quantile(errs, 0.999):.3e}")
    assert np.quantile(errs, 0.999) < 5e-3
    assert errs.max() < 5e-2
    # misses leave their slots untouched (zeros)
    assert not dec[~ok].any()
    pt.set_guide(None)
    g.close()
    pt.close()
