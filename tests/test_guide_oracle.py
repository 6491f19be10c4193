"""Pinning tests of the oracle's guiding restatement (oracle/oracle_guide.c).

What each oracle function is checked against:
  oracle_bf16_round       torch's float32 -> bfloat16 conversion (round to
                          nearest even), ties included
  oracle_guide_encode     sin^2 + cos^2 = 1 per pair, the double-angle
                          identities between consecutive frequencies, the
                          first frequency against numpy, 42 distinct slots
  oracle_guide_pdf        integral over the sphere = 1 (quadrature), the
                          textbook vMF value at the mean, uniform limit
  oracle_guide_sample     chi-square of a histogram against the pdf
  oracle_guide_loss_grad  central finite differences of the loss (FP64), for
                          weights and biases of every layer, small and large
                          concentrations included; the BF16-emulating mode
                          stays close to the FP64 gradient
  oracle_guide_adam       first step from zero moments moves each parameter
                          by lr * sign(gradient)
"""
import numpy as np
import pytest

import oracle_bind as ob

NP = ob.GUIDE_PARAMS


from guide_util import (b_off, random_outputs, random_params, random_samples, row,
                        sphere_grid, w_off)


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=2000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 2000),
                        # exact ties: low 16 bits 0x8000, odd and even upper halves
                        np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x40490000 | 0x8000],
                                 np.uint32).view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([ob.bf16_round(float(v)) for v in x], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def _feature(g, p, odd):
    """index e of the feature of lane group g, pair p, cosine if odd"""
    return 16 * (p // 2) + 4 * g + 2 * (p % 2) + odd


def test_encode_identities():
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = np.concatenate([rng.uniform(0, 1, 3), rng.normal(size=6)]).astype(np.float32)
        x = ob.guide_encode(q)
        seen = np.zeros(64, bool)
        for d in range(3):
            s = np.array([x[_feature(d, f, 0)] for f in range(6)])
            c = np.array([x[_feature(d, f, 1)] for f in range(6)])
            assert np.allclose(s * s + c * c, 1.0, atol=1e-12)
            # frequency f + 1 doubles the angle of frequency f
            assert np.allclose(s[1:], 2 * s[:-1] * c[:-1], atol=1e-12)
            assert np.allclose(c[1:], c[:-1] ** 2 - s[:-1] ** 2, atol=1e-12)
            assert abs(s[0] - np.sin(np.pi * float(q[d]))) < 1e-12
            assert abs(c[0] - np.cos(np.pi * float(q[d]))) < 1e-12
            assert x[_feature(d, 6, 0)] == q[3 + d] and x[_feature(d, 6, 1)] == q[6 + d]
            for f in range(7):
                seen[_feature(d, f, 0)] = seen[_feature(d, f, 1)] = True
        # 42 distinct features, the rest zero
        assert seen.sum() == 42
        assert np.all(x[~seen] == 0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mixture_pdf_integrates_to_one(seed):
    out = random_outputs(seed)
    dirs, dA = sphere_grid(1200, 2400)
    total = ob.guide_pdf(out, dirs).sum() * dA
    assert abs(total - 1.0) < 2e-4


def test_vmf_textbook_values():
    # one lobe carrying all the weight: value at the mean is the textbook
    # kappa / (2 pi (1 - exp(-2 kappa))), antipode exp(-2 kappa) times that
    for kraw in (-12.0, -2.0, 0.5, 3.0, 8.0):
        out = np.zeros(32)
        for j in range(8):
            out[row(j, 0)] = -200.0
        out[row(3, 0)] = 0.0
        out[row(3, 1)] = kraw
        out[row(3, 2)] = 0.7
        out[row(3, 3)] = 2.1
        kappa = np.log1p(np.exp(kraw))
        mu = np.array([np.sin(0.7) * np.cos(2.1), np.sin(0.7) * np.sin(2.1), np.cos(0.7)])
        at_mu = ob.guide_pdf(out, mu)[0]
        ref = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) if kappa > 1e-6 else 1 / (4 * np.pi)
        assert at_mu == pytest.approx(ref, rel=1e-6)
        assert ob.guide_pdf(out, -mu)[0] == pytest.approx(ref * np.exp(-2 * kappa), rel=1e-6)


@pytest.mark.parametrize("seed", [3, 4])
def test_mixture_sampler_chi2(seed):
    out = random_outputs(seed, kappa_raw=(-1.0, 3.0))
    rng = np.random.default_rng(seed)
    n = 400000
    w = ob.guide_sample(out, rng.uniform(size=(n, 3)))
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-9)
    nz, nphi = 12, 24
    zi = np.clip(((w[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    pi_ = np.clip((np.mod(np.arctan2(w[:, 1], w[:, 0]), 2 * np.pi) / (2 * np.pi) * nphi).astype(int),
                  0, nphi - 1)
    hist = np.bincount(zi * nphi + pi_, minlength=nz * nphi).astype(np.float64)
    # expected bin masses from a 20 x 20 sub-grid quadrature of the pdf
    sub = 20
    dirs, dA = sphere_grid(nz * sub, nphi * sub)
    pdf = ob.guide_pdf(out, dirs).reshape(nz, sub, nphi, sub)
    expected = pdf.sum(axis=(1, 3)).reshape(-1) * dA * n
    mask = expected > 5
    chi2 = (((hist - expected) ** 2 / expected)[mask]).sum()
    dof = mask.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    assert hist[~mask].sum() < 50 + 2 * expected[~mask].sum()


def fd_check(P, S, idx, h=1e-6, rtol=2e-5):
    _, g = ob.guide_loss_grad(P, S, emulate_bf16=False)
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        lp, _ = ob.guide_loss_grad(Pp, S)
        lm, _ = ob.guide_loss_grad(Pm, S)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= rtol * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def test_gradient_matches_finite_differences():
    P = random_params(0)
    S = random_samples(1, 6)
    rng = np.random.default_rng(2)
    idx = []
    for l in range(5):
        n_out = 64 if l < 4 else 32
        idx += list(w_off(l) + rng.integers(0, n_out * 64, 12))
        idx += list(b_off(l) + rng.integers(0, n_out, 4))
    fd_check(P, S, idx)


def test_gradient_small_and_large_kappa():
    P = random_params(5)
    # output biases: lobes alternate tiny and large concentration
    for j in range(8):
        P[b_off(4) + row(j, 1)] = -12.0 if j % 2 else 6.0
    S = random_samples(6, 5)
    idx = [b_off(4) + r for r in range(32)]
    fd_check(P, S, idx)


def test_emulated_gradient_close_to_fp64():
    # BF16 operands perturb the network itself (weights and activations
    # rounded to 8 mantissa bits); the gradient stays within a few percent,
    # most for the layers furthest from the loss.
    P = random_params(7)
    S = random_samples(8, 256)
    l64, g64 = ob.guide_loss_grad(P, S, emulate_bf16=False)
    lbf, gbf = ob.guide_loss_grad(P, S, emulate_bf16=True)
    assert lbf == pytest.approx(l64, rel=1e-2)
    for l, tol in zip(range(5), (0.15, 0.1, 0.05, 0.05, 0.02)):
        n_out = 64 if l < 4 else 32
        sl = slice(w_off(l), b_off(l) + n_out)
        err = np.linalg.norm(gbf[sl] - g64[sl]) / np.linalg.norm(g64[sl])
        assert err < tol, (l, err)


def test_adam_first_step_is_sign_step():
    rng = np.random.default_rng(9)
    P = rng.normal(size=NP)
    g = rng.normal(size=NP) * 10.0 ** rng.uniform(-3, 2, NP)
    P0 = P.copy()
    m, v = np.zeros(NP), np.zeros(NP)
    ob.guide_adam(P, m, v, g, 1, 0.01, 0.9, 0.99, 1e-20)
    assert np.allclose(P - P0, -0.01 * np.sign(g), rtol=1e-6, atol=1e-12)
