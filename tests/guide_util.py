"""Helpers shared by the guiding tests: the parameter layout of the guiding
network and seeded random parameters, samples and network outputs."""
import numpy as np

NP = 18720


def w_off(l):
    return l * (64 * 64 + 64)


def b_off(l):
    return w_off(l) + (64 if l < 4 else 32) * 64


def row(lobe, param):
    return 16 * (lobe % 2) + 4 * (lobe // 2) + param


def random_params(seed, out_scale=0.3):
    rng = np.random.default_rng(seed)
    P = np.zeros(NP)
    for l in range(5):
        n_out = 64 if l < 4 else 32
        s = np.sqrt(6.0 / 64.0) * (1.0 if l < 4 else out_scale)
        P[w_off(l):w_off(l) + n_out * 64] = rng.uniform(-s, s, n_out * 64)
        P[b_off(l):b_off(l) + n_out] = rng.uniform(-0.1, 0.1, n_out)
    return P


def random_samples(seed, n):
    rng = np.random.default_rng(seed)
    S = np.zeros((n, 16), np.float32)
    S[:, 0:3] = rng.uniform(0, 1, (n, 3))
    nrm = rng.normal(size=(n, 3))
    S[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    S[:, 6:9] = wo / np.linalg.norm(wo, axis=1, keepdims=True)
    om = rng.normal(size=(n, 3))
    S[:, 9:12] = om / np.linalg.norm(om, axis=1, keepdims=True)
    S[:, 12] = rng.uniform(0.1, 3.0, n)
    return S


def random_outputs(seed, kappa_raw=(-1.0, 4.0)):
    rng = np.random.default_rng(seed)
    out = np.zeros(32)
    for j in range(8):
        out[row(j, 0)] = rng.normal()
        out[row(j, 1)] = rng.uniform(*kappa_raw)
        out[row(j, 2)] = rng.uniform(-4, 4)
        out[row(j, 3)] = rng.uniform(-7, 7)
    return out


def sphere_grid(nz, nphi):
    z = -1.0 + (np.arange(nz) + 0.5) * (2.0 / nz)
    phi = (np.arange(nphi) + 0.5) * (2 * np.pi / nphi)
    Z, PHI = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - Z * Z)
    dirs = np.stack([r * np.cos(PHI), r * np.sin(PHI), Z], -1).reshape(-1, 3)
    return dirs, (2.0 / nz) * (2 * np.pi / nphi)
