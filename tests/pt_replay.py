"""Replay of path-tracer samples that differ from the oracle: the GPU's
per-vertex record of a path (wl_pt_trace_samples) against the oracle's
(oracle_pt_trace_f32), classified by where and how they first diverge.

Rows of a record: depth, mesh, pdf, mean f|cos|/pdf, mean throughput
reaching the vertex, -, -, trans (one row per surface vertex that sampled a
direction)."""
import numpy as np

import oracle_bind as ob


def close(a, b, tol=1e-4):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-6)


def records(tr):
    """Rows of a per-vertex record up to its first empty slot."""
    n = 0
    while n < tr.shape[0] and tr[n, 0] != 0:
        n += 1
    return tr[:n]


def classify(gt, ot, g_rgb, o_rgb):
    """First diverging vertex of two per-vertex records -> (kind, vertex):
    "hit"   a different surface (or depth) at this vertex,
    "lobe"  the same vertex, but the sampled direction's pdf or weight jumps,
    "end"   one path ends earlier (Russian roulette, a zero or NaN pdf, a
            light, the depth cap),
    "drift" the same branches so far with throughputs more than 1e-4 apart
            (an arithmetic difference, not a branch flip),
    and with equal records: "visibility" (the samples differ by more than
    1e-3: a whole contribution, a shadow ray or the last extension ray
    hitting on one side only) or "rounding" (less: rounding along a long
    path)."""
    n = min(len(gt), len(ot))
    for k in range(n):
        g, o = gt[k], ot[k]
        if int(g[0]) != int(o[0]) or int(g[1]) != int(o[1]):
            return "hit", k
        if not close(g[4], o[4]):
            return "drift", k
        if not (close(g[2], o[2]) and close(g[3], o[3])):
            return "lobe", k
    if len(gt) != len(ot):
        return "end", n
    rel = np.abs(g_rgb - o_rgb) / np.maximum(np.abs(o_rgb), 1e-6)
    return ("visibility" if rel.max() > 1e-3 else "rounding"), n


def replay_frames(pt, opt, cfg, ocfg, frames, maxv=24, rel_bar=1e-5):
    """Per-sample GPU radiance vs the oracle over `frames`; every sample
    outside rel_bar replayed and classified.  Returns (kinds: kind -> list of
    (frame, x, y, vertex, meshes of the oracle path's vertices up to and
    including that vertex), number of samples, per-sample differences GPU -
    oracle as an (N, 3) float64 array)."""
    kinds, diffs, total = {}, [], 0
    for frame in frames:
        g, tr = pt.trace_samples(cfg, frame, maxv)
        g = g.cpu().numpy()
        tr = tr.cpu().numpy()
        o, _ = opt.render_frame(ocfg, frame)
        assert np.isfinite(g).all()
        diffs.append((g.astype(np.float64) - o).reshape(-1, 3))
        total += g.shape[0] * g.shape[1]
        rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-6)
        ys, xs = np.nonzero(rel.max(-1) > rel_bar)
        for y, x in zip(ys, xs):
            _, otr = opt.trace(ocfg, int(x), int(y), frame, maxv)
            kind, k = classify(records(tr[y, x]), otr, g[y, x], o[y, x])
            meshes = tuple(int(m) for m in otr[:k + 1, 1])
            kinds.setdefault(kind, []).append((frame, int(x), int(y), k, meshes))
    return kinds, total, np.concatenate(diffs)


def paired_z(d):
    """Per-channel z of the mean GPU - oracle difference."""
    return d.mean(0) / np.maximum(d.std(0) / np.sqrt(len(d)), 1e-30)
