"""CPU check of the batched-serving schedule and of the index arithmetic the
multi-request kernels derive from it (wayland_amd/csrc/serve_plan.h compiled
for the host in libwl_hostemu.so): every path and pixel of every megabatch
maps to an item that contains it, every derived index stays inside the batch
buffers and the request's film, and every (pixel, sample) of every request
is rendered exactly once, with a pixel's samples reaching its film in
ascending frame order (what makes each film equal its standalone render)."""
import numpy as np
import pytest

import hostemu_bind as hb
from wayland_amd import scenes


def _paths(reqs):
    return sum(w * h * s for w, h, s, _ in reqs)


@pytest.mark.parametrize("capacity", [1 << 12, 1 << 14, 40_000])
def test_random_request_mixes(capacity):
    rng = np.random.default_rng(capacity)
    for trial in range(20):
        n = int(rng.integers(1, 30))
        reqs = []
        for k in range(n):
            w = int(rng.integers(1, 64))
            h = int(rng.integers(1, min(64, capacity // w) + 1))
            reqs.append((w, h, int(rng.integers(0, 12)), int(rng.integers(0, 1 << 20))))
        bad, nb, items, paths, maxp = hb.serve_check(reqs, capacity)
        assert bad == 0, (trial, reqs)
        assert paths == _paths(reqs)
        assert maxp <= capacity
        assert nb >= -(-paths // capacity)


def test_edge_cases():
    # a request filling the whole megabatch; one pixel; zero samples; many
    # one-pixel requests; requests that only fit one sample per megabatch
    cap = 4096
    cases = [
        [(64, 64, 3, 0)],
        [(1, 1, 1, 7)],
        [(8, 8, 0, 0), (8, 8, 2, 100)],
        [(1, 1, 5, k) for k in range(300)],
        [(63, 65, 4, 0), (2, 2, 9, 1), (64, 63, 2, 2)],
    ]
    for reqs in cases:
        bad, nb, items, paths, maxp = hb.serve_check(reqs, cap)
        assert bad == 0, reqs
        assert paths == _paths(reqs)


def test_serving_workload_schedule():
    """The 64-request serving mix at the batch sizes the autotuner tries."""
    reqs = scenes.serving_requests(np.random.default_rng(5), 64, 128, 1024, 64, 0)
    req4 = [(r["width"], r["height"], r["spp"], r["frame_offset"]) for r in reqs]
    # schedule shape only (the full walk of 826M paths is left to the
    # smaller capacities above); the largest request needs 1024^2 pixels
    small = [(w // 16, h // 16, s, f) for w, h, s, f in req4]
    for cap in (1 << 12, 1 << 14, 1 << 16):
        bad, nb, items, paths, maxp = hb.serve_check(small, cap)
        assert bad == 0
        assert paths == _paths(small)
