"""The run-time chunk schedule of the raw trace kernels (traverse_persistent.h,
trav_persistent with DYN = 2) against a brute-force simulation of its fetch
rule.  XCD x owns rays [x * ceil(n / 8), (x + 1) * ceil(n / 8)) clipped to n;
its W waves (W = 4 per workgroup, grid / 8 workgroups per XCD) start on
chunks 0 .. W - 1 of 256 rays and, each time a wave's chunk is used up, it
takes chunk W + k, k the value its fetch returned from the XCD's counter
(less the launch's base); a fetch past the last chunk ends the wave.  For
every interleaving of the fetches, each ray must be taken exactly once and
the counter must advance by the library's dyn_xcd_chunks(n, x, 256), the
amount the host adds to the base it keeps for the next launch on the stream.
"""
import numpy as np
import pytest

import hostemu_bind as hb

CHUNK = 256
WAVES_PER_BLOCK = 4


def simulate(n, grid, rng):
    n8 = (n + 7) // 8
    taken = np.zeros(n, np.int32)
    advance = []
    for x in range(8):
        lo = min(x * n8, n)
        hi = min(lo + n8, n)
        wavesx = (grid // 8) * WAVES_PER_BLOCK
        counter = 0
        working = []
        for wx in range(wavesx):  # static first chunks
            s0 = lo + wx * CHUNK
            if s0 < hi:
                taken[s0:min(s0 + CHUNK, hi)] += 1
                working.append(wx)
        while working:  # fetches in a random order of the waves
            i = int(rng.integers(len(working)))
            k = counter
            counter += 1
            s0 = lo + (wavesx + k) * CHUNK
            if s0 < hi:
                taken[s0:min(s0 + CHUNK, hi)] += 1
            else:
                working.pop(i)
        advance.append(counter)
    return taken, advance


@pytest.mark.parametrize("grid", [8, 16, 1536, 1544])
def test_every_ray_once_and_counter_advance(grid):
    L = hb.lib()
    rng = np.random.default_rng(grid)
    for n in (1, 7, 8, 255, 256, 257, 2047, 2048, 2049, 65537, 1 << 20, 3 << 21):
        taken, advance = simulate(n, grid, rng)
        assert (taken == 1).all(), (n, grid)
        for x in range(8):
            assert advance[x] == L.wlh_dyn_chunks(n, x, CHUNK), (n, grid, x)
