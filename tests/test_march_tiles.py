"""The fused step kernel's tiling on the CPU (march_body, march_ti and march_ok in
inflation-easy_amd/csrc/ie_lattice.hip): workgroup b runs in the slab of planes of
XCD b mod 8, 8 owner waves (rows j0..j0+7) and a ninth that recomputes row j0+8,
marching TI + 1 plane steps from plane i0.  Restated here to check the invariants
the kernel's sums rely on: every site is stored by exactly one wave, and the
k, j and i pairs of -sum f lapl f = sum_x sum_d (f(x + e_d) - f(x))^2 are each
counted exactly once; the plane loop, unrolled by three, needs TI + 1 to be a
multiple of three."""
from __future__ import annotations

import numpy as np
import pytest

TILE_J, XCD = 8, 8


def march_ti(N: int) -> int:
    return N // XCD if N >= 256 else 8


@pytest.mark.parametrize("N", [128, 256])
def test_march_tiles_cover_sites_and_pairs_once(N):
    TI = march_ti(N)
    assert (TI + 1) % 3 == 0
    jblocks = N // TILE_J
    G = jblocks * (N // TI)
    assert G % XCD == 0
    rng = np.random.default_rng(N)
    f = rng.standard_normal((N, N, N))
    stored = np.zeros((N, N), dtype=np.int32)
    pairs = {d: np.zeros((N, N), dtype=np.int32) for d in "ijk"}
    total = 0.0
    for blk in range(G):
        xcd, lb = blk % XCD, blk // XCD
        i0 = xcd * (N // XCD) + (lb // jblocks) * TI
        assert 0 <= i0 and i0 + TI <= N
        for w in range(TILE_J + 1):
            j = ((lb % jblocks) * TILE_J + w) % N
            if w == TILE_J:
                continue   # the ninth wave stores nothing and sums nothing
            for t in range(TI + 1):
                i = (i0 + t) % N
                if t < TI:
                    stored[i, j] += 1
                    pairs["k"][i, j] += 1
                    pairs["j"][i, j] += 1
                    total += np.sum((np.roll(f[i, j], -1) - f[i, j]) ** 2)
                    total += np.sum((f[i, (j + 1) % N] - f[i, j]) ** 2)
                if t >= 1:
                    pairs["i"][(i - 1) % N, j] += 1
                    total += np.sum((f[i, j] - f[(i - 1) % N, j]) ** 2)
    assert (stored == 1).all()
    for d in "ijk":
        assert (pairs[d] == 1).all(), d
    lap = sum(np.roll(f, s, a) for a in range(3) for s in (1, -1)) - 6 * f
    ref = -np.sum(f * lap)
    assert abs(total - ref) <= 1e-10 * abs(ref)
