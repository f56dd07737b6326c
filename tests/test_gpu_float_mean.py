"""saveN's float mean of a deltaN map (src/output.cpp:1527-1537) on the GPU.

The reference adds the sites with deltaN <= cutoff into a float in lattice order
(float += double) and divides by N^3 in float; the rounding of every step is part
of the result, and the mean decides which sites are zeroed as unfinished.  The
engine takes the steps of each binade of the accumulator as an integer prefix sum
and every other step as a float step; ie_deltaN_zeta's mean must equal the
oracle's serial walk (ieo_float_mean) bit for bit, and the map must come back as
deltaN - mean.  Maps: uniform and lognormal values with some sites above the
cutoff, values on a power-of-two grid (ties of the rounding) mixed with negative
values, zeros, NaN and large jumps, a run of excluded sites ahead of the rest,
and a run of zero sites (+0 and -0) ahead of the rest.
"""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def maps(gs, rng):
    yield "uniform", rng.uniform(0.0, 1e-3, gs), 9.5e-4
    yield "lognormal", rng.lognormal(-7.0, 1.0, gs), 2e-3
    t = rng.integers(0, 64, gs) * 2.0 ** -20
    t[::997] = -1e-4
    t[::1499] = 0.0
    t[::2003] = np.nan
    t[::50001] = 3.0
    yield "ties and specials", t, 1.0
    h = rng.uniform(0.0, 1e-3, gs)
    h[: gs // 2] = 5.0
    yield "excluded head", h, 1e-3
    # sites finished before the first step (deltaN = 0) ahead of the rest: a zero
    # accumulator takes them without a step, -0 included
    z = rng.uniform(0.0, 1e-3, gs)
    z[: gs // 2] = 0.0
    z[1: gs // 2: 7] = -0.0
    yield "zero head", z, 1e-3


@pytest.mark.parametrize("N", [64, 128, 256])
def test_float_mean_equals_serial_walk(N):
    rng = np.random.default_rng(100 + N)
    cfg = E.make_config(N, precision=E.IE_FP64, V0=3.338e-13, ns=0.97, dt=0.001)
    run = E.Run(cfg)
    try:
        for name, dn, cutoff in maps(N ** 3, rng):
            print(f"float mean N={N} {name}", flush=True)
            run.upload_field(E.IE_FIELD_DELTAN, dn)
            got = np.float32(run.deltaN_zeta(cutoff))
            want = O.float_mean(dn, cutoff)
            assert got.tobytes() == want.tobytes(), (name, got, want)
            zeta = run.field(E.IE_FIELD_DELTAN)
            assert np.array_equal(zeta, dn - np.float64(want), equal_nan=True)
    finally:
        run.close()
