"""Tabulated potential on the fused-step path against the oracle, where the
lookups leave the common case:

  - in the dense stretch of the inputs/ grid (phi near 2.683, spacing 3.6e-6
    against a field spread of about 2e-5): a wave's sites lie in tens of
    intervals, so most values take the per-value path through the index the step
    kernel stages in LDS (N = 128: the records read from L2; N = 256: the records
    staged in LDS too);
  - with a table too fine for that LDS copy (the inputs/ table resampled on a
    uniform grid of 40001 points: its index exceeds the 48 KB the kernel reserves),
    so the index is read from global memory.

Same tolerances as the other leapfrog parity tests: a, a', t to 1e-13 relative,
the fluctuation f - <f> to 1e-9 of its rms.
"""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E
from conftest import golden_config, potential_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fluct_err(a, b):
    da = a - a.mean()
    return np.max(np.abs(da - (b - b.mean()))) / np.sqrt(np.mean(da ** 2))


def run_pair(table, phi0: float, N: int = 128, nsteps: int = 20, precision=E.IE_FP64, variant=None):
    _, params, _ = golden_config("tabulated_n16")
    params = dict(params, initial_field=phi0)
    cfg = O.make_config(N, 1, params, table, sigw=0)
    st = O.State(cfg, table)
    st.initialize()
    c = st.cfg
    ecfg = E.make_config(N, precision=precision, numerical=1, sigw=0, table=table, V0=c.V0, ns=c.ns, L=c.L,
                         dt=c.dt, af=c.af, rescale_s=c.rescale_s, rescale_B=c.rescale_B, dN=c.dN, Nend=c.Nend)
    run = E.Run(ecfg, table)
    if variant is not None:
        run.set_shapes(0, 0, variant)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], astep=bg["astep"], hubble_init=bg["hubble_init"])
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * c.dt)
    run.evolve_fields(0.5 * c.dt)
    n_o = O.lib().ieo_leapfrog_steps(st.ptr, nsteps)
    b = run.leapfrog_steps(nsteps)
    return st, run, b, n_o


def check(st, run, b, n_o, nsteps, label, split=False):
    # split storage keeps the fluctuations in FP32: their bound, and the
    # background's through the lattice sums
    bo = st.background()
    assert b["steps"] == n_o == nsteps
    for key in ("a", "ad", "t"):
        assert b[key] == pytest.approx(bo[key], rel=1e-12 if split else 1e-13), key
    ef, ev = fluct_err(st.field(0), run.field(E.IE_FIELD_F)), fluct_err(st.field(1), run.field(E.IE_FIELD_FD))
    print(f"{label}: a rel {abs(b['a'] - bo['a']) / bo['a']:.2e}, f {ef:.2e}, fd {ev:.2e}", flush=True)
    tol = 1e-5 if split else 1e-9
    assert ef < tol and ev < tol


# N = 256 with each fused-step variant pinned: rows of old f from L2 (1, the
# records in LDS beside the index) or through LDS (0, the index alone)
@pytest.mark.parametrize("N,variant,precision", [(128, None, E.IE_FP64), (256, None, E.IE_FP64),
                                                 (256, 0, E.IE_FP64), (256, 0, E.IE_SPLIT),
                                                 (256, 1, E.IE_SPLIT)])
def test_dense_stretch_of_the_table_matches_oracle(N, variant, precision):
    table = potential_table()
    x = table.x
    spacing = -np.diff(x)[np.searchsorted(-x, -2.6827)]
    assert spacing < 5e-6          # the dense stretch this test is about
    print(f"dense stretch N={N} variant={variant} precision={precision}", flush=True)
    st, run, b, n_o = run_pair(table, 2.6827, N=N, nsteps=20 if N <= 128 else 8, precision=precision,
                               variant=variant)
    f = st.field(0)
    print(f"phi0 2.6827: field spread {f.std():.2e}, grid spacing {spacing:.2e}", flush=True)
    check(st, run, b, n_o, 20 if N <= 128 else 8, f"dense stretch N={N}", split=precision == E.IE_SPLIT)
    run.close()


def test_table_beyond_the_lds_index_matches_oracle():
    t = potential_table()
    xf = np.linspace(t.x[0], t.x[-1], 40001)
    # the inputs/ table resampled (ascending for numpy.interp, then descending
    # like the reference's grid)
    v = np.interp(xf[::-1], t.x[::-1], t.v[::-1])[::-1]
    dv = np.interp(xf[::-1], t.x[::-1], t.dv[::-1])[::-1]
    table = O.Table(xf, v, dv)
    st, run, b, n_o = run_pair(table, 2.9181235049318586)
    check(st, run, b, n_o, 20, "40001-point table")
    run.close()
