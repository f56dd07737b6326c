"""GPU parity of the RK4 and RK45 (Dormand-Prince) inflation integrators
(src/evolution.cpp:286-904) against the oracle, from the same initial conditions.

Tolerances as for the FP64 leapfrog: background a, ad to 1e-13 relative, fields
as the fluctuation f - <f> to 1e-9 of its rms (the energy sums that feed a'' are
summed in another order), tensors to 1e-4 of their rms (stored in FP32, source
rounding differs).  RK45 must take the same accept/reject decisions and propose
the same next step to 1e-10 relative.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import ie_engine as E
from oracle import oracle as O
from test_gpu_lattice import engine_for, fluct_err, oracle_state

pytestmark = pytest.mark.gpu


def compare(st, run, sigw):
    bo, bg = st.background(), run.background()
    assert bg["t"] == pytest.approx(bo["t"], rel=1e-14)
    assert bg["a"] == pytest.approx(bo["a"], rel=1e-13)
    assert bg["ad"] == pytest.approx(bo["ad"], rel=1e-13)
    assert fluct_err(st.field(0), run.field(E.IE_FIELD_F)) < 1e-9
    assert fluct_err(st.field(1), run.field(E.IE_FIELD_FD)) < 1e-9
    if sigw:
        for comp in (0, 4, 7, 11):
            ho, hg = st.tensor(comp).astype(np.float64), run.tensor(comp).astype(np.float64)
            scale = np.sqrt(np.mean(ho ** 2))
            assert scale > 0
            assert np.max(np.abs(ho - hg)) < 1e-4 * scale, comp


@pytest.mark.parametrize("case,sigw", [("hilltop_n16", 1), ("tabulated_n16", 0), ("hilltop_n32", 0)])
def test_rk4_matches_oracle(case, sigw):
    st, _, table = oracle_state(case, calculate_sigw=sigw, integrator=1)
    run = engine_for(st, table, sigw=sigw)
    h = st.cfg.dt
    for _ in range(25):
        O.lib().ieo_rk4_step(st.ptr, h)
        run.rk4_step(h)
    compare(st, run, sigw)
    assert run.background()["steps"] == 25


@pytest.mark.parametrize("case,sigw", [("hilltop_n16", 1), ("tabulated_n16", 0)])
def test_rk45_matches_oracle(case, sigw):
    st, _, table = oracle_state(case, calculate_sigw=sigw, integrator=2)
    run = engine_for(st, table, sigw=sigw)
    c = st.cfg
    h, hmax = 4.0 * c.dt, c.rk45_max_dt * 4.0
    decisions = []
    for _ in range(20):
        hn_o = ctypes.c_double()
        acc_o = O.lib().ieo_rk45_step(st.ptr, h, hmax, ctypes.byref(hn_o))
        acc_g, hn_g = run.rk45_step(h, hmax)
        assert acc_g == bool(acc_o)
        assert hn_g == pytest.approx(hn_o.value, rel=1e-10)
        decisions.append(acc_g)
        h = hn_o.value
    assert any(decisions)
    compare(st, run, sigw)


def test_rk_requires_fp64_storage():
    st, _, table = oracle_state("hilltop_n16", calculate_sigw=0)
    run = engine_for(st, table, precision=E.IE_SPLIT, sigw=0)
    with pytest.raises(E.EngineError) as ei:
        run.rk4_step(st.cfg.dt)
    assert ei.value.status == E.IE_ERR_STATE
