"""GPU parity of the deltaN phase (initializeN, the phiref surface, the
register-resident e-fold integrator and the saveN zeta map) against the oracle,
from the same end-of-inflation state.  The site arithmetic keeps the reference's
operation order without FMA contraction, so the deltaN map is compared bit for
bit; the zeta spectrum to 1e-12 relative per bin (FFT and sum order differ)."""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E
from conftest import assert_rel, golden_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def end_of_inflation(case: str, Nend: float):
    cfg, params, table = golden_config(case)
    cfg.Nend = Nend
    st = O.State(cfg, table)
    st.initialize()
    dt = cfg.dt
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * dt)
    O.lib().ieo_leapfrog_steps(st.ptr, 10 ** 7)
    bg = st.background()
    O.lib().ieo_evolve_fields(st.ptr, -0.5 * dt * bg["astep"] ** (cfg.rescale_s - 1.0))
    return st, table


def float_mean(dn, cutoff):
    acc = np.float32(0.0)
    for v in dn:
        if v <= cutoff:
            acc = np.float32(np.float64(acc) + v)
    return np.float32(acc / np.float32(dn.size))


@pytest.mark.parametrize("case,Nend", [("tabulated_n16", 0.6), ("hilltop_n16", 0.0002)])
def test_deltaN_map_bitwise(case, Nend):
    st, table = end_of_inflation(case, Nend)
    c = st.cfg
    run = E.Run(E.make_config(c.N, numerical=c.numerical_potential, sigw=0, table=table, V0=c.V0, ns=c.ns,
                              L=c.L, dt=c.dt, af=c.af, dN=c.dN, Nend=c.Nend, monotonic=c.monotonic), table)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], astep=bg["astep"], phiref=0.0)

    # oracle: run_deltaN_loop (src/evolution.cpp:1296-1328)
    O.lib().ieo_initializeN(st.ptr)
    O.lib().ieo_evolve_fieldsN(st.ptr, 0.5 * c.dN)
    phiref = O.lib().ieo_get_phiref(st.ptr)
    st.set_background(phiref=phiref)
    Ne, nsteps = 0.0, 0
    while Ne <= c.Nend:
        O.lib().ieo_evolve_derivsN(st.ptr, c.dN)
        O.lib().ieo_evolve_fieldsN(st.ptr, c.dN)
        Ne += c.dN
        nsteps += 1

    run.deltaN_init()
    run.evolve_fieldsN(0.5 * c.dN)
    assert run.get_phiref() == phiref
    run.set_background(phiref=phiref)
    run.deltaN_steps(c.dN, nsteps)

    dn_o, dn_g = st.field(2).copy(), run.field(E.IE_FIELD_DELTAN)
    assert np.array_equal(dn_o, dn_g)
    assert np.array_equal(st.field(0), run.field(E.IE_FIELD_F))
    # fd of finished sites keeps evolving (evolve_derivsN runs on every site)
    assert np.array_equal(st.field(1), run.field(E.IE_FIELD_FD))
    finished = dn_o < dn_o.max()
    assert 0 < finished.sum()

    # saveN (src/output.cpp:1522-1556)
    O.lib().ieo_evolve_fieldsN(st.ptr, -0.5 * c.dN)
    run.evolve_fieldsN(-0.5 * c.dN)
    cutoff = c.Nend - c.dN
    dn_o = st.field(2).copy()
    m_o = float_mean(dn_o, cutoff)
    m_g = run.deltaN_zeta(cutoff)
    assert np.float32(m_g) == m_o
    dn_o = dn_o - np.float64(m_o)
    assert np.array_equal(run.field(E.IE_FIELD_DELTAN), dn_o)
    run.deltaN_zero_unfinished(cutoff)
    dn_o[dn_o > cutoff - np.float64(m_o)] = 0.0
    assert np.array_equal(run.field(E.IE_FIELD_DELTAN), dn_o)
    run.deltaN_center()
    dn_o -= dn_o.sum() / dn_o.size
    norm1 = (c.L / c.rescale_B) ** 3 / c.N ** 6
    F, ny = O.fftrn(dn_o, +1)
    po, co, vo = O.bin_spectrum(c.N, c.L, F, ny, norm1)
    p, cc, v = run.spectrum(E.IE_FIELD_DELTAN, norm1)
    assert np.array_equal(cc, co)
    # shell 0 holds the k = 0 mode of a mean-subtracted map: round-off, bounded
    # against the spectrum's scale
    live = vo[1:] > 0
    assert_rel(v[1:][live], vo[1:][live], 1e-13, "P_deltaN shells 1..")
    assert abs(v[0] - vo[0]) <= 1e-12 * vo[1:].max()


@pytest.mark.parametrize("case,Nend", [("tabulated_n16", 0.3), ("hilltop_n16", 0.0001)])
def test_deltaN_rk4_map_bitwise(case, Nend):
    """RK4 deltaN integrator (src/evolution.cpp:1176-1189) kept in registers."""
    st, table = end_of_inflation(case, Nend)
    c = st.cfg
    run = E.Run(E.make_config(c.N, numerical=c.numerical_potential, sigw=0, table=table, V0=c.V0, ns=c.ns,
                              L=c.L, dt=c.dt, af=c.af, dN=c.dN, Nend=c.Nend, monotonic=c.monotonic), table)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], astep=bg["astep"], phiref=0.0)
    O.lib().ieo_initializeN(st.ptr)
    phiref = O.lib().ieo_get_phiref(st.ptr)
    st.set_background(phiref=phiref)
    Ne, nsteps = 0.0, 0
    while Ne <= c.Nend:
        O.lib().ieo_rk4_stepN(st.ptr, c.dN)
        Ne += c.dN
        nsteps += 1
    run.deltaN_init()
    assert run.get_phiref() == phiref
    run.set_background(phiref=phiref)
    run.deltaN_rk4_steps(c.dN, nsteps)
    assert np.array_equal(st.field(2), run.field(E.IE_FIELD_DELTAN))
    assert np.array_equal(st.field(0), run.field(E.IE_FIELD_F))
    # fd of finished sites keeps evolving (evolve_derivsN runs on every site)
    assert np.array_equal(st.field(1), run.field(E.IE_FIELD_FD))
    assert np.array_equal(st.field(1), run.field(E.IE_FIELD_FD))
    assert run.background()["a"] == st.background()["a"]


@pytest.mark.parametrize("case,Nend", [("tabulated_n16", 0.3), ("hilltop_n16", 0.0001)])
def test_deltaN_rk45_matches_oracle(case, Nend):
    """Dormand-Prince deltaN trial steps (src/evolution.cpp:1191-1267).  Both sides
    take the oracle's step sequence, so the per-site maps agree bit for bit; the
    accept decisions must agree and the proposed steps to 1e-10 (the error norm is
    summed in another order)."""
    import ctypes
    st, table = end_of_inflation(case, Nend)
    c = st.cfg
    run = E.Run(E.make_config(c.N, numerical=c.numerical_potential, sigw=0, table=table, V0=c.V0, ns=c.ns,
                              L=c.L, dt=c.dt, af=c.af, dN=c.dN, Nend=c.Nend, monotonic=c.monotonic), table)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], astep=bg["astep"], phiref=0.0)
    O.lib().ieo_initializeN(st.ptr)
    phiref = O.lib().ieo_get_phiref(st.ptr)
    st.set_background(phiref=phiref)
    run.deltaN_init()
    assert run.get_phiref() == phiref
    run.set_background(phiref=phiref)
    h, hmax = c.dN, 8.0 * c.dN
    accepted = 0
    for _ in range(30):
        hn = ctypes.c_double()
        acc_o = O.lib().ieo_rk45_stepN(st.ptr, h, hmax, ctypes.byref(hn))
        acc_g, hn_g = run.deltaN_rk45_step(h, hmax)
        assert acc_g == bool(acc_o)
        assert hn_g == pytest.approx(hn.value, rel=1e-10)
        accepted += acc_g
        h = hn.value
    assert accepted > 0
    assert np.array_equal(st.field(2), run.field(E.IE_FIELD_DELTAN))
    assert np.array_equal(st.field(0), run.field(E.IE_FIELD_F))
    # fd of finished sites keeps evolving (evolve_derivsN runs on every site)
    assert np.array_equal(st.field(1), run.field(E.IE_FIELD_FD))
    assert np.array_equal(st.field(1), run.field(E.IE_FIELD_FD))
    assert run.background()["a"] == st.background()["a"]


def test_histogram_matches_reference_rule():
    st, table = end_of_inflation("hilltop_n16", 0.0002)
    c = st.cfg
    run = E.Run(E.make_config(c.N, sigw=0, V0=c.V0, ns=c.ns, dt=c.dt), None)
    run.upload_fields(st.field(0), st.field(1))
    lo, hi, counts = run.histogram(E.IE_FIELD_F, 256)
    f = st.field(0)
    assert lo == f.min() and hi == f.max()
    df = (hi - lo) / 256
    b = ((f - lo) / df).astype(int)
    b[f == hi] = 255
    ref = np.bincount(b[(b >= 0) & (b < 256)], minlength=256)
    assert np.array_equal(counts, ref)
