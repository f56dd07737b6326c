"""GPU parity of the diagnostics that are off by default in the reference: the
equilateral bispectrum, the k-space forcing cutoff and the log-mapped zeta,
against the oracle from the same uploaded state.

Tolerances: the bispectrum sums the reference's terms over the same pair set in
another order; the terms cancel to a part in 1e3 in some shells, so the shells
agree to 1e-6 relative (plus 1e-7 of the largest shell); shell 0 holds only
round-off of the removed mean and is bounded by 1e-6 of the largest shell.  The
cutoff's FFT round trip differs at 1e-15 of the field.  The LOG map uses the
device log (1 ulp) and tree sums for its two means: 1e-10; its spectrum's k = 0
shell is the squared sum of the map, which cancels, and gets 1e-4."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ie_engine as E
from conftest import golden_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu
LIB = O.lib()


def start(case: str, precision: int = E.IE_FP64, steps: int = 0, sigw: int = 0, **over):
    cfg, params, table = golden_config(case)
    for k, v in over.items():
        setattr(cfg, k, v)
    st = O.State(cfg, table)
    st.initialize()
    if steps:
        LIB.ieo_evolve_fields(st.ptr, 0.5 * cfg.dt)
        LIB.ieo_leapfrog_steps(st.ptr, steps)
    c = st.cfg
    run = E.Run(E.make_config(c.N, precision=precision, numerical=c.numerical_potential, sigw=sigw, table=table,
                              V0=c.V0, ns=c.ns, L=c.L, dt=c.dt, af=c.af, dN=c.dN, Nend=c.Nend,
                              monotonic=c.monotonic), table)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], ad2=bg["ad2"], astep=bg["astep"])
    return st, run


@pytest.mark.parametrize("case,precision", [("tabulated_n16", E.IE_FP64), ("hilltop_n32", E.IE_FP64),
                                            ("tabulated_n16", E.IE_SPLIT)])
def test_bispectrum_matches_oracle(case, precision):
    st, run = start(case, precision, steps=20)
    p, cnt, re, im = run.bispectrum()
    po, co, reo, imo = st.bispectrum()
    assert np.array_equal(cnt, co) and np.array_equal(p, po)
    assert cnt.sum() > 0
    scale = max(np.abs(reo[1:]).max(), np.abs(imo[1:]).max())
    rtol, atol = (1e-6, 1e-7 * scale) if precision == E.IE_FP64 else (1e-3, 1e-4 * scale)
    assert np.allclose(re[1:], reo[1:], rtol=rtol, atol=atol)
    assert np.allclose(im[1:], imo[1:], rtol=rtol, atol=atol)
    assert abs(re[0] - reo[0]) <= 1e-6 * scale and abs(im[0] - imo[0]) <= 1e-6 * scale


@pytest.mark.parametrize("case,precision,low", [("hilltop_n16", E.IE_FP64, 1.2), ("tabulated_n16", E.IE_FP64, 0.0),
                                                ("hilltop_n32", E.IE_SPLIT, 1.2),
                                                ("tabulated_n16", E.IE_SPLIT, 0.0)])
def test_forcing_cutoff_matches_oracle(case, precision, low):
    st, run = start(case, precision, steps=10, high_cutoff_index=5.5, low_cutoff_index=low, forcing_cutoff=1)
    f0 = st.field(0).copy()
    nb = LIB.ieo_spectrum_numbins(st.N)
    # spectraf with the cutoff (src/output.cpp:297-406)
    st.spectrum_f()
    run.forcing_cutoff(5.5, low)
    for which in (E.IE_FIELD_F, E.IE_FIELD_FD):
        ref = st.field(which)
        got = run.field(which)
        tol = 1e-13 if precision == E.IE_FP64 else 2e-7
        scale = np.abs(f0).max() if which == E.IE_FIELD_F else np.abs(ref).max()
        assert np.max(np.abs(got - ref)) <= tol * scale, which
    # the filtered field has no power above the cutoff shell (and none at k = 0
    # with the low cutoff)
    F, ny = O.fftrn(run.field(E.IE_FIELD_F), 1)
    pk, ck, vk = O.bin_spectrum(st.N, st.cfg.L, F, ny, 1.0)
    # (split storage: FP32 rounding of the stored fluctuation refills every mode
    # at 1e-7 of its amplitude)
    assert np.all(vk[7:] <= (1e-20 if precision == E.IE_FP64 else 1e-12) * vk.max())
    assert nb == len(vk)


def test_forcing_cutoff_post_phase_matches_oracle():
    st, run = start("tabulated_n16", E.IE_FP64, steps=10, sigw=1, high_cutoff_index=5.5, forcing_cutoff=1,
                    post_inflation=1)
    run.deltaN_init()
    run.upload_field(E.IE_FIELD_F, st.field(0))
    run.upload_field(E.IE_FIELD_FD, st.field(1))
    rng = np.random.default_rng(1)
    st.field(2)[:] = 0.3 + 1e-3 * rng.standard_normal(st.N ** 3)
    run.upload_field(E.IE_FIELD_DELTAN, st.field(2))
    LIB.ieo_post_initialize(st.ptr)
    c = st.cfg
    run.post_init(omega=c.omega, horizon_factor=c.horizon_factor, dt=c.dt_post, af=c.af_post)
    st.spectrum_f()
    run.forcing_cutoff(5.5, 0.0)
    for which in (E.IE_FIELD_F, E.IE_FIELD_FD):
        ref, got = st.field(which), run.field(which)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max(), which


@pytest.mark.parametrize("precision,eta", [(E.IE_FP64, -0.5), (E.IE_SPLIT, -0.5), (E.IE_FP64, 0.7)])
def test_log_map_matches_oracle(precision, eta):
    st, run = start("tabulated_n16", precision, steps=30, eta_log=eta, nbins=64)
    nbins = 64
    valid, bmin, df, freq = run.deltaN_log(eta, nbins)
    assert valid
    bmo, dfo = C.c_double(), C.c_double()
    fro = np.zeros(nbins)
    assert LIB.ieo_histogram_log(st.ptr, C.byref(bmo), C.byref(dfo), O._dp(fro)) == 1
    tol = 1e-10 if precision == E.IE_FP64 else 1e-5
    assert abs(bmin - bmo.value) <= tol * abs(bmo.value) and abs(df - dfo.value) <= tol * dfo.value
    assert np.max(np.abs(freq - fro)) <= 2.0 / st.N ** 3 + 1e-12
    LIB.ieo_deltaN_log(st.ptr)
    dn_o = st.field(2).copy()
    dn_g = run.field(E.IE_FIELD_DELTAN)
    assert np.allclose(dn_g, dn_o, rtol=tol, atol=tol * np.abs(dn_o).max())
    norm1 = (st.cfg.L / st.cfg.rescale_B) ** 3 / st.N ** 6
    nb = LIB.ieo_spectrum_numbins(st.N)
    po, co, vo = np.zeros(nb), np.zeros(nb, dtype=np.int32), np.zeros(nb)
    LIB.ieo_spectrum_log(st.ptr, O._dp(po), O._ip(co), O._dp(vo))
    p, cnt, v = run.spectrum(E.IE_FIELD_DELTAN, norm1)
    assert np.array_equal(cnt, co)
    assert np.allclose(v[1:], vo[1:], rtol=1e-8 if precision == E.IE_FP64 else 1e-3, atol=0)
    assert abs(v[0] - vo[0]) <= 1e-4 * abs(vo[0])


def test_log_map_invalid_factor():
    st, run = start("tabulated_n16", E.IE_FP64)
    run.upload_fields(st.field(0), np.zeros(st.N ** 3))
    valid, bmin, df, freq = run.deltaN_log(-0.5, 16)
    assert not valid and bmin == 0.0 and df == 1.0 and np.all(freq == 0)
    assert np.all(run.field(E.IE_FIELD_DELTAN) == 0.0)


@pytest.mark.parametrize("precision", [E.IE_FP64, E.IE_SPLIT])
def test_plane_download_is_a_slice_of_the_field(precision):
    st, run = start("tabulated_n16", precision, steps=5)
    N = st.N
    for phase in ("inflation", "deltaN"):
        if phase == "deltaN":
            run.deltaN_init()
            run.upload_field(E.IE_FIELD_DELTAN, np.arange(N ** 3, dtype=np.float64) * 1e-3)
        for which in (E.IE_FIELD_F, E.IE_FIELD_FD) + ((E.IE_FIELD_DELTAN,) if phase == "deltaN" else ()):
            full = run.field(which).reshape(N, N * N)
            for i in (0, 5, N - 1):
                assert np.array_equal(run.plane(which, i), full[i]), (phase, which, i)
    lib = E.lib()
    assert lib.ie_download_plane(run.h, E.IE_FIELD_F, N, E._dp(np.zeros(N * N))) == E.IE_ERR_ARG
