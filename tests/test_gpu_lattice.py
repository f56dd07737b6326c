"""GPU parity of the lattice time step, energies and spectra against the oracle.

Both sides start from the same initial conditions (the oracle's, uploaded through
the C ABI) and take the same steps.  Tolerances:
  - FP64 storage: background a, ad to 1e-13 relative; fields compared as the
    fluctuation f - <f>, to 1e-9 of its rms (sum order and FMA contraction differ
    from the serial reference; both are ulp-level per step);
  - energies of identical fields: 1e-12 relative; spectra: 1e-8 per shell (the
    k = 0 power of f = phi0 + fluctuation dominates and its FFT rounding leaks
    into the other shells, in the reference as here);
  - GW spectra: 2e-5 relative per bin (the oracle, like the reference, sums shells
    in float, the engine in FP64; measured <= 8e-6);
  - split storage (FP32 fluctuation about an FP64 homogeneous mode) against FP64:
    P(k) per bin to 1e-6 relative after 200 steps (measured <= 3e-7).
Spectra are compared with no absolute floor (assert_rel prints the largest error).
"""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E
from conftest import assert_rel, golden_config, potential_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_state(case: str, N: int | None = None, **over):
    cfg, params, table = golden_config(case)
    if N is not None and N != cfg.N:
        cfg = O.make_config(N, cfg.numerical_potential, params, table)
    for k, v in over.items():
        setattr(cfg, k, v)
    st = O.State(cfg, table)
    st.initialize()
    return st, params, table


def engine_for(st: O.State, table, precision=E.IE_FP64, sigw=None):
    c = st.cfg
    cfg = E.make_config(c.N, precision=precision, numerical=c.numerical_potential,
                        sigw=c.calculate_sigw if sigw is None else sigw, table=table,
                        V0=c.V0, ns=c.ns, L=c.L, dt=c.dt, af=c.af, rescale_s=c.rescale_s, rescale_B=c.rescale_B,
                        dN=c.dN, Nend=c.Nend)
    run = E.Run(cfg, table)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], astep=bg["astep"], hubble_init=bg["hubble_init"])
    return run


def fluct_err(a, b):
    da = a - a.mean()
    return np.max(np.abs(da - (b - b.mean()))) / np.sqrt(np.mean(da ** 2))


@pytest.mark.parametrize("case", ["hilltop_n32", "tabulated_n16"])
def test_energies_match_oracle(case):
    st, _, table = oracle_state(case)
    run = engine_for(st, table)
    g_o = O.lib().ieo_gradient_energy(st.ptr)
    k_o = O.lib().ieo_kin_energy(st.ptr)
    p_o = O.lib().ieo_potential_energy(st.ptr)
    assert run.gradient_energy() == pytest.approx(g_o, rel=1e-12)
    assert run.kin_energy() == pytest.approx(k_o, rel=1e-12)
    assert run.potential_energy() == pytest.approx(p_o, rel=1e-13)
    s = run.field_sums()
    f = st.field(0)
    assert s[0] == pytest.approx(f.sum(), rel=1e-13)
    assert s[2] == pytest.approx((f * f).sum(), rel=1e-13)


# N < 64: site kernels; N = 64: row kernels (one site per lane); 128 <= N <= 256:
# fused steps (march kernel with the next step's sums, two and four sites per lane)
@pytest.mark.parametrize("case,N,nsteps", [("hilltop_n32", 32, 200), ("tabulated_n16", 16, 200),
                                           ("hilltop_n32", 64, 100), ("tabulated_n16", 64, 40),
                                           ("tabulated_n16", 128, 20), ("hilltop_n32", 128, 40),
                                           ("hilltop_n32", 256, 6), ("tabulated_n16", 256, 4)])
def test_leapfrog_fp64_matches_oracle(case, N, nsteps):
    print(f"leapfrog parity {case} N={N} steps={nsteps}", flush=True)
    st, _, table = oracle_state(case, N=N, calculate_sigw=0)
    run = engine_for(st, table, sigw=0)
    dt = st.cfg.dt
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * dt)
    run.evolve_fields(0.5 * dt)
    n_o = O.lib().ieo_leapfrog_steps(st.ptr, nsteps)
    b = run.leapfrog_steps(nsteps)
    bo = st.background()
    assert b["steps"] == n_o == nsteps
    for key in ("a", "ad", "t", "astep"):
        assert b[key] == pytest.approx(bo[key], rel=1e-13), key
    assert b["ad2"] == pytest.approx(bo["ad2"], rel=1e-9)
    assert fluct_err(st.field(0), run.field(E.IE_FIELD_F)) < 1e-9
    assert fluct_err(st.field(1), run.field(E.IE_FIELD_FD)) < 1e-9


def test_explicit_kick_and_drift_match_oracle():
    st, _, table = oracle_state("hilltop_n16", calculate_sigw=0)
    run = engine_for(st, table, sigw=0)
    for d in (0.0005, 0.0007):
        O.lib().ieo_evolve_derivs(st.ptr, d)
        run.evolve_derivs(d)
        O.lib().ieo_evolve_fields(st.ptr, -0.5 * d)
        run.evolve_fields(-0.5 * d)
    bo, b = st.background(), run.background()
    for key in ("a", "ad", "ad2", "t"):
        assert b[key] == pytest.approx(bo[key], rel=1e-13), key
    assert fluct_err(st.field(1), run.field(E.IE_FIELD_FD)) < 1e-10


# N = 128: fused steps stop when a step kernel closes the record past af
@pytest.mark.parametrize("N", [16, 128])
def test_loop_stops_at_af(N):
    st, _, table = oracle_state("hilltop_n16", N=N, calculate_sigw=0, af=1.01)
    run = engine_for(st, table, sigw=0)
    run.evolve_fields(0.5 * st.cfg.dt)
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * st.cfg.dt)
    n_o = O.lib().ieo_leapfrog_steps(st.ptr, 100000)
    b = run.leapfrog_steps(n_o + 50)
    assert b["done"] == 1 and b["steps"] == n_o
    assert b["a"] == pytest.approx(st.background()["a"], rel=1e-13)


@pytest.mark.parametrize("case", ["hilltop_n32", "tabulated_n16"])
def test_spectrum_matches_oracle(case):
    st, _, table = oracle_state(case)
    run = engine_for(st, table)
    norm1 = (st.cfg.L / st.cfg.rescale_B) ** 3 / st.cfg.N ** 6
    p, c, v = run.spectrum(E.IE_FIELD_F, norm1)
    po, co, vo = st.spectrum_f()
    assert np.array_equal(c, co)
    assert np.allclose(p, po, rtol=1e-15)
    # FP64 FFTs of f = phi0 + fluctuation: the k = 0 power exceeds the others by
    # up to 1e14, and its rounding leaks into every shell at ~1e-16 x that ratio
    # in amplitude (in the reference's FFTEASY as here), so shells agree to 1e-8
    assert v[0] == pytest.approx(vo[0], rel=1e-13)
    assert np.allclose(v[1:], vo[1:], rtol=1e-8, atol=0)


# N = 16: site kernels; N = 64: row kernels (one site per lane); N = 128: fused
# steps with the tensor step taking the march kernel's scalars
@pytest.mark.parametrize("N,nsteps", [(16, 100), (64, 30), (128, 8)])
def test_sigw_tensors_and_gw_spectra_match_oracle(N, nsteps):
    print(f"sigw parity N={N} steps={nsteps}", flush=True)
    st, _, table = oracle_state("hilltop_n16", N=N)
    run = engine_for(st, table, sigw=1)
    dt = st.cfg.dt
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * dt)
    run.evolve_fields(0.5 * dt)
    O.lib().ieo_leapfrog_steps(st.ptr, nsteps)
    run.leapfrog_steps(nsteps)
    for comp in (0, 3, 5, 6, 9, 11):
        ho, hg = st.tensor(comp).astype(np.float64), run.tensor(comp).astype(np.float64)
        scale = np.sqrt(np.mean(ho ** 2))
        assert scale > 0
        assert np.max(np.abs(ho - hg)) < 1e-4 * scale, comp
    for hdot in (0, 1):
        po, co, vo = st.gw_spectrum(hdot)
        p, c, v = run.gw_spectrum(hdot)
        assert np.array_equal(c, co)
        assert np.array_equal(p, po)
        live = vo > 0
        assert live.sum() > 5
        assert_rel(v[live], vo[live], 2e-5, f"GW spectrum hdot={hdot}")
        assert np.all(v[~live] == 0)


def test_upload_modes_matches_oracle_initial_conditions():
    cfg, _, table = golden_config("hilltop_n32")
    st = O.State(cfg, table)
    assert O.lib().ieo_fill_modes(st.ptr) == 0
    fk, fny, fdk, fdny = (st.field(q).copy() for q in (0, 3, 1, 4))
    st2 = O.State(cfg, table)
    st2.initialize()
    run = E.Run(E.make_config(32, sigw=0, V0=cfg.V0, ns=cfg.ns, dt=cfg.dt), None)
    run.upload_modes(fk, fny, fdk, fdny, cfg.initial_field, cfg.initial_derivative)
    f, fd = run.field(E.IE_FIELD_F), run.field(E.IE_FIELD_FD)
    # FFTEASY advances its twiddles by a rotation recurrence, the GPU transform
    # uses exact twiddles: 1e-11 of the fluctuation rms apart
    assert fluct_err(st2.field(0), f) < 1e-10
    assert fluct_err(st2.field(1), fd) < 1e-10


@pytest.mark.parametrize("N", [32, 64, 128])
def test_split_storage_matches_fp64_spectra(N):
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=0)
    r64 = engine_for(st, table, precision=E.IE_FP64, sigw=0)
    r32 = engine_for(st, table, precision=E.IE_SPLIT, sigw=0)
    dt = st.cfg.dt
    for r in (r64, r32):
        r.evolve_fields(0.5 * dt)
        r.leapfrog_steps(200)
    b64, b32 = r64.background(), r32.background()
    assert b32["a"] == pytest.approx(b64["a"], rel=1e-12)
    norm1 = (st.cfg.L / st.cfg.rescale_B) ** 3 / N ** 6
    _, _, v64 = r64.spectrum(E.IE_FIELD_F, norm1)
    _, _, v32 = r32.spectrum(E.IE_FIELD_F, norm1)
    assert_rel(v32[1:], v64[1:], 1e-6, f"split vs FP64 spectrum N={N}")
    assert v32[0] == pytest.approx(v64[0], rel=1e-12)


@pytest.mark.parametrize("precision", [E.IE_FP64, E.IE_SPLIT])
def test_segmented_rows_match_oracle_n512(precision):
    """N = 512: each row is two 256-site segments per wave; the k +- 1 neighbours
    across segment edges come from the row itself.  One explicit kick and drift of
    seeded random fields against the oracle (FP64: 1e-10 of the fluctuation rms;
    split storage: FP32 rounding of the stored fluctuation, 1e-6)."""
    N = 512
    print(f"segmented rows N={N} precision={precision}", flush=True)
    cfg, params, _ = golden_config("hilltop_n32")
    cfg = O.make_config(N, 0, params, None)
    cfg.calculate_sigw = 0
    st = O.State(cfg, None)
    rng = np.random.default_rng(512)
    f0 = cfg.initial_field + 1e-5 * rng.standard_normal(N ** 3)
    fd0 = cfg.initial_derivative + 1e-5 * rng.standard_normal(N ** 3)
    st.field(0)[:] = f0
    st.field(1)[:] = fd0
    st.set_background(t=0.0, a=1.0, ad=0.01, astep=1.0)
    run = E.Run(E.make_config(N, precision=precision, sigw=0, V0=cfg.V0, ns=cfg.ns, L=cfg.L, dt=cfg.dt,
                              af=cfg.af, rescale_s=cfg.rescale_s, rescale_B=cfg.rescale_B), None)
    run.upload_fields(f0, fd0)
    run.set_background(t=0.0, a=1.0, ad=0.01, astep=1.0)
    d = cfg.dt
    O.lib().ieo_evolve_derivs(st.ptr, d)
    O.lib().ieo_evolve_fields(st.ptr, d)
    run.evolve_derivs(d)
    run.evolve_fields(d)
    bo, b = st.background(), run.background()
    for key in ("a", "ad", "t"):
        assert b[key] == pytest.approx(bo[key], rel=1e-12), key
    tol = 1e-10 if precision == E.IE_FP64 else 1e-6
    assert fluct_err(st.field(0), run.field(E.IE_FIELD_F)) < tol
    assert fluct_err(st.field(1), run.field(E.IE_FIELD_FD)) < tol
    run.close()


# tensor shape codes the tuner tries: row kernel 1/2/4 sites per lane, plane
# march 16 + E (E = 4: one row per wave, FP32 rows), early-load march 24 + E
TUNER_TENSOR_CODES = {128: [1, 2, 17, 18, 25, 26], 256: [1, 2, 4, 17, 18, 20, 25, 26]}


@pytest.mark.parametrize("N", [128, 256])
def test_tuned_shapes_leave_results_unchanged(N):
    """The tuner times real loop steps with the tensor step as the row kernel and
    as the plane-marching kernels (with and without early loads), and both
    variants of the fused step kernel; none changes an operation or a sum order,
    so a run that changes shape between steps matches one that never does, bit
    for bit."""
    print(f"tuned shapes N={N}", flush=True)
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=1)
    ra = engine_for(st, table, precision=E.IE_SPLIT, sigw=1)
    rb = engine_for(st, table, precision=E.IE_SPLIT, sigw=1)
    dt = st.cfg.dt
    for r in (ra, rb):
        r.evolve_fields(0.5 * dt)
    t = ra.tune(4)
    codes = TUNER_TENSOR_CODES[N]
    assert t["tensor_codes"] == codes and t["steps"] == 4 * len(codes)
    assert t["tensor_code"] in codes and t["march_variant"] in (0, 1)
    assert all(x > 0 for x in t["tensor_ms"] + t["march_ms"]) and len(t["march_ms"]) == 2
    ra.leapfrog_steps(12)
    rb.leapfrog_steps(12 + 4 * len(codes))
    assert ra.background()["steps"] == rb.background()["steps"]
    assert np.array_equal(ra.field(E.IE_FIELD_F), rb.field(E.IE_FIELD_F))
    assert np.array_equal(ra.field(E.IE_FIELD_FD), rb.field(E.IE_FIELD_FD))
    for comp in range(12):
        assert np.array_equal(ra.tensor(comp), rb.tensor(comp)), comp


@pytest.mark.parametrize("N,chunks", [(64, (10, 10, 10)), (128, (10, 10, 10)), (128, (7, 1, 6, 5)),
                                      (256, (3, 4))])
def test_loop_in_chunks_equals_one_call(N, chunks):
    """A loop split into several ie_leapfrog_steps calls gives the same bits as one
    call: fused steps (N >= 128) start each call with the sums of the current f in
    a step's own partition and order, and end it by closing the last step record
    (odd and even chunk lengths end in either record slot)."""
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=0)
    ra = engine_for(st, table, precision=E.IE_SPLIT, sigw=0)
    rb = engine_for(st, table, precision=E.IE_SPLIT, sigw=0)
    for r in (ra, rb):
        r.evolve_fields(0.5 * st.cfg.dt)
    for n in chunks:
        ra.leapfrog_steps(n)
    total = sum(chunks)
    rb.leapfrog_steps(total)
    ba, bb = ra.background(), rb.background()
    assert ba["steps"] == bb["steps"] == total
    assert ba["a"] == bb["a"] and ba["ad"] == bb["ad"]
    assert np.array_equal(ra.field(E.IE_FIELD_F), rb.field(E.IE_FIELD_F))
    assert np.array_equal(ra.field(E.IE_FIELD_FD), rb.field(E.IE_FIELD_FD))


@pytest.mark.parametrize("N,sigw", [(64, 0), (128, 0), (128, 1)])
def test_profile_steps_are_loop_steps(N, sigw):
    """ie_profile_steps runs the loop's own launches, twice n steps (events around
    each launch, then around the loop alone), so it leaves the state of 2n loop
    steps, bit for bit, and reports positive times."""
    print(f"profile steps N={N} sigw={sigw}", flush=True)
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=sigw)
    ra = engine_for(st, table, precision=E.IE_SPLIT, sigw=sigw)
    rb = engine_for(st, table, precision=E.IE_SPLIT, sigw=sigw)
    for r in (ra, rb):
        r.evolve_fields(0.5 * st.cfg.dt)
    p = ra.profile_steps(5)
    rb.leapfrog_steps(10)
    assert p["steps"] == 10 and ra.background()["steps"] == 10
    assert p["step_ms"] > 0 and p["kick_drift_ms"] > 0 and p["loop_ms"] > 0
    assert (p["tensor_ms"] > 0) == bool(sigw)
    assert (p["reduce_ms"] > 0) == (N < 128)
    assert np.array_equal(ra.field(E.IE_FIELD_F), rb.field(E.IE_FIELD_F))
    assert np.array_equal(ra.field(E.IE_FIELD_FD), rb.field(E.IE_FIELD_FD))
    for comp in range(12 if sigw else 0):
        assert np.array_equal(ra.tensor(comp), rb.tensor(comp)), comp
