"""GPU parity of the post-inflation phase against the oracle (which reproduces
the reference's results/post_inflation files bit for bit).

Both sides start from the same state (f, fd and the deltaN map uploaded to the
engine's FP64 deltaN-phase state); the post-inflation set-up, the leapfrog steps
(Phi kick, tensor kick with the Hessian source, drifts) and the RK steps keep the
reference's operation order without FMA contraction, so Phi, Phi', h_ij, h_ij'
and the background are compared bit for bit, for FP64 and split runs (whose
post-inflation state is FP64 too).  The outputs (sums, spectra) are compared to
round-off: their summation order differs."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ie_engine as E
from conftest import assert_rel, load_golden, potential_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIB = O.lib()


def pre_post_state(N: int, seed: int = 3):
    """Oracle state at the tabulated preset's parameters with a synthetic deltaN
    map: a smooth part (the initial f fluctuations) plus white noise."""
    g = load_golden("tabulated_n16_post")
    params = O.parse_params_text(str(g["params_text"]))
    table = potential_table()
    cfg = O.make_config(N, 1, params, table, post=1)
    st = O.State(cfg, table)
    st.initialize()
    f = st.field(0)
    rng = np.random.default_rng(seed)
    df = f - f.mean()
    dn = 0.7 + 2e-3 * df / (np.abs(df).max() + 1e-300) + 1e-4 * rng.standard_normal(f.size)
    st.field(2)[:] = dn
    return st, table


def engine_twin(st: O.State, table, precision: int) -> E.Run:
    c = st.cfg
    run = E.Run(E.make_config(c.N, precision=precision, numerical=1, sigw=1, table=table, V0=c.V0, ns=c.ns,
                              L=c.L, dt=c.dt, af=c.af, dN=c.dN, Nend=c.Nend, monotonic=c.monotonic), table)
    run.upload_fields(st.field(0), st.field(1))
    bg = st.background()
    run.set_background(t=bg["t"], a=bg["a"], ad=bg["ad"], ad2=bg["ad2"], astep=bg["astep"])
    run.deltaN_init()
    run.upload_field(E.IE_FIELD_F, st.field(0))
    run.upload_field(E.IE_FIELD_FD, st.field(1))
    run.upload_field(E.IE_FIELD_DELTAN, st.field(2))
    return run


def assert_same_state(st, run, label):
    assert np.array_equal(run.field(E.IE_FIELD_F), st.field(0)), f"{label}: Phi"
    assert np.array_equal(run.field(E.IE_FIELD_FD), st.field(1)), f"{label}: Phi'"
    for comp in range(12):
        assert np.array_equal(run.tensor(comp), st.tensor(comp)), f"{label}: tensor {comp}"
    bo, bg = st.background(), run.background()
    for k in ("t", "a", "ad", "ad2"):
        assert bg[k] == bo[k], f"{label}: {k} {bg[k]!r} != {bo[k]!r}"


def start_post(N, precision, from_deltaN=1):
    st, table = pre_post_state(N)
    if not from_deltaN:
        # the oracle's set-up branch follows its perform_deltaN toggle
        cfg = st.cfg
        cfg.perform_deltaN = 0
        st2 = O.State(cfg, table)
        for w in (0, 1, 2):
            st2.field(w)[:] = st.field(w)
        st2.set_background(**st.background())
        st = st2
    run = engine_twin(st, table, precision)
    c = st.cfg
    LIB.ieo_post_initialize(st.ptr)
    run.post_init(omega=c.omega, horizon_factor=c.horizon_factor, dt=c.dt_post, af=c.af_post, integrator=0,
                  from_deltaN=from_deltaN)
    return st, run


# N = 16, 32: site kernels; 64, 128, 256: the tensor march (1, 2, 4 sites per
# lane); rows=True pins a row tensor code, which keeps the row tensor kernel (the
# path of N >= 512: 2 sites per lane, two segments per row at N = 256)
@pytest.mark.parametrize("N,precision,rows", [(16, E.IE_FP64, False), (32, E.IE_FP64, False),
                                              (64, E.IE_FP64, False), (128, E.IE_FP64, False),
                                              (256, E.IE_FP64, False), (16, E.IE_SPLIT, False),
                                              (128, E.IE_SPLIT, False), (64, E.IE_FP64, True),
                                              (256, E.IE_FP64, True)])
def test_post_leapfrog_bitwise(N, precision, rows):
    st, run = start_post(N, precision)
    if rows:
        run.set_shapes(tensor_code=1)
    assert_same_state(st, run, "set-up")
    assert np.abs(st.field(0)).max() > 0
    c = st.cfg
    dt = c.dt_post
    # run_post_inflation_loop (src/evolution.cpp:1596-1604)
    LIB.ieo_evolve_fields(st.ptr, 0.5 * dt)
    run.evolve_fields(0.5 * dt)
    nsteps = 4
    for _ in range(nsteps):
        LIB.ieo_post_evolve_derivs(st.ptr, dt)
        LIB.ieo_evolve_fields(st.ptr, dt)
    bg = run.post_leapfrog_steps(nsteps)
    assert bg["steps"] == nsteps and not bg["done"]
    assert_same_state(st, run, f"after {nsteps} steps")
    assert np.abs(st.tensor(6)).max() > 0 and np.abs(st.tensor(11)).max() > 0
    # one kick alone (evolve_derivs_post_inflation)
    LIB.ieo_post_evolve_derivs(st.ptr, dt)
    run.post_evolve_derivs(dt)
    assert_same_state(st, run, "kick")


def test_post_setup_without_deltaN_bitwise():
    st, run = start_post(16, E.IE_FP64, from_deltaN=0)
    assert_same_state(st, run, "set-up from f")
    assert np.abs(st.field(0)).max() > 0


# Maps for the set-up's lattice-order mean (k_serial_sums), which takes runs of
# steps inside one binade of the accumulator as a prefix scan: many binade
# changes (centered maps), rounding ties (values on a 2^-10 grid once the sum
# reaches 2^48; a centred map of dN-grid values, like saveN's zeta), sign flips,
# zero and non-finite accumulators, huge and tiny values.
def serial_map(name: str, gs: int, rng) -> np.ndarray:
    if name == "centered":
        return rng.standard_normal(gs) * 1e-5
    if name == "centered drift":
        return rng.standard_normal(gs) * 1e-5 + 3e-12
    if name == "ties":
        return rng.integers(0, 2 ** 40, gs) / 1024.0
    if name == "ties signed":
        return rng.integers(-2 ** 40, 2 ** 40, gs) / 1024.0
    if name == "grid halves":
        return rng.integers(0, 64, gs) * 2.0 ** -20 - 2.0 ** -15
    if name == "alternating":
        return np.where(np.arange(gs) % 2 == 0, 1.0, -1.0)
    if name == "zeros then values":
        return np.concatenate([np.zeros(gs // 2), rng.uniform(0, 1e-3, gs - gs // 2)])
    if name == "huge and tiny":
        return np.where(rng.random(gs) < 0.001, 1e300, rng.uniform(-1, 1, gs) * 1e-310)
    if name == "lognormal":
        return rng.lognormal(-7, 3, gs)
    if name == "dN grid centred":
        # like saveN's zeta: deltaN values taken from an accumulated dN = 1e-4
        # grid (plus the half step), each many times, less a float mean
        grid = np.add.accumulate(np.full(14000, 1e-4)) + 5e-5
        return grid[rng.integers(0, 14000, gs)] - np.float64(np.float32(0.5352471470832825))
    m = rng.standard_normal(gs)
    m[gs - 1000] = np.nan if name == "nan late" else np.inf
    return m


SERIAL_MAPS = ["centered", "centered drift", "ties", "ties signed", "grid halves", "alternating",
               "zeros then values", "huge and tiny", "lognormal", "dN grid centred", "nan late", "inf late"]


@pytest.mark.parametrize("name", SERIAL_MAPS)
def test_post_setup_mean_bitwise_on_crafted_maps(name):
    N = 64
    st, table = pre_post_state(N)
    st.field(2)[:] = serial_map(name, N ** 3, np.random.default_rng(11))
    print(f"set-up mean N={N} map {name}", flush=True)
    run = engine_twin(st, table, E.IE_FP64)
    c = st.cfg
    LIB.ieo_post_initialize(st.ptr)
    run.post_init(omega=c.omega, horizon_factor=c.horizon_factor, dt=c.dt_post, af=c.af_post, integrator=0,
                  from_deltaN=1)
    assert np.array_equal(run.field(E.IE_FIELD_F), st.field(0), equal_nan=True), name
    run.close()


def test_post_setup_two_means_bitwise_on_crafted_maps():
    # Phi from f without deltaN: lattice-order means of f and f' (two sums)
    N = 64
    st, table = pre_post_state(N)
    rng = np.random.default_rng(12)
    cfg = st.cfg
    cfg.perform_deltaN = 0
    st2 = O.State(cfg, table)
    # f stays inside the potential table's range (the deltaN set-up evaluates
    # it); f' takes the crafted values
    st2.field(0)[:] = st.field(0) + serial_map("centered drift", N ** 3, rng)
    st2.field(1)[:] = serial_map("ties signed", N ** 3, rng)
    st2.field(2)[:] = st.field(2)
    st2.set_background(**st.background())
    run = engine_twin(st2, table, E.IE_FP64)
    c = st2.cfg
    LIB.ieo_post_initialize(st2.ptr)
    run.post_init(omega=c.omega, horizon_factor=c.horizon_factor, dt=c.dt_post, af=c.af_post, integrator=0,
                  from_deltaN=0)
    assert np.array_equal(run.field(E.IE_FIELD_F), st2.field(0))
    run.close()


def test_post_loop_stops_at_af():
    st, table = pre_post_state(16)
    run = engine_twin(st, table, E.IE_FP64)
    c = st.cfg
    af = 1.05
    LIB.ieo_post_initialize(st.ptr)
    run.post_init(omega=c.omega, horizon_factor=c.horizon_factor, dt=c.dt_post, af=af)
    LIB.ieo_evolve_fields(st.ptr, 0.5 * c.dt_post)
    run.evolve_fields(0.5 * c.dt_post)
    n = 0
    while st.background()["a"] <= af:
        LIB.ieo_post_evolve_derivs(st.ptr, c.dt_post)
        LIB.ieo_evolve_fields(st.ptr, c.dt_post)
        n += 1
    bg = run.post_leapfrog_steps(10 ** 6)
    assert 0 < n == bg["steps"] and bg["done"] == 1
    assert bg["a"] == st.background()["a"]
    assert_same_state(st, run, "loop end")


@pytest.mark.parametrize("N", [16, 64])
def test_post_rk4_bitwise(N):
    st, run = start_post(N, E.IE_FP64 if N == 16 else E.IE_SPLIT)
    h = st.cfg.dt_post
    for _ in range(2):
        LIB.ieo_post_rk4_step(st.ptr, h)
        run.rk4_step(h)
    assert_same_state(st, run, "rk4")


def test_post_rk45_bitwise_on_the_oracle_steps():
    st, run = start_post(32, E.IE_FP64)
    c = st.cfg
    h = c.dt_post
    hmax = min(c.rk45_max_dt, max(c.rk45_min_dt, 2.0 * c.dt_post))
    accepted = 0
    for _ in range(3):
        hn = C.c_double()
        acc_o = LIB.ieo_post_rk45_step(st.ptr, h, hmax, C.byref(hn))
        acc_g, _hn_g = run.rk45_step(h, hmax)
        assert bool(acc_o) == acc_g
        accepted += acc_g
        h = hn.value
    assert accepted >= 1
    assert_same_state(st, run, "rk45")


def test_post_outputs_match_oracle():
    st, run = start_post(32, E.IE_FP64)
    c = st.cfg
    LIB.ieo_evolve_fields(st.ptr, 0.5 * c.dt_post)
    run.evolve_fields(0.5 * c.dt_post)
    for _ in range(3):
        LIB.ieo_post_evolve_derivs(st.ptr, c.dt_post)
        LIB.ieo_evolve_fields(st.ptr, c.dt_post)
    run.post_leapfrog_steps(3)
    f = st.field(0).copy()
    s = run.field_sums()
    ref = [f.sum(), st.field(1).sum(), (f * f).sum()]
    assert np.allclose(s, ref, rtol=1e-12, atol=0)
    # histogram of Phi (src/output.cpp:1768-1836): min, max and the bin counts
    nb = 64
    lo, hi, counts = run.histogram(E.IE_FIELD_F, nb)
    f = st.field(0)
    assert lo == f.min() and hi == f.max()
    df = (hi - lo) / nb
    b = ((f - lo) / df).astype(np.int64)
    b[f == hi] = nb - 1
    assert np.array_equal(counts, np.bincount(b[(b >= 0) & (b < nb)], minlength=nb))
    # spectra (after the histogram: the oracle's in-place transforms perturb f
    # and h at round-off): the scalar one of Phi, the tensor ones with the float
    # factor of spectraGWdot_post_inflation
    norm1 = (c.L / c.rescale_B) ** 3 / c.N ** 6
    po, co, vo = st.spectrum_f()
    p, cnt, v = run.spectrum(E.IE_FIELD_F, norm1)
    assert np.array_equal(cnt, co) and np.allclose(p, po, rtol=1e-15)
    assert np.allclose(v, vo, rtol=1e-9, atol=0)
    for hdot, ohdot in ((0, 0), (1, 2)):
        po, co, vo = st.gw_spectrum(ohdot)
        p, cg, vg = run.gw_spectrum(hdot)
        assert np.array_equal(cg, co) and np.array_equal(p, po)
        live = vo > 0
        assert live.sum() > 5
        assert_rel(vg[live], vo[live], 1e-5, f"post GW spectrum hdot={hdot}")


def test_post_phase_refuses_inflation_steps():
    st, run = start_post(16, E.IE_FP64)
    lib = E.lib()
    b = E.Background()
    assert lib.ie_leapfrog_steps(run.h, 1, C.byref(b)) == E.IE_ERR_STATE
    assert lib.ie_evolve_derivs(run.h, 0.001) == E.IE_ERR_STATE
    assert lib.ie_deltaN_steps(run.h, 0.001, 1) == E.IE_ERR_STATE
