"""GPU parity at the sizes of the benchmark's own workloads, against the oracle.

Both sides start from the oracle's initial conditions (the reference's
initializer restated, seed 8 of the presets, dt = 0.001) unless stated.
  - 256^3 hilltop, split storage (the benchmark's headline loop: the fused step
    kernel with four sites per lane, both variants run by the tuner): 20 steps;
    a, ad, t to 1e-12; P(k) in the 222 shells to 1e-4 per shell against the
    oracle's FP64 run; the fluctuation within the FP32 storage bound;
  - 256^3 hilltop, FP64 storage: 20 steps, the FP64 tolerances of the smaller
    cases (a to 1e-13, fields to 1e-9 of the fluctuation rms);
  - 256^3 hilltop, split storage with SIGW, after the tuner (every tensor shape
    of the tuner, its choice included, takes steps): the six h_ij and their
    momenta and the TT spectrum of h;
  - 128^3 tabulated USR preset, FP64 and split storage, 50 steps; then from the
    oracle's end state the deltaN map bit for bit (100 deltaN steps), the saveN
    zeta map and P_zeta in the 111 shells and in 64 log-spaced bins; P_zeta of
    the split run's own deltaN phase against the oracle's;
  - 256 runs of 64^3 in one batch (seeds 1000 + r, V0 log-uniform in
    [1e-11, 1e-9], ns uniform in [0.95, 0.98]; the preset's deltaN window scaled
    by sqrt(V0 / V0_preset)), whole runs with per-run P_zeta:
    8 sampled runs bitwise equal to the same runs alone, one against the oracle;
  - rescale_s = 1 (the step's scale-factor powers through the device pow).
Each test prints the errors it measures.
"""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E
from oracle import oracle as O
from test_gpu_lattice import engine_for, fluct_err, oracle_state

pytestmark = pytest.mark.gpu


def shell_rel(v, vo, first=1):
    live = vo[first:] > 0
    return np.abs(v[first:][live] - vo[first:][live]) / vo[first:][live]


def log_bins(k_index_values: np.ndarray, nbins: int = 64):
    """Means of shell values (shells 1 .. n-1) over nbins log-spaced bins in the
    shell index; bins narrower than a shell stay empty (returned as NaN)."""
    n = len(k_index_values)
    shell = np.arange(1, n)
    which = np.clip(np.digitize(shell, np.logspace(0.0, np.log10(n), nbins + 1)) - 1, 0, nbins - 1)
    out = np.full(nbins, np.nan)
    for q in range(nbins):
        sel = shell[which == q]
        if sel.size:
            out[q] = k_index_values[sel].mean()
    return out


def step_both(st, run, nsteps, tune_steps_per_candidate=0):
    dt = st.cfg.dt
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * dt)
    run.evolve_fields(0.5 * dt)
    taken = 0
    if tune_steps_per_candidate:
        t = run.tune(tune_steps_per_candidate)
        taken = t["steps"]
    if nsteps > taken:
        run.leapfrog_steps(nsteps - taken)
    O.lib().ieo_leapfrog_steps(st.ptr, max(nsteps, taken))
    return max(nsteps, taken)


def test_split_256_headline_loop_matches_oracle():
    N = 256
    print(f"split 256 headline loop vs oracle", flush=True)
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=0)
    run = engine_for(st, table, precision=E.IE_SPLIT, sigw=0)
    n = step_both(st, run, 20, tune_steps_per_candidate=2)
    bo, b = st.background(), run.background()
    assert b["steps"] == n == 20
    for key in ("a", "ad", "t"):
        assert b[key] == pytest.approx(bo[key], rel=1e-12), key
    ef, efd = fluct_err(st.field(0), run.field(E.IE_FIELD_F)), fluct_err(st.field(1), run.field(E.IE_FIELD_FD))
    norm1 = (st.cfg.L / st.cfg.rescale_B) ** 3 / N ** 6
    _, c, v = run.spectrum(E.IE_FIELD_F, norm1)
    _, co, vo = st.spectrum_f()
    rel = shell_rel(v, vo)
    print(f"split 256: fluct err f {ef:.3e} fd {efd:.3e}; P(k) max rel {rel.max():.3e} over {rel.size} shells",
          flush=True)
    assert np.array_equal(c, co) and len(c) == 222
    assert rel.max() < 1e-4
    assert ef < 1e-5 and efd < 1e-5


def test_fp64_256_loop_matches_oracle():
    N = 256
    print(f"fp64 256 loop vs oracle", flush=True)
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=0)
    run = engine_for(st, table, precision=E.IE_FP64, sigw=0)
    step_both(st, run, 20)
    bo, b = st.background(), run.background()
    for key in ("a", "ad", "t", "astep"):
        assert b[key] == pytest.approx(bo[key], rel=1e-13), key
    ef, efd = fluct_err(st.field(0), run.field(E.IE_FIELD_F)), fluct_err(st.field(1), run.field(E.IE_FIELD_FD))
    print(f"fp64 256: fluct err f {ef:.3e} fd {efd:.3e}", flush=True)
    assert ef < 1e-9 and efd < 1e-9


def test_sigw_split_256_matches_oracle():
    N = 256
    print(f"sigw split 256 vs oracle", flush=True)
    st, _, table = oracle_state("hilltop_n32", N=N, calculate_sigw=1)
    run = engine_for(st, table, precision=E.IE_SPLIT, sigw=1)
    dt = st.cfg.dt
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * dt)
    run.evolve_fields(0.5 * dt)
    t = run.tune(1)
    run.leapfrog_steps(4)
    O.lib().ieo_leapfrog_steps(st.ptr, t["steps"] + 4)
    assert run.background()["steps"] == t["steps"] + 4
    assert 20 in t["tensor_codes"]
    worst = 0.0
    for comp in range(12):
        ho, hg = st.tensor(comp).astype(np.float64), run.tensor(comp).astype(np.float64)
        scale = np.sqrt(np.mean(ho ** 2))
        assert scale > 0
        worst = max(worst, np.max(np.abs(ho - hg)) / scale)
    po, co, vo = st.gw_spectrum(0)
    p, c, v = run.gw_spectrum(0)
    live = vo > 0
    rel = np.abs(v[live] - vo[live]) / vo[live]
    print(f"sigw 256 (tensor shape {t['tensor_code']}): h max err {worst:.3e} of rms; GW P_h max rel "
          f"{rel.max():.3e} over {live.sum()} shells", flush=True)
    assert np.array_equal(c, co) and np.array_equal(p, po)
    assert live.sum() >= 170 and np.all(v[~live] == 0)
    assert worst < 1e-3
    assert rel.max() < 2e-4


def oracle_deltaN(st):
    """run_deltaN_loop and saveN's map in the oracle (src/evolution.cpp:1296-1328,
    src/output.cpp:1522-1556): the centred zeta map, phiref, steps."""
    n = O.lib().ieo_deltaN_zeta_map(st.ptr, None)
    return st.field(2).copy(), st.background()["phiref"], n


def engine_deltaN(run, c):
    run.deltaN_init()
    run.set_background(phiref=0.0)
    run.evolve_fieldsN(0.5 * c.dN)
    phiref = run.get_phiref()
    run.set_background(phiref=phiref)
    Ne, n = 0.0, 0
    while Ne <= c.Nend:
        Ne += c.dN
        n += 1
    run.deltaN_steps(c.dN, n)
    run.evolve_fieldsN(-0.5 * c.dN)
    cutoff = c.Nend - c.dN
    run.deltaN_zeta(cutoff)
    run.deltaN_zero_unfinished(cutoff)
    run.deltaN_center()
    return phiref, n


def pzeta(N, L, B, vals):
    k = O.modes(N, L, B)
    m = min(len(k), len(vals))
    return O.pzeta_deltaN(k[:m], vals[:m])


def test_tabulated_128_loop_and_deltaN_match_oracle():
    N = 128
    print("tabulated 128: loop (FP64, split) and deltaN vs oracle", flush=True)
    st, _, table = oracle_state("tabulated_n16", N=N, calculate_sigw=0, Nend=0.01)
    r64 = engine_for(st, table, precision=E.IE_FP64, sigw=0)
    r32 = engine_for(st, table, precision=E.IE_SPLIT, sigw=0)
    step_both(st, r64, 50)
    dt = st.cfg.dt
    r32.evolve_fields(0.5 * dt)
    r32.leapfrog_steps(50)
    bo = st.background()
    for r, tol_f in ((r64, 1e-9), (r32, 1e-5)):
        b = r.background()
        for key in ("a", "ad", "t"):
            assert b[key] == pytest.approx(bo[key], rel=1e-12 if r is r32 else 1e-13), key
        ef = fluct_err(st.field(0), r.field(E.IE_FIELD_F))
        print(f"tabulated 128 {'split' if r is r32 else 'fp64'}: fluct err f {ef:.3e}", flush=True)
        assert ef < tol_f
    # the end of inflation as the loop leaves it (src/evolution.cpp:993-995)
    c = st.cfg
    back = -0.5 * dt * bo["astep"] ** (c.rescale_s - 1.0)
    O.lib().ieo_evolve_fields(st.ptr, back)
    for r in (r64, r32):
        r.evolve_fields(back)
    # deltaN from the oracle's own end state, uploaded (FP64 storage): bit for bit
    rd = engine_for(st, table, precision=E.IE_FP64, sigw=0)
    zeta_o, phiref_o, n_o = oracle_deltaN(st)
    phiref_g, n_g = engine_deltaN(rd, c)
    assert phiref_g == phiref_o and n_g == n_o and n_o > 90
    zg = rd.field(E.IE_FIELD_DELTAN)
    assert np.count_nonzero(zeta_o) > 0
    # the engine's map before centring is bitwise the oracle's; centring sums in
    # another order (1 ulp of the mean)
    assert np.max(np.abs(zg - zeta_o)) <= 1e-15 * max(1.0, np.abs(zeta_o).max())
    norm1 = (c.L / c.rescale_B) ** 3 / N ** 6
    F, ny = O.fftrn(zeta_o, +1)
    po, co, vo = O.bin_spectrum(N, c.L, F, ny, norm1)
    p, cc, v = rd.spectrum(E.IE_FIELD_DELTAN, norm1)
    assert np.array_equal(cc, co) and len(cc) == 111
    pz_o, pz_g = pzeta(N, c.L, c.rescale_B, vo), pzeta(N, c.L, c.rescale_B, v)
    rel = shell_rel(pz_g, pz_o)
    lo, lg = log_bins(pz_o), log_bins(pz_g)
    ok = np.isfinite(lo) & (lo > 0)
    lrel = np.abs(lg[ok] - lo[ok]) / lo[ok]
    print(f"tabulated 128 deltaN: P_zeta max rel {rel.max():.3e} (111 shells), {lrel.max():.3e} "
          f"({ok.sum()} of 64 log bins)", flush=True)
    assert rel.max() < 1e-10 and lrel.max() < 1e-10
    # the split run's own deltaN phase (FP64 state rebuilt from its FP32
    # fluctuation) against the oracle's FP64 run
    engine_deltaN(r32, c)
    _, _, v32 = r32.spectrum(E.IE_FIELD_DELTAN, norm1)
    rel32 = shell_rel(pzeta(N, c.L, c.rescale_B, v32), pz_o)
    l32 = log_bins(pzeta(N, c.L, c.rescale_B, v32))
    lrel32 = np.abs(l32[ok] - lo[ok]) / lo[ok]
    print(f"tabulated 128 split deltaN: P_zeta max rel {rel32.max():.3e} (shells), {lrel32.max():.3e} "
          f"(log bins)", flush=True)
    assert lrel32.max() < 1e-2


PHI0, DPHI0 = 0.0935, 0.000796


def deltaN_window(V0):
    """dN and Nend of the hilltop preset (1e-7, 1e-3 at V0 = 3.338e-13) scaled by
    sqrt(V0 / 3.338e-13): V0 only rescales the fluctuations (V / B^2 does not
    depend on it), so the spread of the sites' e-folds, and with it the window
    the deltaN phase needs, grows as sqrt(V0)."""
    s = float(np.sqrt(V0 / 3.338e-13))
    return 1e-7 * s, 1e-3 * s


def ensemble_run(r, V0, ns, af):
    dN, Nend = deltaN_window(V0[r])
    cfg = E.make_config(64, precision=E.IE_SPLIT, V0=float(V0[r]), ns=float(ns[r]), dt=0.001, L=10.0, af=af,
                        dN=dN, Nend=Nend)
    run = E.Run(cfg)
    run.initialize(1000 + r, PHI0, DPHI0)
    run.evolve_fields(0.0005)
    return run


def step_until_done(batch, runs, chunk=256):
    for _ in range(200):
        if all(r.background()["done"] for r in runs):
            return
        batch.steps(chunk, graph_steps=64)
    raise AssertionError("batch did not reach af")


def test_ensemble_256x64_per_run_pzeta():
    nruns, af = 256, 4.0
    print(f"ensemble {nruns} x 64^3 to af={af}", flush=True)
    rng = np.random.default_rng(20261203)
    V0 = 10.0 ** rng.uniform(-11, -9, nruns)
    ns = rng.uniform(0.95, 0.98, nruns)
    runs = [ensemble_run(r, V0, ns, af) for r in range(nruns)]
    batch = E.Batch(runs)
    step_until_done(batch, runs)
    batch.close()
    outs = [r.outputs(1) for r in runs]
    assert all(o["nbins"] == 56 for o in outs)
    bad = [q for q, o in enumerate(outs) if not np.all(o["pzeta"][1:] > 0)]
    if bad:
        o = outs[bad[0]]
        print(f"ensemble: {len(bad)} runs with P_zeta <= 0 in some shell; run {bad[0]} (V0 {V0[bad[0]]:.3e}): "
              f"{np.count_nonzero(o['pzeta'][1:] <= 0)} shells, P_zeta {o['pzeta'][:8]}", flush=True)
    assert not bad
    for r in runs[16:]:
        r.close()
    sample = [0, 37, 64, 101, 128, 170, 211, 255]
    # alone, a batch of one takes the resident path only when forced (the
    # default gives small batches the per-step kernels)
    E.Batch.set_resident_min(1, 1)
    try:
        alones = []
        for q in sample:
            alone = ensemble_run(q, V0, ns, af)
            b1 = E.Batch([alone])
            step_until_done(b1, [alone])
            b1.close()
            alones.append(alone)
    finally:
        E.Batch.set_resident_min(0, 0)
    for q, alone in zip(sample, alones):
        o = alone.outputs(1)
        alone.close()
        assert o["steps"] == outs[q]["steps"] and o["a_end"] == outs[q]["a_end"], q
        assert np.array_equal(o["pzeta"], outs[q]["pzeta"]), q
    # run 0 against the oracle (FP64) from the same initial conditions
    dN0, Nend0 = deltaN_window(V0[0])
    cfg = O.make_config(64, 0, {"dt": 0.001, "V0": float(V0[0]), "ns": float(ns[0]), "seed": 1000,
                                "initial_field": PHI0, "initial_derivative": DPHI0, "L": 10.0, "af": af,
                                "dN": dN0, "Nend": Nend0}, None, sigw=0)
    st = O.State(cfg, None)
    st.initialize()
    O.lib().ieo_evolve_fields(st.ptr, 0.0005)
    n_o = O.lib().ieo_leapfrog_steps(st.ptr, 10 ** 7)
    bo = st.background()
    O.lib().ieo_evolve_fields(st.ptr, -0.0005 * bo["astep"] ** (cfg.rescale_s - 1.0))
    assert outs[0]["steps"] == n_o
    assert outs[0]["a_end"] == pytest.approx(st.background()["a"], rel=1e-12)
    zeta_o, _, _ = oracle_deltaN(st)
    norm1 = (cfg.L / cfg.rescale_B) ** 3 / 64 ** 6
    F, ny = O.fftrn(zeta_o, +1)
    _, _, vo = O.bin_spectrum(64, cfg.L, F, ny, norm1)
    pz_o = pzeta(64, cfg.L, cfg.rescale_B, vo)
    rel = shell_rel(outs[0]["pzeta"], pz_o[:56])
    lo, lg = log_bins(pz_o[:56]), log_bins(outs[0]["pzeta"])
    ok = np.isfinite(lo) & (lo > 0)
    lrel = np.abs(lg[ok] - lo[ok]) / lo[ok]
    print(f"ensemble run 0 (split) vs oracle (FP64): P_zeta max rel {rel.max():.3e} (shells 1..55), "
          f"{lrel.max():.3e} (log bins)", flush=True)
    assert lrel.max() < 1e-2


def test_rescale_s_1_matches_oracle():
    print("rescale_s = 1 vs oracle", flush=True)
    st, _, table = oracle_state("hilltop_n32", N=32, calculate_sigw=0, rescale_s=1.0)
    run = engine_for(st, table, precision=E.IE_FP64, sigw=0)
    step_both(st, run, 100)
    bo, b = st.background(), run.background()
    for key in ("a", "ad", "t", "astep"):
        assert b[key] == pytest.approx(bo[key], rel=1e-13), key
    ef = fluct_err(st.field(0), run.field(E.IE_FIELD_F))
    print(f"rescale_s = 1: fluct err f {ef:.3e}", flush=True)
    assert ef < 1e-9
