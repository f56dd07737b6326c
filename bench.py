#!/usr/bin/env python3
"""Benchmark: lattice site-updates/s of the leapfrog step at 256^3 on one MI355X.

Workload (one "step" = one leapfrog step of run_evolution_loop over the whole
lattice: Friedmann reduction, kick, drift; src/evolution.cpp:937-945):
  256^3 lattice, analytic hilltop preset (params.analytic.txt: V0 = 3.338e-13,
  ns = 0.97, phi0 = 0.0935, dphi0 = 0.000796, L = 10, dt = 0.001, seed 8),
  Bunch-Davies initial conditions from the engine's own initializer.
  value: split storage (FP32 fluctuation about the FP64 homogeneous mode, every
  right-hand side in FP64), SIGW off.  Also reported: the FP64-storage path at
  the same size and the split path with the six SIGW tensors on.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "inflation-easy_amd"))
sys.path.insert(0, str(REPO))

METRIC = "lattice site-updates/s at 256³ (split-FP32 vs FP64 GPU); p50/p99 run latency"
HBM_PEAK_GBS = 8000.0
PRESET = dict(V0=3.338e-13, ns=0.97, L=10.0, dt=0.001)
PHI0, DPHI0, SEED = 0.0935, 0.000796, 8


def make_run(N: int, precision: int, sigw: int):
    import ie_engine as E
    cfg = E.make_config(N, precision=precision, sigw=sigw, **PRESET)
    run = E.Run(cfg)
    run.initialize(SEED, PHI0, DPHI0)
    run.evolve_fields(0.5 * PRESET["dt"])   # leapfrog stagger (src/evolution.cpp:935)
    return run


def time_steps(run, steps: int, warmup: int) -> tuple[float, dict]:
    if warmup:
        run.leapfrog_steps(warmup)
    t0 = time.perf_counter()
    bg = run.leapfrog_steps(steps)
    t1 = time.perf_counter()
    return (t1 - t0), bg


def bench_case(N: int, precision: int, sigw: int, steps: int, warmup: int, profile_steps: int):
    run = make_run(N, precision, sigw)
    # kernel shapes of the kick / tensor launches picked by the engine's tuner on
    # real loop steps, before the untimed warmup
    tuning = run.tune(8)
    secs, bg = time_steps(run, steps, warmup)
    if bg["steps"] != steps + warmup + tuning["steps"] or bg["done"]:
        raise RuntimeError(f"unexpected loop state {bg}")
    prof = run.profile_steps(profile_steps) if profile_steps else None
    run.close()
    return secs, prof, tuning


def tabulated_case(N: int = 128, steps: int = 1000, warmup: int = 50, dense_from: int = 5500) -> dict:
    """Split against FP64 storage with the tabulated SR -> USR -> SR potential
    (inputs/ tables, linear interpolation), N^3, the USR preset's initial field:
    loop steps per second of each, after the tuner and an untimed warmup; then
    `steps` more from loop step `dense_from`, where the field has reached the
    dense stretch of the grid (phi near 2.683: spacing 3.6e-6 against a spread of
    2e-5, so a wave's sites lie in tens of intervals)."""
    import numpy as np

    import ie_engine as E
    with np.load(REPO / "tests" / "golden" / "potential_table.npz") as z:
        class T:  # table kept alive for the run
            x = np.ascontiguousarray(z["field"])
            v = np.ascontiguousarray(z["potential"])
            dv = np.ascontiguousarray(z["potential_derivative"])
    out = {"N": N, "steps": steps, "potential": "tabulated SR-USR-SR (inputs/ tables)"}
    for name, prec in (("split", E.IE_SPLIT), ("fp64", E.IE_FP64)):
        cfg = E.make_config(N, precision=prec, numerical=1, table=T, V0=3e-9, dt=0.001)
        run = E.Run(cfg, T)
        run.initialize(SEED, 2.9181235049318586, -0.06727651095116181)
        run.evolve_fields(0.0005)
        run.tune(8)
        secs, bg = time_steps(run, steps, warmup)
        if bg["done"]:
            raise RuntimeError(f"unexpected loop state {bg}")
        out[f"{name}_ms_per_step"] = secs / steps * 1e3
        out[f"{name}_site_updates_per_s"] = N ** 3 * steps / secs
        run.leapfrog_steps(max(0, dense_from - bg["steps"]))
        secs, bg = time_steps(run, steps, 0)
        if bg["done"]:
            raise RuntimeError(f"unexpected loop state {bg}")
        out[f"{name}_dense_ms_per_step"] = secs / steps * 1e3
        run.close()
    out["split_over_fp64"] = out["split_site_updates_per_s"] / out["fp64_site_updates_per_s"]
    out["dense_from_step"] = dense_from
    out["dense_split_over_fp64"] = out["fp64_dense_ms_per_step"] / out["split_dense_ms_per_step"]
    # the hilltop split step at the same N and dt: the cost of the table lookups
    run = make_run(N, E.IE_SPLIT, 0)
    run.tune(8)
    secs, _ = time_steps(run, steps, warmup)
    run.close()
    out["hilltop_split_ms_per_step"] = secs / steps * 1e3
    out["split_over_hilltop_time"] = out["split_ms_per_step"] / out["hilltop_split_ms_per_step"]
    return out


def read_traffic(name: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/traffic.json)."""
    p = REPO / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get(name)
    except (ValueError, OSError):
        return None


def host_cpu() -> dict:
    """The host's CPU model, logical CPU count and the CPUs this process may use."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "allowed_cpus": allowed}


def cpu_baseline(N: int = 256, target_steps_af: float = 1.25) -> dict | None:
    """The reference program (built from its sources by oracle/build_ref.sh, OpenMP)
    timed on this host with one OpenMP thread per CPU this process may use
    (OMP_NUM_THREADS when set, which caps it to the host share given to this
    job): loop time = wall(af = target) - wall(af = 1, no step), the initial
    conditions and outputs being identical in both runs."""
    binary = REPO / "oracle" / "_ref" / f"ie_ref_n{N}_p0_g0_d0_o1"
    if not binary.exists():
        return None
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cpu["allowed_cpus"]
    base = (f"seed = {SEED}\ndt = {PRESET['dt']}\nV0 = {PRESET['V0']}\nns = {PRESET['ns']}\n"
            f"initial_field = {PHI0}\ninitial_derivative = {DPHI0}\nL = {PRESET['L']}\n"
            "output_freq = 100000000\noutput_infrequent_freq = 100000000\noutput_spectra = 0\n"
            "output_histogram = 0\noutput_energy = 0\nscreen_updates = 0\n")

    def one(af: float):
        d = Path(tempfile.mkdtemp(prefix="ie_cpu_"))
        try:
            (d / "params.txt").write_text(base + f"af = {af}\n")
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            t0 = time.perf_counter()
            subprocess.run([str(binary)], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, timeout=600)
            wall = time.perf_counter() - t0
            text = (d / "results" / "output.txt").read_text()
            found = re.findall(r"numsteps (\d+)", text)
            return wall, int(found[-1]) if found else 0
        finally:
            shutil.rmtree(d, ignore_errors=True)

    print(f"# cpu baseline: reference {binary.name}, {threads} threads, af=1.0", file=sys.stderr, flush=True)
    w0, n0 = one(1.0)
    print(f"# cpu baseline: af=1.0 took {w0:.1f} s; af={target_steps_af}", file=sys.stderr, flush=True)
    w1, n1 = one(target_steps_af)
    print(f"# cpu baseline: af={target_steps_af} took {w1:.1f} s, {n1 - n0} steps", file=sys.stderr, flush=True)
    steps = n1 - n0
    if steps <= 0 or w1 <= w0:
        return None
    rate = N ** 3 * steps / (w1 - w0)
    return {"value": rate, "unit": "site-updates/s", "cores": threads, "kind": "reference",
            "host_cpu": cpu,
            "sample": f"{N}^3 hilltop, FP64, {steps} leapfrog steps (af={target_steps_af}) of the reference "
                      f"OpenMP build on {threads} threads; loop time = wall(af={target_steps_af}) - wall(af=1), "
                      f"{w1 - w0:.1f} s of CPU loop time"}


def load_table():
    import numpy as np
    with np.load(REPO / "tests" / "golden" / "potential_table.npz") as z:
        class T:  # table kept alive for the request pointers
            x = np.ascontiguousarray(z["field"])
            v = np.ascontiguousarray(z["potential"])
            dv = np.ascontiguousarray(z["potential_derivative"])
    return T


USR = dict(V0=3e-9, dt=0.001, dN=1e-4, Nend=2.0)           # params.numerical.txt
USR_PHI0, USR_DPHI0 = 2.9181235049318586, -0.06727651095116181
HILL = dict(V0=3.338e-13, ns=0.97, dt=0.001, dN=1e-7, Nend=0.001)   # params.analytic.txt


def request_stream(nreq: int, seed: int):
    """The serving mix: N in {32, 64, 128, 256} with weights 0.4/0.3/0.2/0.1,
    half hilltop / half tabulated USR, 20% with SIGW, every 20th with an FP64
    twin (5%)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = rng.choice([32, 64, 128, 256], size=nreq, p=[0.4, 0.3, 0.2, 0.1])
    tab = rng.random(nreq) < 0.5
    sigw = rng.random(nreq) < 0.2
    check = np.arange(nreq) % 20 == 0
    return [(int(sizes[n]), bool(tab[n]), int(sigw[n]), int(check[n])) for n in range(nreq)], rng


def submit_request(srv, E, T, n, spec):
    N, tab, sigw, check = spec
    if tab:
        cfg = E.make_config(N, precision=E.IE_SPLIT, numerical=1, sigw=sigw, table=T, **USR)
        return srv.submit(cfg, 1000 + n, USR_PHI0, USR_DPHI0, 1, check, T)
    cfg = E.make_config(N, precision=E.IE_SPLIT, sigw=sigw, **HILL)
    return srv.submit(cfg, 1000 + n, PHI0, DPHI0, 1, check)


def serve_stream(nreq: int, load: float = 0.8, burst: int = 24, seed: int = 20261203,
                 graph_steps: int | None = None) -> dict:
    """Serving: each request is a whole run (initial conditions, inflation to
    af = 2N, Omega_GW for SIGW runs, the deltaN map, P_zeta).  Capacity is
    measured first: `burst` requests of the mix submitted at once, runs/s over
    their makespan.  Then `nreq` requests of the same mix (another draw) arrive
    as a Poisson stream at load x capacity while the server keeps admitting by
    footprint (240 GiB budget) and stepping whatever is resident."""
    import numpy as np

    import ie_engine as E
    T = load_table()

    def run_stream(specs, arrivals, label):
        srv = E.Server(240 << 30, 100, 8)
        if graph_steps is not None:
            srv.set_graph_steps(graph_steps)
        results, n = [], 0
        t0 = srv.clock()
        while len(results) < len(specs):
            now = srv.clock() - t0
            while n < len(specs) and arrivals[n] <= now:
                submit_request(srv, E, T, n, specs[n])
                n += 1
            srv.pump()
            for r in srv.results():
                results.append(r)
                if len(results) % 10 == 0 or len(results) == len(specs):
                    print(f"# serve {label}: {len(results)}/{len(specs)} done, last N={r['N']} "
                          f"latency={r['latency_s']:.2f}s", file=sys.stderr, flush=True)
        makespan = max(r["done_s"] for r in results) - t0
        srv.close()
        return results, makespan

    bspecs, _ = request_stream(burst, seed + 1)
    bres, bspan = run_stream(bspecs, np.zeros(burst), "capacity burst")
    capacity = burst / bspan
    rate = load * capacity
    specs, rng = request_stream(nreq, seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    results, makespan = run_stream(specs, arrivals, f"poisson {rate:.3f}/s")
    lat = np.array([r["latency_s"] for r in results])
    errs = [r["fp64_max_rel_err"] for r in results if r["fp64_max_rel_err"] >= 0]
    gerrs = [r["gw_fp64_max_rel_err"] for r in results if r["gw_fp64_max_rel_err"] >= 0]
    return {"requests": nreq, "capacity_runs_per_s": capacity, "capacity_burst": burst,
            "offered_load": load, "arrival_rate_per_s": rate, "runs_per_s": nreq / makespan,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "latency_samples": int(lat.size), "max_latency_s": float(lat.max()),
            "failed": int(sum(r["status"] != 0 for r in results) + sum(r["status"] != 0 for r in bres)),
            "sigw_runs_with_omega_gw": int(sum(r["gw_nbins"] > 0 for r in results)),
            "fp64_checks": len(errs), "fp64_max_rel_err_pzeta": max(errs) if errs else None,
            "fp64_max_rel_err_omega_gw": max(gerrs) if gerrs else None,
            "mix": "N 32/64/128/256 (0.4/0.3/0.2/0.1), 50% tabulated USR, 20% SIGW, every 20th with an FP64 twin; "
                   "each run: ICs, inflation to af=2N, Omega_GW (SIGW), deltaN, P_zeta"}


def shell_modes(N: int, n: int):
    """Weighted mode count per reference shell (bin = lround |k|, weight 2 for
    0 < kz < N/2, 1 for kz = 0 and N/2; src/output.cpp:63-151)."""
    import numpy as np
    i = np.arange(N)
    p = np.where(i <= N // 2, i, i - N)
    px, py = np.meshgrid(p, p, indexing="ij")
    cnt = np.zeros(n, dtype=np.int64)
    for z in range(N // 2 + 1):
        b = np.floor(np.sqrt(px * px + py * py + z * z) + 0.5).astype(np.int64).ravel()
        keep = b < n
        np.add.at(cnt, b[keep], 1 if z in (0, N // 2) else 2)
    return cnt


def compare_spectra(pa, pb, nlog: int, N: int | None = None, notebook_cut: bool = False) -> dict:
    """Split (pa) against FP64 (pb) spectra on the reference shells: the largest
    relative difference over shells 1 .. n-1 with pb > 0, over the complete
    spherical shells 1 .. N/2 (when n > N/2), over the shells the reference's
    P_zeta plot shows (notebook_cut: notebooks/plot.ipynb cell 9 plots k_list up
    to len(k_list) - 10) and over nlog log-spaced bins in k (each the mean of the
    shell values whose index falls in it; bins narrower than a shell stay empty);
    the three worst shells with their weighted mode counts."""
    import numpy as np
    n = min(len(pa), len(pb))
    pa, pb = np.asarray(pa[:n]), np.asarray(pb[:n])
    rel = np.where(pb > 0, np.abs(pa - pb) / np.where(pb > 0, pb, 1.0), 0.0)
    shell = np.arange(1, n)
    which = np.clip(np.digitize(shell, np.logspace(0.0, np.log10(n), nlog + 1)) - 1, 0, nlog - 1)
    lrel = []
    for q in range(nlog):
        sel = shell[which == q]
        if sel.size and pb[sel].mean() > 0:
            lrel.append(abs(pa[sel].mean() - pb[sel].mean()) / pb[sel].mean())
    out = {"shells": n, "max_rel_err_shells": float(rel[1:].max()), "log_bins": nlog,
           "log_bins_filled": len(lrel), "max_rel_err_log_bins": float(max(lrel)) if lrel else None}
    if N is not None and n > N // 2:
        out["max_rel_err_shells_1_to_N_over_2"] = float(rel[1:N // 2 + 1].max())
    if notebook_cut and n > 11:
        out["max_rel_err_shells_notebook_range"] = float(rel[1:n - 10].max())
        out["notebook_range"] = f"shells 1..{n - 11}"
    if N is not None:
        modes = shell_modes(N, n)
        worst = np.argsort(rel)[::-1][:3]
        out["worst_shells"] = [[int(b), float(rel[b]), int(modes[b])] for b in worst]
    return out


def whole_run(E, N, precision, sigw, numerical, T=None, seed=SEED):
    """One whole run on its own: initial conditions, the leapfrog loop to
    af = 2N, then ie_run_outputs (final sync, Omega_GW, deltaN, P_zeta)."""
    if numerical:
        cfg = E.make_config(N, precision=precision, numerical=1, sigw=sigw, table=T, **USR)
        run = E.Run(cfg, T)
        run.initialize(seed, USR_PHI0, USR_DPHI0)
    else:
        cfg = E.make_config(N, precision=precision, sigw=sigw, **HILL)
        run = E.Run(cfg)
        run.initialize(seed, PHI0, DPHI0)
    run.evolve_fields(0.5 * cfg.dt)
    t0 = time.perf_counter()
    bg = run.leapfrog_steps(10 ** 8)
    loop_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = run.outputs(1)
    out_s = time.perf_counter() - t0
    run.close()
    if not bg["done"]:
        raise RuntimeError(f"loop did not reach af: {bg}")
    out["loop_s"], out["outputs_s"] = loop_s, out_s
    return out


def c2_tabulated(N: int = 128) -> dict:
    """The tabulated SR -> USR -> SR potential at N^3 (params.numerical.txt:
    dt = 0.001, dN = 1e-4, Nend = 2.0, seed 8): whole runs in split and FP64
    storage from the same initial conditions; P_zeta in the reference shells and
    in 64 log bins, split against FP64."""
    import ie_engine as E
    T = load_table()
    a = whole_run(E, N, E.IE_SPLIT, 0, 1, T)
    b = whole_run(E, N, E.IE_FP64, 0, 1, T)
    res = {"N": N, "potential": "tabulated SR-USR-SR (inputs/ tables)", "steps": a["steps"],
           "split_loop_s": a["loop_s"], "fp64_loop_s": b["loop_s"], "outputs_s": a["outputs_s"]}
    res["pzeta"] = compare_spectra(a["pzeta"], b["pzeta"], 64, N, notebook_cut=True)
    return res


def c3_sigw(N: int = 256) -> dict:
    """The hilltop preset at N^3 with SIGW and deltaN (params.analytic.txt:
    dt = 0.001, dN = 1e-7, Nend = 1e-3, seed 8): a whole split run and its FP64
    twin from the same initial conditions; P_zeta in the 222 reference shells and
    128 log bins, Omega_GW over the kept shells (floor(0.8 numbins)), each against
    the twin."""
    import ie_engine as E
    a = whole_run(E, N, E.IE_SPLIT, 1, 0)
    b = whole_run(E, N, E.IE_FP64, 1, 0)
    res = {"N": N, "potential": "hilltop", "sigw": True, "steps": a["steps"], "a_end": a["a_end"],
           "split_loop_s": a["loop_s"], "fp64_loop_s": b["loop_s"], "outputs_s": a["outputs_s"]}
    res["pzeta"] = compare_spectra(a["pzeta"], b["pzeta"], 128, N, notebook_cut=True)
    res["omega_gw"] = compare_spectra(a["omega_gw"], b["omega_gw"], 128, N)
    res["omega_gw"]["kept_shells"] = int(a["gw_nbins"])
    return res


def ensemble(nruns: int = 256, N: int = 64, steps: int = 1000, seed: int = 20261203, graph_steps: int = 50,
             sequential: bool = True, whole: bool = True, seq_steps: int = 200) -> dict:
    """Batched ensemble: nruns hilltop lattices of N^3, seeds 1000 + r, V0
    log-uniform in [1e-11, 1e-9], ns uniform in [0.95, 0.98], split storage,
    shapes tuned before the batch is built, the deltaN window scaled with V0.  `steps` steps are timed as one
    batch (one launch per kernel per step, HIP graph replay; the call's wall time, the waits and the
    256 runs' background read-back included) and `seq_steps` as sequential per-run launches, each
    reported per step; then the batch runs every lattice to
    af = 2N and each run's P_zeta is computed (ie_run_outputs).  sequential /
    whole: False leaves out the per-run comparison / the whole runs (profiling
    runs of the batched step alone)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import ie_engine as E
    rng = np.random.default_rng(seed)
    V0 = 10.0 ** rng.uniform(-11, -9, nruns)
    ns = rng.uniform(0.95, 0.98, nruns)

    def make(r):
        # the preset's deltaN window (dN = 1e-7, Nend = 1e-3 at V0 = 3.338e-13)
        # scaled by sqrt(V0 / 3.338e-13): V0 only rescales the fluctuations, so
        # the spread of the sites' e-folds grows as sqrt(V0)
        w = float(np.sqrt(V0[r] / HILL["V0"]))
        cfg = E.make_config(N, precision=E.IE_SPLIT, V0=float(V0[r]), ns=float(ns[r]), dt=0.001, L=10.0,
                            dN=1e-7 * w, Nend=1e-3 * w)
        run = E.Run(cfg)
        run.initialize(1000 + r, PHI0, DPHI0)
        run.evolve_fields(0.0005)
        return run

    probe = make(0)
    tuning = probe.tune(8)    # shapes of this (N, storage), taken by the batch
    probe.close()
    workers = min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        runs = list(ex.map(make, range(nruns)))
    batch = E.Batch(runs)
    batch.steps(20, graph_steps=min(20, graph_steps))   # warm-up, graph instantiation
    t0 = time.perf_counter()
    batch.steps(steps, graph_steps=graph_steps)
    tb = time.perf_counter() - t0
    sites = nruns * N ** 3 * steps
    res = {"runs": nruns, "N": N, "steps": steps, "tuning": tuning,
           "batched_site_updates_per_s": sites / tb, "batched_ms_per_step": 1e3 * tb / steps,
           "graph_steps": graph_steps}
    if sequential:
        for r in runs[:4]:
            r.leapfrog_steps(5)                          # warm-up of the per-run path
        t0 = time.perf_counter()
        for r in runs[4:]:
            r.leapfrog_steps(seq_steps)
        ts = (time.perf_counter() - t0) * nruns / (nruns - 4) * steps / seq_steps   # for `steps` steps
        res.update({"sequential_site_updates_per_s": sites / ts, "batched_over_sequential": ts / tb,
                    "sequential_ms_per_step": 1e3 * ts / steps, "sequential_steps_timed": seq_steps})
    batch.close()
    if not whole:
        for r in runs:
            r.close()
        return res
    # whole runs: a fresh batch of the same lattices to af, then per-run P_zeta
    for r in runs:
        r.close()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        runs = list(ex.map(make, range(nruns)))
    batch = E.Batch(runs)
    t0 = time.perf_counter()
    rounds = 0
    while not all(r.background()["done"] for r in runs):
        batch.steps(500, graph_steps=graph_steps)
        rounds += 1
        if rounds > 100:
            raise RuntimeError("ensemble did not reach af")
    t_whole = time.perf_counter() - t0
    batch.close()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        outs = list(ex.map(lambda r: r.outputs(1), runs))
    t_out = time.perf_counter() - t0
    for r in runs:
        r.close()
    pz = np.array([o["pzeta"] for o in outs])
    res["whole_runs"] = {"steps_per_run": int(outs[0]["steps"]), "loop_s": t_whole, "outputs_s": t_out,
                           "pzeta_returned": int(sum(o["nbins"] > 0 for o in outs)), "pzeta_shells": int(pz.shape[1]),
                           "pzeta_positive_runs": int(sum(np.all(p[1:] > 0) for p in pz)),
                           "pzeta_shell8_min": float(pz[:, 8].min()), "pzeta_shell8_max": float(pz[:, 8].max()),
                           "pzeta_run0": [float(x) for x in pz[0]]}
    return res


def post_inflation(N: int = 256, steps: int = 200, warmup: int = 10, precision: int | None = None) -> dict:
    """Post-inflation leapfrog (run_post_inflation_loop, src/evolution.cpp:1596-1604)
    at N^3: FP64 Phi, Phi' and the six FP32 h_ij, h_ij' from a deltaN map of the
    hilltop initial field's fluctuations (w = 1/3, horizon_factor 1, dt 0.001).
    One step = the Phi kick (24 B/site) and the tensor kick with the Hessian source
    plus both drifts (120 B/site); wall time over `steps` device-resident steps."""
    import numpy as np

    import ie_engine as E
    cfg = E.make_config(N, precision=E.IE_SPLIT if precision is None else precision, sigw=1, **PRESET)
    run = E.Run(cfg)
    run.initialize(SEED, PHI0, DPHI0)
    run.deltaN_init()
    f = run.field(E.IE_FIELD_F)
    dn = 0.5 + 2e-3 * (f - f.mean()) / (np.abs(f - f.mean()).max() + 1e-300)
    run.upload_field(E.IE_FIELD_DELTAN, dn)
    del f, dn
    run.post_init(omega=1.0 / 3.0, horizon_factor=1.0, dt=0.001, af=1e9)
    run.evolve_fields(0.0005)
    run.post_leapfrog_steps(warmup)
    t0 = time.perf_counter()
    bg = run.post_leapfrog_steps(steps)
    secs = time.perf_counter() - t0
    run.close()
    if bg["steps"] != steps + warmup:
        raise RuntimeError(f"unexpected post-inflation loop state {bg}")
    ms = 1e3 * secs / steps
    alg = 144 * N ** 3
    return {"N": N, "steps": steps, "ms_per_step": ms, "site_updates_per_s": N ** 3 * steps / secs,
            "algorithmic_bytes_per_step": alg, "achieved_GBps": alg / (ms * 1e-3) / 1e9,
            "kernels": "k_post_kick_rows (24 B/site) + k_post_tensor_march (120 B/site; k_post_tensor_rows from N = 512)",
            "storage": "FP64 Phi, Phi'; FP32 h, hdot"}


def copy_calibration(nbytes: int = 1 << 30) -> None:
    """A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings (run under the profiler)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--post-steps", type=int, default=200, help="post-inflation steps timed (0: skip)")
    ap.add_argument("--pzeta-check", type=int, default=1,
                    help="whole runs, split vs FP64: 128^3 tabulated and 256^3 SIGW + deltaN (0: skip)")
    ap.add_argument("--only", choices=["split", "fp64", "sigw", "post", "c2", "c3", "serve", "tab", "ensemble",
                                       "batch"],
                    default=None, help="time a single case (for profiling runs)")
    ap.add_argument("--ensemble-runs", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson arrival rate as a fraction of the capacity measured first")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="loop steps per HIP graph launch of the batched legs (0: direct launches, as for a "
                         "kernel-traced run; default: 50 for the ensemble, the quantum for serving)")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; runs on several GPUs are not supported "
              "(ensembles and serving would shard by run, with no communication).", file=sys.stderr)
        return 2
    import ie_engine as E
    if E.lib().ie_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 3
    N, K, W = args.N, args.steps, args.warmup
    sites = N ** 3
    if args.only == "post":
        print(json.dumps({"post_inflation": post_inflation(N, args.post_steps)}), flush=True)
        return 0
    if args.only == "c2":
        print(json.dumps({"c2_tabulated_128": c2_tabulated(128)}), flush=True)
        return 0
    if args.only == "c3":
        print(json.dumps({"c3_sigw_256": c3_sigw(N)}), flush=True)
        return 0
    if args.only in ("ensemble", "batch"):
        # batch: the batched steps alone (no per-run comparison, no whole runs)
        gs = 50 if args.graph_steps is None else args.graph_steps
        full = args.only == "ensemble"
        print(json.dumps({"ensemble": ensemble(args.ensemble_runs, steps=args.steps if not full else 1000,
                                               graph_steps=gs, sequential=full, whole=full)}), flush=True)
        return 0
    if args.only == "serve":
        print(json.dumps({"serving": serve_stream(args.serve_requests, args.serve_load,
                                                  graph_steps=args.graph_steps)}), flush=True)
        return 0
    if args.only == "tab":
        print(json.dumps({f"tabulated_{N}": tabulated_case(N)}), flush=True)
        return 0
    cases = {"split": (E.IE_SPLIT, 0), "fp64": (E.IE_FP64, 0), "sigw": (E.IE_SPLIT, 1)}
    todo = [args.only] if args.only else ["split", "fp64", "sigw"]
    res = {}
    for name in todo:
        prec, sigw = cases[name]
        secs, prof, tuning = bench_case(N, prec, sigw, K, W, args.profile_steps)
        res[name] = {"secs": secs, "sups": sites * K / secs, "ms_per_step": 1e3 * secs / K, "profile": prof,
                     "tuning": tuning}
        print(f"# {name}: {res[name]['sups']:.4e} site-updates/s, {res[name]['ms_per_step']:.4f} ms/step, "
              f"profile {prof}, tuning {tuning}", file=sys.stderr, flush=True)
    head_name = "split" if "split" in res else next(iter(res))
    head = res[head_name]
    out = {
        "metric": METRIC,
        "value": head["sups"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{N}^3 hilltop lattice, leapfrog, split storage (FP32 fluctuation about FP64 "
                               f"homogeneous mode, FP64 arithmetic), SIGW off; params.analytic.txt preset, seed 8",
                   "N": N, "potential": "hilltop", "integrator": "leapfrog", "storage": "split-fp32"},
    }
    if "fp64" in res:
        out["fp64_value"] = res["fp64"]["sups"]
        out["fp64_ms_per_step"] = res["fp64"]["ms_per_step"]
        if "split" in res:
            out["split_over_fp64"] = res["split"]["sups"] / res["fp64"]["sups"]
    if "sigw" in res:
        out["sigw_split_value"] = res["sigw"]["sups"]
        out["sigw_ms_per_step"] = res["sigw"]["ms_per_step"]
        if res["sigw"]["profile"]:
            out["sigw_tensor_ms"] = res["sigw"]["profile"]["tensor_ms"]
    out["autotune"] = {k: v["tuning"] for k, v in res.items()}
    prof = head["profile"]
    if prof:
        # dominant kernel: the fused step kernel (k_march: the previous step's
        # close, kick, drift and the next step's Friedmann sums, 128 <= N <= 256)
        # or the row kick; algorithmic bytes per site: read dphi and dpi, write dpi
        # and the next dphi (4 B each in split storage, 8 B in FP64); the sums and
        # the close read nothing more per site
        per_site = 16 if cases[head_name][0] == E.IE_SPLIT else 32
        alg = per_site * sites
        fused = 128 <= N <= 256
        # a fused step without SIGW is one launch of the step kernel: events around
        # the loop alone give its mean launch duration (the per-launch events of
        # the profile pass hold the launches apart and read higher)
        one_launch = fused and cases[head_name][1] == 0
        ms = prof["loop_ms"] if one_launch else prof["kick_drift_ms"]
        achieved = alg / (ms * 1e-3) / 1e9
        variant = head["tuning"].get("march_variant", 0)
        split_head = cases[head_name][0] == E.IE_SPLIT
        targs = "float, true" if split_head else "double, false"
        kname = f"{'k_march_pf' if variant == 1 else 'k_march'}<{targs}, 4, 0>" + \
            " (previous step's close + kick + drift + next-step sums: the whole loop step)"
        storage = "split" if split_head else "fp64"
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS,
                           "traffic": read_traffic(f"{'march' if fused else 'kick_drift'}_{storage}_{N}"),
                           "kernel": kname if fused else "k_row_kick (kick + drift)", "kernel_ms": ms,
                           "algorithmic_bytes": alg, "reduce_ms": prof["reduce_ms"],
                           "kernel_ms_per_launch_events": prof["kick_drift_ms"],
                           "step_ms_events": prof["step_ms"]}
    if not args.only:
        for tn in (128, N):
            out[f"tabulated_{tn}"] = tabulated_case(tn)
            print(f"# tabulated {tn}: {out[f'tabulated_{tn}']}", file=sys.stderr, flush=True)
    if args.pzeta_check and not args.only:
        out["c2_tabulated_128"] = c2_tabulated(128)
        print(f"# c2: {out['c2_tabulated_128']}", file=sys.stderr, flush=True)
        out["c3_sigw_256"] = c3_sigw(N)
        print(f"# c3: {out['c3_sigw_256']}", file=sys.stderr, flush=True)
    if args.post_steps > 0 and not args.only:
        out["post_inflation"] = post_inflation(N, args.post_steps)
        print(f"# post-inflation: {out['post_inflation']}", file=sys.stderr, flush=True)
    if args.ensemble_runs > 0 and not args.only:
        out["ensemble"] = ensemble(args.ensemble_runs, graph_steps=50 if args.graph_steps is None else args.graph_steps)
        print(f"# ensemble: {out['ensemble']}", file=sys.stderr, flush=True)
    if args.serve_requests > 0 and not args.only:
        out["serving"] = serve_stream(args.serve_requests, args.serve_load, graph_steps=args.graph_steps)
        out["p50_latency_s"] = out["serving"]["p50_latency_s"]
        out["p99_latency_s"] = out["serving"]["p99_latency_s"]
    if not args.no_cpu_baseline:
        cb = cpu_baseline(N)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
