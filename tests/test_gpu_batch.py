"""Batched stepping (ie_batch_*): many lattices of uneven N advanced by one launch
per kernel per step, captured in a HIP graph; 64^3 lattices without tensors by
one resident workgroup each, every step of a launch in one kernel.

  - a run's result does not depend on the other runs of its batch: alone (batch of
    one) and in a mix of sizes and potentials it is bitwise identical;
  - graph replay equals direct launches bitwise, over several replays with host
    work in between;
  - against the single-run path (ie_leapfrog_steps, another workgroup count and so
    another sum order) to 1e-13 relative in a and 1e-9 of the fluctuation rms.
"""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E
from conftest import potential_table

pytestmark = pytest.mark.gpu

HILL = dict(V0=3.338e-13, ns=0.97, dt=0.001)
TAB = dict(V0=3e-9, dt=0.001)


def new_run(N, seed, numerical=0, sigw=0, precision=E.IE_SPLIT, table=None, ns=0.97):
    if numerical:
        cfg = E.make_config(N, precision=precision, numerical=1, sigw=sigw, table=table, **TAB)
        run = E.Run(cfg, table)
        run.initialize(seed, 2.9181235049318586, -0.06727651095116181)
    else:
        cfg = E.make_config(N, precision=precision, sigw=sigw, **dict(HILL, ns=ns))
        run = E.Run(cfg)
        run.initialize(seed, 0.0935, 0.000796)
    run.evolve_fields(0.5 * cfg.dt)
    return run


def state(run):
    return run.field(E.IE_FIELD_F), run.field(E.IE_FIELD_FD), run.background()


@pytest.fixture
def resident_always():
    """The resident path for every 64^3 lattice without tensors (the default
    takes it only for batches of many such lattices)."""
    E.Batch.set_resident_min(1, 1)
    yield
    E.Batch.set_resident_min(0, 0)


def test_alone_equals_in_mix_bitwise():
    table = potential_table()
    specs = [(32, 11, 0, 0, 0.97), (16, 12, 1, 0, 0.97), (128, 13, 0, 1, 0.96), (32, 14, 0, 0, 0.95),
             (64, 15, 0, 0, 0.96), (64, 16, 1, 0, 0.97), (64, 17, 0, 1, 0.97)]
    mix = [new_run(N, s, numerical=p, sigw=g, table=table, ns=ns) for N, s, p, g, ns in specs]
    alone = [new_run(N, s, numerical=p, sigw=g, table=table, ns=ns) for N, s, p, g, ns in specs]
    E.Batch(mix).steps(60, graph_steps=20)
    for r in alone:
        E.Batch([r]).steps(60, graph_steps=0)
    for a, b in zip(mix, alone):
        fa, fda, ba = state(a)
        fb, fdb, bb = state(b)
        assert np.array_equal(fa, fb) and np.array_equal(fda, fdb)
        assert ba["a"] == bb["a"] and ba["ad"] == bb["ad"] and ba["steps"] == bb["steps"] == 60


def test_batches_running_together_equal_batches_alone():
    """Two batches enqueued on their own streams before either is waited for
    (ie_batch_launch, ie_batch_wait: the serving rounds) give each run the bits
    it gets with the batches stepped one after the other."""
    table = potential_table()
    specs = [[(32, 21, 0, 0), (32, 22, 1, 0)], [(128, 23, 0, 1)]]
    together = [[new_run(N, s, numerical=p, sigw=g, table=table) for N, s, p, g in grp] for grp in specs]
    apart = [[new_run(N, s, numerical=p, sigw=g, table=table) for N, s, p, g in grp] for grp in specs]
    bt = [E.Batch(g) for g in together]
    for q, b in zip((40, 10), bt):
        b.launch(q, graph_steps=10)
    for b in bt:
        b.wait()
    for q, g in zip((40, 10), apart):
        E.Batch(g).steps(q, graph_steps=10)
    for ga, gb in zip(together, apart):
        for a, b in zip(ga, gb):
            fa, fda, ba = state(a)
            fb, fdb, bb = state(b)
            assert np.array_equal(fa, fb) and np.array_equal(fda, fdb)
            assert ba["a"] == bb["a"] and ba["steps"] == bb["steps"]
    assert together[0][0].background()["steps"] == 40 and together[1][0].background()["steps"] == 10


@pytest.mark.parametrize("nsteps,graph_steps", [(25, 25), (23, 7)])
def test_graph_replay_equals_direct_launches(nsteps, graph_steps):
    """(23, 7): three replays per call starting at even and odd steps of the loop
    (graphs captured for both step parities), then two direct steps."""
    r1 = [new_run(32, 21), new_run(16, 22), new_run(64, 24), new_run(128, 23, sigw=1)]
    r2 = [new_run(32, 21), new_run(16, 22), new_run(64, 24), new_run(128, 23, sigw=1)]
    b1, b2 = E.Batch(r1), E.Batch(r2)
    for _ in range(4):   # several replays with host work in between
        b1.steps(nsteps, graph_steps=graph_steps)
        b2.steps(nsteps, graph_steps=0)
        for a, b in zip(r1, r2):
            assert np.array_equal(a.field(E.IE_FIELD_F), b.field(E.IE_FIELD_F))
            assert a.background()["a"] == b.background()["a"]


def test_batch_matches_single_run_path():
    rb = [new_run(32, 31, precision=E.IE_FP64), new_run(64, 33, precision=E.IE_FP64),
          new_run(128, 32, precision=E.IE_FP64)]
    rs = [new_run(32, 31, precision=E.IE_FP64), new_run(64, 33, precision=E.IE_FP64),
          new_run(128, 32, precision=E.IE_FP64)]
    E.Batch(rb).steps(100, graph_steps=50)
    for r in rs:
        r.leapfrog_steps(100)
    for a, b in zip(rb, rs):
        fa, _, ba = state(a)
        fb, _, bb = state(b)
        assert ba["a"] == pytest.approx(bb["a"], rel=1e-13)
        d = fb - fb.mean()
        assert np.max(np.abs((fa - fa.mean()) - d)) < 1e-9 * np.sqrt(np.mean(d ** 2))


def test_batch_stops_each_run_at_its_af(resident_always):
    cfg_short = E.make_config(16, precision=E.IE_SPLIT, af=1.005, **HILL)
    short = E.Run(cfg_short)
    short.initialize(41, 0.0935, 0.000796)
    short.evolve_fields(0.5 * HILL["dt"])
    cfg_short64 = E.make_config(128, precision=E.IE_SPLIT, af=1.005, **HILL)
    short64 = E.Run(cfg_short64)
    short64.initialize(43, 0.0935, 0.000796)
    short64.evolve_fields(0.5 * HILL["dt"])
    long_ = new_run(16, 42)
    cfg_res = E.make_config(64, precision=E.IE_SPLIT, af=1.005, **HILL)
    short_res = E.Run(cfg_res)
    short_res.initialize(44, 0.0935, 0.000796)
    short_res.evolve_fields(0.5 * HILL["dt"])
    E.Batch([short, long_, short64, short_res]).steps(40, graph_steps=10)
    bs, bl, b64 = short.background(), long_.background(), short64.background()
    br = short_res.background()
    assert br["done"] == 1 and br["steps"] < 40 and br["a"] > 1.005
    assert bs["done"] == 1 and bs["steps"] < 40 and bs["a"] > 1.005
    assert b64["done"] == 1 and b64["steps"] < 40 and b64["a"] > 1.005
    assert bl["done"] == 0 and bl["steps"] == 40


def test_fp64_tabulated_64_takes_the_step_path_beside_resident_runs(resident_always):
    """An FP64 tabulated 64^3 lattice (its rows and table exceed a resident
    workgroup's LDS) steps by the per-step kernels in a batch with resident split
    lattices, and each run gets the bits it gets alone."""
    table = potential_table()
    specs = [(64, 51, 1, E.IE_FP64), (64, 52, 1, E.IE_SPLIT), (64, 53, 0, E.IE_FP64)]
    for prec in (E.IE_FP64, E.IE_SPLIT):
        grp = [(N, s, p, q) for N, s, p, q in specs if q == prec]
        mix = [new_run(N, s, numerical=p, precision=q, table=table) for N, s, p, q in grp]
        alone = [new_run(N, s, numerical=p, precision=q, table=table) for N, s, p, q in grp]
        E.Batch(mix).steps(30, graph_steps=10)
        for r in alone:
            E.Batch([r]).steps(30, graph_steps=0)
        for a, b in zip(mix, alone):
            fa, fda, ba = state(a)
            fb, fdb, bb = state(b)
            assert np.array_equal(fa, fb) and np.array_equal(fda, fdb)
            assert ba["a"] == bb["a"] and ba["steps"] == bb["steps"] == 30


@pytest.mark.parametrize("chunks", [(10, 10, 10), (7, 1, 6, 16)])
def test_resident_launches_in_chunks_equal_one(chunks, resident_always):
    """A resident 64^3 lattice starts every launch with the sums of its current f
    in a step's own partition and order, so launches of any lengths give the bits
    of one launch."""
    ra, rb = new_run(64, 61), new_run(64, 61)
    ba = E.Batch([ra])
    for n in chunks:
        ba.steps(n, graph_steps=0)
    E.Batch([rb]).steps(sum(chunks), graph_steps=0)
    fa, fda, bga = state(ra)
    fb, fdb, bgb = state(rb)
    assert np.array_equal(fa, fb) and np.array_equal(fda, fdb)
    assert bga["a"] == bgb["a"] and bga["ad"] == bgb["ad"] and bga["steps"] == bgb["steps"] == sum(chunks)


def test_resident_tabulated_split_matches_single_run_path(resident_always):
    """The resident kernel against the single-run row kernels (another sum order):
    a to 1e-13 relative, the FP32 fluctuation to 1e-5 of its rms over 200 steps."""
    table = potential_table()
    rb = new_run(64, 71, numerical=1, table=table)
    rs = new_run(64, 71, numerical=1, table=table)
    E.Batch([rb]).steps(200, graph_steps=50)
    rs.leapfrog_steps(200)
    fa, _, ba = state(rb)
    fb, _, bb = state(rs)
    assert ba["steps"] == bb["steps"] == 200
    assert ba["a"] == pytest.approx(bb["a"], rel=1e-13)
    d = fb - fb.mean()
    err = np.max(np.abs((fa - fa.mean()) - d)) / np.sqrt(np.mean(d ** 2))
    print(f"resident tabulated 64 vs single run: fluct err {err:.3e}", flush=True)
    assert err < 1e-5


def test_resident_path_taken_by_lattice_count():
    """ie_batch_set_resident_min: a batch takes the resident path for its 64^3
    lattices of a potential kind only when it holds at least the minimum number of
    them; each path gives the bits it gives when forced."""
    table = potential_table()
    specs = [(64, 81, 0), (64, 82, 0), (64, 83, 1), (64, 84, 1)]

    def run_batch(min_hill, min_tab, which):
        E.Batch.set_resident_min(min_hill, min_tab)
        try:
            rr = [new_run(N, s, numerical=p, table=table) for N, s, p in (specs[q] for q in which)]
            E.Batch(rr).steps(30, graph_steps=10)
            return [state(r)[:2] for r in rr]
        finally:
            E.Batch.set_resident_min(0, 0)

    big = 1 << 20
    forced_res = run_batch(1, 1, range(4))
    forced_steps = run_batch(big, big, range(4))
    # two hilltop lattices meet a minimum of 2, two tabulated ones miss 3
    mixed = run_batch(2, 3, range(4))
    for q in range(4):
        want = forced_res[q] if q < 2 else forced_steps[q]
        assert np.array_equal(mixed[q][0], want[0]) and np.array_equal(mixed[q][1], want[1]), q
    # the two paths differ only by the sum order: same physics
    for q in range(4):
        d = forced_steps[q][0] - forced_steps[q][0].mean()
        err = np.max(np.abs((forced_res[q][0] - forced_res[q][0].mean()) - d)) / np.sqrt(np.mean(d ** 2))
        assert err < 1e-5, (q, err)
