"""Serving loop (ie_server_*): whole runs (initial conditions, inflation to af,
Omega_GW of SIGW runs, deltaN, P_zeta) admitted by device footprint and stepped
together.

  - a run's P_zeta (and Omega_GW) is bitwise the same alone and in a mix of
    sizes, potentials and SIGW settings (per-run reductions never mix runs),
    small lattices and a mix holding a 128^3 tabulated and a 256^3 SIGW run;
  - a request whose field leaves the tabulated range while stepping fails with
    IE_ERR_TABLE_RANGE and the requests beside it complete;
  - with a budget that holds one run at a time, admission is serial and every
    request still completes;
  - the FP64 twin of a split run reports its P_zeta difference (checked to 1e-3
    per shell here: split storage keeps fluctuations in FP32).
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import ie_engine as E
from conftest import potential_table

pytestmark = pytest.mark.gpu

HILL = dict(V0=3.338e-13, ns=0.97, dt=0.001, dN=1e-7, Nend=0.0005)
TAB = dict(V0=3e-9, dt=0.001, dN=1e-4, Nend=0.5)


def cfg_for(N, numerical, sigw=0, precision=E.IE_SPLIT, table=None, af=None):
    kw = dict(TAB if numerical else HILL)
    if af is not None:
        kw["af"] = af
    return E.make_config(N, precision=precision, numerical=numerical, sigw=sigw, table=table, **kw)


USR_PHI0, USR_DPHI0 = 2.9181235049318586, -0.06727651095116181


def submit(srv, spec, table):
    N, numerical, sigw, seed, fp64 = spec[:5]
    phi = spec[5] if len(spec) > 5 else None
    cfg = cfg_for(N, numerical, sigw, table=table if numerical else None, af=4.0)
    if numerical:
        phi0, dphi0 = phi or (USR_PHI0, USR_DPHI0)
        return srv.submit(cfg, seed, phi0, dphi0, 1, fp64, table)
    return srv.submit(cfg, seed, 0.0935, 0.000796, 1, fp64)


def serve(specs, budget, quantum=50, timeout=240.0):
    table = potential_table()
    srv = E.Server(budget, quantum, 4)
    ids = [submit(srv, s, table) for s in specs]
    got = {}
    t0 = time.time()
    while len(got) < len(ids):
        srv.pump()
        for r in srv.results():
            got[r["id"]] = r
        assert time.time() - t0 < timeout, "serving loop did not finish"
    srv.close()
    return [got[i] for i in ids]


def same_result(a, b):
    assert a["status"] == 0 and b["status"] == 0
    assert a["nbins"] > 0 and np.array_equal(a["pzeta"], b["pzeta"])
    assert a["gw_nbins"] == b["gw_nbins"] and np.array_equal(a["omega_gw"], b["omega_gw"])
    assert a["steps"] == b["steps"] and a["a_end"] == b["a_end"]


def test_alone_equals_in_mix():
    a, g = (16, 0, 0, 101, 0), (16, 0, 1, 103, 0)
    alone = serve([a], 1 << 34)[0]
    alone_g = serve([g], 1 << 34)[0]
    mix = serve([(32, 1, 0, 102, 0), a, g, (32, 0, 0, 104, 0)], 1 << 34)
    assert all(r["status"] == 0 for r in mix)
    same_result(alone, mix[1])
    same_result(alone_g, mix[2])
    assert alone["gw_nbins"] == 0
    # SIGW runs: Omega_GW over the kept shells floor(0.8 numbins), positive past k = 0
    nb = E.lib().ie_spectrum_numbins(16)
    assert alone_g["gw_nbins"] == int(0.8 * nb) and np.all(alone_g["omega_gw"][1:] > 0)


def test_large_runs_alone_equal_in_mix():
    """A 128^3 tabulated run and a 256^3 SIGW run, alone and beside each other and
    two small runs (both batches of a round on their own streams)."""
    big = [(128, 1, 0, 401, 0), (256, 0, 1, 402, 0)]
    alone = [serve([b], 1 << 36, timeout=600.0)[0] for b in big]
    mix = serve([big[0], (32, 0, 0, 403, 0), big[1], (64, 1, 1, 404, 0)], 1 << 36, timeout=600.0)
    same_result(alone[0], mix[0])
    same_result(alone[1], mix[2])
    assert mix[1]["status"] == 0 and mix[3]["status"] == 0
    assert alone[1]["gw_nbins"] == int(0.8 * E.lib().ie_spectrum_numbins(256))


def test_failing_request_does_not_stall_the_others():
    """A tabulated run started near the table's upper end with a large positive
    velocity leaves the table within a few dozen steps: the kernels flag the
    range error, that request returns IE_ERR_TABLE_RANGE, and the others (one
    admitted beside it, two waiting for the budget it holds) complete.  The
    budget holds two runs, so the last requests are admitted only if the failed
    run's footprint came back."""
    cfg = cfg_for(32, 1, table=potential_table(), af=4.0)
    one = E.lib().ie_footprint_bytes(cfg)
    specs = [(32, 1, 0, 501, 0, (3.49, 0.5)), (32, 1, 0, 502, 0), (32, 1, 0, 503, 0), (32, 0, 0, 504, 0)]
    res = serve(specs, int(one * 2.5))
    assert res[0]["status"] == E.IE_ERR_TABLE_RANGE
    assert all(r["status"] == 0 and r["nbins"] > 0 for r in res[1:])


def test_admission_by_footprint_is_serial_when_budget_holds_one():
    cfg = cfg_for(32, 0)
    one = E.lib().ie_footprint_bytes(cfg)
    res = serve([(32, 0, 0, 201 + q, 0) for q in range(3)], int(one * 1.5))
    assert all(r["status"] == 0 for r in res)
    for prev, nxt in zip(res, res[1:]):
        assert nxt["admit_s"] >= prev["done_s"] - 1e-3


def test_fp64_spot_check_reports_small_error():
    res = serve([(32, 0, 0, 301, 1)], 1 << 34)[0]
    assert res["status"] == 0
    assert 0.0 <= res["fp64_max_rel_err"] < 1e-3
