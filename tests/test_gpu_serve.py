"""Serving loop (ie_server_*): whole runs (initial conditions, inflation to af,
deltaN, P_zeta) admitted by device footprint and stepped together.

  - a run's P_zeta is bitwise the same alone and in a mix of sizes, potentials
    and SIGW settings (per-run reductions never mix runs);
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


def submit(srv, spec, table):
    N, numerical, sigw, seed, fp64 = spec
    cfg = cfg_for(N, numerical, sigw, table=table if numerical else None, af=4.0)
    if numerical:
        return srv.submit(cfg, seed, 2.9181235049318586, -0.06727651095116181, 1, fp64, table)
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


def test_alone_equals_in_mix():
    a = (16, 0, 0, 101, 0)
    alone = serve([a], 1 << 34)[0]
    mix = serve([(32, 1, 0, 102, 0), a, (16, 0, 1, 103, 0), (32, 0, 0, 104, 0)], 1 << 34)
    assert all(r["status"] == 0 for r in mix) and alone["status"] == 0
    assert alone["nbins"] > 0
    assert np.array_equal(alone["pzeta"], mix[1]["pzeta"])
    assert alone["steps"] == mix[1]["steps"] and alone["a_end"] == mix[1]["a_end"]


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
