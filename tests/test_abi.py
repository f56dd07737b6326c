"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/ie_api.h declares; host-only entry points run here."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ie_engine as E
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(E.LIB_PATH))
    names = E.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    E.lib()  # every binding resolves
    bound = set()
    for n in E.declared_symbols():
        assert hasattr(E.lib(), n)
        bound.add(n)
    assert "ie_leapfrog_steps" in bound and "ie_deltaN_steps" in bound


@pytest.mark.parametrize("N", [16, 32, 64])
def test_modes_match_oracle_bitwise(N):
    """modes.dat closed form (src/output.cpp:674-738), host side of the library."""
    B = np.sqrt(3e-9)
    assert np.array_equal(E.modes(N, 10.0, B), O.modes(N, 10.0, B))


@pytest.mark.parametrize("case", ["hilltop_n16", "tabulated_n32"])
def test_host_mode_fill_matches_oracle_bitwise(case):
    """The host k-space fill of initializef (Park-Miller draw order, Hermitian
    mirrors, Nyquist plane) equals the oracle's, which the golden fixtures pin
    to the reference's box.dat."""
    from conftest import golden_config
    ocfg, params, table = golden_config(case)
    st = O.State(ocfg, table)
    assert O.lib().ieo_fill_modes(st.ptr) == 0
    cfg = E.make_config(ocfg.N, numerical=ocfg.numerical_potential, table=table, V0=ocfg.V0, ns=ocfg.ns,
                        L=ocfg.L, dt=ocfg.dt)
    fk, fny, fdk, fdny, H = E.fill_modes(cfg, ocfg.seed, ocfg.initial_field, ocfg.initial_derivative)
    assert np.array_equal(fk, st.field(0))
    assert np.array_equal(fdk, st.field(1))
    assert np.array_equal(fny, st.field(3))
    assert np.array_equal(fdny, st.field(4))
    assert H == st.background()["hubble_init"]


def test_fill_modes_rejects_courant_violation():
    cfg = E.make_config(16, dt=0.5)
    with pytest.raises(E.EngineError) as e:
        E.fill_modes(cfg, 8, 0.0935, 0.000796)
    assert e.value.status == 2


def test_footprint_and_numbins():
    cfg = E.make_config(256, precision=E.IE_SPLIT, sigw=1)
    fp = E.lib().ie_footprint_bytes(C.byref(cfg))
    assert 1.0e9 < fp < 3.0e9
    assert E.lib().ie_spectrum_numbins(256) == 222
    assert E.lib().ie_spectrum_numbins(64) == O.lib().ieo_spectrum_numbins(64)
