"""Pins of the oracle's log-mapped zeta (build_deltaN_log, histogramsLOG,
spectraLOG; src/output.cpp:167-198, 1280-1403).

The reference cannot produce fixtures for these outputs: with output_LOG = 1 its
save_last writes the map into the deltaN vector before initializeN allocates it
(src/output.cpp:1511-1516, src/initialize.cpp:221) and the program dies.  The
oracle's restatement is pinned instead by what the mathematics fixes: the
small-eta limit (the linear map -(f - <f>)/factt), the inverse map, the mask of
non-positive arguments, the invalid-factor branch, and the spectrum as the
composition of the FFT and shell binning pinned elsewhere."""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import oracle as O

N = 16


def make_state(eta, f, fd, a=2.0, ad=3.0, nbins=32):
    cfg = O.make_config(N, 0, {"eta_log": eta, "nbins": nbins})
    st = O.State(cfg)
    st.field(0)[:] = f
    st.field(1)[:] = fd
    st.set_background(a=a, ad=ad)
    return st


def fields(seed=5, amp=1e-3, v=0.02):
    rng = np.random.default_rng(seed)
    f = 0.09 + amp * rng.standard_normal(N ** 3)
    fd = v * (1.0 + 0.1 * rng.standard_normal(N ** 3))
    return f, fd


def factt_of(fd, a=2.0, ad=3.0, s=0.0):
    return float(np.mean(fd * a ** (s - 1.0) / (ad * a ** (s - 2.0))))


def test_small_eta_is_the_linear_map():
    f, fd = fields()
    eta = 1e-5
    st = make_state(eta, f, fd)
    assert O.lib().ieo_deltaN_log(st.ptr) == 1
    x = (f - f.mean()) / factt_of(fd)
    # (1/eta) log(1 - eta x) = -x - eta x^2 / 2 - eta^2 x^3 / 3 - ...
    series = -x - eta * x ** 2 / 2
    # round-off of log(1 - eta x) / eta: ~1e-16 / eta
    assert np.allclose(st.field(2), series, rtol=1e-9, atol=1e-10 * np.abs(x).max())
    assert not np.allclose(st.field(2), -x, rtol=1e-9, atol=0)   # the second order is seen


def test_inverse_map_recovers_the_field():
    f, fd = fields()
    eta = -0.5
    st = make_state(eta, f, fd)
    O.lib().ieo_deltaN_log(st.ptr)
    dn = st.field(2)
    back = f.mean() + factt_of(fd) * (1.0 - np.exp(eta * dn)) / eta
    assert np.allclose(back - f.mean(), f - f.mean(), rtol=1e-9, atol=1e-15)


def test_mask_and_histogram():
    f, fd = fields(amp=0.05, v=0.02)
    eta = 0.9
    st = make_state(eta, f, fd, nbins=32)
    arg = 1.0 - eta * (f - f.mean()) / factt_of(fd)
    live = arg > 0
    assert 0 < live.sum() < f.size           # some sites are masked
    O.lib().ieo_deltaN_log(st.ptr)
    dn = st.field(2).copy()
    assert np.all(dn[~live] == 0.0)
    bmin, df = C.c_double(), C.c_double()
    freq = np.zeros(32)
    assert O.lib().ieo_histogram_log(st.ptr, C.byref(bmin), C.byref(df), O._dp(freq)) == 1
    assert abs(freq.sum() - 1.0) < 1e-12
    lo = min(dn[0], dn[live].min())
    hi = max(dn[0], dn[live].max())
    assert bmin.value == lo and abs(df.value - (hi - lo) / 32) <= 1e-15 * abs(hi - lo)
    # every live site lands in a bin: counts = freq * points
    assert np.allclose(freq * live.sum(), np.round(freq * live.sum()), atol=1e-9)


def test_invalid_factor_gives_zeros():
    f, _ = fields()
    st = make_state(-0.5, f, np.zeros(N ** 3))
    assert O.lib().ieo_deltaN_log(st.ptr) == 0
    assert np.all(st.field(2) == 0.0)
    bmin, df = C.c_double(), C.c_double()
    freq = np.ones(32)
    assert O.lib().ieo_histogram_log(st.ptr, C.byref(bmin), C.byref(df), O._dp(freq)) == 0
    assert bmin.value == 0.0 and df.value == 1.0 and np.all(freq == 0.0)


def test_spectrum_log_is_fft_then_binning():
    f, fd = fields()
    st = make_state(-0.5, f, fd)
    nb = O.lib().ieo_spectrum_numbins(N)
    p, c, v = np.zeros(nb), np.zeros(nb, dtype=np.int32), np.zeros(nb)
    O.lib().ieo_spectrum_log(st.ptr, O._dp(p), O._ip(c), O._dp(v))
    st2 = make_state(-0.5, f, fd)
    O.lib().ieo_deltaN_log(st2.ptr)
    F, ny = O.fftrn(st2.field(2).copy(), 1)
    norm1 = (st.cfg.L / st.cfg.rescale_B) ** 3 / N ** 6
    p2, c2, v2 = O.bin_spectrum(N, st.cfg.L, F, ny, norm1)
    assert np.array_equal(c, c2) and np.array_equal(p, p2)
    assert np.allclose(v, v2, rtol=1e-13, atol=0)
