"""Independent pins of the oracle: results fixed by mathematics or by library
routines, chosen so that a plausible restatement mistake (dropped term, wrong sign,
transposed index, wrong convention) fails one of them."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import oracle as O


def _rand_field(N, seed, dtype=np.float64):
    return np.random.default_rng(seed).standard_normal(N ** 3).astype(dtype)


@pytest.mark.parametrize("N", [4, 8, 16])
def test_fft_forward_is_conjugate_numpy_rfftn(N):
    """FFTEASY forward kernel is e^{+i}: F = conj(numpy.fft.rfftn) with the
    kz = N/2 plane in the separate Nyquist array (src/ffteasy.hpp:127-190)."""
    x = _rand_field(N, N)
    F, ny = O.fftrn(x, +1)
    ref = np.conj(np.fft.rfftn(x.reshape(N, N, N)))
    Fc = F.reshape(N, N, N // 2, 2)
    assert np.allclose(Fc[..., 0], ref[..., : N // 2].real, rtol=0, atol=1e-12 * N ** 3)
    assert np.allclose(Fc[..., 1], ref[..., : N // 2].imag, rtol=0, atol=1e-12 * N ** 3)
    nyc = ny.reshape(N, N, 2)
    assert np.allclose(nyc[..., 0], ref[..., N // 2].real, atol=1e-12 * N ** 3)
    assert np.allclose(nyc[..., 1], ref[..., N // 2].imag, atol=1e-12 * N ** 3)
    # inverse restores the field (1/N per transformed dimension)
    G = F.copy()
    nyg = ny.copy()
    O.lib().ieo_fftrn_d(O._dp(G), O._dp(nyg), N, -1)
    assert np.allclose(G, x, atol=1e-12)


def test_fft_float_matches_double():
    N = 16
    x = _rand_field(N, 3, np.float32)
    Ff, nyf = O.fftrn(x, +1)
    Fd, nyd = O.fftrn(x.astype(np.float64), +1)
    scale = np.max(np.abs(Fd))
    assert np.max(np.abs(Ff - Fd)) < 2e-5 * scale
    assert np.max(np.abs(nyf - nyd)) < 2e-5 * scale


def test_park_miller_known_answer():
    """Park & Miller (CACM 1988): from seed 1 the 10000th state is 1043618065.
    The reference discards 99 warm-up draws (src/initialize.cpp:26), so the
    9901st returned value is that state divided by 2^31 - 1."""
    cfg = O.make_config(4, 0, {"seed": 1})
    st = O.State(cfg)
    v = None
    for _ in range(9901):
        v = O.lib().ieo_rand_uniform(st.ptr)
    assert v == 1043618065 / 2147483647


def _state_with_field(N, f, a=1.0):
    cfg = O.make_config(N, 0, {})
    st = O.State(cfg)
    st.field(0)[:] = f
    st.set_background(a=a)
    return st


@pytest.mark.parametrize("n", [(1, 0, 0), (0, 2, 1), (3, 1, 2)])
def test_laplacian_plane_wave_eigenvalue(n):
    """cos(2 pi n.x/N) is an eigenvector of the 7-point Laplacian with eigenvalue
    -4 sum_d sin^2(pi n_d/N): gradient_energy = 0.5 lambda <f^2>/(a dx)^2."""
    N = 8
    i, j, k = np.meshgrid(*(np.arange(N),) * 3, indexing="ij")
    f = np.cos(2 * np.pi * (n[0] * i + n[1] * j + n[2] * k) / N).ravel()
    st = _state_with_field(N, f, a=1.3)
    lam = 4 * sum(np.sin(np.pi * q / N) ** 2 for q in n)
    dx = 10.0 / N
    expect = 0.5 * lam * np.mean(f ** 2) / (1.3 * dx) ** 2
    got = O.lib().ieo_gradient_energy(st.ptr)
    assert got == pytest.approx(expect, rel=1e-12)


def test_gradient_energy_forward_difference_identity():
    """-sum f lapl f == sum_d sum_x (f(x+e_d) - f(x))^2 on the periodic lattice."""
    N = 8
    f = _rand_field(N, 11)
    st = _state_with_field(N, f)
    g = f.reshape(N, N, N)
    s = sum(np.sum((np.roll(g, -1, axis=d) - g) ** 2) for d in range(3))
    dx = 10.0 / N
    expect = 0.5 * s / dx ** 2 / N ** 3
    assert O.lib().ieo_gradient_energy(st.ptr) == pytest.approx(expect, rel=1e-12)


def test_spectrum_parseval_and_weights():
    """Shell binning: weights sum to N^3 less the modes beyond the last shell, and
    sum count*P/norm1 equals N^3 sum f^2 less the power of those modes."""
    N = 8
    x = _rand_field(N, 5)
    F, ny = O.fftrn(x, +1)
    norm1 = 0.37
    p, cnt, vals = O.bin_spectrum(N, 10.0, F, ny, norm1)
    full = np.fft.fftn(x.reshape(N, N, N))
    n = np.fft.fftfreq(N, 1.0 / N)
    nx, nyy, nz = np.meshgrid(n, n, n, indexing="ij")
    mag = np.sqrt(nx ** 2 + nyy ** 2 + np.abs(nz) ** 2)
    # the r2c layout counts kz = -N/2 as +N/2: same |n|
    bins = np.floor(mag + 0.5).astype(int)
    kept = bins < len(p)
    assert cnt.sum() == kept.sum()
    power = np.abs(full) ** 2
    assert np.sum(cnt * vals) / norm1 == pytest.approx(np.sum(power[kept]), rel=1e-12)
    assert np.sum(power) == pytest.approx(N ** 3 * np.sum(x ** 2), rel=1e-12)
    for b in range(len(p)):
        sel = kept & (bins == b)
        if sel.any():
            assert vals[b] == pytest.approx(norm1 * power[sel].mean(), rel=1e-12)
        assert p[b] == pytest.approx(2 * np.pi / 10.0 * b)


def _gw_state(N):
    cfg = O.make_config(N, 0, {})
    return O.State(cfg)


def test_gw_pure_trace_projects_to_zero():
    """Lambda projects out a pure-trace tensor: P(s 1)P - 1/2 P tr(P s) = 0."""
    N = 8
    st = _gw_state(N)
    s = _rand_field(N, 7, np.float32)
    for c in range(3):
        st.tensor(c)[:] = s
    p, cnt, vals = st.gw_spectrum(0)
    ref_scale = np.mean(s.astype(np.float64) ** 2) * N ** 3
    assert np.max(np.abs(vals)) < 1e-5 * ref_scale * 0.5 * (10.0 / O.make_config(N, 0).rescale_B) ** 3 / N ** 6


def test_gw_single_tt_plane_wave():
    """h_xy = A cos(2 pi m z / N) is transverse-traceless for k along z: its
    projected power is 2 |h_xy(k)|^2 per mode, all in shell m."""
    N, m, A = 8, 2, 0.5
    st = _gw_state(N)
    k = np.arange(N)
    wave = (A * np.cos(2 * np.pi * m * k / N)).astype(np.float32)
    h = np.broadcast_to(wave, (N, N, N)).ravel()
    st.tensor(3)[:] = h
    p, cnt, vals = st.gw_spectrum(0)
    B = st.cfg.rescale_B
    norm1 = np.float32(0.5 * (10.0 / B) ** 3 / N ** 6)
    amp2 = (A * N ** 3 / 2) ** 2
    expect = norm1 * (2 * 2 * amp2) / cnt[m]
    assert vals[m] == pytest.approx(float(expect), rel=1e-5)
    others = np.delete(vals, m)
    assert np.max(np.abs(others)) < 1e-6 * vals[m]


def test_tabulated_interpolation_matches_numpy_interp(table):
    cfg = O.make_config(4, 1, {}, table)
    st = O.State(cfg, table)
    rng = np.random.default_rng(0)
    xs = rng.uniform(table.x[-1], table.x[0], 200)
    B2 = cfg.rescale_B ** 2
    want_v = np.interp(xs, table.x[::-1], table.v[::-1]) / B2
    want_dv = np.interp(xs, table.x[::-1], table.dv[::-1]) / B2
    got_v = np.array([O.lib().ieo_potential(st.ptr, v) for v in xs])
    got_dv = np.array([O.lib().ieo_potential_derivative_value(st.ptr, v) for v in xs])
    assert np.allclose(got_v, want_v, rtol=1e-12)
    assert np.allclose(got_dv, want_dv, rtol=1e-12)
    err = C.c_int()
    for v in xs[:50]:
        l1 = O.lib().ieo_find_bracketing_index(st.ptr, v, 1, C.byref(err))
        l2 = O.lib().ieo_find_bracketing_index(st.ptr, v, 400, C.byref(err))
        assert l1 == l2 and table.x[l1 - 1] >= v > table.x[l1]
    O.lib().ieo_find_bracketing_index(st.ptr, table.x[0] + 1e-3, 1, C.byref(err))
    assert err.value == 2


def test_hilltop_closed_form():
    cfg = O.make_config(4, 0, {})
    st = O.State(cfg)
    for v in (0.0, 0.0935, -0.3):
        V = O.lib().ieo_potential(st.ptr, v)
        dV = O.lib().ieo_potential_derivative_value(st.ptr, v)
        assert V == pytest.approx(1.0 - (1.0 - 0.97) / 4.0 * v * v, rel=1e-14)
        assert dV == pytest.approx(-(1.0 - 0.97) / 2.0 * v, rel=1e-14, abs=1e-300)


def test_homogeneous_limit_matches_background_ode():
    """With no fluctuations the lattice reduces to the FLRW background
    (src/evolution.cpp:150-284 vs the RHS of :509-516): with rescale_s = 1 the
    step is constant and the staggered leapfrog is second order, so
    phi'' = -3 (a'/a) phi' - V'(phi), a'' = a V - 2 a'^2/a must be matched to O(dt^2)."""
    N = 4
    dt = 0.002
    cfg = O.make_config(N, 0, {"dt": dt, "af": 5.0, "rescale_s": 1.0})
    st = O.State(cfg)
    phi0, dphi0 = 0.0935, 0.000796
    st.field(0)[:] = phi0
    st.field(1)[:] = dphi0
    ns = 0.97
    V = lambda p: 1 - (1 - ns) / 4 * p * p
    dV = lambda p: -(1 - ns) / 2 * p
    H0 = np.sqrt((0.5 * dphi0 ** 2 + V(phi0)) / 3)
    st.set_background(a=1.0, ad=H0, astep=1.0, t=0.0)

    def lattice(nsteps):
        st.field(0)[:] = phi0
        st.field(1)[:] = dphi0
        st.set_background(a=1.0, ad=H0, astep=1.0, t=0.0)
        O.lib().ieo_evolve_fields(st.ptr, 0.5 * dt)
        assert O.lib().ieo_leapfrog_steps(st.ptr, nsteps) == nsteps
        O.lib().ieo_evolve_fields(st.ptr, -0.5 * dt)
        bg = st.background()
        return bg["t"], bg["a"], st.field(0)[0]

    def rhs(t, y):
        p, pd, a, ad = y
        return [pd, -3 * ad / a * pd - dV(p), ad, a * V(p) - 2 * ad * ad / a]

    t1, a1, p1 = lattice(500)
    sol = solve_ivp(rhs, (0.0, t1), [phi0, dphi0, 1.0, H0], rtol=1e-12, atol=1e-15)
    p_ref, _, a_ref, _ = sol.y[:, -1]
    err_a = abs(a1 - a_ref) / a_ref
    err_p = abs(p1 - p_ref) / abs(p_ref)
    assert err_a < 1e-6 and err_p < 1e-6
    assert a1 > 1.5  # the run covered a real expansion
