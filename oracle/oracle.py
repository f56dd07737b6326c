"""ctypes binding of the oracle library (liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the InflationEasy hot path used as
the parity checker and the CPU baseline.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module.

Also holds small helpers the checker needs: a reader for the reference's
``params.txt`` format (src/runtime_parameters.cpp:195-365) and for its numeric
``results/*.dat`` files.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

LEAPFROG, RK4, RK45 = 0, 1, 2
GENERIC, MONOTONIC, ANTIMONOTONIC = 0, 1, -1


class Config(C.Structure):
    """Mirror of ``ieo_config`` in ie_oracle.h (field order matters)."""

    _fields_ = [
        ("N", C.c_int),
        ("numerical_potential", C.c_int),
        ("calculate_sigw", C.c_int),
        ("perform_deltaN", C.c_int),
        ("monotonic", C.c_int),
        ("seed", C.c_int),
        ("rescale_s", C.c_double),
        ("af", C.c_double),
        ("V0", C.c_double),
        ("rescale_B", C.c_double),
        ("ns", C.c_double),
        ("initial_field", C.c_double),
        ("initial_derivative", C.c_double),
        ("L", C.c_double),
        ("dt", C.c_double),
        ("integrator", C.c_int),
        ("deltaN_integrator", C.c_int),
        ("rk45_abs_tol", C.c_double),
        ("rk45_rel_tol", C.c_double),
        ("rk45_min_dt", C.c_double),
        ("rk45_max_dt", C.c_double),
        ("rk45_safety", C.c_double),
        ("output_freq", C.c_int),
        ("output_infrequent_freq", C.c_int),
        ("dN", C.c_double),
        ("Nend", C.c_double),
        ("use_phiref_manual", C.c_int),
        ("phiref_manual_value", C.c_double),
        ("high_cutoff_index", C.c_double),
        ("low_cutoff_index", C.c_double),
        ("forcing_cutoff", C.c_int),
        ("output_spectra", C.c_int),
        ("output_histogram", C.c_int),
        ("output_energy", C.c_int),
        ("output_box3D", C.c_int),
        ("output_box2D", C.c_int),
        ("nbins", C.c_int),
        ("int_err", C.c_int),
        ("int_errN", C.c_int),
        ("tab_n", C.c_int),
        ("tab_x", C.POINTER(C.c_double)),
        ("tab_v", C.POINTER(C.c_double)),
        ("tab_dv", C.POINTER(C.c_double)),
        ("post_inflation", C.c_int),
        ("horizon_factor", C.c_double),
        ("omega", C.c_double),
        ("dt_post", C.c_double),
        ("af_post", C.c_double),
        ("post_integrator", C.c_int),
        ("output_bispectrum", C.c_int),
        ("output_LOG", C.c_int),
        ("eta_log", C.c_double),
    ]


_lib = None


def build() -> None:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        sig = {
            "ieo_config_defaults": (None, [C.POINTER(Config), C.c_int, C.c_int]),
            "ieo_config_finalize": (None, [C.POINTER(Config), C.c_int, C.c_int]),
            "ieo_create": (P, [C.POINTER(Config)]),
            "ieo_destroy": (None, [P]),
            "ieo_field": (dp, [P, C.c_int]),
            "ieo_tensor": (fp, [P, C.c_int]),
            "ieo_lstart": (ip, [P]),
            "ieo_get_background": (None, [P, dp]),
            "ieo_set_background": (None, [P, dp]),
            "ieo_rand_uniform": (C.c_double, [P]),
            "ieo_find_bracketing_index": (C.c_int, [P, C.c_double, C.c_int, ip]),
            "ieo_potential": (C.c_double, [P, C.c_double]),
            "ieo_potential_derivative_value": (C.c_double, [P, C.c_double]),
            "ieo_initialize": (C.c_int, [P]),
            "ieo_fill_modes": (C.c_int, [P]),
            "ieo_gradient_energy": (C.c_double, [P]),
            "ieo_kin_energy": (C.c_double, [P]),
            "ieo_potential_energy": (C.c_double, [P]),
            "ieo_evolve_derivs": (None, [P, C.c_double]),
            "ieo_evolve_fields": (None, [P, C.c_double]),
            "ieo_leapfrog_steps": (C.c_int, [P, C.c_int]),
            "ieo_rk4_step": (None, [P, C.c_double]),
            "ieo_rk45_step": (C.c_int, [P, C.c_double, C.c_double, dp]),
            "ieo_fftrn_d": (None, [dp, dp, C.c_int, C.c_int]),
            "ieo_fftrn_f": (None, [fp, fp, C.c_int, C.c_int]),
            "ieo_spectrum_numbins": (C.c_int, [C.c_int]),
            "ieo_bin_spectrum": (None, [C.c_int, C.c_double, dp, dp, C.c_double, dp, ip, dp]),
            "ieo_spectrum_f": (None, [P, dp, ip, dp]),
            "ieo_gw_spectrum": (None, [P, C.c_int, dp, ip, dp]),
            "ieo_modes": (None, [C.c_int, C.c_double, C.c_double, dp]),
            "ieo_initializeN": (None, [P]),
            "ieo_evolve_derivsN": (None, [P, C.c_double]),
            "ieo_evolve_fieldsN": (None, [P, C.c_double]),
            "ieo_get_phiref": (C.c_double, [P]),
            "ieo_deltaN_zeta_map": (C.c_int, [P, C.POINTER(C.c_float)]),
            "ieo_float_mean": (C.c_float, [C.POINTER(C.c_double), C.c_size_t, C.c_double]),
            "ieo_rk4_stepN": (None, [P, C.c_double]),
            "ieo_rk45_stepN": (C.c_int, [P, C.c_double, C.c_double, dp]),
            "ieo_post_initialize": (None, [P]),
            "ieo_post_evolve_derivs": (None, [P, C.c_double]),
            "ieo_post_rk4_step": (None, [P, C.c_double]),
            "ieo_post_rk45_step": (C.c_int, [P, C.c_double, C.c_double, dp]),
            "ieo_cutoff_modes": (None, [P]),
            "ieo_bispectrum": (None, [P, dp, ip, dp, dp]),
            "ieo_deltaN_log": (C.c_int, [P]),
            "ieo_spectrum_log": (None, [P, dp, ip, dp]),
            "ieo_histogram_log": (C.c_int, [P, dp, dp, dp]),
            "ieo_program": (C.c_int, [C.POINTER(Config), C.c_char_p]),
            "ieo_num_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# --------------------------------------------------------------- parameters

_INT_KEYS = {
    "seed", "output_freq", "output_infrequent_freq", "use_phiref_manual", "forcing_cutoff",
    "output_spectra", "output_histogram", "output_energy", "output_box3D", "output_box2D",
    "output_bispectrum", "output_LOG", "screen_updates", "nbins", "int_err", "int_errN",
}
_INTEGRATORS = {"leapfrog": 0, "lf": 0, "rk4": 1, "rk45": 2, "rkf45": 2, "dopri5": 2}


def parse_params_text(text: str) -> dict:
    """key = value lines, '#' comments (src/runtime_parameters.cpp:223-240)."""
    out = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        k, v = (s.strip() for s in line.split("=", 1))
        if k.endswith("integrator"):
            s = v.lower()
            out[k] = _INTEGRATORS[s] if s in _INTEGRATORS else int(s)
        elif k in _INT_KEYS:
            out[k] = int(v)
        else:
            try:
                out[k] = float(v)
            except ValueError:
                out[k] = v
    return out


class Table:
    """Tabulated potential, kept alive for the pointers handed to C."""

    def __init__(self, x, v, dv):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.v = np.ascontiguousarray(v, dtype=np.float64)
        self.dv = np.ascontiguousarray(dv, dtype=np.float64)


_POST_KEYS = ("horizon_factor", "omega", "dt_post", "af_post", "post_integrator")


def make_config(N: int, numerical: int, params: dict | None = None, table: Table | None = None,
                sigw: int = 1, deltaN: int = 1, post: int = 0) -> Config:
    """Oracle config; the post-inflation keys of params.txt are read only with
    post = 1, as the reference ignores them when built without that phase
    (src/runtime_parameters.cpp:327-332)."""
    cfg = Config()
    lib().ieo_config_defaults(C.byref(cfg), N, numerical)
    cfg.calculate_sigw = sigw
    cfg.perform_deltaN = deltaN
    cfg.post_inflation = post
    params = dict(params or {})
    rescale_B_set = "rescale_B" in params
    af_set = "af" in params
    remap = {"inflation_integrator": "integrator", "dt_post_inflation": "dt_post",
             "af_post_inflation": "af_post", "post_inflation_integrator": "post_integrator"}
    for k, v in params.items():
        k = remap.get(k, k)
        if k == "screen_updates":
            continue
        if k in _POST_KEYS and not post:
            continue
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    if table is not None:
        cfg.tab_n = len(table.x)
        cfg.tab_x = table.x.ctypes.data_as(C.POINTER(C.c_double))
        cfg.tab_v = table.v.ctypes.data_as(C.POINTER(C.c_double))
        cfg.tab_dv = table.dv.ctypes.data_as(C.POINTER(C.c_double))
    lib().ieo_config_finalize(C.byref(cfg), int(rescale_B_set), int(af_set))
    return cfg


# ------------------------------------------------------------------- state

class State:
    """Owning wrapper of an ``ieo_state``."""

    def __init__(self, cfg: Config, table: Table | None = None):
        self.cfg = cfg
        self.table = table
        self.N = cfg.N
        self.ptr = lib().ieo_create(C.byref(cfg))
        if not self.ptr:
            raise MemoryError("ieo_create failed")

    def close(self):
        if self.ptr:
            lib().ieo_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def _view(self, p, n, ctype):
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), shape=(n,))

    def field(self, which: int) -> np.ndarray:
        """0 f, 1 fd, 2 deltaN (N^3 view); 3/4 Nyquist planes (N*2N view)."""
        n = self.N ** 3 if which < 3 else self.N * 2 * self.N
        return self._view(lib().ieo_field(self.ptr, which), n, C.c_double)

    def tensor(self, comp: int) -> np.ndarray:
        return self._view(lib().ieo_tensor(self.ptr, comp), self.N ** 3, C.c_float)

    def background(self) -> dict:
        b = (C.c_double * 8)()
        lib().ieo_get_background(self.ptr, b)
        return dict(zip(["t", "a", "ad", "ad2", "astep", "Ne", "phiref", "hubble_init"], list(b)))

    def set_background(self, **kw):
        cur = self.background()
        cur.update(kw)
        b = (C.c_double * 8)(*[cur[k] for k in ["t", "a", "ad", "ad2", "astep", "Ne", "phiref", "hubble_init"]])
        lib().ieo_set_background(self.ptr, b)

    def initialize(self):
        rc = lib().ieo_initialize(self.ptr)
        if rc:
            raise RuntimeError(f"ieo_initialize failed ({rc})")

    def spectrum_f(self):
        nb = lib().ieo_spectrum_numbins(self.N)
        p = np.zeros(nb)
        c = np.zeros(nb, dtype=np.int32)
        v = np.zeros(nb)
        lib().ieo_spectrum_f(self.ptr, _dp(p), _ip(c), _dp(v))
        return p, c, v

    def bispectrum(self):
        nb = int(np.sqrt(3.0) * (self.N // 2)) + 1
        p, re, im = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        c = np.zeros(nb, dtype=np.int32)
        lib().ieo_bispectrum(self.ptr, _dp(p), _ip(c), _dp(re), _dp(im))
        return p, c, re, im

    def gw_spectrum(self, hdot: int):
        nb = int(np.sqrt(3.0) * (self.N // 2)) + 1
        p = np.zeros(nb)
        c = np.zeros(nb, dtype=np.int32)
        v = np.zeros(nb)
        lib().ieo_gw_spectrum(self.ptr, hdot, _dp(p), _ip(c), _dp(v))
        return p, c, v


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def float_mean(dn: np.ndarray, cutoff: float) -> np.float32:
    """saveN's float mean of a deltaN map (ieo_float_mean)."""
    a = np.ascontiguousarray(dn, dtype=np.float64).ravel()
    return np.float32(lib().ieo_float_mean(_dp(a), a.size, cutoff))


def fftrn(field: np.ndarray, forward: int):
    """FFTEASY-layout real 3-D FFT (returns transformed copy and Nyquist plane)."""
    N = round(field.size ** (1 / 3))
    if field.dtype == np.float32:
        f = np.ascontiguousarray(field, dtype=np.float32).copy()
        ny = np.zeros(N * 2 * N, dtype=np.float32)
        lib().ieo_fftrn_f(_fp(f), _fp(ny), N, forward)
    else:
        f = np.ascontiguousarray(field, dtype=np.float64).copy()
        ny = np.zeros(N * 2 * N)
        lib().ieo_fftrn_d(_dp(f), _dp(ny), N, forward)
    return f, ny


def bin_spectrum(N: int, L: float, F: np.ndarray, ny: np.ndarray, norm1: float):
    nb = lib().ieo_spectrum_numbins(N)
    p = np.zeros(nb)
    c = np.zeros(nb, dtype=np.int32)
    v = np.zeros(nb)
    lib().ieo_bin_spectrum(N, L, _dp(np.ascontiguousarray(F, dtype=np.float64)),
                           _dp(np.ascontiguousarray(ny, dtype=np.float64)), norm1, _dp(p), _ip(c), _dp(v))
    return p, c, v


def modes(N: int, L: float, B: float) -> np.ndarray:
    nb = int(np.sqrt(3.0) * (N // 2)) + 1
    out = np.zeros(nb)
    lib().ieo_modes(N, L, B, _dp(out))
    return out


def run_program(cfg: Config, results_dir: str | os.PathLike) -> int:
    return lib().ieo_program(C.byref(cfg), str(results_dir).encode())


# ------------------------------------------------------------ results files

def read_blocks(path: str | os.PathLike) -> list[np.ndarray]:
    """Read a results/*.dat file as a list of blank-line separated numeric blocks."""
    blocks, cur = [], []
    with open(path) as fh:
        for line in fh:
            s = line.split()
            if not s:
                if cur:
                    blocks.append(np.array(cur, dtype=np.float64))
                    cur = []
                continue
            cur.append([float(x) for x in s])
    if cur:
        blocks.append(np.array(cur, dtype=np.float64))
    return blocks


# ------------------------------------------------------------------ analysis
# The reference's analysis notebook (notebooks/plot.ipynb) turns the written
# spectra into curvature and tensor power spectra; each spectrum row pairs with
# the modes.dat row of the same index.

def pzeta_linear(k, p_phi, phidot, hubble):
    """P_zeta = P_phi k^3/(2 pi^2) (phidot/H)^-2 (notebooks/plot.ipynb cell 9, with
    phidot = velocity.dat[-1, 2] and H = sf.dat[-1, 2])."""
    k = np.asarray(k, dtype=np.float64)
    return np.asarray(p_phi, dtype=np.float64) * k ** 3 / (2 * np.pi ** 2) * (phidot / hubble) ** (-2)


def pzeta_deltaN(k, p_dn):
    """P_zeta = P_dN k^3/(2 pi^2) (notebooks/plot.ipynb cell 9)."""
    k = np.asarray(k, dtype=np.float64)
    return np.asarray(p_dn, dtype=np.float64) * k ** 3 / (2 * np.pi ** 2)


def delta_h2(k, p_h):
    """Delta_h^2 = k^3/(2 pi^2) P_h (notebooks/plot.ipynb cell 23)."""
    k = np.asarray(k, dtype=np.float64)
    return (k ** 3 / (2.0 * np.pi ** 2)) * np.asarray(p_h, dtype=np.float64)


def omega_gw(k, p_h, p_hdot, a, hubble):
    """Omega_GW = 1/(12 H^2) k^3/(2 pi^2) (P_hdot + (k/a)^2 P_h) (notebooks/plot.ipynb
    cell 23, with a = sf.dat[t, 1], H = sf.dat[t, 2])."""
    k = np.asarray(k, dtype=np.float64)
    term2 = (k / a) ** 2 * np.asarray(p_h, dtype=np.float64)
    return (1.0 / (12.0 * hubble ** 2)) * (k ** 3 / (2.0 * np.pi ** 2)) * (np.asarray(p_hdot, dtype=np.float64) + term2)
