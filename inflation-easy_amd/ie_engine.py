"""ctypes binding of the gfx950 lattice engine (libie_amd.so, C ABI in include/ie_api.h).

Used by the tests, bench.py and __graft_entry__.  It loads the in-tree library
and fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libie_amd.so"
HEADER = HERE.parent / "include" / "ie_api.h"

IE_FP64, IE_SPLIT = 0, 1
IE_LEAPFROG, IE_RK4, IE_RK45 = 0, 1, 2
IE_GENERIC, IE_MONOTONIC, IE_ANTIMONOTONIC = 0, 1, -1
IE_FIELD_F, IE_FIELD_FD, IE_FIELD_DELTAN = 0, 1, 2

STATUS = {0: "IE_OK", 1: "IE_ERR_ARG", 2: "IE_ERR_COURANT", 3: "IE_ERR_TABLE_RANGE", 4: "IE_ERR_NONFINITE",
          5: "IE_ERR_RK45", 6: "IE_ERR_OOM", 7: "IE_ERR_HIP", 8: "IE_ERR_STATE"}
(IE_OK, IE_ERR_ARG, IE_ERR_COURANT, IE_ERR_TABLE_RANGE, IE_ERR_NONFINITE, IE_ERR_RK45, IE_ERR_OOM, IE_ERR_HIP,
 IE_ERR_STATE) = range(9)


class EngineError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Config(C.Structure):
    """Mirror of ``ie_config``."""

    _fields_ = [
        ("N", C.c_int), ("precision", C.c_int), ("numerical_potential", C.c_int), ("calculate_sigw", C.c_int),
        ("monotonic", C.c_int), ("integrator", C.c_int), ("deltaN_integrator", C.c_int),
        ("rescale_s", C.c_double), ("af", C.c_double), ("V0", C.c_double), ("rescale_B", C.c_double),
        ("ns", C.c_double), ("L", C.c_double), ("dt", C.c_double),
        ("rk45_abs_tol", C.c_double), ("rk45_rel_tol", C.c_double), ("rk45_min_dt", C.c_double),
        ("rk45_max_dt", C.c_double), ("rk45_safety", C.c_double),
        ("dN", C.c_double), ("Nend", C.c_double), ("use_phiref_manual", C.c_int),
        ("phiref_manual_value", C.c_double),
        ("tab_n", C.c_int), ("tab_x", C.POINTER(C.c_double)), ("tab_v", C.POINTER(C.c_double)),
        ("tab_dv", C.POINTER(C.c_double)),
    ]


class IcParams(C.Structure):
    """Mirror of ``ie_ic_params``."""

    _fields_ = [("seed", C.c_int), ("initial_field", C.c_double), ("initial_derivative", C.c_double),
                ("high_cutoff_index", C.c_double), ("low_cutoff_index", C.c_double)]


class Request(C.Structure):
    """Mirror of ``ie_request``."""

    _fields_ = [("cfg", Config), ("ic", IcParams), ("deltaN", C.c_int), ("fp64_check", C.c_int)]


class Result(C.Structure):
    """Mirror of ``ie_result``."""

    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("N", C.c_int32), ("submit_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double), ("steps", C.c_int64), ("a_end", C.c_double),
                ("nbins", C.c_int32), ("pzeta", C.c_double * 512), ("fp64_max_rel_err", C.c_double),
                ("gw_nbins", C.c_int32), ("omega_gw", C.c_double * 512), ("gw_fp64_max_rel_err", C.c_double)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pzeta", "omega_gw")}
        d["pzeta"] = np.array(self.pzeta[: self.nbins])
        d["omega_gw"] = np.array(self.omega_gw[: self.gw_nbins])
        d["latency_s"] = self.done_s - self.submit_s
        return d


class Profile(C.Structure):
    """Mirror of ``ie_profile``."""

    _fields_ = [("step_ms", C.c_double), ("reduce_ms", C.c_double), ("kick_drift_ms", C.c_double),
                ("tensor_ms", C.c_double), ("steps", C.c_int64), ("loop_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    """Mirror of ``ie_tuning``."""

    _fields_ = [("ncand", C.c_int), ("elems", C.c_int * 3), ("kick_ms", C.c_double * 3),
                ("ntensor", C.c_int), ("tensor_code", C.c_int * 8), ("tensor_ms", C.c_double * 8),
                ("nmarch", C.c_int), ("march_ms", C.c_double * 2),
                ("kick_elems", C.c_int), ("tensor_code_chosen", C.c_int), ("march_variant", C.c_int),
                ("steps", C.c_int64)]

    def as_dict(self) -> dict:
        n, t = self.ncand, self.ntensor
        return {"candidates": list(self.elems[:n]), "kick_ms": list(self.kick_ms[:n]),
                "tensor_codes": list(self.tensor_code[:t]), "tensor_ms": list(self.tensor_ms[:t]),
                "march_ms": list(self.march_ms[:self.nmarch]), "march_variant": self.march_variant,
                "kick_elems": self.kick_elems, "tensor_code": self.tensor_code_chosen, "steps": self.steps}


class PostParams(C.Structure):
    """Mirror of ``ie_post_params``."""
    _fields_ = [("horizon_factor", C.c_double), ("omega", C.c_double), ("dt", C.c_double), ("af", C.c_double),
                ("integrator", C.c_int), ("from_deltaN", C.c_int)]


class Background(C.Structure):
    """Mirror of ``ie_background``."""

    _fields_ = [("t", C.c_double), ("a", C.c_double), ("ad", C.c_double), ("ad2", C.c_double),
                ("astep", C.c_double), ("Ne", C.c_double), ("phiref", C.c_double), ("hubble_init", C.c_double),
                ("dt_last", C.c_double), ("steps", C.c_int64), ("done", C.c_int32), ("error", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        st = C.c_int
        sig = {
            "ie_create": (st, [C.POINTER(Config), C.POINTER(P)]),
            "ie_destroy": (st, [P]),
            "ie_last_error": (C.c_char_p, [P]),
            "ie_footprint_bytes": (C.c_size_t, [C.POINTER(Config)]),
            "ie_fill_modes": (st, [C.POINTER(Config), C.POINTER(IcParams), dp, dp, dp, dp, dp]),
            "ie_initialize": (st, [P, C.POINTER(IcParams), C.POINTER(Background)]),
            "ie_potential_value": (st, [C.POINTER(Config), C.c_double, dp, dp]),
            "ie_upload_fields": (st, [P, dp, dp]),
            "ie_upload_modes": (st, [P, dp, dp, dp, dp, C.c_double, C.c_double]),
            "ie_download_field": (st, [P, C.c_int, dp]),
            "ie_upload_field": (st, [P, C.c_int, dp]),
            "ie_download_plane": (st, [P, C.c_int, C.c_int, dp]),
            "ie_post_init": (st, [P, C.POINTER(PostParams), C.POINTER(Background)]),
            "ie_forcing_cutoff": (st, [P, C.c_double, C.c_double]),
            "ie_bispectrum": (st, [P, dp, ip, dp, dp]),
            "ie_deltaN_log": (st, [P, C.c_double, C.c_int, dp, dp, dp, ip]),
            "ie_post_evolve_derivs": (st, [P, C.c_double]),
            "ie_post_leapfrog_steps": (st, [P, C.c_int, C.POINTER(Background)]),
            "ie_upload_tensor": (st, [P, C.c_int, fp]),
            "ie_download_tensor": (st, [P, C.c_int, fp]),
            "ie_get_background": (st, [P, C.POINTER(Background)]),
            "ie_set_background": (st, [P, C.POINTER(Background)]),
            "ie_get_config": (st, [P, C.POINTER(Config)]),
            "ie_run_outputs": (st, [P, C.c_int, C.POINTER(Result)]),
            "ie_gradient_energy": (st, [P, dp]),
            "ie_kin_energy": (st, [P, dp]),
            "ie_potential_energy": (st, [P, dp]),
            "ie_field_sums": (st, [P, dp]),
            "ie_histogram": (st, [P, C.c_int, C.c_int, dp, dp, dp]),
            "ie_evolve_derivs": (st, [P, C.c_double]),
            "ie_evolve_fields": (st, [P, C.c_double]),
            "ie_leapfrog_steps": (st, [P, C.c_int, C.POINTER(Background)]),
            "ie_profile_steps": (st, [P, C.c_int, C.POINTER(Profile)]),
            "ie_tune": (st, [P, C.c_int, C.POINTER(Tuning)]),
            "ie_set_shapes": (st, [P, C.c_int, C.c_int, C.c_int]),
            "ie_rk4_step": (st, [P, C.c_double]),
            "ie_rk45_step": (st, [P, C.c_double, C.c_double, ip, dp]),
            "ie_spectrum_numbins": (C.c_int, [C.c_int]),
            "ie_spectrum": (st, [P, C.c_int, C.c_double, dp, ip, dp]),
            "ie_spectrum_noalloc": (st, [C.c_void_p, C.c_int, C.c_double, C.c_double, dp, ip, dp, C.c_void_p]),
            "ie_gw_spectrum": (st, [P, C.c_int, dp, ip, dp]),
            "ie_deltaN_init": (st, [P]),
            "ie_evolve_fieldsN": (st, [P, C.c_double]),
            "ie_get_phiref": (st, [P, dp]),
            "ie_deltaN_steps": (st, [P, C.c_double, C.c_int64]),
            "ie_deltaN_rk4_steps": (st, [P, C.c_double, C.c_int64]),
            "ie_deltaN_rk45_step": (st, [P, C.c_double, C.c_double, ip, dp]),
            "ie_deltaN_zeta": (st, [P, C.c_double, C.POINTER(C.c_float)]),
            "ie_deltaN_zero_unfinished": (st, [P, C.c_double]),
            "ie_deltaN_center": (st, [P]),
            "ie_modes": (st, [C.c_int, C.c_double, C.c_double, dp]),
            "ie_batch_create": (st, [C.POINTER(P), C.c_int, C.POINTER(P)]),
            "ie_batch_steps": (st, [P, C.c_int, C.c_int]),
            "ie_batch_set_resident_min": (st, [C.c_int, C.c_int]),
            "ie_batch_launch": (st, [P, C.c_int, C.c_int]),
            "ie_batch_wait": (st, [P]),
            "ie_batch_destroy": (st, [P]),
            "ie_batch_last_error": (C.c_char_p, [P]),
            "ie_server_create": (st, [C.c_size_t, C.c_int, C.c_int, C.POINTER(P)]),
            "ie_server_submit": (st, [P, C.POINTER(Request), C.POINTER(C.c_int64)]),
            "ie_server_pump": (st, [P, ip, ip]),
            "ie_server_pop_result": (st, [P, C.POINTER(Result)]),
            "ie_server_clock": (C.c_double, [P]),
            "ie_server_set_graph_steps": (st, [P, C.c_int]),
            "ie_server_destroy": (st, [P]),
            "ie_pzeta_linear": (st, [C.c_int, dp, dp, C.c_double, C.c_double, dp]),
            "ie_pzeta_deltaN": (st, [C.c_int, dp, dp, dp]),
            "ie_delta_h2": (st, [C.c_int, dp, dp, dp]),
            "ie_omega_gw": (st, [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp]),
            "ie_version": (C.c_char_p, []),
            "ie_device_count": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Function names declared in include/ie_api.h."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ie_[A-Za-z0-9_]+)\s*\(", text)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_config(N: int, *, precision: int = IE_FP64, numerical: int = 0, sigw: int = 0, table=None,
                **params) -> Config:
    """Config with the reference defaults for the potential kind
    (src/runtime_parameters.cpp:17-112), overridden by params."""
    c = Config()
    c.N = N
    c.precision = precision
    c.numerical_potential = numerical
    c.calculate_sigw = sigw
    c.monotonic = IE_MONOTONIC if numerical else IE_ANTIMONOTONIC
    c.integrator = IE_LEAPFROG
    c.deltaN_integrator = IE_LEAPFROG
    c.rescale_s = 0.0
    c.af = 2.0 * N
    c.V0 = 3e-9 if numerical else 3.338e-13
    c.ns = 0.97
    c.L = 10.0
    c.dt = 0.001 if numerical else 0.0005
    c.rk45_abs_tol, c.rk45_rel_tol, c.rk45_min_dt, c.rk45_max_dt, c.rk45_safety = 1e-8, 1e-6, -1.0, -1.0, 0.9
    c.dN = 0.0001 if numerical else 1e-7
    c.Nend = 5.0 if numerical else 0.001
    rescale_B_set = "rescale_B" in params
    for k, v in params.items():
        setattr(c, k, v)
    if not rescale_B_set:
        c.rescale_B = float(np.sqrt(c.V0))
    if table is not None:
        c.tab_n = len(table.x)
        c.tab_x = table.x.ctypes.data_as(C.POINTER(C.c_double))
        c.tab_v = table.v.ctypes.data_as(C.POINTER(C.c_double))
        c.tab_dv = table.dv.ctypes.data_as(C.POINTER(C.c_double))
    if not (c.rk45_min_dt > 0):
        c.rk45_min_dt = abs(c.dt) * 1e-6
    if not (c.rk45_max_dt > 0):
        c.rk45_max_dt = abs(c.dt)
    return c


class Run:
    """One lattice run on the GPU."""

    def __init__(self, cfg: Config, table=None):
        self.cfg = cfg
        self.table = table          # keeps the table arrays alive
        self.N = cfg.N
        self.gs = cfg.N ** 3
        h = C.c_void_p()
        self._check(lib().ie_create(C.byref(cfg), C.byref(h)), None)
        self.h = h

    def _check(self, status: int, handle=...):
        if status != 0:
            hnd = self.h if handle is ... else handle
            msg = lib().ie_last_error(hnd)
            raise EngineError(status, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            lib().ie_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- state
    def initialize(self, seed: int, phi0: float, dphi0: float, high_cutoff: float = 0.0,
                   low_cutoff: float = 0.0) -> dict:
        ic = IcParams(seed, phi0, dphi0, high_cutoff, low_cutoff)
        b = Background()
        self._check(lib().ie_initialize(self.h, C.byref(ic), C.byref(b)))
        return b.as_dict()

    def upload_fields(self, f: np.ndarray, fd: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64)
        fd = np.ascontiguousarray(fd, dtype=np.float64)
        assert f.size == self.gs and fd.size == self.gs
        self._check(lib().ie_upload_fields(self.h, _dp(f), _dp(fd)))

    def upload_modes(self, fk, fny, fdk, fdny, phi0, dphi0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (fk, fny, fdk, fdny)]
        assert arrs[0].size == self.gs and arrs[1].size == 2 * self.N * self.N
        self._check(lib().ie_upload_modes(self.h, *[_dp(a) for a in arrs], phi0, dphi0))

    def field(self, which: int) -> np.ndarray:
        out = np.empty(self.gs)
        self._check(lib().ie_download_field(self.h, which, _dp(out)))
        return out

    def plane(self, which: int, i: int) -> np.ndarray:
        out = np.empty(self.N * self.N)
        self._check(lib().ie_download_plane(self.h, which, i, _dp(out)))
        return out

    def tensor(self, comp: int) -> np.ndarray:
        out = np.empty(self.gs, dtype=np.float32)
        self._check(lib().ie_download_tensor(self.h, comp, _fp(out)))
        return out

    def upload_tensor(self, comp: int, h: np.ndarray):
        h = np.ascontiguousarray(h, dtype=np.float32)
        self._check(lib().ie_upload_tensor(self.h, comp, _fp(h)))

    def background(self) -> dict:
        b = Background()
        self._check(lib().ie_get_background(self.h, C.byref(b)))
        return b.as_dict()

    def set_background(self, **kw):
        b = Background()
        self._check(lib().ie_get_background(self.h, C.byref(b)))
        for k, v in kw.items():
            setattr(b, k, v)
        self._check(lib().ie_set_background(self.h, C.byref(b)))

    # ---- reductions
    def _scalar(self, fn) -> float:
        out = C.c_double()
        self._check(fn(self.h, C.byref(out)))
        return out.value

    def gradient_energy(self):
        return self._scalar(lib().ie_gradient_energy)

    def kin_energy(self):
        return self._scalar(lib().ie_kin_energy)

    def potential_energy(self):
        return self._scalar(lib().ie_potential_energy)

    def field_sums(self):
        out = (C.c_double * 3)()
        self._check(lib().ie_field_sums(self.h, out))
        return list(out)

    def histogram(self, which: int, nbins: int):
        counts = np.zeros(nbins)
        lo, hi = C.c_double(), C.c_double()
        self._check(lib().ie_histogram(self.h, which, nbins, C.byref(lo), C.byref(hi), _dp(counts)))
        return lo.value, hi.value, counts

    # ---- stepping
    def evolve_derivs(self, d: float):
        self._check(lib().ie_evolve_derivs(self.h, d))

    def evolve_fields(self, d: float):
        self._check(lib().ie_evolve_fields(self.h, d))

    def leapfrog_steps(self, n: int) -> dict:
        b = Background()
        self._check(lib().ie_leapfrog_steps(self.h, n, C.byref(b)))
        return b.as_dict()

    def rk4_step(self, h: float):
        self._check(lib().ie_rk4_step(self.h, h))

    def rk45_step(self, h: float, hmax: float) -> tuple[bool, float]:
        """One Dormand-Prince trial step; returns (accepted, proposed next step)."""
        acc, hn = C.c_int(), C.c_double()
        self._check(lib().ie_rk45_step(self.h, h, hmax, C.byref(acc), C.byref(hn)))
        return bool(acc.value), hn.value

    def tune(self, steps_per_candidate: int = 8) -> dict:
        t = Tuning()
        self._check(lib().ie_tune(self.h, steps_per_candidate, C.byref(t)))
        return t.as_dict()

    def set_shapes(self, kick_elems: int = 0, tensor_code: int = 0, march_variant: int = 0) -> None:
        """Pin this run's step kernel shapes (ie_tuning's codes); the run is no
        longer tuned by leapfrog_steps."""
        self._check(lib().ie_set_shapes(self.h, kick_elems, tensor_code, march_variant))

    def profile_steps(self, n: int) -> dict:
        p = Profile()
        self._check(lib().ie_profile_steps(self.h, n, C.byref(p)))
        return p.as_dict()

    # ---- spectra
    def spectrum(self, which: int, norm1: float):
        nb = lib().ie_spectrum_numbins(self.N)
        p = np.zeros(nb)
        c = np.zeros(nb, dtype=np.int32)
        v = np.zeros(nb)
        self._check(lib().ie_spectrum(self.h, which, norm1, _dp(p), _ip(c), _dp(v)))
        return p, c, v

    def gw_spectrum(self, hdot: int):
        nb = int(np.sqrt(3.0) * (self.N // 2)) + 1
        p = np.zeros(nb)
        c = np.zeros(nb, dtype=np.int32)
        v = np.zeros(nb)
        self._check(lib().ie_gw_spectrum(self.h, hdot, _dp(p), _ip(c), _dp(v)))
        return p, c, v

    def outputs(self, deltaN: int = 1) -> dict:
        """End of the run (ie_run_outputs): final sync, Omega_GW of a SIGW run,
        the deltaN phase and P_zeta."""
        r = Result()
        self._check(lib().ie_run_outputs(self.h, deltaN, C.byref(r)))
        d = r.as_dict()
        return {k: d[k] for k in ("steps", "a_end", "nbins", "pzeta", "gw_nbins", "omega_gw")}

    # ---- deltaN
    def deltaN_init(self):
        self._check(lib().ie_deltaN_init(self.h))

    def evolve_fieldsN(self, d: float):
        self._check(lib().ie_evolve_fieldsN(self.h, d))

    def get_phiref(self) -> float:
        return self._scalar(lib().ie_get_phiref)

    def deltaN_steps(self, dN: float, n: int):
        self._check(lib().ie_deltaN_steps(self.h, dN, n))

    def deltaN_rk4_steps(self, h: float, n: int):
        self._check(lib().ie_deltaN_rk4_steps(self.h, h, n))

    def deltaN_rk45_step(self, h: float, hmax: float) -> tuple[bool, float]:
        """One Dormand-Prince trial step of the deltaN system; (accepted, next step)."""
        acc, hn = C.c_int(), C.c_double()
        self._check(lib().ie_deltaN_rk45_step(self.h, h, hmax, C.byref(acc), C.byref(hn)))
        return bool(acc.value), hn.value

    def deltaN_zeta(self, cutoff: float) -> float:
        m = C.c_float()
        self._check(lib().ie_deltaN_zeta(self.h, cutoff, C.byref(m)))
        return m.value

    def deltaN_zero_unfinished(self, cutoff: float):
        self._check(lib().ie_deltaN_zero_unfinished(self.h, cutoff))

    def deltaN_center(self):
        self._check(lib().ie_deltaN_center(self.h))

    def upload_field(self, which: int, src: np.ndarray):
        """FP64 upload: the deltaN map, or f / fd of the deltaN and post-inflation phases."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        assert src.size == self.gs
        self._check(lib().ie_upload_field(self.h, which, _dp(src)))

    # ---- diagnostics off by default in the reference
    def forcing_cutoff(self, high: float, low: float = 0.0):
        self._check(lib().ie_forcing_cutoff(self.h, high, low))

    def bispectrum(self):
        nb = int(np.sqrt(3.0) * (self.N // 2)) + 1
        p, re, im = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        c = np.zeros(nb, dtype=np.int32)
        self._check(lib().ie_bispectrum(self.h, _dp(p), _ip(c), _dp(re), _dp(im)))
        return p, c, re, im

    def deltaN_log(self, eta: float, nbins: int):
        """LOG map into the deltaN buffer; returns (valid, bmin, df, freq)."""
        freq = np.zeros(nbins)
        bmin, df, valid = C.c_double(), C.c_double(), C.c_int()
        self._check(lib().ie_deltaN_log(self.h, eta, nbins, C.byref(bmin), C.byref(df), _dp(freq), C.byref(valid)))
        return bool(valid.value), bmin.value, df.value, freq

    # ---- post-inflation
    def post_init(self, omega: float = 1.0 / 3.0, horizon_factor: float = 1.0, dt: float = 0.001,
                  af: float = 0.0, integrator: int = 0, from_deltaN: int = 1) -> dict:
        p = PostParams(horizon_factor, omega, dt, af, integrator, from_deltaN)
        b = Background()
        self._check(lib().ie_post_init(self.h, C.byref(p), C.byref(b)))
        return b.as_dict()

    def post_evolve_derivs(self, d: float):
        self._check(lib().ie_post_evolve_derivs(self.h, d))

    def post_leapfrog_steps(self, n: int) -> dict:
        b = Background()
        self._check(lib().ie_post_leapfrog_steps(self.h, n, C.byref(b)))
        return b.as_dict()


class Batch:
    """Runs stepped together, one launch per kernel per step (ie_batch_*)."""

    def __init__(self, runs: list[Run]):
        self.runs = list(runs)
        arr = (C.c_void_p * len(runs))(*[r.h.value for r in runs])
        h = C.c_void_p()
        st = lib().ie_batch_create(arr, len(runs), C.byref(h))
        if st:
            raise EngineError(st, lib().ie_batch_last_error(None).decode())
        self.h = h

    def steps(self, n: int, graph_steps: int = 0):
        st = lib().ie_batch_steps(self.h, n, graph_steps)
        if st:
            raise EngineError(st, lib().ie_batch_last_error(self.h).decode())

    def launch(self, n: int, graph_steps: int = 0):
        """Enqueue n steps on the batch's stream and return (ie_batch_launch)."""
        st = lib().ie_batch_launch(self.h, n, graph_steps)
        if st:
            raise EngineError(st, lib().ie_batch_last_error(self.h).decode())

    def wait(self):
        st = lib().ie_batch_wait(self.h)
        if st:
            raise EngineError(st, lib().ie_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().ie_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def set_resident_min(hilltop: int = 0, tabulated: int = 0) -> None:
        """Minimum 64^3 lattices of a potential kind for the resident path in
        batches created afterwards (ie_batch_set_resident_min; 0: the default)."""
        st = lib().ie_batch_set_resident_min(hilltop, tabulated)
        if st:
            raise EngineError(st, lib().ie_batch_last_error(None).decode())


class Server:
    """Request loop (ie_server_*): FIFO admission by device footprint, continuous
    batching of resident runs, finishing on worker threads."""

    def __init__(self, hbm_budget_bytes: int, quantum_steps: int = 100, finish_threads: int = 4):
        h = C.c_void_p()
        st = lib().ie_server_create(hbm_budget_bytes, quantum_steps, finish_threads, C.byref(h))
        if st:
            raise EngineError(st, "ie_server_create failed")
        self.h = h
        self._keep = []

    def submit(self, cfg: Config, seed: int, phi0: float, dphi0: float, deltaN: int = 1, fp64_check: int = 0,
               table=None) -> int:
        req = Request()
        req.cfg = cfg
        req.ic = IcParams(seed, phi0, dphi0, 0.0, 0.0)
        req.deltaN = deltaN
        req.fp64_check = fp64_check
        rid = C.c_int64()
        st = lib().ie_server_submit(self.h, C.byref(req), C.byref(rid))
        if st:
            raise EngineError(st, "ie_server_submit failed")
        return rid.value

    def pump(self) -> tuple[int, int]:
        a, p = C.c_int(), C.c_int()
        st = lib().ie_server_pump(self.h, C.byref(a), C.byref(p))
        if st:
            raise EngineError(st, "ie_server_pump failed")
        return a.value, p.value

    def results(self) -> list[dict]:
        out = []
        r = Result()
        while lib().ie_server_pop_result(self.h, C.byref(r)) == 0:
            out.append(r.as_dict())
        return out

    def clock(self) -> float:
        return lib().ie_server_clock(self.h)

    def set_graph_steps(self, n: int):
        """Steps per HIP graph launch of the batches (0: direct launches)."""
        st = lib().ie_server_set_graph_steps(self.h, n)
        if st:
            raise EngineError(st, "ie_server_set_graph_steps failed")

    def close(self):
        if getattr(self, "h", None):
            lib().ie_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def fill_modes(cfg: Config, seed: int, phi0: float, dphi0: float, high_cutoff: float = 0.0,
               low_cutoff: float = 0.0):
    """Host-side k-space fill of the initial conditions (no GPU needed)."""
    N = cfg.N
    fk, fdk = np.zeros(N ** 3), np.zeros(N ** 3)
    fny, fdny = np.zeros(2 * N * N), np.zeros(2 * N * N)
    H = C.c_double()
    ic = IcParams(seed, phi0, dphi0, high_cutoff, low_cutoff)
    st = lib().ie_fill_modes(C.byref(cfg), C.byref(ic), _dp(fk), _dp(fny), _dp(fdk), _dp(fdny), C.byref(H))
    if st:
        raise EngineError(st, "ie_fill_modes failed")
    return fk, fny, fdk, fdny, H.value


def modes(N: int, L: float, B: float) -> np.ndarray:
    nb = int(np.sqrt(3.0) * (N // 2)) + 1
    out = np.zeros(nb)
    st = lib().ie_modes(N, L, B, _dp(out))
    if st:
        raise EngineError(st, lib().ie_last_error(None).decode())
    return out


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float64)


def _spectrum_call(fn, k, *arrays_and_scalars) -> np.ndarray:
    k = _f64(k)
    out = np.empty_like(k)
    args = [len(k), _dp(k)]
    keep = [k]
    for a in arrays_and_scalars:
        if isinstance(a, (int, float, np.floating)):
            args.append(float(a))
        else:
            arr = _f64(a)
            keep.append(arr)
            args.append(_dp(arr))
    args.append(_dp(out))
    status = fn(*args)
    if status != 0:
        raise EngineError(status, "bad argument")
    return out


def pzeta_linear(k, p_phi, phidot: float, hubble: float) -> np.ndarray:
    """ie_pzeta_linear: P_phi k^3/(2 pi^2) (phidot/H)^-2."""
    return _spectrum_call(lib().ie_pzeta_linear, k, p_phi, phidot, hubble)


def pzeta_deltaN(k, p_dn) -> np.ndarray:
    """ie_pzeta_deltaN: P_dN k^3/(2 pi^2)."""
    return _spectrum_call(lib().ie_pzeta_deltaN, k, p_dn)


def delta_h2(k, p_h) -> np.ndarray:
    """ie_delta_h2: k^3/(2 pi^2) P_h."""
    return _spectrum_call(lib().ie_delta_h2, k, p_h)


def omega_gw(k, p_h, p_hdot, a: float, hubble: float) -> np.ndarray:
    """ie_omega_gw: k^3/(2 pi^2) (P_hdot + (k/a)^2 P_h) / (12 H^2)."""
    return _spectrum_call(lib().ie_omega_gw, k, p_h, p_hdot, a, hubble)
