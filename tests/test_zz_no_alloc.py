"""ie_spectrum_noalloc, the entry point that allocates nothing: the spectrum of an
FP64 field in caller-owned device memory, its workspace a static device array of
the library.  Against ie_spectrum (the same kernels over the run's own
workspace) on the same field, bit for bit, on two calls in a row, the second
with a smaller lattice (so the second call reuses the first call's workspace
contents); and a lattice above the workspace's size is refused on the host."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ie_engine as E

pytestmark = pytest.mark.gpu


def run_spectrum(N, seed):
    cfg = E.make_config(N, precision=E.IE_FP64, sigw=0, V0=3.338e-13, ns=0.97, dt=0.001)
    run = E.Run(cfg)
    run.initialize(seed, 0.0935, 0.000796)
    run.leapfrog_steps(10)
    f = run.field(E.IE_FIELD_F)
    norm1 = (cfg.L / cfg.rescale_B) ** 3 / N ** 6
    p, c, v = run.spectrum(E.IE_FIELD_F, norm1)
    run.close()
    return cfg, f, norm1, (p, c, v)


def hip_runtime():
    """The HIP runtime the engine library loaded (the same handle), for the
    caller-side device buffers of the test."""
    E.lib()
    for line in open("/proc/self/maps"):
        path = line.split()[-1]
        if "libamdhip64.so" in path:
            h = C.CDLL(path)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            return h
    raise RuntimeError("the engine library has no HIP runtime mapped")


def noalloc_spectrum(f, N, L, norm1):
    hip = hip_runtime()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), f.nbytes) == 0
    try:
        assert hip.hipMemcpy(dev, f.ctypes.data, f.nbytes, 1) == 0   # hipMemcpyHostToDevice
        nb = E.lib().ie_spectrum_numbins(N)
        p, c, v = np.zeros(nb), np.zeros(nb, dtype=np.int32), np.zeros(nb)
        st = E.lib().ie_spectrum_noalloc(dev, N, L, norm1, E._dp(p), E._ip(c), E._dp(v), None)
        assert st == 0, E.lib().ie_last_error(None).decode()
    finally:
        hip.hipFree(dev)
    return p, c, v


def test_noalloc_spectrum_equals_ie_spectrum_twice():
    for N, seed in ((128, 21), (64, 22)):
        print(f"no-alloc spectrum N={N}", flush=True)
        cfg, f, norm1, ref = run_spectrum(N, seed)
        got = noalloc_spectrum(f, N, cfg.L, norm1)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), N


def test_noalloc_spectrum_refuses_large_lattice():
    nb = E.lib().ie_spectrum_numbins(256)
    p, c, v = np.zeros(nb), np.zeros(nb, dtype=np.int32), np.zeros(nb)
    st = E.lib().ie_spectrum_noalloc(C.c_void_p(16), 256, 10.0, 1.0, E._dp(p), E._ip(c), E._dp(v), None)
    assert st == E.IE_ERR_ARG
    assert "power of two in [8, 128]" in E.lib().ie_last_error(None).decode()
