"""Pin the oracle to the reference: fixtures written by the compiled reference
(tests/golden/make_golden.py) must be reproduced exactly, file by file.

The reference prints %e / %f / %.17g text; both sides are parsed back to doubles,
so equality of the arrays is equality of the printed text.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_toggles, load_golden, potential_table
from oracle import oracle as O
from oracle.oracle import read_blocks

CASES = ["hilltop_n16", "tabulated_n16", "tabulated_n16_rk4", "tabulated_n16_rk45",
         "hilltop_n32", "tabulated_n32"]
# post-inflation phase (results/post_inflation/*.dat): leapfrog, RK4, RK45, Phi
# from f without deltaN, and the analytic potential
POST_CASES = ["tabulated_n16_post", "tabulated_n16_post_rk4", "tabulated_n16_post_rk45",
              "tabulated_n16_post_nodn", "hilltop_n16_post"]
# outputs off by default: the equilateral bispectrum, the forcing cutoff (inflation
# and post-inflation)
EXTRA_CASES = ["tabulated_n16_extras", "hilltop_n16_extras", "tabulated_n32_extras", "tabulated_n16_cutoff",
               "hilltop_n16_cutoff", "tabulated_n16_cutoff_post"]
META = ("params_text", "N", "numerical_potential", "perform_deltaN", "post_inflation")


@pytest.mark.parametrize("case", CASES + POST_CASES + EXTRA_CASES)
def test_oracle_reproduces_reference(case, tmp_path):
    g = load_golden(case)
    N = int(g["N"])
    numerical = int(g["numerical_potential"])
    params = O.parse_params_text(str(g["params_text"]))
    table = potential_table() if numerical else None
    cfg = O.make_config(N, numerical, params, table, **golden_toggles(g))
    assert O.run_program(cfg, tmp_path) == 0
    checked = 0
    for key, ref in g.items():
        if key in META:
            continue
        path = tmp_path / "post_inflation" / f"{key[5:]}.dat" if key.startswith("post_") else tmp_path / f"{key}.dat"
        blocks = read_blocks(path)
        if key == "box":
            blocks = blocks[:1]
        if ref.ndim == 2 and key not in ("box",):
            got = np.concatenate(blocks, axis=0)
        else:
            got = np.stack(blocks, axis=0)
        assert got.shape == ref.shape, key
        assert np.array_equal(got, ref), f"{case}/{key}: max |diff| {np.max(np.abs(got - ref))}"
        checked += 1
    assert checked >= (10 if golden_toggles(g)["post"] else 14)


@pytest.mark.parametrize("case", ["hilltop_n16", "tabulated_n16", "hilltop_n32", "tabulated_n32"])
def test_oracle_deltaN_zeta_map_gives_reference_spectraN(case):
    """ieo_deltaN_zeta_map (the deltaN phase and saveN's map, no files) after the
    inflation loop and its final sync, then the FFT and shell binning, reproduces
    the reference's spectraN.dat as printed (%e: 7 significant digits)."""
    g = load_golden(case)
    N = int(g["N"])
    numerical = int(g["numerical_potential"])
    params = O.parse_params_text(str(g["params_text"]))
    table = potential_table() if numerical else None
    cfg = O.make_config(N, numerical, params, table, **golden_toggles(g))
    st = O.State(cfg, table)
    st.initialize()
    O.lib().ieo_evolve_fields(st.ptr, 0.5 * cfg.dt)
    O.lib().ieo_leapfrog_steps(st.ptr, 10 ** 8)
    O.lib().ieo_evolve_fields(st.ptr, -0.5 * cfg.dt * st.background()["astep"] ** (cfg.rescale_s - 1.0))
    n = O.lib().ieo_deltaN_zeta_map(st.ptr, None)
    assert n > 0
    norm1 = (cfg.L / cfg.rescale_B) ** 3 / N ** 6
    F, ny = O.fftrn(st.field(2), +1)
    p, c, v = O.bin_spectrum(N, cfg.L, F, ny, norm1)
    ref = g["spectraN"]
    ref = ref[-1] if ref.ndim == 3 else ref
    assert np.array_equal(c, ref[:, 1].astype(int))
    got = np.array([float(f"{x:e}") for x in v])
    assert np.array_equal(got, ref[:, 2])
