"""Interval index of the tabulated potential (inflation-easy_amd/csrc/ie_tabindex.h)
on the CPU: the header the lattice kernels use, compiled with g++ and driven by
tests/host/tabindex_test.cpp over the reference's inputs/ grid and synthetic
grids.  A candidate interval that passes its check must be the bisection's
(src/potential.cpp:28-37); fewer than 1e-4 of random values may fall back to the
bisection; edge values (grid points, their neighbouring doubles, sub-bucket
edges) are probed for out-of-range candidates.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_interval_index_matches_bisection(tmp_path):
    exe = tmp_path / "tabindex_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    f"-I{REPO / 'inflation-easy_amd' / 'csrc'}", str(REPO / "tests" / "host" / "tabindex_test.cpp"),
                    "-o", str(exe)], check=True)
    with np.load(REPO / "tests" / "golden" / "potential_table.npz") as z:
        x = np.ascontiguousarray(z["field"], dtype=np.float64)
    grid = tmp_path / "grid.bin"
    x.tofile(grid)
    res = subprocess.run([str(exe), str(grid)], capture_output=True, text=True, timeout=120,
                         env={"ASAN_OPTIONS": "halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1"})
    print(res.stdout)
    assert res.returncode == 0, res.stderr[-3000:]
    assert "tabindex_test ok" in res.stdout
