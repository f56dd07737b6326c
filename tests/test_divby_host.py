"""Correctly rounded division through a known reciprocal
(inflation-easy_amd/csrc/ie_numeric.h, used by the deltaN kernels to keep the
reference's divisions by B^2 and by the table's interval widths bit for bit):
tests/host/divby_test.cpp, built with g++ without FP contraction, compares it with
IEEE division on random operands over many decades and on the reference's
inputs/ table (interval widths as divisors, V and V' as numerators).
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_div_by_equals_ieee_division(tmp_path):
    exe = tmp_path / "divby_test"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{REPO / 'inflation-easy_amd' / 'csrc'}",
                    str(REPO / "tests" / "host" / "divby_test.cpp"), "-o", str(exe)], check=True)
    with np.load(REPO / "tests" / "golden" / "potential_table.npz") as z:
        t = np.concatenate([z["field"], z["potential"], z["potential_derivative"]]).astype(np.float64)
    grid = tmp_path / "table.bin"
    t.tofile(grid)
    res = subprocess.run([str(exe), str(grid)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stderr[-2000:]
    assert " 0 differ" in res.stdout
