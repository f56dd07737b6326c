"""Shared test configuration: the `gpu` marker and fixture helpers."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
GOLDEN = REPO / "tests" / "golden"
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_golden(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def potential_table():
    from oracle import oracle as O
    with np.load(GOLDEN / "potential_table.npz", allow_pickle=False) as z:
        return O.Table(z["field"], z["potential"], z["potential_derivative"])


@pytest.fixture(scope="session")
def table():
    return potential_table()

ENGINE_DIR = REPO / "inflation-easy_amd"
if str(ENGINE_DIR) not in sys.path:
    sys.path.insert(0, str(ENGINE_DIR))


def golden_toggles(g: dict) -> dict:
    """Build toggles of a golden case (fixtures without the keys: deltaN on, post off)."""
    return {"deltaN": int(g.get("perform_deltaN", 1)), "post": int(g.get("post_inflation", 0))}


def golden_config(case: str):
    """Oracle config and parsed params of a golden case."""
    from oracle import oracle as O
    g = load_golden(case)
    N = int(g["N"])
    numerical = int(g["numerical_potential"])
    params = O.parse_params_text(str(g["params_text"]))
    table = potential_table() if numerical else None
    return O.make_config(N, numerical, params, table, **golden_toggles(g)), params, table
