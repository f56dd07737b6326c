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


def assert_rel(got, want, rtol: float, what: str = "") -> float:
    """Largest relative error of got against want (want nonzero), printed and
    bounded by rtol with no absolute floor: spectra of fields and tensors are far
    below numpy's default atol, which would pass any value."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    assert np.all(want != 0), what
    rel = np.abs(got - want) / np.abs(want)
    worst = float(np.max(rel)) if want.size else 0.0
    print(f"{what}: max rel err {worst:.3e} (bound {rtol:g})", flush=True)
    if worst > rtol:
        idx = np.argsort(-rel.ravel())[:5]
        print("  worst (flat index, got, want):",
              [(int(i), float(got.ravel()[i]), float(want.ravel()[i])) for i in idx], flush=True)
    assert worst <= rtol, (what, worst, rtol)
    return worst


def assert_spectrum(got, want, rtol: float, what: str = "", floor: float = 1e-12) -> float:
    """A binned spectrum against its reference: relative error (assert_rel) over
    the bins above floor x the largest, and an absolute bound of floor x the
    largest on the others (the k = 0 bin of a mean-subtracted field, modes zeroed
    by a cutoff: round-off in both)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    scale = float(np.max(np.abs(want)))
    big = np.abs(want) > floor * scale
    assert np.all(np.abs(got[~big] - want[~big]) <= floor * scale), what
    return assert_rel(got[big], want[big], rtol, what)


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
