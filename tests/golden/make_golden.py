#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference program.

Runs only in a container that has the reference sources at /root/reference: each
case compiles the reference for its compile-time configuration with
oracle/build_ref.sh (binaries land in oracle/_ref/, never committed), runs it in a
scratch directory on one of the reference's own presets (params.analytic.txt,
params.numerical.txt) with the overrides listed below, and stores the numeric
results/*.dat files as arrays in a compressed .npz next to this script.  The
fixtures are data only: the parameters used, and the reference's outputs.

Cases that build the reference with its post-inflation phase on (src/parameters.h:31)
store only the results/post_inflation/*.dat files, under keys prefixed "post_"; the
files of the earlier phases are the same as those of the case without it.

It also stores the tabulated-potential input tables (inputs/*.dat of the
reference) as tests/golden/potential_table.npz, since tests on machines without
the reference need them.

usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
REF = Path(os.environ.get("IE_REFERENCE_DIR", "/root/reference"))

sys.path.insert(0, str(REPO))
from oracle.oracle import read_blocks  # noqa: E402

# name: (N, numerical_potential, preset, overrides[, build toggles])
CASES = {
    "hilltop_n16": (16, 0, "params.analytic.txt", {"output_box3D": 1, "output_box2D": 1}),
    "tabulated_n16": (16, 1, "params.numerical.txt", {"output_box3D": 1, "output_box2D": 1}),
    "tabulated_n16_rk4": (16, 1, "params.numerical.txt",
                          {"output_box2D": 1, "inflation_integrator": "rk4", "deltaN_integrator": "rk4",
                           "af": 4, "Nend": 0.5}),
    "tabulated_n16_rk45": (16, 1, "params.numerical.txt",
                           {"output_box2D": 1, "inflation_integrator": "rk45", "deltaN_integrator": "rk45",
                            "af": 4, "Nend": 0.5}),
    "hilltop_n32": (32, 0, "params.analytic.txt", {"output_box2D": 1}),
    "tabulated_n32": (32, 1, "params.numerical.txt", {"output_box2D": 1}),
    "tabulated_n16_post": (16, 1, "params.numerical.txt", {}, {"post": 1}),
    "tabulated_n16_post_rk4": (16, 1, "params.numerical.txt",
                               {"post_inflation_integrator": "rk4", "af_post_inflation": 4}, {"post": 1}),
    "tabulated_n16_post_rk45": (16, 1, "params.numerical.txt",
                                {"post_inflation_integrator": "rk45", "af_post_inflation": 4}, {"post": 1}),
    "tabulated_n16_post_nodn": (16, 1, "params.numerical.txt", {"af_post_inflation": 8},
                                {"post": 1, "deltaN": 0}),
    "hilltop_n16_post": (16, 0, "params.analytic.txt", {"af_post_inflation": 8}, {"post": 1}),
    # the equilateral bispectrum of save_last (off by default).  output_LOG = 1 is
    # not run: save_last writes the log map into the deltaN vector before
    # initializeN allocates it (src/output.cpp:1511-1516, src/initialize.cpp:221)
    # and the program dies with a segmentation fault.
    "tabulated_n16_extras": (16, 1, "params.numerical.txt", {"output_bispectrum": 1}),
    "hilltop_n16_extras": (16, 0, "params.analytic.txt", {"output_bispectrum": 1}),
    "tabulated_n32_extras": (32, 1, "params.numerical.txt", {"output_bispectrum": 1, "af": 4, "Nend": 0.3}),
    # the k-space forcing cutoff of spectraf and spectraf_post_inflation.  A low
    # cutoff above 0 removes the k = 0 mode too (the field mean), which takes the
    # tabulated potential out of its table, so it is run on the hilltop only.
    "tabulated_n16_cutoff": (16, 1, "params.numerical.txt", {"high_cutoff_index": 5.5, "forcing_cutoff": 1}),
    "hilltop_n16_cutoff": (16, 0, "params.analytic.txt",
                           {"high_cutoff_index": 6.5, "low_cutoff_index": 1.2, "forcing_cutoff": 1}),
    "tabulated_n16_cutoff_post": (16, 1, "params.numerical.txt",
                                  {"high_cutoff_index": 5.5, "forcing_cutoff": 1, "af_post_inflation": 8},
                                  {"post": 1}),
}

# files whose first block only is kept (large text dumps)
FIRST_BLOCK_ONLY = {"box"}


def params_text(preset: str, overrides: dict) -> str:
    lines = []
    for raw in (REF / preset).read_text().splitlines():
        key = raw.split("#", 1)[0].split("=", 1)[0].strip()
        if key and key in overrides:
            continue
        lines.append(raw)
    for k, v in overrides.items():
        lines.append(f"{k} = {v}")
    return "\n".join(lines) + "\n"


def build(N: int, numerical: int, deltaN: int = 1, post: int = 0) -> Path:
    out = subprocess.run([str(REPO / "oracle" / "build_ref.sh"), str(N), str(numerical), "1", str(deltaN), "0",
                          str(post)], check=True, capture_output=True, text=True).stdout.strip().splitlines()[-1]
    return Path(out)


SERIES = ("means", "variance", "velocity", "sf", "energy", "conservation", "spectratimes",
          "histogramtimes", "histogramtimesN")


def store(arrays: dict, folder: Path, prefix: str) -> None:
    for f in sorted(folder.glob("*.dat")):
        stem = f.stem
        if stem == "info":
            continue
        blocks = read_blocks(f)
        if stem in FIRST_BLOCK_ONLY:
            blocks = blocks[:1]
        if stem in SERIES:
            arrays[prefix + stem] = np.concatenate(blocks, axis=0)
        else:
            arrays[prefix + stem] = np.stack(blocks, axis=0)


def run_case(name: str) -> None:
    N, numerical, preset, overrides = CASES[name][:4]
    toggles = CASES[name][4] if len(CASES[name]) > 4 else {}
    post = toggles.get("post", 0)
    binary = build(N, numerical, toggles.get("deltaN", 1), post)
    text = params_text(preset, overrides)
    with tempfile.TemporaryDirectory(prefix=f"ie_golden_{name}_") as tmp:
        tmp = Path(tmp)
        (tmp / "params.txt").write_text(text)
        os.symlink(REF / "inputs", tmp / "inputs")
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([str(binary)], cwd=tmp, env=env, check=True, stdout=subprocess.DEVNULL)
        arrays = {"params_text": np.array(text), "N": np.array(N), "numerical_potential": np.array(numerical),
                  "perform_deltaN": np.array(toggles.get("deltaN", 1)), "post_inflation": np.array(post)}
        if post:
            store(arrays, tmp / "results" / "post_inflation", "post_")
        else:
            store(arrays, tmp / "results", "")
    np.savez_compressed(HERE / f"{name}.npz", **arrays)
    print(f"{name}: {len(arrays) - 5} files")


def write_table() -> None:
    x = np.loadtxt(REF / "inputs" / "field_values.dat")
    v = np.loadtxt(REF / "inputs" / "potential.dat")
    dv = np.loadtxt(REF / "inputs" / "potential_derivative.dat")
    np.savez_compressed(HERE / "potential_table.npz", field=x, potential=v, potential_derivative=dv)


def main(argv: list[str]) -> None:
    if not (REF / "src" / "parameters.h").exists():
        sys.exit(f"reference not found under {REF}")
    write_table()
    for name in argv or list(CASES):
        run_case(name)


if __name__ == "__main__":
    main(sys.argv[1:])
