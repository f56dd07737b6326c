"""bench.py's command line on the CPU: more than one GPU is refused before any
GPU work, with a message and a non-zero status."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]


def test_bench_refuses_several_gpus():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    p = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], cwd=REPO, env=env,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)
    assert p.stdout.strip() == ""   # no JSON line
