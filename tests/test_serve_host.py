"""Host logic of the serving loop (inflation-easy_amd/csrc/ie_serve.cpp) on the CPU.

The loop is compiled with g++ against a stand-in of the engine's C ABI
(tests/host/serve_stub.cpp: runs in host memory, sleeps for compute, injected
failures) under ThreadSanitizer and under AddressSanitizer + UBSan, and driven
with a stream of the benchmark's request mix.  The stub aborts when:
  - a lifecycle call (create, destroy, batch create/launch/wait/destroy) comes
    from a thread other than the pump's;
  - a batch is launched on a run it was not built from (a destroyed run, or a
    new run at a destroyed run's address);
  - a compute call overlaps a batch stepping the same run, or a destroy;
  - the live runs' footprint exceeds the server's budget;
and the driver checks that every request returns with its expected status (a
failed initial condition, a kernel table-range error while stepping, a failed
deltaN phase; the other requests unaffected), that SIGW runs carry Omega_GW,
and that the footprint returns to zero.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
SRCS = [REPO / "inflation-easy_amd" / "csrc" / "ie_serve.cpp", REPO / "inflation-easy_amd" / "csrc" / "ie_post.cpp",
        REPO / "tests" / "host" / "serve_stub.cpp"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("sanitizer", ["thread", "address,undefined"])
def test_serving_loop_under_sanitizer(tmp_path, sanitizer):
    exe = tmp_path / "serve_stub"
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
                    f"-I{REPO / 'include'}", *map(str, SRCS), "-o", str(exe), "-pthread"], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    res = subprocess.run([str(exe), "300"], env=env, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-4000:]
    assert "serve_stub ok: 300 requests" in res.stdout
