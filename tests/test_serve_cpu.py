"""The serving scheduler's logic on the CPU: tests/serve_core_check.cpp
drives gx_serve_core.hpp with a fake executor (allocator, FIFO and backfill
admission, refusal above the budget, cancellation, injected failures,
results held until released, latency percentiles, the threaded loop),
compiled here with the address and undefined-behaviour sanitizers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_scheduler_checks(tmp_path):
    exe = str(tmp_path / "serve_core_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Werror",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "guanaco_amd", "csrc"),
           os.path.join(ROOT, "tests", "serve_core_check.cpp"), "-o", exe,
           "-pthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
