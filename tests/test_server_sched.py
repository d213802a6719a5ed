"""Served-form scheduler (csrc/server_core.h) on the CPU with a test double in
place of the device backend: admission by device-memory budget, continuous
batching over vote rounds, row coverage of every vote pass, overflow reruns,
error isolation, cancellation, timeouts, shutdown.  Built once plainly and
once under ThreadSanitizer."""
import os
import subprocess

import pytest

from conftest import PKG, ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "server_sched_test.cpp")


@pytest.mark.parametrize("sanitize", [[], ["-fsanitize=thread"]], ids=["plain", "tsan"])
def test_scheduler(tmp_path, sanitize):
    exe = str(tmp_path / "server_sched_test")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror", "-pthread", *sanitize,
           "-I" + os.path.join(PKG, "csrc"), SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=180, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout and "WARNING" not in r.stderr, r.stdout + r.stderr
