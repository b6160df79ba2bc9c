"""CPU unit tests of host-side library code that needs no GPU: the device
pool's reuse rules (phj-bench_amd/csrc/arena.h), built with g++ against a
recording backend (tests/arena_selftest.cpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_device_pool_reuse_rules():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "arena_selftest")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "phj-bench_amd", "csrc"),
                        os.path.join(ROOT, "tests", "arena_selftest.cpp"), "-o", exe], check=True, timeout=120)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
