"""Host test of the factorized pair form's detection (pair_factor.h, used
by the coefficient builder for the pair-space H.v of schedule 7): the
C++ test program tests/c/test_pair_factor.cpp compiled with g++ and run.
No GPU: the detection is host code."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pair_factorize_host():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "test_pair_factor")
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "mcpi_amd", "csrc"),
               os.path.join(ROOT, "tests", "c", "test_pair_factor.cpp"), "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "all passed" in r.stdout
