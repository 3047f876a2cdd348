"""A C program written against the reference's interface (tests/c/
lieb_ground.c: NC, setupNCmat, setupFocks, OneOneMap, TwoTwoMap, OneTwoMap,
setupHo, setupHint, carrMod, ground) compiles with gcc against
include/mcpi_hip.h and links libmcpi_hip.so; on the GPU it reproduces the
reference's liebGS.dat values (N = 5, M = 11, nc/8 Lanczos steps from the
deterministic start vector) to the printed precision."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "mcpi_amd")
FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "liebGS.json")))


def _build(tmp_path) -> str:
    exe = str(tmp_path / "lieb_ground")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", os.path.join(ROOT, "tests", "c", "lieb_ground.c"),
           "-I", os.path.join(ROOT, "include"), "-L", LIBDIR, "-lmcpi_hip", f"-Wl,-rpath,{LIBDIR}", "-lm",
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_driver_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libmcpi_hip.so")):
        pytest.skip("libmcpi_hip.so not built")
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
@pytest.mark.parametrize("g", [10.0, 100.0])
def test_c_driver_reproduces_liebgs(tmp_path, g):
    from tests.gpu_util import note
    row = [r for r in FX["rows"] if r["g"] == g][0]
    exe = _build(tmp_path)
    note(f"C driver lieb_ground 5 11 {g}")
    r = subprocess.run([exe, "5", "11", str(g)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert abs(float(r.stdout.strip()) - row["E_M11"]) <= 5e-6 + 1e-9, (r.stdout, row)
