"""The package's groundStateLieb program (mcpi_amd/apps/groundStateLieb.c,
built by mcpi_amd.build): the reference's command line and output
(groundStateLieb.c:475-645), checked against the reference's published
liebGS.dat, both columns (N = 5; M = 11, nc 3,003, and M = 21, nc 53,130;
nc/8 Lanczos steps from the deterministic start vector; E0/N printed to 5
decimals)."""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mcpi_amd", "bin", "groundStateLieb")
FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "liebGS.json")))


def _exe():
    if not os.path.exists(EXE):
        pytest.skip("groundStateLieb not built (python -m mcpi_amd.build)")
    return EXE


def test_usage_message_and_exit_status():
    """Wrong arguments: the reference's message and EXIT_FAILURE, before any
    GPU call."""
    r = subprocess.run([_exe(), "5", "11"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "./exe_name Npar Norb g" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("M,col", [(11, "E_M11"), (21, "E_M21")])
def test_regenerates_liebgs(M, col):
    from tests.gpu_util import note
    for row in FX["rows"]:
        note(f"groundStateLieb 5 {M} {row['g']}")
        r = subprocess.run([_exe(), "5", str(M), repr(row["g"])], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"g = ([-0-9.]+) \| Eo = ([-0-9.]+)", r.stdout)
        assert m, r.stdout
        assert abs(float(m.group(1)) - row["g"]) < 1e-9
        assert abs(float(m.group(2)) - row[col]) <= 5e-6 + 1e-9, (M, row, m.group(0))
        assert f"{row[col]:.5f}" == m.group(2)
