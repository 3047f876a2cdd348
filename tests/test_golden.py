"""The oracle against fixtures written from the reference's own code
(tests/golden/make_golden.py).  These run everywhere, including machines
without the reference sources."""
import glob
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = sorted(glob.glob(os.path.join(GOLD, "*.npz")))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_oracle_matches_golden(oracle_lib, path):
    g = np.load(path)
    N, M = int(g["N"]), int(g["M"])
    b = oracle.Basis(N, M)
    if "IF" in g:
        for name in ("NCmat", "IF", "Map", "MapOT", "MapTT"):
            assert np.array_equal(getattr(b, name), g[name]), name
        assert np.array_equal(b.strideOT, g["strideOT"]) and np.array_equal(b.strideTT, g["strideTT"])
    out = b.acth(g["C"], g["Ho"], g["Hint"])
    assert np.linalg.norm(out - g["out"]) <= 1e-13 * np.linalg.norm(g["out"])
    assert np.linalg.norm(out - g["out_hash"]) <= 1e-13 * np.linalg.norm(g["out"])
    assert np.abs(b.rho1(g["C"]) - g["rho1"]).max() <= 1e-13
    assert np.abs(b.rho2(g["C"]) - g["rho2"]).max() <= 1e-12
