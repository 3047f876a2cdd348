"""C-ABI checks that need no GPU: the library loads, exports every function
include/mcpi_hip.h declares, and its host-only entry points (hashing helpers,
the reference's INT32 guards, the tridiagonal solver) behave like the
reference / numpy."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import mcpi_amd
import mcpi_amd.reference as R
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "mcpi_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
        name = m.group(1)
        before = txt[: m.start()].rstrip()
        if name in {"defined", "if", "sizeof"} or before.endswith("#define") or before.endswith("define"):
            continue
        # declarations start a line with a return type; skip macro bodies
        line = txt[txt.rfind("\n", 0, m.start()) + 1: m.start()]
        if line.strip().startswith("#"):
            continue
        names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert "applyHconf_omp" in names and "mcpi_hip_applyH" in names and "lanczos" in names
    L = mcpi_amd.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", mcpi_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported, set(names) - exported


def test_host_hashing_entry_points():
    for (N, M) in [(5, 5), (3, 7), (12, 12), (16, 12)]:
        assert R.NC(N, M) == oracle.nc(N, M)
    N, M = 5, 5
    T = oracle.ncmat(N, M)
    IF = oracle.setup_focks(N, M).reshape(-1, M)
    for k in range(oracle.nc(N, M)):
        assert np.array_equal(R.IndexToFock(k, N, M), IF[k])
        assert R.FockToIndex(N, M, T, IF[k]) == k


def _run(code):
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)


def test_reference_guard_nc_exits():
    """NC overflows int32 -> message + exit(EXIT_FAILURE) (configurationsMap.h:162-168)."""
    r = _run("import mcpi_amd.reference as R; R.NC(30, 30); print('survived')")
    assert r.returncode == 1 and "survived" not in r.stdout
    assert "INTEGER SIZE ERROR" in r.stdout


def test_extended_api_reports_errors_without_exit():
    with pytest.raises(mcpi_amd.MCPIError):
        mcpi_amd._check(mcpi_amd.lib().mcpi_hip_tridiag_lowest(0, np.zeros(1), np.zeros(1),
                                                              None, np.zeros(1)), "tridiag")


def test_setup_lieb_liniger_matches_oracle():
    for M in (5, 11, 12):
        assert np.array_equal(R.setupHo(M, 1.0), oracle.ll_ho(M))
        assert np.array_equal(R.setupHint(M, 2.5), oracle.ll_hint(M, 2.5))
        assert np.array_equal(mcpi_amd.ll_ho(M), oracle.ll_ho(M))
        assert np.array_equal(mcpi_amd.ll_hint(M, 2.5), oracle.ll_hint(M, 2.5))
    for n in (2, 126, 3003):  # n = 1 is 0/0 in the reference formula too
        assert np.allclose(mcpi_amd.start_vector(n), oracle.start_vector(n), rtol=0, atol=1e-15)


def test_host_tridiagonal_solver_matches_numpy():
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 17, 120, 400):
        d = rng.standard_normal(n) * 5
        e = rng.standard_normal(max(n - 1, 0))
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w = np.linalg.eigvalsh(T)
        lam, z = mcpi_amd.tridiag_lowest(d, e)
        assert abs(lam - w[0]) <= 1e-12 * max(1.0, abs(w).max())
        assert abs(np.linalg.norm(z) - 1) < 1e-12
        assert np.linalg.norm(T @ z - lam * z) < 1e-9 * max(1.0, abs(w).max())
    # a Lanczos-like tridiagonal with a tight cluster at the bottom
    d = np.concatenate([[1.0, 1.0 + 1e-9], np.linspace(2, 50, 60)])
    e = np.full(d.size - 1, 1e-3)
    w = np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))
    assert abs(mcpi_amd.tridiag_lowest(d, e)[0] - w[0]) < 1e-12 * 50


def test_tridiagonal_solvers_match_lapack_dstev(oracle_lib):
    """The Ritz step's tridiagonal solve (groundStateLieb.c:435 calls LAPACK
    dstev) against dstev itself, here from SciPy's LAPACK: lowest eigenvalue
    and its eigenvector (up to sign), for the library's host solver and the
    oracle's, on random and Lanczos-shaped tridiagonals."""
    from scipy.linalg.lapack import dstev
    rng = np.random.default_rng(7)
    cases = []
    for n in (1, 2, 3, 17, 120, 200):
        cases.append((rng.standard_normal(n) * 5, rng.standard_normal(max(n - 1, 0))))
    # Lanczos-shaped: decaying off-diagonal, well-separated bottom, growing diagonal
    n = 150
    cases.append((np.linspace(-30.0, 400.0, n) + rng.standard_normal(n), 10.0 * np.exp(-np.arange(n - 1) / 40.0)))
    for d, e in cases:
        w, V, info = dstev(d, e if e.size else np.zeros(1), compute_v=1)   # e: max(1, n - 1) entries
        assert info == 0
        gap = (w[1] - w[0]) if d.size > 1 else 1.0
        scale = max(1.0, float(np.abs(w).max()))
        for lam, z in (mcpi_amd.tridiag_lowest(d, e), oracle.tridiag_min(d, e)):
            assert abs(lam - w[0]) <= 1e-12 * scale
            z = np.asarray(z) * np.sign(np.dot(np.asarray(z), V[:, 0]))
            assert np.linalg.norm(z - V[:, 0]) <= 1e-12 * scale / gap + 1e-13
