"""Hellmann-Feynman adjoint at N=10, M=12 (nc 352,716): dE0/dHo = rho1 and
dE0/dHint = rho2 / 2 from the VJP kernel, against central finite differences
of the ground-state energy along 32 seeded random directions.

Inputs: random Hermitian Ho and fully symmetric Hint with entries ~ N(0, 0.1)
(sigma = 0.1); directions: Hermitian D_o and fully symmetric D_int with unit
normal entries; step h = 1e-6; the + and - solves of every direction start
from the same vector.  The 64 shifted solves run as batched Lanczos calls
(16 instances sharing one basis per call).  Tolerance: 1e-5 relative (the
E0 rounding noise divided by h)."""
import numpy as np
import pytest

import mcpi_amd
import oracle
from tests.gpu_util import assert_close, note, oracle_on_device_tables

pytestmark = pytest.mark.gpu

N, M, LM, H_STEP = 10, 12, 300, 1e-6


def test_vjp_central_differences_n10_m12():
    note("VJP vs central differences N=10 M=12, 32 directions")
    b = mcpi_amd.Basis(N, M, "full")
    _, Ho, Hint = oracle.random_inputs(N, M, seed=1012, sigma=0.1)
    rng = np.random.default_rng(2024)
    C0 = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C0 /= np.linalg.norm(C0)
    dC0 = mcpi_amd.DeviceArray.from_numpy(C0)

    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    E, psi, steps = H.ground(LM, dC0)
    assert steps == LM
    v = psi.numpy()
    r = H(v) - E * v
    res = float(np.linalg.norm(r))
    note(f"E0 = {E:.12f}, residual {res:.2e}")
    assert res < 1e-6, "ground state not converged: the adjoint needs an eigenvector"
    gHo, gHint = b.vjp(psi)
    # energy identity at the eigenvector: E0 = sum Ho.rho1 + 1/2 sum Hint.rho2
    assert abs(np.real(np.sum(gHo * Ho) + np.sum(gHint * Hint)) - E) < 1e-9 * max(1.0, abs(E))
    # the densities themselves against the oracle's OBrho_XM / TBrho_XX
    # restatement on the same (device-built) tables at the same vector
    note("oracle rho1 / rho2 at the ground vector")
    o = oracle_on_device_tables(b)
    assert_close(gHo, o.rho1(v), 1e-12, "dE0/dHo vs oracle rho1")
    assert_close(gHint, 0.5 * o.rho2(v), 1e-12, "dE0/dHint vs oracle rho2 / 2")
    del o

    dirs = [oracle.random_inputs(N, M, seed=7000 + t, sigma=1.0)[1:] for t in range(32)]
    fd = np.zeros(32)
    for chunk in range(4):
        hams = []
        for t in range(8 * chunk, 8 * chunk + 8):
            Do, Dint = dirs[t]
            hams.append(mcpi_amd.Hamiltonian(b, Ho + H_STEP * Do, Hint + H_STEP * Dint))
            hams.append(mcpi_amd.Hamiltonian(b, Ho - H_STEP * Do, Hint - H_STEP * Dint))
        Es, st = mcpi_amd.lanczos_batched(hams, [dC0] * len(hams), LM)
        assert (st == LM).all()
        for k in range(8):
            fd[8 * chunk + k] = (Es[2 * k] - Es[2 * k + 1]) / (2 * H_STEP)
        for h in hams:
            h.close()
    an = np.array([np.real(np.sum(gHo * Do) + np.sum(gHint * Dint)) for (Do, Dint) in dirs])
    err = np.abs(fd - an) / np.maximum(1.0, np.abs(an))
    note(f"max relative FD error {err.max():.2e}")
    assert err.max() <= 1e-5, (fd, an)
