"""Density matrices and the Hellmann-Feynman adjoint on the GPU."""
import os

import numpy as np
import pytest

import mcpi_amd
import mcpi_amd.reference as R
import oracle
from tests.gpu_util import assert_close, note

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,M", [(1, 3), (2, 3), (5, 5), (4, 6), (6, 4), (3, 7), (7, 9), (3, 12), (3, 13), (2, 14),
                                 (2, 17), (2, 24), (2, 32)])
@pytest.mark.parametrize("maps", ["full", "single", "none"])
def test_rho_matches_oracle(oracle_lib, N, M, maps):
    """Every storage mode (the densities read no map).  Tiles of the Gram
    matrix of rho2 (80 pair columns a side): M = 12 fills one tile (78 pairs),
    M = 13 (91) needs two and the off-diagonal tile pair, M = 24 (300) four
    tiles a side, M = 32 (528) seven; N = 1: rho2 = 0, rho1 from the
    one-particle Gram matrix."""
    note(f"rho N={N} M={M} maps={maps}")
    o = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=21 + N)
    b = mcpi_amd.Basis(N, M, maps)
    dC = mcpi_amd.DeviceArray.from_numpy(C)
    assert_close(b.rho1(dC), o.rho1(C), 1e-12, "rho1")
    assert_close(b.rho2(dC), o.rho2(C), 1e-12, "rho2")


@pytest.mark.parametrize("N,M", [(0, 3), (1, 1), (2, 1), (4, 1), (3, 2), (12, 2), (30, 3)])
@pytest.mark.parametrize("maps", ["full", "none"])
def test_rho_edge_shapes(oracle_lib, N, M, maps):
    """No particle (one state, both densities zero), one orbital (rho1 = N,
    rho2 = N(N-1): a single unranked row of N-1 or N-2 particles), two
    orbitals, and many particles in few orbitals (N = 30, M = 3: the rank
    table's widest rows)."""
    note(f"rho edge N={N} M={M} maps={maps}")
    o = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=40 + N)
    b = mcpi_amd.Basis(N, M, maps)
    dC = mcpi_amd.DeviceArray.from_numpy(C)
    r1, r2 = b.rho1(dC), b.rho2(dC)
    ref1, ref2 = o.rho1(C), o.rho2(C)
    if N == 0:
        assert not r1.any() and not r2.any()
        return
    assert_close(r1, ref1, 1e-12, "rho1")
    if N == 1:
        assert not r2.any()
    else:
        assert_close(r2, ref2, 1e-12, "rho2")


def test_rho_golden():
    gd = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1_perf.npz"))
    note("rho golden cfg1")
    b = mcpi_amd.Basis(5, 5, "full")
    dC = mcpi_amd.DeviceArray.from_numpy(gd["C"])
    assert_close(b.rho1(dC), gd["rho1"], 1e-12, "rho1")
    assert_close(b.rho2(dC), gd["rho2"], 1e-12, "rho2")


@pytest.mark.parametrize("N,M", [(5, 5), (2, 17)])
def test_reference_shaped_rho(oracle_lib, N, M):
    """Reference-layout maps passed in; M = 17 reads the double-jump targets
    from them.  Every reference variant: OBrho (no table, onebodyMatrix.h:34),
    OBrho_X (IF, :116), OBrho_XM (IF, Map, :190), TBrho (IF,
    twobodyMatrix.h:38), TBrho_X (IF, Map, MapOT, :502), TBrho_XX (:991)."""
    note(f"OBrho* / TBrho* N={N} M={M}")
    r = R.RefBasis(N, M)
    o = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=1)
    r1, r2 = o.rho1(C), o.rho2(C)
    assert_close(r.OBrho_XM(C), r1, 1e-12, "OBrho_XM")
    assert_close(r.TBrho_XX(C), r2, 1e-12, "TBrho_XX")
    IF0 = r.IF.copy()
    assert_close(r.OBrho(C), r1, 1e-12, "OBrho")
    assert_close(r.OBrho_X(C), r1, 1e-12, "OBrho_X")
    assert_close(r.TBrho(C), r2, 1e-12, "TBrho")
    assert_close(r.TBrho_X(C), r2, 1e-12, "TBrho_X")
    assert np.array_equal(r.IF, IF0)


def test_rho_invariants_n10_m12(oracle_lib):
    """N=10, M=12 (nc 352,716): partial traces, symmetries, energy identity."""
    N, M = 10, 12
    note("rho invariants N=10 M=12")
    b = mcpi_amd.Basis(N, M, "full")
    rng = np.random.default_rng(10)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    _, Ho, Hint = oracle.random_inputs(2, M, seed=3, sigma=0.1)
    dC = mcpi_amd.DeviceArray.from_numpy(C)
    r1, r2 = b.rho1(dC), b.rho2(dC)
    assert abs(np.trace(r1) - N) < 1e-11
    assert abs(np.einsum("ksks->", r2) - N * (N - 1)) < 1e-9
    assert np.allclose(np.einsum("kssl->kl", r2), (N - 1) * r1, atol=1e-11)
    assert np.allclose(r2, np.conj(r2.transpose(3, 2, 1, 0)), atol=1e-12)
    E = np.vdot(C, mcpi_amd.Hamiltonian(b, Ho, Hint)(C))
    E2 = np.sum(Ho * r1) + 0.5 * np.sum(Hint * r2)
    assert abs(E - E2) <= 1e-11 * abs(E)


def test_vjp_finite_differences(oracle_lib):
    """dE0/dHo = rho1, dE0/dHint = rho2/2 at the ground state, against central
    differences of the GPU ground-state energy (step 1e-6) along random
    Hermitian / fully symmetric directions."""
    N, M = 5, 6
    note("vjp finite differences N=5 M=6")
    b = mcpi_amd.Basis(N, M, "full")
    C0, Ho, Hint = oracle.random_inputs(N, M, seed=77, sigma=0.1)
    lm = b.nc

    def e0(ho, hint):
        E, vec, _ = mcpi_amd.Hamiltonian(b, ho, hint).ground(lm, C0)
        return E, vec

    E, psi = e0(Ho, Hint)
    gHo, gHint = b.vjp(psi)
    rng = np.random.default_rng(3)
    h = 1e-6
    for t in range(4):
        _, Do, Dint = oracle.random_inputs(N, M, seed=500 + t, sigma=1.0)
        fd = (e0(Ho + h * Do, Hint + h * Dint)[0] - e0(Ho - h * Do, Hint - h * Dint)[0]) / (2 * h)
        an = np.real(np.sum(gHo * Do) + np.sum(gHint * Dint))
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (t, fd, an)


def test_rho_single_map_n16_m12_invariants(oracle_lib):
    """N=16, M=12 with only Map stored (nc 13,037,895; the Gram matrix of
    rho2 runs over the 4,457,400 states of 14 particles): partial traces,
    contraction, Hermiticity, and the energy identity against H.v of the
    same basis."""
    N, M = 16, 12
    note("rho invariants N=16 M=12 single map")
    b = mcpi_amd.Basis(N, M, "single")
    rng = np.random.default_rng(16)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    _, Ho, Hint = oracle.random_inputs(2, M, seed=5, sigma=0.1)
    dC = mcpi_amd.DeviceArray.from_numpy(C)
    r1, r2 = b.rho1(dC), b.rho2(dC)
    assert abs(np.trace(r1) - N) < 1e-10
    assert abs(np.einsum("ksks->", r2) - N * (N - 1)) < 1e-8
    assert np.allclose(np.einsum("kssl->kl", r2), (N - 1) * r1, atol=1e-10)
    assert np.allclose(r2, np.conj(r2.transpose(3, 2, 1, 0)), atol=1e-12)
    E = np.vdot(C, mcpi_amd.Hamiltonian(b, Ho, Hint)(C))
    E2 = np.sum(Ho * r1) + 0.5 * np.sum(Hint * r2)
    assert abs(E - E2) <= 1e-11 * abs(E)
