"""Lanczos and the ground state on the GPU."""
import json
import os

import numpy as np
import pytest

import mcpi_amd
import mcpi_amd.reference as R
import oracle
from oracle import bethe
from tests.gpu_util import assert_close, note

pytestmark = pytest.mark.gpu


def test_lanczos_coefficients_match_oracle(oracle_lib):
    N, M, lm = 6, 6, 60
    note("lanczos coefficients N=6 M=6")
    o = oracle.Basis(N, M)
    C0, Ho, Hint = oracle.random_inputs(N, M, seed=2)
    n_o, d_o, e_o, V_o = o.lanczos(lm, C0, Ho, Hint)
    b = mcpi_amd.Basis(N, M, "full")
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    lvec = mcpi_amd.DeviceArray(lm * b.nc)
    n_g, d_g, e_g = H.lanczos(lm, mcpi_amd.DeviceArray.from_numpy(C0), lvec)
    assert n_g == n_o == lm
    # the recurrence amplifies rounding slowly; the first steps agree tightly
    assert np.abs(d_g[:20] - d_o[:20]).max() < 1e-11
    assert np.abs(e_g[:20] - e_o[:20]).max() < 1e-11
    V = lvec.numpy().reshape(lm, b.nc)
    G = V.conj() @ V.T
    assert np.abs(G - np.eye(lm)).max() < 1e-12


@pytest.mark.parametrize("N,M,maps", [(5, 5, "full"), (6, 7, "single"), (4, 8, "none")])
def test_ground_matches_oracle(oracle_lib, N, M, maps):
    note(f"ground N={N} M={M} maps={maps}")
    o = oracle.Basis(N, M)
    C0, Ho, Hint = oracle.random_inputs(N, M, seed=N + M)
    lm = min(o.nc, 120)
    E_o, psi_o, _ = o.ground(lm, C0, Ho, Hint)
    b = mcpi_amd.Basis(N, M, maps)
    E_g, psi_g, steps = mcpi_amd.Hamiltonian(b, Ho, Hint).ground(lm, C0)
    assert abs(E_g - E_o) < 1e-10
    v = psi_g.numpy()
    assert abs(abs(np.vdot(psi_o, v)) - 1) < 1e-8


def test_reference_shaped_lanczos_and_ground(oracle_lib):
    N, M, lm = 5, 5, 40
    note("reference-shaped lanczos / ground")
    r = R.RefBasis(N, M)
    o = oracle.Basis(N, M)
    C0, Ho, Hint = oracle.random_inputs(N, M, seed=9)
    n, d, e, V = r.lanczos(lm, C0, Ho, Hint)
    n_o, d_o, e_o, _ = o.lanczos(lm, C0, Ho, Hint)
    assert n == n_o and np.abs(d[:10] - d_o[:10]).max() < 1e-11
    E, C = r.ground(lm, C0, Ho, Hint)
    E_o, _, _ = o.ground(lm, C0, Ho, Hint)
    assert abs(E - E_o) < 1e-10


def test_breakdown_exact_eigenvector():
    N, M = 5, 7
    note("lanczos breakdown")
    b = mcpi_amd.Basis(N, M, "full")
    IF = b.IF().reshape(-1, M)
    i0 = int(np.where(IF[:, M // 2] == N)[0][0])
    C0 = np.zeros(b.nc, complex)
    C0[i0] = 1
    H = mcpi_amd.Hamiltonian(b, oracle.ll_ho(M), oracle.ll_hint(M, 0.0))
    E, psi, steps = H.ground(10, C0)
    assert steps == 1 and E == 0.0


FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "liebGS.json")))


@pytest.mark.parametrize("row", FX["rows"], ids=[f"g{r['g']:g}" for r in FX["rows"]])
def test_lieb_liniger_m11_all(row):
    """liebGS.dat, 11 orbitals: every g, nc/8 steps from the reference's start vector."""
    N, M = 5, 11
    note(f"liebGS M=11 g={row['g']}")
    b = mcpi_amd.Basis(N, M, "full")
    H = mcpi_amd.Hamiltonian(b, mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, row["g"]))
    E, _, steps = H.ground(b.nc // 8, mcpi_amd.start_vector(b.nc))
    assert steps == b.nc // 8
    assert abs(E / N - row["E_M11"]) <= 5e-6 + 1e-9
    assert E / N >= bethe.energy_per_particle(N, row["g"]) - 1e-9


@pytest.fixture(scope="module")
def basis_5_21():
    return mcpi_amd.Basis(5, 21, "full")


@pytest.mark.parametrize("row", FX["rows"], ids=[f"g{r['g']:g}" for r in FX["rows"]])
def test_lieb_liniger_m21_all(row, basis_5_21):
    """liebGS.dat, 21 orbitals: every g (nc 53,130; nc/8 = 6,641 Lanczos
    steps from the reference's start vector)."""
    N, M = 5, 21
    note(f"liebGS M=21 g={row['g']}")
    b = basis_5_21
    H = mcpi_amd.Hamiltonian(b, mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, row["g"]))
    E, _, steps = H.ground(b.nc // 8, mcpi_amd.start_vector(b.nc))
    assert steps == b.nc // 8
    assert abs(E / N - row["E_M21"]) <= 5e-6 + 1e-9
    assert E / N >= bethe.energy_per_particle(N, row["g"]) - 1e-9
    H.close()
