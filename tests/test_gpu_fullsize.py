"""Parity at the sizes of the benchmark workloads.

* A 64-instance-style batched call that holds (10,13) Lieb-Liniger instances
  (nc 646,646) next to (6,9) and (8,11) ones, each against the oracle's
  Lanczos ground state (groundStateLieb.c:217-471 restated) run on the same
  tables: |dE0| < 1e-10, Ritz vectors to 1e-8 in overlap.
* The N=16, M=12 Lieb-Liniger ground state (nc 13,037,895, 200 Lanczos
  steps, g = 50): the Ritz pair's Rayleigh quotient and residual, the energy
  identity of its densities, and 2,000 seeded rows of H psi against the
  oracle's hash-only H.v.
* The start vector written by an H.v on the NULL stream and handed straight
  to the ground-state solver (which runs on a stream of its own).
"""
import numpy as np
import pytest

import mcpi_amd
import oracle
from tests.gpu_util import assert_close, note, oracle_on_device_tables

pytestmark = pytest.mark.gpu


def _start(nc, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    return v / np.linalg.norm(v)


def test_batched_with_10_13_instances_matches_oracle(oracle_lib):
    """(10,13) instances inside a mixed batched call, lm = 40.  The oracle
    runs on the device-built (10,13) tables copied to the host (its own
    builders take minutes at this size; the device builders are checked
    entry by entry against them at smaller sizes); (6,9) and (8,11) use the
    oracle's own tables."""
    spec = [(6, 9), (10, 13), (8, 11), (10, 13), (6, 9)]
    rng = np.random.default_rng(1013)
    gs = 10 ** rng.uniform(-1, 2, len(spec))
    lm = 40
    bases = {nm: mcpi_amd.Basis(nm[0], nm[1], "full") for nm in set(spec)}
    hams, starts, outs, C0s = [], [], [], []
    for x, (N, M) in enumerate(spec):
        b = bases[(N, M)]
        C0 = _start(b.nc, 300 + x)
        C0s.append(C0)
        hams.append(mcpi_amd.Hamiltonian(b, mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, float(gs[x]))))
        starts.append(mcpi_amd.DeviceArray.from_numpy(C0))
        outs.append(mcpi_amd.DeviceArray(b.nc))
    note(f"batched {spec} lm={lm}")
    E, steps = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    assert (steps == lm).all()
    obases = {(6, 9): oracle.Basis(6, 9), (8, 11): oracle.Basis(8, 11)}
    note("oracle on the device-built (10,13) tables")
    obases[(10, 13)] = oracle_on_device_tables(bases[(10, 13)])
    for x, (N, M) in enumerate(spec):
        E_o, psi_o, it = obases[(N, M)].ground(lm, C0s[x], oracle.ll_ho(M), oracle.ll_hint(M, float(gs[x])))
        assert it == lm
        assert abs(E[x] - E_o) < 1e-10, (x, spec[x], E[x], E_o)
        assert abs(abs(np.vdot(psi_o, outs[x].numpy())) - 1) < 1e-8, x


def test_ground_n16_m12_lieb_liniger():
    """N=16, M=12, g=50, 200 Lanczos steps from the reference's start vector
    (groundStateLieb.c:578-586), Map stored and double jumps ranked or read
    from the pattern tables.  The Ritz value equals the Rayleigh quotient of
    the Ritz vector; the densities at that vector give it back through the
    energy identity sum Ho.rho1 + 1/2 sum Hint.rho2; H psi agrees with the
    oracle on 2,000 seeded rows."""
    N, M, g, lm = 16, 12, 50.0, 200
    note("N=16 M=12 ground state, 200 steps")
    b = mcpi_amd.Basis(N, M, "single")
    Ho, Hint = mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, g)
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    dS = mcpi_amd.DeviceArray.from_numpy(mcpi_amd.start_vector(b.nc))
    E0, psi, steps = H.ground(lm, dS)
    assert steps == lm
    v = psi.numpy()
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    dH = mcpi_amd.DeviceArray(b.nc)
    H.apply(psi, dH)
    hv = dH.numpy()
    rq = np.vdot(v, hv)
    assert abs(rq.imag) < 1e-9 * abs(E0)
    assert abs(rq.real - E0) < 1e-10 * abs(E0), (rq, E0)
    res = float(np.linalg.norm(hv - E0 * v))
    note(f"E0/N = {E0 / N:.10f}, residual |H psi - E0 psi| = {res:.3e}")
    assert res < 1e-3 * abs(E0)
    g1, g2 = b.vjp(psi)
    ei = np.real(np.sum(g1 * Ho) + np.sum(g2 * Hint))
    assert abs(ei - E0) < 1e-10 * abs(E0), (ei, E0)
    rows = np.sort(np.random.default_rng(16).choice(b.nc, 2000, replace=False))
    note("oracle hash-only H.v on 2,000 rows")
    ref = oracle.acth_hash_rows(N, M, rows, v, Ho, Hint)
    assert_close(hv[rows], ref, 1e-12, "H psi rows")


def test_start_vector_from_null_stream_h_v(oracle_lib):
    """mcpi_hip_applyH(..., NULL) writes the start vector (here H = number
    operator / N with N = 8, so H C = C exactly: every v_k / 8 and their sum
    are exact in binary) and mcpi_hip_ground reads it at once from its own
    stream: the solver waits for the NULL stream first.  The result equals a
    solve from a host-uploaded copy of the same vector, and the oracle's."""
    N, M, lm = 8, 13, 30
    note("start vector written on the NULL stream, N=8 M=13")
    b = mcpi_amd.Basis(N, M, "full")
    C0 = _start(b.nc, 77)
    Id = mcpi_amd.Hamiltonian(b, np.eye(M) / N, np.zeros((M,) * 4), schedule=2)
    _, Ho, Hint = oracle.random_inputs(2, M, seed=78, sigma=0.1)
    Hl = mcpi_amd.Hamiltonian(b, Ho, Hint)
    dC = mcpi_amd.DeviceArray.from_numpy(C0)
    E_ref, psi_ref, _ = Hl.ground(lm, dC)
    dY = mcpi_amd.DeviceArray(b.nc)
    Id.apply(dC, dY)
    assert np.array_equal(dY.numpy(), C0)
    for _ in range(3):
        dX = mcpi_amd.DeviceArray(b.nc)
        Id.apply(dC, dX)                 # NULL stream, not synchronized
        E, psi, steps = Hl.ground(lm, dX)
        assert E == E_ref
        assert np.array_equal(psi.numpy(), psi_ref.numpy())
        dX.free()
    o = oracle_on_device_tables(b)
    E_o, _, _ = o.ground(lm, C0, Ho, Hint)
    assert abs(E_ref - E_o) < 1e-10
