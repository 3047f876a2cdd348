"""Sector-ordered bases (rows sorted by a conserved charge) on the GPU.

The Lieb-Liniger Hint is non-zero only where momentum is conserved
(groundStateLieb.c:200-207), so H is block-diagonal over the sectors of equal
total momentum sum_j j v_j.  A sector-ordered basis stores its rows sorted by
that charge; every vector at the API stays in the reference order.  Checked
against the oracle and against the plain (reference-order) basis: the row
permutation, IF, the single-jump map, H.v, Lanczos / ground state, batched
Lanczos and the densities."""
import numpy as np
import pytest

import mcpi_amd
import oracle
from tests.gpu_util import assert_close, note

pytestmark = pytest.mark.gpu


def _ll(M, g):
    return mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, g)


def test_conserved_charge_detection():
    M = 7
    Ho, Hint = _ll(M, 3.0)
    assert np.array_equal(mcpi_amd.conserved_charge(Ho, Hint), np.arange(M))
    _, Ho2, Hint2 = oracle.random_inputs(2, M, seed=1)
    assert mcpi_amd.conserved_charge(Ho2, Hint2) is None


@pytest.mark.parametrize("N,M", [(3, 4), (5, 6), (6, 9), (4, 12)])
def test_sector_tables(oracle_lib, N, M):
    """perm / iperm are inverse permutations sorting the rows by charge
    (stable: reference order inside a sector); IF in row order is the
    reference IF permuted; the map holds the row of every target."""
    note(f"sector tables N={N} M={M}")
    b = mcpi_amd.Basis(N, M, "single", charge=np.arange(M))
    o = oracle.Basis(N, M)
    IFr = o.IF.reshape(-1, M)
    Q = IFr @ np.arange(M)
    perm, iperm = b.copy_range("perm", 0, b.nc), b.copy_range("iperm", 0, b.nc)
    assert np.array_equal(perm[iperm], np.arange(b.nc))
    assert np.array_equal(iperm, np.argsort(Q, kind="stable"))
    off = b.sector_offsets()
    assert off[0] == 0 and off[-1] == b.nc and b.nsectors == len(np.unique(Q)) == off.size - 1
    assert b.largest_sector == np.diff(off).max()
    assert np.array_equal(b.IF(), o.IF)                                  # reference order
    assert np.array_equal(b.copy_range("IF", 0, b.nc * M).reshape(-1, M), IFr[iperm])
    Map = b.copy_range("Map", 0, b.map_len).reshape(M, M, b.nc)           # [l][k][row]
    ref = o.Map.reshape(M, M, b.nc)[:, :, iperm]
    want = np.where(ref >= 0, perm[np.maximum(ref, 0)], -1)
    assert np.array_equal(Map, want)


@pytest.mark.parametrize("N,M,g", [(5, 6, 2.0), (6, 9, 0.7), (8, 11, 30.0), (7, 13, 5.0)])
def test_sector_h_v_matches_oracle(oracle_lib, N, M, g):
    """H.v on the sector-ordered basis (reference-order vectors in and out,
    and row-order vectors through apply_rows) against the oracle."""
    note(f"sector H.v N={N} M={M} g={g}")
    Ho, Hint = _ll(M, g)
    o = oracle.Basis(N, M)
    rng = np.random.default_rng(N * M)
    C = rng.standard_normal(o.nc) + 1j * rng.standard_normal(o.nc)
    ref = o.acth(C, Ho, Hint)
    b = mcpi_amd.Basis(N, M, "single", charge=np.arange(M))
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    assert_close(H(C), ref, 1e-12, "reference order")
    assert H.schedule == 4
    iperm = b.copy_range("iperm", 0, b.nc)
    dC, dO = mcpi_amd.DeviceArray.from_numpy(C[iperm]), mcpi_amd.DeviceArray(b.nc)
    H.apply_rows(dC, dO)
    mcpi_amd.synchronize()
    assert_close(dO.numpy(), ref[iperm], 1e-12, "row order")
    # reorder round trip
    dR, dB = mcpi_amd.DeviceArray(b.nc), mcpi_amd.DeviceArray(b.nc)
    b.from_rows(dC, dR)
    b.to_rows(dR, dB)
    mcpi_amd.synchronize()
    assert np.array_equal(dR.numpy(), C) and np.array_equal(dB.numpy(), C[iperm])


def test_sector_refuses_non_conserving_hamiltonian():
    M = 6
    b = mcpi_amd.Basis(4, M, "single", charge=np.arange(M))
    _, Ho, Hint = oracle.random_inputs(4, M, seed=3)
    with pytest.raises(mcpi_amd.MCPIError, match="conserve"):
        mcpi_amd.Hamiltonian(b, Ho, Hint)
    with pytest.raises(mcpi_amd.MCPIError):
        mcpi_amd.Basis(4, M, "full", charge=np.arange(M))
    H = mcpi_amd.Hamiltonian(b, *_ll(M, 1.0))
    with pytest.raises(mcpi_amd.MCPIError):
        H.set_skip_zero(False)
    with pytest.raises(mcpi_amd.MCPIError):
        H.set_schedule(2)


@pytest.mark.parametrize("N,M,g", [(5, 11, 10.0), (6, 9, 0.5)])
def test_sector_ground_and_densities(oracle_lib, N, M, g):
    """Ground state, Ritz vector, Lanczos basis and densities on the
    sector-ordered basis equal the plain basis's (the Krylov vectors come
    back in the reference order)."""
    note(f"sector ground N={N} M={M} g={g}")
    Ho, Hint = _ll(M, g)
    bp = mcpi_amd.Basis(N, M, "full")
    bs = mcpi_amd.Basis(N, M, "single", charge=np.arange(M))
    lm = min(bp.nc, 120)
    C0 = mcpi_amd.start_vector(bp.nc)
    E_p, psi_p, st_p = mcpi_amd.Hamiltonian(bp, Ho, Hint).ground(lm, C0)
    Hs = mcpi_amd.Hamiltonian(bs, Ho, Hint)
    E_s, psi_s, st_s = Hs.ground(lm, C0)
    assert st_s == st_p
    assert abs(E_s - E_p) < 1e-10
    vp, vs = psi_p.numpy(), psi_s.numpy()
    assert abs(abs(np.vdot(vp, vs)) - 1) < 1e-8
    E_o, psi_o, _ = oracle.Basis(N, M).ground(lm, C0, Ho, Hint)
    assert abs(E_s - E_o) < 1e-10
    lvec = mcpi_amd.DeviceArray(20 * bs.nc)
    n, d, e = Hs.lanczos(20, mcpi_amd.DeviceArray.from_numpy(C0), lvec)
    lv_p = mcpi_amd.DeviceArray(20 * bp.nc)
    n_p, d_p, e_p = mcpi_amd.Hamiltonian(bp, Ho, Hint).lanczos(20, mcpi_amd.DeviceArray.from_numpy(C0), lv_p)
    assert n == n_p and np.abs(d - d_p).max() < 1e-10 * np.abs(d_p).max()
    V, Vp = lvec.numpy().reshape(20, -1), lv_p.numpy().reshape(20, -1)
    assert np.abs(V - Vp).max() < 1e-9
    assert_close(bs.rho1(psi_s), bp.rho1(psi_s), 1e-12, "rho1")
    assert_close(bs.rho2(psi_s), bp.rho2(psi_s), 1e-12, "rho2")


def test_sector_batched_lanczos(oracle_lib):
    """Batched Lanczos with instances on sector-ordered bases (several g per
    basis, so the group shares the zero pattern) next to a plain-basis
    instance: E0 and Ritz vectors equal the per-instance plain solves."""
    spec = [(6, 9, 0.3), (6, 9, 30.0), (8, 11, 2.0), (8, 11, 70.0), (6, 9, 5.0)]
    lm = 120
    bs = {(6, 9): mcpi_amd.Basis(6, 9, "single", charge=np.arange(9)),
          (8, 11): mcpi_amd.Basis(8, 11, "single", charge=np.arange(11))}
    bp = {(6, 9): mcpi_amd.Basis(6, 9, "full"), (8, 11): mcpi_amd.Basis(8, 11, "full")}
    hams, starts, outs, ref = [], [], [], []
    rng = np.random.default_rng(9)
    for (N, M, g) in spec:
        b = bs[(N, M)]
        v = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
        v /= np.linalg.norm(v)
        Ho, Hint = _ll(M, g)
        hams.append(mcpi_amd.Hamiltonian(b, Ho, Hint))
        starts.append(mcpi_amd.DeviceArray.from_numpy(v))
        outs.append(mcpi_amd.DeviceArray(b.nc))
        ref.append(mcpi_amd.Hamiltonian(bp[(N, M)], Ho, Hint).ground(lm, v))
    # one plain-basis instance in the same call
    Ho, Hint = _ll(9, 1.0)
    hams.append(mcpi_amd.Hamiltonian(bp[(6, 9)], Ho, Hint))
    starts.append(starts[0])
    outs.append(mcpi_amd.DeviceArray(bp[(6, 9)].nc))
    ref.append(hams[-1].ground(lm, starts[0]))
    note(f"sector batched {len(hams)} instances")
    E, steps = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    for x in range(len(hams)):
        E_r, psi_r, _ = ref[x]
        assert abs(E[x] - E_r) < 1e-10, (x, E[x], E_r)
        assert abs(abs(np.vdot(psi_r.numpy(), outs[x].numpy())) - 1) < 1e-8, x


def test_sector_n16_m12_sampled_rows(oracle_lib):
    """N=16, M=12 (nc 13,037,895), Lieb-Liniger g=50 on the sector-ordered
    basis: 3,000 seeded rows of H C against the oracle's hash-only H.v."""
    N, M = 16, 12
    note("sector H.v N=16 M=12 sampled rows")
    b = mcpi_amd.Basis(N, M, "single", charge=np.arange(M))
    assert b.nsectors == 177 and b.largest_sector == 252117
    Ho, Hint = _ll(M, 50.0)
    rng = np.random.default_rng(161)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    out = mcpi_amd.Hamiltonian(b, Ho, Hint)(C)
    rows = np.sort(rng.choice(b.nc, 3000, replace=False))
    assert_close(out[rows], oracle.acth_hash_rows(N, M, rows, C, Ho, Hint), 1e-12, "rows")
