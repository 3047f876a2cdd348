"""Pins for the CPU oracle (oracle/mcpi_oracle.c).

Each oracle function is checked against something that does not share its
formulas: closed forms, exhaustive brute force, an independent
second-quantization construction (tests/bruteforce.py), invariants of the
density matrices, the reference's own C code compiled unmodified
(oracle/_ref, when the reference sources are present), and the reference's
published Lieb-Liniger energies (liebGS.dat).
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from oracle import bethe
import tests.bruteforce as bf

SMALL = [(1, 1), (3, 1), (0, 3), (1, 4), (2, 3), (3, 3), (2, 5), (4, 3), (3, 4), (5, 5), (4, 6), (6, 4)]


# ------------------------------------------------------------- hashing --

def test_nc_closed_form(oracle_lib):
    for N in range(0, 21):
        for M in range(1, 21):
            assert oracle.nc(N, M) == math.comb(N + M - 1, N)


def test_ncmat_layout(oracle_lib):
    N, M = 7, 6
    t = oracle.ncmat(N, M)
    for n in range(N + 1):
        assert t[n] == 0                       # column 0 holds 0 (configurationsMap.h:228)
        assert t[n + (N + 1)] == 1             # column 1 holds 1
        for m in range(2, M + 1):
            assert t[n + (N + 1) * m] == math.comb(n + m - 1, n)


@pytest.mark.parametrize("N,M", [(N, M) for N in range(0, 7) for M in range(1, 7)])
def test_rank_bijection_closed_form(oracle_lib, N, M):
    IF = oracle.setup_focks(N, M).reshape(-1, M)
    T = oracle.ncmat(N, M)
    nc = IF.shape[0]
    assert nc == math.comb(N + M - 1, N)
    assert (IF.sum(1) == N).all() and (IF >= 0).all()
    assert len({tuple(r) for r in IF}) == nc
    for i in range(nc):
        assert oracle.rank(N, M, T, IF[i]) == i
        assert bf.closed_rank(tuple(IF[i])) == i
    assert tuple(IF[0]) == (N,) + (0,) * (M - 1)
    assert tuple(IF[-1]) == (0,) * (M - 1) + (N,)


def test_rank_order_example(oracle_lib):
    # N = 2 bosons in M = 3 orbitals: colexicographic from the last orbital
    got = [tuple(oracle.unrank(k, 2, 3)) for k in range(6)]
    assert got == [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)]


# ---------------------------------------------------------------- maps --

def brute_maps(N, M):
    """Every map entry from rank(unrank(i) +- unit vectors)."""
    IF = oracle.setup_focks(N, M).reshape(-1, M)
    nc = IF.shape[0]
    idx = {tuple(r): i for i, r in enumerate(IF)}
    Map = -np.ones(nc * M * M, np.int64)
    ot, tt, sot, stt = [], [], [], []
    for i in range(nc):
        v = IF[i]
        for k in range(M):
            if v[k] < 1:
                continue
            for l in range(M):
                w = v.copy(); w[k] -= 1; w[l] += 1
                Map[i + k * nc + l * M * nc] = idx[tuple(w)]
        sot.append(len(ot))
        for k in range(M):
            if v[k] < 2:
                continue
            for l in range(M):
                for q in range(M):
                    w = v.copy(); w[k] -= 2; w[l] += 1; w[q] += 1
                    ot.append(idx[tuple(w)])      # position l + q*M
        stt.append(len(tt))
        for k in range(M):
            for s in range(k + 1, M):
                if v[k] < 1 or v[s] < 1:
                    continue
                for l in range(M):
                    for q in range(M):
                        w = v.copy(); w[k] -= 1; w[s] -= 1; w[l] += 1; w[q] += 1
                        tt.append(idx[tuple(w)])
    return (Map, np.array(ot, np.int64), np.array(tt, np.int64),
            np.array(sot, np.int64), np.array(stt, np.int64))


@pytest.mark.parametrize("N,M", SMALL)
def test_maps_brute_force(oracle_lib, N, M):
    b = oracle.Basis(N, M)
    Map, ot, tt, sot, stt = brute_maps(N, M)
    assert np.array_equal(b.Map, Map)
    assert np.array_equal(b.MapOT, ot)
    assert np.array_equal(b.MapTT, tt)
    assert np.array_equal(b.strideOT, sot)
    assert np.array_equal(b.strideTT, stt)
    # size identities (configurationsMap.h:608-610, 756-759)
    if N >= 2:
        nc2 = math.comb(N - 2 + M - 1, N - 2)
        assert b.MapOT.size == M * nc2 * M * M
        assert b.MapTT.size == math.comb(M, 2) * nc2 * M * M
        assert b.strideOT[-1] + M * M == b.MapOT.size
        assert b.strideTT[-1] == b.MapTT.size
    IFm = b.IF.reshape(-1, M)
    for k in range(M):
        for l in range(M):
            col = b.Map[np.arange(b.nc) + k * b.nc + l * M * b.nc]
            assert ((col == -1) == (IFm[:, k] == 0)).all()


# ------------------------------------------------------ reference build --

needs_ref = pytest.mark.skipif(not __import__("oracle.ref").ref.available(),
                               reason="reference sources not present (oracle/_ref not built)")


@needs_ref
@pytest.mark.parametrize("N,M", [(2, 3), (5, 5), (4, 6), (6, 4), (3, 7), (1, 5), (7, 3)])
def test_oracle_equals_reference_build(oracle_lib, N, M):
    from oracle import ref
    b, r = oracle.Basis(N, M), ref.RefBasis(N, M)
    for name in ("NCmat", "IF", "Map", "MapOT", "MapTT"):
        assert np.array_equal(getattr(b, name), getattr(r, name)), name
    assert np.array_equal(b.strideOT, r.strideOT) and np.array_equal(b.strideTT, r.strideTT)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=N * 31 + M)
    out = b.acth(C, Ho, Hint)
    nrm = np.linalg.norm(out)
    for var in ("omp", "XX", "hash"):
        if var == "X" and N <= 1:
            continue  # applyHconf_X reads MapOT out of bounds for N <= 1
        assert np.linalg.norm(out - r.acth(C, Ho, Hint, var)) <= 1e-13 * nrm, var
    assert np.abs(b.rho1(C) - r.rho1(C, "XM")).max() < 1e-13
    assert np.abs(b.rho2(C) - r.rho2(C, "XX")).max() < 1e-12


@needs_ref
def test_oracle_equals_reference_build_perf_inputs(oracle_lib):
    """The timing harness's deterministic inputs (performanceTest.c:169-219)."""
    from oracle import ref
    N, M = 5, 5
    b, r = oracle.Basis(N, M), ref.RefBasis(N, M)
    C, Ho, Hint = oracle.start_vector(b.nc), oracle.perf_ho(M), oracle.perf_hint(M)
    out = b.acth(C, Ho, Hint)
    assert np.linalg.norm(out - r.acth(C, Ho, Hint, "omp")) <= 1e-13 * np.linalg.norm(out)


# ---------------------------------------------------------- operators --

TINY = [(1, 1), (3, 1), (0, 3), (1, 3), (2, 3), (3, 3), (2, 4), (3, 4), (4, 3), (2, 5)]


@pytest.mark.parametrize("N,M", TINY)
def test_acth_dense_brute_force(oracle_lib, N, M):
    b = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=7 + N + 10 * M)
    H = bf.dense_h(N, M, Ho, Hint)
    assert np.allclose(H, H.conj().T, atol=1e-12)
    ref_out = H @ C
    out = b.acth(C, Ho, Hint)
    assert np.linalg.norm(out - ref_out) <= 1e-12 * max(np.linalg.norm(ref_out), 1e-300) + 1e-14
    rows = np.arange(b.nc)
    out2 = oracle.acth_hash_rows(N, M, rows, C, Ho, Hint)
    assert np.linalg.norm(out2 - ref_out) <= 1e-12 * max(np.linalg.norm(ref_out), 1e-300) + 1e-14


@pytest.mark.parametrize("N,M", [(5, 5), (4, 7), (7, 4)])
def test_acth_hermitian(oracle_lib, N, M):
    b = oracle.Basis(N, M)
    rng = np.random.default_rng(3)
    _, Ho, Hint = oracle.random_inputs(N, M, seed=11)
    x = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    y = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    a = np.vdot(x, b.acth(y, Ho, Hint))
    c = np.vdot(y, b.acth(x, Ho, Hint))
    assert abs(a - np.conj(c)) <= 1e-12 * abs(a)


def test_acth_weak_symmetry_still_hermitian(oracle_lib):
    """Hint with only Hint[k,l,q,s] = Hint[l,k,s,q] = conj(Hint[q,s,k,l])
    (the weaker relation some configurations state): the folded operator stays
    Hermitian."""
    N, M = 4, 5
    b = oracle.Basis(N, M)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((M,) * 4) + 1j * rng.standard_normal((M,) * 4)
    A = 0.5 * (A + A.transpose(1, 0, 3, 2))
    A = 0.5 * (A + np.conj(A.transpose(2, 3, 0, 1)))
    _, Ho, _ = oracle.random_inputs(N, M, seed=2)
    x = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    y = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    a = np.vdot(x, b.acth(y, Ho, A))
    c = np.vdot(y, b.acth(x, Ho, A))
    assert abs(a - np.conj(c)) <= 1e-12 * abs(a)


def test_acth_noninteracting_limit(oracle_lib):
    N, M = 5, 6
    b = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=1)
    e = np.linspace(-1, 2, M)
    out = b.acth(C, np.diag(e).astype(complex), np.zeros((M,) * 4, complex))
    IF = b.IF.reshape(-1, M)
    assert np.allclose(out, (IF @ e) * C, rtol=0, atol=1e-13)


def test_acth_lieb_liniger_conserves_momentum(oracle_lib):
    N, M = 5, 7
    b = oracle.Basis(N, M)
    IF = b.IF.reshape(-1, M)
    P = IF @ (np.arange(M) - M // 2)
    Ho, Hint = oracle.ll_ho(M), oracle.ll_hint(M, 3.0)
    for p0 in (-3, 0, 4):
        C = np.where(P == p0, 1.0 + 0.5j * np.arange(b.nc) / b.nc, 0)
        out = b.acth(C, Ho, Hint)
        assert np.abs(out[P != p0]).max() == 0.0


# --------------------------------------------------- density matrices --

@pytest.mark.parametrize("N,M", [(2, 3), (3, 3), (2, 4), (4, 3), (3, 4)])
def test_rho_brute_force(oracle_lib, N, M):
    b = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=N + M)
    assert np.abs(b.rho1(C) - bf.rho1(N, M, C)).max() < 1e-13
    assert np.abs(b.rho2(C) - bf.rho2(N, M, C)).max() < 1e-13


@pytest.mark.parametrize("N,M", [(5, 5), (4, 7), (6, 4)])
def test_rho_invariants_and_energy(oracle_lib, N, M):
    b = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=9)
    r1, r2 = b.rho1(C), b.rho2(C)
    assert abs(np.trace(r1) - N) < 1e-12
    assert np.allclose(r1, r1.conj().T, atol=1e-13)
    assert abs(np.einsum("ksks->", r2) - N * (N - 1)) < 1e-11
    assert np.allclose(np.einsum("kssl->kl", r2), (N - 1) * r1, atol=1e-12)
    assert np.allclose(r2, r2.transpose(1, 0, 2, 3), atol=1e-13)
    assert np.allclose(r2, r2.transpose(0, 1, 3, 2), atol=1e-13)
    assert np.allclose(r2, np.conj(r2.transpose(3, 2, 1, 0)), atol=1e-13)
    E = np.vdot(C, b.acth(C, Ho, Hint))
    E2 = np.sum(Ho * r1) + 0.5 * np.sum(Hint * r2)
    assert abs(E - E2) < 1e-11 * abs(E)


# ---------------------------------------------------------- Lanczos --

def test_tridiag_min_matches_numpy(oracle_lib):
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 40, 200):
        d = rng.standard_normal(n) * 3
        e = rng.standard_normal(max(n - 1, 0))
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w, V = np.linalg.eigh(T)
        lam, z = oracle.tridiag_min(d, e)
        assert abs(lam - w[0]) < 1e-12 * max(1, abs(w[0]))
        assert np.linalg.norm(T @ z - lam * z) < 1e-10


@pytest.mark.parametrize("N,M", [(4, 5), (3, 6)])
def test_ground_matches_dense(oracle_lib, N, M):
    b = oracle.Basis(N, M)
    C0, Ho, Hint = oracle.random_inputs(N, M, seed=4)
    H = bf.dense_h(N, M, Ho, Hint) if b.nc <= 60 else None
    E0, psi, it = b.ground(b.nc, C0, Ho, Hint)
    if H is None:
        H = np.column_stack([b.acth(np.eye(b.nc)[:, j], Ho, Hint) for j in range(b.nc)])
    w = np.linalg.eigvalsh(H)
    assert abs(E0 - w[0]) < 1e-10 * max(1, abs(w[0]))
    r = b.acth(psi, Ho, Hint) - E0 * psi
    assert np.linalg.norm(r) < 1e-8


def test_lanczos_basis_orthonormal(oracle_lib):
    b = oracle.Basis(5, 5)
    C0, Ho, Hint = oracle.random_inputs(5, 5, seed=8)
    n, diag, off, V = b.lanczos(40, C0, Ho, Hint)
    assert n == 40
    G = V[:n].conj() @ V[:n].T
    assert np.abs(G - np.eye(n)).max() < 1e-12
    assert np.abs(diag.imag).max() < 1e-12


def test_lanczos_breakdown_exact_eigenvector(oracle_lib):
    """Start from an exact eigenvector: the Krylov space has dimension 1 (the
    reference divides 0/0 here, groundStateLieb.c:289-294)."""
    N, M = 5, 7
    b = oracle.Basis(N, M)
    IF = b.IF.reshape(-1, M)
    # all bosons in the zero-momentum orbital, g = 0: H v = 0 exactly
    i0 = int(np.where(IF[:, M // 2] == N)[0][0])
    C0 = np.zeros(b.nc, complex)
    C0[i0] = 1
    n, diag, off, V = b.lanczos(10, C0, oracle.ll_ho(M), oracle.ll_hint(M, 0.0))
    assert n == 1 and diag[0] == 0
    E0, psi, it = b.ground(10, C0, oracle.ll_ho(M), oracle.ll_hint(M, 0.0))
    assert it == 1 and E0 == 0.0


@pytest.mark.parametrize("g", [1.0, 100.0, 1000.0])
def test_lieb_liniger_known_answers_m11(oracle_lib, g):
    """liebGS.dat (11 orbitals, N = 5, nc/8 Lanczos steps, printed %.5lf)."""
    import json, os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "liebGS.json")))
    row = [r for r in fx["rows"] if r["g"] == g][0]
    N, M = 5, 11
    b = oracle.Basis(N, M)
    E0, _, it = b.ground(b.nc // 8, oracle.start_vector(b.nc), oracle.ll_ho(M), oracle.ll_hint(M, g))
    assert it == b.nc // 8
    assert abs(E0 / N - row["E_M11"]) <= 5e-6 + 1e-9
    # variational bound against the Bethe-ansatz energy
    assert E0 / N >= bethe.energy_per_particle(N, g) - 1e-9


def test_bethe_solver(oracle_lib):
    N = 5
    for g in (0.5, 10.0, 300.0):
        k = bethe.quasi_momenta(N, g)
        assert np.abs(bethe.residual(k, g, N)).max() < 1e-12
    # weak coupling: E/N -> g (N-1)/2 / L ; strong coupling -> Tonks-Girardeau
    assert abs(bethe.energy_per_particle(N, 1e-4) - 1e-4 * (N - 1) / 2) < 1e-7
    assert abs(bethe.energy_per_particle(N, 1e7) - bethe.tonks_girardeau(N)) < 1e-4
    assert abs(bethe.tonks_girardeau(5) - 4 * np.pi ** 2) < 1e-12


@pytest.mark.parametrize("N,M", [(4, 5), (6, 4), (3, 7)])
def test_packed_layout_restatement(oracle_lib, N, M):
    """The test-side wave-packed layout holds every reference chunk once:
    E entries per chunk, offsets increasing; a slot keeps the entries q <= l
    outside the removed orbitals, and each one, decoded by position, is the
    rank of v - e_k - e_s + e_q + e_l (brute force through the oracle)."""
    from tests import packed
    o = oracle.Basis(N, M)
    IF = o.IF.reshape(-1, M)
    T = oracle.ncmat(N, M)
    for tt, Mp, st in ((False, o.MapOT, o.strideOT), (True, o.MapTT, o.strideTT)):
        pack, poff = packed.expected_packed(IF, M, Mp, st, tt)
        assert pack.size == Mp.size // (M * M) * packed.entries(M, tt)
        assert poff[0] == 0 and poff[-1] == pack.size and np.all(np.diff(poff) >= 0)
        # wave 0: walk its slots in jump order and check every entry of every
        # active lane against the rank of the moved state
        rows = list(range(min(64, IF.shape[0])))
        E, off = packed.entries(M, tt), 0
        for (k, s) in packed.jumps(M, tt):
            act = [r for r in rows if (IF[r, k] >= 1 and IF[r, s] >= 1 if tt else IF[r, k] >= 2)]
            if not act:
                continue
            pairs = packed.slot_pairs(M, k, s)
            assert len(pairs) == packed.real_entries(M, tt)
            for a, r in enumerate(act):
                for t in range(E):
                    got = pack[off + ((t // 4) * len(act) + a) * 4 + t % 4]
                    if t < len(pairs):
                        q, l = pairs[t]
                        v = IF[r].copy()
                        v[k] -= 1
                        v[s] -= 1
                        v[q] += 1
                        v[l] += 1
                        assert got == oracle.rank(N, M, T, v), (k, s, q, l, r)
                    else:
                        assert got == r
            off += E * len(act)
        assert off == poff[1]


def test_lanczos_tridiagonal_is_projection(oracle_lib):
    """V^H H V = T for the oracle's Lanczos basis (full reorthogonalization):
    the tridiagonal it returns (Re alpha on the diagonal, Re beta beside it)
    is the projection of H onto the Krylov basis, to rounding."""
    N, M, lm = 5, 6, 40
    b = oracle.Basis(N, M)
    C0, Ho, Hint = oracle.random_inputs(N, M, seed=31)
    n, diag, off, V = b.lanczos(lm, C0, Ho, Hint)
    assert n == lm
    HV = np.stack([b.acth(V[t], Ho, Hint) for t in range(n)])
    Tp = V[:n].conj() @ HV.T
    T = np.diag(diag.real) + np.diag(off.real[: n - 1], 1) + np.diag(off.real[: n - 1], -1)
    scale = np.abs(T).max()
    assert np.abs(Tp - T).max() < 1e-11 * scale, np.abs(Tp - T).max() / scale
    assert np.abs(diag.imag).max() < 1e-12 * scale and np.abs(off.imag).max() == 0.0


def test_start_vector_properties(oracle_lib):
    """groundStateLieb.c:578-586: C_i ~ sin(20 i/nc)(i%13) + I ((i%8)-(i%3)),
    normalized.  Pinned by properties a transcription slip breaks: unit
    norm; Re C vanishes at multiples of 13; the imaginary part has period 24,
    vanishes at multiples of 24 and takes the integer values -2..7 times one
    unit; Re C_i / (i%13) is A sin(20 i/nc), so
    f(i+d) + f(i-d) = 2 f(i) cos(20 d/nc)."""
    nc = 3003
    C = oracle.start_vector(nc)
    assert abs(np.linalg.norm(C) - 1) < 1e-14
    assert np.all(C.real[::13] == 0.0)
    assert np.array_equal(C.imag[: nc - 24], C.imag[24:])
    assert np.all(C.imag[::24] == 0.0)
    u = np.abs(C.imag[C.imag != 0]).min()      # the imaginary parts are integers times u
    r = C.imag / u
    assert np.abs(r - np.round(r)).max() < 1e-9 and round(r.min()) == -2 and round(r.max()) == 7
    d = 7
    i = np.array([i for i in range(d + 1, nc - d) if i % 13 and (i + d) % 13 and (i - d) % 13])
    f = lambda j: C.real[j] / (j % 13)  # noqa: E731
    lhs = f(i + d) + f(i - d)
    rhs = 2 * f(i) * np.cos(20 * d / nc)
    assert np.abs(lhs - rhs).max() < 1e-14


def test_perf_inputs_properties(oracle_lib):
    """performanceTest.c:169-219's deterministic Ho and Hint, pinned by what
    they must satisfy: Ho Hermitian with diagonal (i%4)-1 (period 4, values
    -1..2); Hint fully bosonic-symmetric (creator swap, annihilator swap,
    conjugate pair exchange), equal to the 1.234 background wherever the
    index pairs are not four distinct orbitals, and off it on some disjoint
    sets."""
    M = 7
    Ho = oracle.perf_ho(M)
    assert np.array_equal(Ho, Ho.conj().T)
    dg = np.diag(Ho)
    assert np.all(dg.imag == 0) and np.array_equal(dg.real[:4], [-1, 0, 1, 2])
    assert np.array_equal(dg[: M - 4], dg[4:])
    H = oracle.perf_hint(M)
    assert np.array_equal(H, H.transpose(1, 0, 2, 3))
    assert np.array_equal(H, H.transpose(0, 1, 3, 2))
    assert np.array_equal(H, np.conj(H.transpose(2, 3, 0, 1)))
    idx = np.indices((M,) * 4).reshape(4, -1).T
    distinct = np.array([len(set(r)) == 4 for r in idx])
    flatH = H.reshape(-1)
    assert np.all(flatH[~distinct] == 1.234)
    assert np.mean(flatH[distinct] != 1.234) > 0.9
