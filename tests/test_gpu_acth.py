"""H.v on the GPU against the oracle: every storage mode, zero-skipping on
and off, edge-case bases, the reference-shaped and CUDA-shaped entry points,
the golden fixtures, and the full N=12/M=12 and N=16/M=12 shapes through
sampled rows and size-independent properties.  Tolerance: complex128 within
1e-12 relative (normwise and elementwise against the largest entry)."""
import glob
import os

import numpy as np
import pytest

import mcpi_amd
import mcpi_amd.reference as R
import oracle
from tests.gpu_util import assert_close, note

pytestmark = pytest.mark.gpu

CASES = [(1, 1), (3, 1), (0, 3), (1, 3), (2, 3), (5, 5), (4, 6), (6, 4), (3, 7), (8, 5), (2, 9), (6, 9)]


def schedules(M, maps):
    """1 generic; 2 specialised (reads every stored map); 3 specialised,
    double jumps ranked on the fly (differs from 2 only with full maps);
    4 specialised with the pattern tables of the zero pattern (zero-skipping
    on a sparse Hamiltonian with Map stored; otherwise it runs as 2);
    5 every term with each block's double jumps compacted into full
    wavefronts (full maps; otherwise it runs as 2); 6 every term through
    the pair space (two gathers around a dense product with the pair
    matrix, any storage mode, M <= 12; otherwise it runs as 2); 7 the pair
    space in the factorized form (diagonal one-body part and a pair matrix
    with rank-one blocks, as Lieb-Liniger's; otherwise it runs as 2)."""
    out = [1]
    if 3 <= M <= 13:
        out.append(2)
        if maps == "full":
            out.append(3)
        if maps != "none":
            out.append(4)
        out.append(5)
        out.append(6)
        out.append(7)
    return out


@pytest.mark.parametrize("N,M", CASES)
@pytest.mark.parametrize("maps", ["full", "single", "none"])
def test_apply_random(oracle_lib, N, M, maps):
    note(f"acth random N={N} M={M} maps={maps}")
    o = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=N * 100 + M)
    ref = o.acth(C, Ho, Hint)
    g = mcpi_amd.Basis(N, M, maps)
    for skip in (True, False):
        for sched in schedules(M, maps):
            H = mcpi_amd.Hamiltonian(g, Ho, Hint, skip_zero=skip, schedule=sched)
            assert_close(H(C), ref, 1e-12, f"{maps} skip={skip} schedule={sched}")
        H = mcpi_amd.Hamiltonian(g, Ho, Hint, skip_zero=skip)   # autotuned
        assert_close(H(C), ref, 1e-12, f"{maps} skip={skip} auto")
        assert H.schedule in schedules(M, maps)


def test_apply_sparse_hint_all_schedules(oracle_lib):
    """A Hint with scattered exact zeros and a non-diagonal Ho: the
    zero-skipping lists must drop exactly the zero terms."""
    N, M = 6, 7
    o = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=99)
    rng = np.random.default_rng(4)
    mask = rng.random((M,) * 4) < 0.6
    mask = mask & mask.transpose(1, 0, 2, 3) & mask.transpose(0, 1, 3, 2) & mask.transpose(2, 3, 0, 1)
    Hint = np.where(mask, Hint, 0)
    Ho = np.where(rng.random((M, M)) < 0.5, Ho, 0)
    Ho = np.triu(Ho) + np.triu(Ho, 1).conj().T
    ref = o.acth(C, Ho, Hint)
    for maps in ("full", "single", "none"):
        g = mcpi_amd.Basis(N, M, maps)
        for sched in schedules(M, maps):
            for skip in (True, False):
                note(f"sparse hint maps={maps} schedule={sched} skip={skip}")
                H = mcpi_amd.Hamiltonian(g, Ho, Hint, skip_zero=skip, schedule=sched)
                assert_close(H(C), ref, 1e-12, f"{maps} {sched} {skip}")


@pytest.mark.parametrize("sched", [6, 7])
def test_pair_schedule_two_streams(oracle_lib, sched):
    """Schedules 6 and 7 keep one workspace per stream: two Hamiltonians on
    one basis (random ones for 6, Lieb-Liniger with two couplings for 7),
    their H.v enqueued alternately on two non-blocking streams without
    synchronization in between (so that they may overlap), each match the
    oracle; then the same on the default stream."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    N, M = 9, 10
    o = oracle.Basis(N, M)
    C1, Ho1, Hint1 = oracle.random_inputs(N, M, seed=61)
    C2, Ho2, Hint2 = oracle.random_inputs(N, M, seed=62)
    if sched == 7:
        Ho1, Hint1 = oracle.ll_ho(M), oracle.ll_hint(M, 2.0)
        Ho2, Hint2 = oracle.ll_ho(M), oracle.ll_hint(M, 30.0)
    ref1, ref2 = o.acth(C1, Ho1, Hint1), o.acth(C2, Ho2, Hint2)
    b = mcpi_amd.Basis(N, M, "single")
    H1 = mcpi_amd.Hamiltonian(b, Ho1, Hint1, skip_zero=False, schedule=sched)
    H2 = mcpi_amd.Hamiltonian(b, Ho2, Hint2, skip_zero=False, schedule=sched)
    d1, d2 = mcpi_amd.DeviceArray.from_numpy(C1), mcpi_amd.DeviceArray.from_numpy(C2)
    o1, o2 = mcpi_amd.DeviceArray(b.nc), mcpi_amd.DeviceArray(b.nc)
    streams = [ct.c_void_p(), ct.c_void_p()]
    for s in streams:
        assert hip.hipStreamCreateWithFlags(ct.byref(s), ct.c_uint(1)) == 0
    try:
        note(f"pair schedule {sched} two streams N={N} M={M}")
        for _ in range(4):
            H1.apply(d1, o1, streams[0])
            H2.apply(d2, o2, streams[1])
        for s in streams:
            assert hip.hipStreamSynchronize(s) == 0
        assert H1.schedule == sched and H2.schedule == sched
        assert_close(o1.numpy(), ref1, 1e-12, "stream 0")
        assert_close(o2.numpy(), ref2, 1e-12, "stream 1")
        H1.apply(d2, o1)
        H2.apply(d1, o2)
        mcpi_amd.synchronize()
        assert_close(o1.numpy(), o.acth(C2, Ho1, Hint1), 1e-12, "default stream, H1 C2")
        assert_close(o2.numpy(), o.acth(C1, Ho2, Hint2), 1e-12, "default stream, H2 C1")
    finally:
        for s in streams:
            hip.hipStreamDestroy(s)


def test_pair_workspace_sizes(oracle_lib):
    """Schedule 7 allocates only its Z rows; schedule 6 on the same stream
    then needs a larger workspace (allocated beside the first), and 7 after
    it reuses either: every H.v matches the oracle.  The autotuner, with
    zero-skipping on a Hamiltonian that factorizes, never picks 6."""
    N, M = 9, 10
    o = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=63)
    Ho, Hint = oracle.ll_ho(M), oracle.ll_hint(M, 5.0)
    ref = o.acth(C, Ho, Hint)
    b = mcpi_amd.Basis(N, M, "none")
    H = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=False, schedule=7)
    note(f"pair workspace sizes N={N} M={M}")
    for sched in (7, 6, 7, 6):
        H.set_schedule(sched)
        assert_close(H(C), ref, 1e-12, f"schedule {sched}")
        assert H.schedule == sched
    Ht = mcpi_amd.Hamiltonian(mcpi_amd.Basis(N, M, "none"), Ho, Hint, skip_zero=True)
    assert_close(Ht(C), ref, 1e-12, "autotuned")
    assert Ht.schedule != 6


@pytest.mark.parametrize("N,M", [(2, 3), (3, 5), (7, 8), (9, 12), (6, 13), (4, 16), (3, 21), (2, 24)])
def test_factorized_pair_form(oracle_lib, N, M):
    """Schedule 7 on Hamiltonians that factorize (a contact interaction
    conserving momentum with a random real strength per total momentum K,
    and a random real diagonal Ho) and on ones that do not (an off-diagonal
    Ho, a dense Hint: the schedule falls back to 2), each against the
    oracle."""
    note(f"factorized pair form N={N} M={M}")
    o = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=N * 31 + M)
    rng = np.random.default_rng(M)
    gK = rng.standard_normal(2 * M) + 1j * rng.standard_normal(2 * M)
    gK = gK + np.conj(gK)            # real: Hint Hermitian, W_K = g_K / 2 m m^T
    k = np.arange(M)
    K = k[:, None, None, None] + k[None, :, None, None]
    Ksum = k[None, None, :, None] + k[None, None, None, :]
    Hint = np.where(K == Ksum, gK[np.minimum(K, 2 * M - 1)], 0).astype(complex)
    Hd = np.diag(rng.standard_normal(M)).astype(complex)
    _, Ho_dense, Hint_dense = oracle.random_inputs(N, M, seed=5)
    b = mcpi_amd.Basis(N, M, "single")
    for label, Ho, Hi in (("factorizes", Hd, Hint), ("off-diagonal Ho", Ho_dense, Hint),
                          ("dense Hint", Hd, Hint_dense)):
        ref = o.acth(C, Ho, Hi)
        for skip in (True, False):
            H = mcpi_amd.Hamiltonian(b, Ho, Hi, skip_zero=skip, schedule=7)
            assert_close(H(C), ref, 1e-12, f"{label} skip={skip}")
        H = mcpi_amd.Hamiltonian(b, Ho, Hi)   # autotuned
        assert_close(H(C), ref, 1e-12, f"{label} auto")


@pytest.mark.parametrize("maps", ["full", "single", "none"])
def test_large_m_schedules(oracle_lib, maps):
    """M = 15 (beyond the kernels specialised on M <= 13): every schedule
    either applies (1; 7 for the Lieb-Liniger Hamiltonian) or falls back
    to the generic kernel, and all agree with the oracle."""
    N, M = 3, 15
    note(f"large M schedules N={N} M={M} maps={maps}")
    o = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=315)
    b = mcpi_amd.Basis(N, M, maps)
    for label, ho, hi in (("random", Ho, Hint), ("Lieb-Liniger", oracle.ll_ho(M), oracle.ll_hint(M, 3.0))):
        ref = o.acth(C, ho, hi)
        for sched in range(1, 8):
            H = mcpi_amd.Hamiltonian(b, ho, hi, schedule=sched)
            assert_close(H(C), ref, 1e-12, f"{label} schedule {sched}")
        H = mcpi_amd.Hamiltonian(b, ho, hi)
        assert_close(H(C), ref, 1e-12, f"{label} auto")
        assert H.schedule in ((1, 7) if label == "Lieb-Liniger" else (1,))


@pytest.mark.parametrize("N,M,g", [(5, 11, 10.0), (8, 9, 0.3), (6, 12, 50.0)])
@pytest.mark.parametrize("maps", ["full", "single", "none"])
def test_apply_lieb_liniger(oracle_lib, N, M, g, maps):
    note(f"acth LL N={N} M={M} g={g} maps={maps}")
    o = oracle.Basis(N, M)
    C, _, _ = oracle.random_inputs(N, M, seed=3)
    Ho, Hint = oracle.ll_ho(M), oracle.ll_hint(M, g)
    ref = o.acth(C, Ho, Hint)
    b = mcpi_amd.Basis(N, M, maps)
    for skip in (True, False):
        for sched in schedules(M, maps):
            H = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=skip, schedule=sched)
            assert_close(H(C), ref, 1e-12, f"skip={skip} schedule={sched}")


GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_apply_golden(path):
    gd = np.load(path)
    N, M = int(gd["N"]), int(gd["M"])
    note(f"acth golden {os.path.basename(path)}")
    for maps in ("full", "single", "none"):
        b = mcpi_amd.Basis(N, M, maps)
        assert_close(mcpi_amd.Hamiltonian(b, gd["Ho"], gd["Hint"])(gd["C"]), gd["out"], 1e-12, maps)


def test_reference_shaped_apply(oracle_lib):
    for (N, M) in [(5, 5), (4, 6), (1, 4)]:
        note(f"reference-shaped applyHconf_omp / applyHconf N={N} M={M}")
        r = R.RefBasis(N, M)
        o = oracle.Basis(N, M)
        C, Ho, Hint = oracle.random_inputs(N, M, seed=5)
        ref = o.acth(C, Ho, Hint)
        assert_close(r.applyHconf_omp(C, Ho, Hint), ref, 1e-12, "applyHconf_omp")
        assert_close(r.applyHconf_omp(C, Ho, Hint), ref, 1e-12, "cached tables")
        assert_close(r.applyHconf(C, Ho, Hint), ref, 1e-12, "applyHconf")


def test_reference_shaped_variants_and_cache(oracle_lib):
    """applyHconf_X (hamiltonianMatrix.h:431) and the repeated-call cache:
    the same arrays again reuse the device tables and the tuned Hamiltonian;
    a changed Hint (same array, new values) or new arrays are picked up; the
    cache can be dropped."""
    N, M = 5, 6
    note("reference-shaped applyHconf_X and the call cache")
    r = R.RefBasis(N, M)
    o = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=15)
    ref = o.acth(C, Ho, Hint)
    assert_close(r.applyHconf_X(C, Ho, Hint), ref, 1e-12, "applyHconf_X")
    for _ in range(3):
        assert_close(r.applyHconf_omp(C, Ho, Hint), ref, 1e-12, "applyHconf_omp repeated")
    Hint2 = Hint.copy()
    Hint2 *= 0.5
    assert_close(r.applyHconf_omp(C, Ho, Hint2), o.acth(C, Ho, Hint2), 1e-12, "Hint changed")
    Ho2 = Ho + np.eye(M)
    assert_close(r.applyHconf_omp(C, Ho2, Hint2), o.acth(C, Ho2, Hint2), 1e-12, "Ho changed")
    R.release_cached()
    assert_close(r.applyHconf_omp(C, Ho, Hint), ref, 1e-12, "after release")
    r2 = R.RefBasis(4, 7)
    o2 = oracle.Basis(4, 7)
    C2, Ho3, Hint3 = oracle.random_inputs(4, 7, seed=16)
    assert_close(r2.applyHconf_omp(C2, Ho3, Hint3), o2.acth(C2, Ho3, Hint3), 1e-12, "other basis")


def test_reference_shaped_weak_symmetry_fixture():
    """cfg1 with a Hint of only the weaker symmetry
    Hint[k,l,q,s] = Hint[l,k,s,q] = conj(Hint[q,s,k,l]): the reference's
    applyHconf_omp output (written by its own build) through the
    reference-shaped entry point, every schedule of the extended API, and
    the CUDA-shaped one."""
    gd = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1_weaksym.npz"))
    N, M = int(gd["N"]), int(gd["M"])
    note("weak-symmetry cfg1 fixture")
    r = R.RefBasis(N, M)
    assert_close(r.applyHconf_omp(gd["C"], gd["Ho"], gd["Hint"]), gd["out"], 1e-12, "applyHconf_omp")
    assert_close(r.applyHconf_X(gd["C"], gd["Ho"], gd["Hint"]), gd["out"], 1e-12, "applyHconf_X")
    b = mcpi_amd.Basis(N, M, "full")
    for sched in schedules(M, "full"):
        for skip in (True, False):
            H = mcpi_amd.Hamiltonian(b, gd["Ho"], gd["Hint"], skip_zero=skip, schedule=sched)
            assert_close(H(gd["C"]), gd["out"], 1e-12, f"schedule {sched} skip {skip}")


def test_cuda_shaped_dev_entry_point(oracle_lib):
    N, M = 6, 5
    note("applyHconf_dev N=6 M=5")
    o = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=6)
    D = mcpi_amd.DeviceArray.from_numpy
    s_ot = np.concatenate([o.strideOT, [o.MapOT.size]]).astype(np.int64)
    s_tt = np.concatenate([o.strideTT, [o.MapTT.size]]).astype(np.int64)
    out = mcpi_amd.DeviceArray(o.nc)
    mcpi_amd.applyHconf_dev(N, M, D(o.Map), D(o.MapOT), D(o.MapTT), D(s_ot), D(s_tt), D(o.IF),
                            D(C), D(oracle.flat(Ho)), D(oracle.flat(Hint)), out)
    mcpi_amd.synchronize()
    assert_close(out.numpy(), o.acth(C, Ho, Hint), 1e-12)


@pytest.mark.parametrize("N,M", [(6, 5), (3, 14)])
def test_cuda_shaped_dev32_entry_point(oracle_lib, N, M):
    """int32 strides as the CUDA kernel takes them (cuda/hamiltonianMatrix.cuh:
    61-63); M = 14 runs the generic kernel, which reads the widened strides.
    Repeated calls reuse the folded coefficients; new Ho/Hint values in the
    same device arrays are picked up."""
    note(f"applyHconf_dev32 N={N} M={M}")
    o = oracle.Basis(N, M)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=60 + M)
    D = mcpi_amd.DeviceArray.from_numpy
    out = mcpi_amd.DeviceArray(o.nc)
    dHo, dHint = D(oracle.flat(Ho)), D(oracle.flat(Hint))
    args = (D(o.Map), D(o.MapOT), D(o.MapTT), D(o.strideOT.astype(np.int32)), D(o.strideTT.astype(np.int32)),
            D(o.IF), D(C))
    for _ in range(2):
        mcpi_amd.applyHconf_dev32(N, M, *args, dHo, dHint, out)
        mcpi_amd.synchronize()
        assert_close(out.numpy(), o.acth(C, Ho, Hint), 1e-12, "dev32")
    dHint.upload(oracle.flat(2.0 * Hint))
    mcpi_amd.applyHconf_dev32(N, M, *args, dHo, dHint, out)
    mcpi_amd.synchronize()
    assert_close(out.numpy(), o.acth(C, Ho, 2.0 * Hint), 1e-12, "dev32 new Hint")


def _hermitian_check(H, nc, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    y = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    a = np.vdot(x, H(y))
    c = np.vdot(y, H(x))
    assert abs(a - np.conj(c)) <= 1e-11 * abs(a)


@pytest.mark.parametrize("maps", ["full", "single"])
def test_n12_m12_lieb_liniger_sampled_rows(oracle_lib, maps):
    """N=12, M=12, Lieb-Liniger g=10: 4000 seeded rows recomputed by the
    oracle's hash-only H.v; Hermiticity <x,Hy> = conj<y,Hx>."""
    N, M = 12, 12
    note(f"N=12 M=12 LL maps={maps}")
    b = mcpi_amd.Basis(N, M, maps)
    rng = np.random.default_rng(1212)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    Ho, Hint = oracle.ll_ho(M), oracle.ll_hint(M, 10.0)
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    out = H(C)
    rows = np.concatenate([[0, 1, b.nc - 1], rng.integers(0, b.nc, 4000)])
    ref = oracle.acth_hash_rows(N, M, rows, C, Ho, Hint)
    assert_close(out[rows], ref, 1e-12, "rows")
    for sched in schedules(M, maps):
        for skip in (False, True):
            H2 = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=skip, schedule=sched)
            assert_close(H2(C), out, 1e-12, f"schedule {sched} skip {skip}")
    _hermitian_check(H, b.nc, 7)


def test_n12_m12_generic_sampled_rows(oracle_lib):
    N, M = 12, 12
    note("N=12 M=12 generic Hint, full maps")
    b = mcpi_amd.Basis(N, M, "full")
    _, Ho, Hint = oracle.random_inputs(3, M, seed=44, sigma=0.1)
    rng = np.random.default_rng(5)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    out = H(C)
    rows = np.concatenate([[0, b.nc - 1], rng.integers(0, b.nc, 1500)])
    assert_close(out[rows], oracle.acth_hash_rows(N, M, rows, C, Ho, Hint), 1e-12, "rows")
    for sched in (1, 2, 3, 5, 6, 7):
        H2 = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=False, schedule=sched)
        assert_close(H2(C), out, 1e-12, f"schedule {sched}")


def test_n16_m12_single_map_sampled_rows(oracle_lib):
    """N=16, M=12, LL g=50, Map + on-the-fly double jumps (the large shape)."""
    N, M = 16, 12
    note("N=16 M=12 LL single map")
    b = mcpi_amd.Basis(N, M, "single")
    rng = np.random.default_rng(1612)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    Ho, Hint = oracle.ll_ho(M), oracle.ll_hint(M, 50.0)
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    out = H(C)
    rows = np.concatenate([[0, b.nc - 1], rng.integers(0, b.nc, 3000)])
    assert_close(out[rows], oracle.acth_hash_rows(N, M, rows, C, Ho, Hint), 1e-12, "rows")
    _hermitian_check(H, b.nc, 9)
    for sched in (2, 4, 6, 7):
        H2 = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=True, schedule=sched)
        assert_close(H2(C), out, 1e-12, f"schedule {sched}")
