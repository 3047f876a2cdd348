"""Batched Lanczos on the GPU (mcpi_hip_lanczos_batched): many instances of
uneven (N, M, g) advanced together, against the oracle's Lanczos ground state
(groundStateLieb.c:217-471 restated) and against the one-instance solver.
Tolerance: |dE0| < 1e-10 absolute; Ritz vectors to 1e-8 in overlap."""
import numpy as np
import pytest

import mcpi_amd
import oracle
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


def _start(nc, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    return v / np.linalg.norm(v)


def _ll_instances(spec, seed):
    """spec: list of (N, M); g log-uniform in [0.1, 100] from the seed."""
    rng = np.random.default_rng(seed)
    return [(N, M, float(10 ** rng.uniform(-1, 2))) for (N, M) in spec]


@pytest.fixture(params=[0, 1, 2, 3], ids=["timed", "batched_kernel", "multi_vector", "factorized_multi"])
def batched_kernel(request):
    mcpi_amd.set_batched_kernel(request.param)
    yield request.param
    mcpi_amd.set_batched_kernel(0)


def test_batched_matches_oracle_ll_mixed_bases(oracle_lib, batched_kernel):
    """Lieb-Liniger instances over (6,9) and (8,11) with 120 steps each
    (the small bases of the 64-instance workload), one call; every H.v
    kernel choice for the batched groups (the multi-vector kernel applies:
    Lieb-Liniger couplings share the zero pattern)."""
    spec = [(6, 9), (8, 11), (6, 9), (8, 11), (6, 9), (6, 9), (8, 11), (8, 11)]
    inst = _ll_instances(spec, 3)
    lm = 120
    bases = {nm: mcpi_amd.Basis(nm[0], nm[1], "full") for nm in set(spec)}
    obases = {nm: oracle.Basis(nm[0], nm[1]) for nm in set(spec)}
    hams, starts, outs, ref = [], [], [], []
    for x, (N, M, g) in enumerate(inst):
        b = bases[(N, M)]
        Ho, Hint = oracle.ll_ho(M), oracle.ll_hint(M, g)
        C0 = _start(b.nc, 100 + x)
        hams.append(mcpi_amd.Hamiltonian(b, Ho, Hint))
        starts.append(mcpi_amd.DeviceArray.from_numpy(C0))
        outs.append(mcpi_amd.DeviceArray(b.nc))
        ref.append(obases[(N, M)].ground(lm, C0, Ho, Hint))
    note(f"batched LL {len(inst)} instances lm={lm}")
    E, steps = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    for x in range(len(inst)):
        E_o, psi_o, _ = ref[x]
        assert steps[x] == lm
        assert abs(E[x] - E_o) < 1e-10, (x, inst[x], E[x], E_o)
        assert abs(abs(np.vdot(psi_o, outs[x].numpy())) - 1) < 1e-8


def test_batched_equals_single_instance_generic(batched_kernel):
    """Generic complex Ho/Hint (no zero terms to skip), three bases, mixed
    sizes: every instance's E0, steps and Ritz vector equal the one-instance
    solver's."""
    spec = [(4, 5), (5, 6), (4, 5), (3, 7), (5, 6), (4, 5)]
    lm = 40
    bases = {nm: mcpi_amd.Basis(nm[0], nm[1], "full") for nm in set(spec)}
    hams, starts, outs, single = [], [], [], []
    for x, (N, M) in enumerate(spec):
        b = bases[(N, M)]
        C0, Ho, Hint = oracle.random_inputs(N, M, seed=20 + x)
        H = mcpi_amd.Hamiltonian(b, Ho, Hint)
        hams.append(H)
        starts.append(mcpi_amd.DeviceArray.from_numpy(C0))
        outs.append(mcpi_amd.DeviceArray(b.nc))
        single.append(H.ground(lm, C0))
    note("batched vs single, generic Hamiltonians")
    E, steps = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    for x in range(len(spec)):
        E1, psi1, st1 = single[x]
        assert steps[x] == st1
        assert abs(E[x] - E1) < 1e-12
        assert abs(abs(np.vdot(psi1.numpy(), outs[x].numpy())) - 1) < 1e-10


def test_batched_breakdown_and_unbatched_groups(oracle_lib, batched_kernel):
    """Bases smaller than lm (N=1, M=6: nc = 6; M=1: nc = 1) run next to
    full-length instances; a group whose members disagree on zero-skipping
    runs its instances one by one; an M outside the specialised kernels
    (M=15) mixes in.  Ritz outputs may be omitted per instance.  Once the
    Krylov space is exhausted the next beta is rounding noise, and whether
    it passes the reference's relative breakdown test (beta_i / max beta <
    1e-15, groundStateLieb.c:287-292) depends on that noise, so only E0 is
    compared for the small bases (the spurious Ritz values sit within
    rounding of the true ones)."""
    cases = [(1, 6, True), (4, 6, True), (4, 6, False), (2, 15, True), (3, 1, True), (4, 6, True)]
    lm = 30
    bases = {}
    hams, starts, outs, refs = [], [], [], []
    for x, (N, M, skip) in enumerate(cases):
        b = bases.setdefault((N, M), mcpi_amd.Basis(N, M, "full"))
        C0, Ho, Hint = oracle.random_inputs(N, M, seed=40 + x)
        hams.append(mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=skip))
        starts.append(mcpi_amd.DeviceArray.from_numpy(C0))
        outs.append(mcpi_amd.DeviceArray(b.nc) if x % 2 == 0 else None)
        o = oracle.Basis(N, M)
        refs.append(o.ground(lm, C0, Ho, Hint))
    note("batched breakdown + unbatched groups")
    E, steps = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    for x, (N, M, _) in enumerate(cases):
        E_o, psi_o, it_o = refs[x]
        nc = bases[(N, M)].nc
        if nc < lm:
            assert 1 <= steps[x] <= lm
        else:
            assert steps[x] == lm == it_o
        assert abs(E[x] - E_o) < 1e-10, (x, cases[x], E[x], E_o)
        if outs[x] is not None and nc >= lm:
            assert abs(abs(np.vdot(psi_o, outs[x].numpy())) - 1) < 1e-8


def test_batched_empty_and_argument_errors():
    E, steps = mcpi_amd.lanczos_batched([], [], 10)
    assert E.size == 0 and steps.size == 0
    b = mcpi_amd.Basis(3, 4, "full")
    H = mcpi_amd.Hamiltonian(b, oracle.ll_ho(4), oracle.ll_hint(4, 1.0))
    with pytest.raises(ValueError):
        mcpi_amd.lanczos_batched([H], [mcpi_amd.DeviceArray(b.nc + 1)], 10)
    with pytest.raises(mcpi_amd.MCPIError):
        mcpi_amd.lanczos_batched([H], [mcpi_amd.DeviceArray.from_numpy(_start(b.nc, 1))], 0)


def test_multi_vector_sparse_generic(oracle_lib, batched_kernel):
    """Generic complex Hamiltonians with a common pattern of exact zeros in
    Hint and Ho (so the multi-vector kernel applies to non-real
    coefficients), five instances (one partial group of four), full and
    single maps."""
    N, M = 5, 7
    rng = np.random.default_rng(8)
    mask = rng.random((M,) * 4) < 0.6
    mask = mask & mask.transpose(1, 0, 2, 3) & mask.transpose(0, 1, 3, 2) & mask.transpose(2, 3, 0, 1)
    mho = rng.random((M, M)) < 0.5
    mho = mho & mho.T
    o = oracle.Basis(N, M)
    for maps in ("full", "single"):
        b = mcpi_amd.Basis(N, M, maps)
        hams, starts, ref = [], [], []
        for x in range(5):
            C0, Ho, Hint = oracle.random_inputs(N, M, seed=60 + x)
            Hint = np.where(mask, Hint, 0)
            Ho = np.where(mho, Ho, 0)
            hams.append(mcpi_amd.Hamiltonian(b, Ho, Hint))
            starts.append(mcpi_amd.DeviceArray.from_numpy(C0))
            ref.append(o.ground(40, C0, Ho, Hint)[0])
        note(f"multi-vector sparse generic maps={maps}")
        E, steps = mcpi_amd.lanczos_batched(hams, starts, 40)
        for x in range(5):
            assert abs(E[x] - ref[x]) < 1e-10, (maps, x, E[x], ref[x])


def test_factorized_pair_form_several_members(oracle_lib):
    """Batched groups whose members all have the factorized pair form with
    the same blocks (Lieb-Liniger couplings of one basis, a random diagonal
    Ho per member) run it four members at a time, then two, then one:
    fifteen members of (8,11) and three of (6,9), with kind 3 forcing that
    path.  A
    random Hamiltonian in a third basis's group keeps it from applying
    there.  Every E0 and Ritz vector equals the one-instance solver's (the
    same arithmetic per member), two instances are also checked against the
    oracle."""
    rng = np.random.default_rng(77)
    spec = [(8, 11)] * 15 + [(6, 9)] * 3 + [(5, 6)] * 2
    lm = 60
    bases = {nm: mcpi_amd.Basis(nm[0], nm[1], "full") for nm in set(spec)}
    hams, starts, outs, mats = [], [], [], []
    for x, (N, M) in enumerate(spec):
        b = bases[(N, M)]
        if (N, M) == (5, 6) and x == len(spec) - 1:
            _, Ho, Hint = oracle.random_inputs(N, M, seed=5)
        else:
            Ho = np.diag(oracle.ll_ho(M).diagonal() + rng.uniform(-1, 1, M)).astype(complex)
            Hint = oracle.ll_hint(M, float(10 ** rng.uniform(-1, 2)))
        mats.append((Ho, Hint))
        hams.append(mcpi_amd.Hamiltonian(b, Ho, Hint))
        starts.append(mcpi_amd.DeviceArray.from_numpy(_start(b.nc, 300 + x)))
        outs.append(mcpi_amd.DeviceArray(b.nc))
    note(f"factorized pair form several members, {len(spec)} instances lm={lm}")
    mcpi_amd.set_batched_kernel(3)
    try:
        E, steps = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    finally:
        mcpi_amd.set_batched_kernel(0)
    for x in range(len(spec)):
        single = mcpi_amd.DeviceArray(hams[x].basis.nc)
        E1, _, st1 = hams[x].ground(lm, starts[x], out=single)
        assert steps[x] == st1 == lm
        assert abs(E[x] - E1) < 1e-10, (x, spec[x], E[x], E1)
        assert abs(abs(np.vdot(single.numpy(), outs[x].numpy())) - 1) < 1e-8
    for x in (0, 14):
        N, M = spec[x]
        E_o, psi_o, _ = oracle.Basis(N, M).ground(lm, starts[x].numpy(), *mats[x])
        assert abs(E[x] - E_o) < 1e-10, (x, E[x], E_o)
        assert abs(abs(np.vdot(psi_o, outs[x].numpy())) - 1) < 1e-8


def test_batched_repeat_calls_reuse_tuning():
    """A second call on the same group reuses the first call's H.v kernel
    choice (same kernels: identical results); after the handles are closed,
    new Hamiltonians of another kind on the same basis (their handles may
    reuse the old addresses) still get a choice that applies to them:
    every E0 equals the one-instance solver's."""
    b = mcpi_amd.Basis(8, 11, "full")
    lm = 40
    starts = [mcpi_amd.DeviceArray.from_numpy(_start(b.nc, 500 + x)) for x in range(5)]
    hams = [mcpi_amd.Hamiltonian(b, oracle.ll_ho(11), oracle.ll_hint(11, 0.5 + 3 * x)) for x in range(5)]
    note("batched repeat calls")
    E1, _ = mcpi_amd.lanczos_batched(hams, starts, lm)
    E2, _ = mcpi_amd.lanczos_batched(hams, starts, lm)
    assert np.array_equal(E1, E2)
    for H in hams:
        H.close()
    hams = []
    for x in range(5):
        _, Ho, Hint = oracle.random_inputs(8, 11, seed=600 + x)
        hams.append(mcpi_amd.Hamiltonian(b, Ho, Hint))
    E3, _ = mcpi_amd.lanczos_batched(hams, starts, lm)
    for x in range(5):
        E_single = hams[x].ground(lm, starts[x])[0]
        assert abs(E3[x] - E_single) < 1e-10, (x, E3[x], E_single)
