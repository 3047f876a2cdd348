#!/usr/bin/env python3
"""Benchmark: single-vector H.v throughput on one MI355X.

Workload (the configuration the headline metric is quoted on): N = 12 bosons
in M = 12 plane-wave orbitals (nc = C(23,12) = 1,352,078), Lieb-Liniger
Hamiltonian on a ring of length 1 (Ho diagonal (2 pi j)^2 / 2, j = -6..5;
Hint = g delta(k+l = q+s), g = 10), complex128, all three jump maps stored on
the device (int64 offsets).  The input vector is seeded random normalized
complex128.  One step = one H.v over the whole basis, inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints ONE JSON line.  ``value`` is H.v throughput in GB/s of algorithmic
bytes: B = 16 nc (C) + 16 nc (out) + 4 M nc (IF) + 4 M^2 nc (Map)
+ 4 |MapOT| + 4 |MapTT| + 4 nc (strideOT) + 8 nc (strideTT), i.e. every
stored table and vector counted once per H.v.  The timed kernel evaluates
every rule term and reads every map entry (zero coefficients are not
skipped); the Lieb-Liniger-aware path that skips exact zeros is reported
separately under ``extra``.  ``roofline.achieved`` uses the H.v kernel's own
duration from HIP events recorded on its launch stream.  ``cpu_baseline``
is the oracle's C/OpenMP restatement of the reference's applyHconf_omp, full
H.v passes repeated for about --cpu-seconds on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_, M_, G_ = 12, 12, 10.0


def algorithmic_bytes(nc: int, M: int, ot_len: int, tt_len: int) -> int:
    return 16 * nc + 16 * nc + 4 * M * nc + 4 * M * M * nc + 4 * ot_len + 4 * tt_len + 4 * nc + 8 * nc


# sources of the headline H.v kernel: the recorded PMC traffic is only used
# while they are unchanged since it was measured
KERNEL_SOURCES = ["mcpi_amd/csrc/acth_comp.hip", "mcpi_amd/csrc/acth_fast.h", "mcpi_amd/csrc/mcpi_internal.h"]


def density_model(N: int, M: int, rho2_ms: float) -> dict:
    """Work of the Gram form of rho2: Phi (rows: the states of N-2
    particles, columns: the M(M+1)/2 orbital pairs) gathered from C, and the
    upper triangle of Phi^H Phi at 8 real flops per complex multiply-add;
    rate per host call against the FP64 peak."""
    from math import comb
    nu = comb(N - 2 + M - 1, M - 1) if N >= 2 else 0
    npair = M * (M + 1) // 2
    flops = 8 * nu * npair * (npair + 1) // 2
    return {"rho2_rows": nu, "rho2_pairs": npair, "rho2_gathers": nu * npair, "rho2_flops": flops,
            "rho2_tflops_per_call": round(flops / (rho2_ms * 1e-3) / 1e12, 2),
            "rho2_fp64_peak_tflops": 78.6,
            "rho2_frac_fp64_per_call": round(flops / (rho2_ms * 1e-3) / 78.6e12, 3)}


def pair_product_flops(N: int, M: int) -> int:
    """Real FP64 flops of the pair-space product Y2 = W X of one H.v:
    6 NP^2 nu2 (three real products of 2 flops per complex entry), NP =
    M(M+1)/2 orbital pairs, nu2 = C(N-2+M-1, N-2) states of N-2 particles."""
    NP = M * (M + 1) // 2
    return 6 * NP * NP * math.comb(N - 2 + M - 1, N - 2)


def kernel_sources_sha() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update(open(os.path.join(ROOT, f), "rb").read())
    return h.hexdigest()[:16]


def traffic_from_profiles():
    """HBM bytes per H.v launch measured with rocprofv3 PMC counters
    (profiles/acth_full_n12_m12_traffic.json), with the record's kernel name
    and whether the kernel sources still hash to what was measured.  A stale
    record gives traffic None (the old figure is kept beside it)."""
    p = os.path.join(ROOT, "profiles", "acth_full_n12_m12_traffic.json")
    if not os.path.exists(p):
        return None, {"traffic_record": None}
    try:
        rec = json.load(open(p))
    except Exception:
        return None, {"traffic_record": "unreadable"}
    fresh = rec.get("kernel_sources_sha16") == kernel_sources_sha()
    info = {"traffic_kernel": rec.get("kernel"), "traffic_sources_sha16": rec.get("kernel_sources_sha16"),
            "traffic_matches_sources": fresh}
    if not fresh:
        info["traffic_stale_bytes"] = rec.get("hbm_bytes_per_launch")
    return (rec.get("hbm_bytes_per_launch") if fresh else None), info


def host_cores():
    """Logical CPUs this process may use, all logical CPUs, and physical cores
    (distinct (core, socket) pairs in /proc/cpuinfo)."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    phys = set()
    try:
        core = sock = None
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("physical id"):
                sock = ln.split(":")[1].strip()
            elif ln.startswith("core id"):
                core = ln.split(":")[1].strip()
            elif not ln.strip() and core is not None:
                phys.add((sock, core))
                core = sock = None
        if core is not None:
            phys.add((sock, core))
    except OSError:
        pass
    return {"usable_logical": usable, "logical": os.cpu_count(), "physical": len(phys) or None}


def cpu_baseline(b, C, Ho, Hint, gpu_out, target_s: float = 12.0, gpu_rho=None):
    """Time the oracle's OpenMP H.v (reference applyHconf_omp restated, full
    maps with int64 offsets) over the whole basis on the host cores, repeated
    until about target_s seconds have run; check each result against the GPU.
    With gpu_rho = (rho1, rho2) from the GPU, also time the oracle's density
    matrices (OBrho_XM / TBrho_XX restated) once each on the same tables and
    check them."""
    import oracle
    L = oracle.lib()
    nc, M = b.nc, b.M
    t0 = time.time()
    IF = b.copy_range("IF", 0, nc * M)
    Map = b.copy_range("Map", 0, b.map_len)
    sot = b.copy_range("strideOT", 0, nc)
    stt = b.copy_range("strideTT", 0, nc)
    MapOT = b.copy_range("MapOT", 0, b.ot_len)
    MapTT = b.copy_range("MapTT", 0, b.tt_len)
    copy_s = time.time() - t0
    Hof, Hif = oracle.flat(Ho), oracle.flat(Hint)
    threads = oracle.num_threads()
    out = np.zeros(nc, np.complex128)
    times, rel = [], 0.0
    while sum(times) < target_s or len(times) < 2:
        t = time.perf_counter()
        L.orc_acth_range(N_, M, 0, nc, Map, MapOT, MapTT, sot, stt, IF, C, Hof, Hif, out)
        times.append(time.perf_counter() - t)
        rel = max(rel, float(np.linalg.norm(out - gpu_out) / np.linalg.norm(gpu_out)))
    t_hv = float(np.median(times))
    B = algorithmic_bytes(nc, M, b.ot_len, b.tt_len)
    dens = None
    if gpu_rho is not None:
        r1 = np.zeros(M * M, np.complex128)
        r2 = np.zeros(M ** 4, np.complex128)
        t = time.perf_counter()
        L.orc_rho1(N_, M, Map, IF, C, r1)
        t1 = time.perf_counter() - t
        t = time.perf_counter()
        L.orc_rho2(N_, M, Map, MapOT, MapTT, sot, stt, IF, C, r2)
        t2 = time.perf_counter() - t
        d1 = float(np.abs(oracle.unflat(r1, M, 2) - gpu_rho[0]).max() / np.abs(gpu_rho[0]).max())
        d2 = float(np.abs(oracle.unflat(r2, M, 4) - gpu_rho[1]).max() / np.abs(gpu_rho[1]).max())
        dens = {"rho1_ms": round(t1 * 1e3, 1), "rho2_ms": round(t2 * 1e3, 1), "cores": threads,
                "max_rel_diff_vs_gpu": max(d1, d2)}
    return {
        "value": round(B / t_hv / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "host_cores": host_cores(),
        "kind": "port",
        "sample": (f"oracle C/OpenMP restatement of applyHconf_omp (full maps, int64 offsets), {len(times)} "
                   f"full H.v over all {nc} rows, {sum(times):.1f} s in total, median {t_hv * 1e3:.1f} ms "
                   f"per H.v; results match the GPU to {rel:.1e} (relative 2-norm). The port computes "
                   f"each two-orbital chunk ordinal once per (row, pair) where the reference recounts it "
                   f"per term (hamiltonianMatrix.h:1927-1941), so it is faster than the reference's "
                   f"code; the reference build itself is timed at N=11 under reference_build"),
        "ms_per_hv": round(t_hv * 1e3, 2),
        "table_copy_s": round(copy_s, 2),
        "densities": dens,
    }


def reference_build_baseline(target_s: float = 20.0):
    """The reference's own applyHconf_omp (hamiltonianMatrix.h:1431, compiled
    unmodified into oracle/_ref) timed on the host cores at N=11, M=12
    (nc 705,432, Lieb-Liniger g=10): the largest M=12 basis its INT32 guards
    admit (configurationsMap.h:462-467).  Tables are the device-built ones
    copied to the host (int32 strides); next to it the oracle's restatement
    and the GPU on the same basis, all three results compared."""
    from oracle import ref
    if not ref.available():
        return None
    import mcpi_amd
    import oracle
    N, M = 11, 12
    b = mcpi_amd.Basis(N, M, "full")
    nc = b.nc
    IF = b.copy_range("IF", 0, nc * M)
    Map = b.copy_range("Map", 0, b.map_len)
    sot = b.copy_range("strideOT", 0, nc)
    stt = b.copy_range("strideTT", 0, nc)
    MapOT = b.copy_range("MapOT", 0, b.ot_len)
    MapTT = b.copy_range("MapTT", 0, b.tt_len)
    assert sot.max() < 2**31 and stt.max() < 2**31
    rng = np.random.default_rng(1112)
    C = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    C /= np.linalg.norm(C)
    Ho, Hint = mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, G_)
    dC, dO = mcpi_amd.DeviceArray.from_numpy(C), mcpi_amd.DeviceArray(nc)
    H = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=False, schedule=5)
    H.apply(dC, dO)
    gpu_ms = H.apply_timed(dC, dO, 20)
    gpu_out = dO.numpy()
    H.close()
    times, rel = [], 0.0
    while sum(times) < target_s and len(times) < 3 or not times:
        t = time.perf_counter()
        out = ref.acth_omp_tables(N, M, Map, MapOT, MapTT, sot, stt, IF, C, Ho, Hint)
        times.append(time.perf_counter() - t)
        rel = max(rel, float(np.linalg.norm(out - gpu_out) / np.linalg.norm(gpu_out)))
    t_ref = float(np.median(times))
    L = oracle.lib()
    out = np.zeros(nc, np.complex128)
    t = time.perf_counter()
    L.orc_acth_range(N, M, 0, nc, Map, MapOT, MapTT, sot.astype(np.int64), stt.astype(np.int64), IF, C,
                     oracle.flat(Ho), oracle.flat(Hint), out)
    t_port = time.perf_counter() - t
    rel_port = float(np.linalg.norm(out - gpu_out) / np.linalg.norm(gpu_out))
    Bn = algorithmic_bytes(nc, M, b.ot_len, b.tt_len)
    dC.free()
    dO.free()
    b.close()
    return {"workload": "H.v N=11, M=12, Lieb-Liniger g=10, full maps (every term)", "nc": nc,
            "bytes_per_hv": Bn, "threads": oracle.num_threads(), "cores": host_cores(),
            "reference_ms_per_hv": round(t_ref * 1e3, 1), "reference_calls": len(times),
            "reference_GB_s": round(Bn / t_ref / 1e9, 3),
            "port_ms_per_hv": round(t_port * 1e3, 1), "gpu_ms_per_hv": round(gpu_ms, 4),
            "gpu_vs_reference": round(t_ref * 1e3 / gpu_ms, 1),
            "rel_diff_reference_vs_gpu": rel, "rel_diff_port_vs_gpu": rel_port}


def reference_shaped_calls(reps: int = 20):
    """Per-call cost of the drop-in entry point applyHconf_omp
    (hamiltonianMatrix.h:1431 shape: host tables, host vectors) at N=10,
    M=12 (nc 352,716, Lieb-Liniger g=10), as a reference program calling it
    in a loop pays it: the first call uploads the tables, folds the rules and
    tunes the H.v schedule; later calls with the same tables and Ho/Hint
    reuse all of that and pay the vector copies plus one H.v.  Beside it the
    H.v alone on device vectors (mcpi_hip_applyH)."""
    import mcpi_amd
    import mcpi_amd.reference as R
    N, M = 10, 12
    t = time.perf_counter()
    r = R.RefBasis(N, M)
    build_s = time.perf_counter() - t
    rng = np.random.default_rng(1012)
    C = rng.standard_normal(r.nc) + 1j * rng.standard_normal(r.nc)
    C /= np.linalg.norm(C)
    Ho, Hint = mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, G_)
    t = time.perf_counter()
    first = r.applyHconf_omp(C, Ho, Hint)
    first_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    for _ in range(reps):
        out = r.applyHconf_omp(C, Ho, Hint)
    call_ms = (time.perf_counter() - t) / reps * 1e3
    b = mcpi_amd.Basis(N, M, "full")
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    dC, dO = mcpi_amd.DeviceArray.from_numpy(C), mcpi_amd.DeviceArray(b.nc)
    H.apply(dC, dO)
    dev_ms = H.apply_timed(dC, dO, reps)
    same = float(np.abs(out - dO.numpy()).max() / np.abs(out).max())
    R.release_cached()
    return {"workload": "applyHconf_omp, N=10, M=12, Lieb-Liniger g=10, host tables from the "
                        "reference-shaped builders", "nc": r.nc, "table_build_s": round(build_s, 2),
            "first_call_ms": round(first_ms, 2), "repeated_call_ms": round(call_ms, 3),
            "device_hv_ms": round(dev_ms, 4), "host_vector_copy_share": round(1 - dev_ms / call_ms, 3),
            "max_rel_diff_vs_device_api": same, "first_vs_repeated_max_diff": float(np.abs(first - out).max())}


def random_hamiltonian(M: int, seed: int, sigma: float = 0.1):
    """Random Hermitian Ho and a random bosonic Hint, symmetric under k <-> l
    and q <-> s and with Hint[k,l,q,s] = conj(Hint[q,s,k,l]); entries ~ N(0,
    sigma) (re and im) before the symmetrization.  (With this symmetry the
    reference's rules read the operator 1/2 sum Hint a*_k a*_l a_q a_s, so
    E0 = sum Ho rho1 + 1/2 sum Hint rho2 holds at the eigenvector.)"""
    rng = np.random.default_rng(seed)
    A = sigma * (rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M)))
    Ho = 0.5 * (A + A.conj().T)
    B = sigma * (rng.standard_normal((M,) * 4) + 1j * rng.standard_normal((M,) * 4))
    B = 0.5 * (B + B.transpose(1, 0, 2, 3))
    B = 0.5 * (B + B.transpose(0, 1, 3, 2))
    Hint = 0.5 * (B + np.conj(B.transpose(2, 3, 0, 1)))
    return Ho, Hint


def adjoint_n10_m12(lm: int = 300, seed: int = 1012):
    """The adjoint configuration: N=10, M=12 (nc = 352,716), random Hermitian
    Ho and bosonic random Hint (entries ~ N(0, 0.1)), full maps;
    Lanczos ground state (lm steps, full reorthogonalization) from a seeded
    random vector, then the VJP (dE0/dHo = rho1, dE0/dHint = rho2 / 2).  The
    Hamiltonian is dense, so the autotuner's H.v is the pair-space one."""
    import mcpi_amd
    N, M = 10, 12
    t = time.perf_counter()
    b = mcpi_amd.Basis(N, M, "full")
    build_s = time.perf_counter() - t
    Ho, Hint = random_hamiltonian(M, seed)
    rng = np.random.default_rng(seed + 1)
    C0 = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    dC0 = mcpi_amd.DeviceArray.from_numpy(C0 / np.linalg.norm(C0))
    H = mcpi_amd.Hamiltonian(b, Ho, Hint)
    dO = mcpi_amd.DeviceArray(b.nc)
    H.apply(dC0, dO)                            # schedule tuning
    ms_hv = H.apply_timed(dC0, dO, 20)
    H.ground(lm, dC0)                           # untimed first call (graph capture, allocations)
    mcpi_amd.synchronize()
    t = time.perf_counter()
    E, psi, steps = H.ground(lm, dC0)
    t_ground = time.perf_counter() - t
    t = time.perf_counter()
    g1, g2 = b.vjp(psi)
    t_vjp = time.perf_counter() - t
    v = psi.numpy()
    res = float(np.linalg.norm(H(v) - E * v))
    out = {"workload": "N=10, M=12, random Hermitian Ho / bosonic symmetric Hint (sigma 0.1), full maps",
           "nc": b.nc, "basis_build_s": round(build_s, 3), "schedule": H.schedule, "ms_per_hv": round(ms_hv, 4),
           "lm": lm, "steps": steps, "E0": E, "ground_s": round(t_ground, 4), "vjp_s": round(t_vjp, 4),
           "residual_norm": res,
           "energy_identity_residual": float(abs(np.real(np.sum(g1 * Ho) + np.sum(g2 * Hint)) - E))}
    H.close()
    b.close()
    return out


def batched_lanczos(n_inst: int = 64, lm: int = 120, seed: int = 64120, serial: bool = True,
                    cpu_sample: bool = True):
    """64 Lieb-Liniger instances, (N, M) drawn from the seed over (6,9),
    (8,11), (10,13), g log-uniform in [0.1, 100], lm = 120 Lanczos steps from
    seeded random normalized complex128 start vectors already on the device.
    Ground states per second end to end (start vectors on the device ->
    E0 on the host and the ground-state vector on the device), batched (one call, every instance advanced together)
    against per-instance GPU solves one after another.  Basis (map) builds
    and Hamiltonian set-up (coefficient tables) are outside both timings."""
    import mcpi_amd
    rng = np.random.default_rng(seed)
    shapes = [(6, 9), (8, 11), (10, 13)]
    picks = rng.integers(0, 3, n_inst)
    gs = 10 ** rng.uniform(-1, 2, n_inst)
    bases = {}
    t0 = time.time()
    for k in sorted(set(int(p) for p in picks)):
        N, M = shapes[k]
        bases[k] = mcpi_amd.Basis(N, M, "full")
    build_s = time.time() - t0
    hams, starts = [], []
    for x in range(n_inst):
        b = bases[int(picks[x])]
        M = b.M
        hams.append(mcpi_amd.Hamiltonian(b, mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, float(gs[x]))))
        v = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
        starts.append(mcpi_amd.DeviceArray.from_numpy(v / np.linalg.norm(v)))
    outs = [mcpi_amd.DeviceArray(h.basis.nc) for h in hams]   # ground-state vectors
    # untimed first calls (per-Hamiltonian H.v schedule tuning, allocator)
    mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    if serial:
        for x in range(n_inst):
            hams[x].ground(lm, starts[x], out=outs[x])
    mcpi_amd.synchronize()
    t = time.perf_counter()
    E_b, steps_b = mcpi_amd.lanczos_batched(hams, starts, lm, outs)
    t_batched = time.perf_counter() - t
    E_s = np.full(n_inst, np.nan)
    t_serial = float("nan")
    t_each = np.zeros(n_inst)
    if serial:
        t = time.perf_counter()
        for x in range(n_inst):
            t1 = time.perf_counter()
            E_s[x] = hams[x].ground(lm, starts[x], out=outs[x])[0]
            t_each[x] = time.perf_counter() - t1
        t_serial = time.perf_counter() - t
    # the same split by basis: each basis's instances batched alone against
    # their share of the one-by-one solves
    per_shape = {}
    if serial:
        for k in sorted(set(int(p) for p in picks)):
            xs = [x for x in range(n_inst) if picks[x] == k]
            mcpi_amd.synchronize()
            t = time.perf_counter()
            mcpi_amd.lanczos_batched([hams[x] for x in xs], [starts[x] for x in xs], lm, [outs[x] for x in xs])
            tb = time.perf_counter() - t
            ts = float(t_each[xs].sum())
            per_shape[f"({shapes[k][0]},{shapes[k][1]})"] = {"instances": len(xs), "s_batched": round(tb, 4),
                                                            "s_per_instance_total": round(ts, 4),
                                                            "speedup": round(ts / tb, 3)}
    dE = float(np.abs(E_b - E_s).max()) if serial else None
    counts = {f"({shapes[k][0]},{shapes[k][1]})": int((picks == k).sum()) for k in range(3)}
    cpu = None
    if serial and cpu_sample:
        # the oracle's C/OpenMP restatement of lanczos + ground
        # (groundStateLieb.c:217-471) on the host cores, one instance of each
        # of the two smaller bases (the (10,13) tables take minutes to build on
        # the host); compared with the same instances' GPU solves
        import oracle
        cpu = {"cores": oracle.num_threads(), "kind": "port", "per_basis": {}}
        for k in (0, 1):
            xs = [x for x in range(n_inst) if picks[x] == k]
            if not xs:
                continue
            x = xs[0]
            N, M = shapes[k]
            o = oracle.Basis(N, M)
            v0 = starts[x].numpy()
            t = time.perf_counter()
            E_cpu = o.ground(lm, v0, mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, float(gs[x])))[0]
            t_cpu = time.perf_counter() - t
            mcpi_amd.synchronize()
            t = time.perf_counter()
            E_gpu = hams[x].ground(lm, starts[x], out=outs[x])[0]
            t_gpu = time.perf_counter() - t
            cpu["per_basis"][f"({N},{M})"] = {"cpu_s_per_gs": round(t_cpu, 4), "gpu_s_per_gs": round(t_gpu, 5),
                                            "abs_dE0_cpu_vs_gpu": float(abs(E_cpu - E_gpu))}
    # Bytes the two classical Gram-Schmidt passes must move in any solver
    # that reorthogonalizes like the reference (groundStateLieb.c:316-332):
    # step i reads the i+2 basis vectors once for the dots and once for the
    # update.  Batched and per-instance solves both pay them, so they bound
    # the batched speed-up from above.
    ncs = np.array([hams[x].basis.nc for x in range(n_inst)], dtype=np.float64)
    gs_bytes = float(32.0 * ncs.sum() * (lm * (lm + 1) / 2 - 1))
    gs_floor_s = gs_bytes / (HBM_PEAK_GBS * 1e9)
    return {
        "gram_schmidt_bytes": gs_bytes,
        "gram_schmidt_floor_s_at_hbm_peak": round(gs_floor_s, 4),
        "speedup_ceiling_at_hbm_peak": (round(t_serial / gs_floor_s, 2) if serial else None),
        "workload": (f"{n_inst} Lieb-Liniger instances, (N,M) from the seed over (6,9)/(8,11)/(10,13) "
                     f"{counts}, g log-uniform [0.1,100], lm={lm}, full maps, complex128"),
        "gs_per_s_batched": round(n_inst / t_batched, 3),
        "gs_per_s_per_instance": round(n_inst / t_serial, 3) if serial else None,
        "speedup": round(t_serial / t_batched, 3) if serial else None,
        "s_batched": round(t_batched, 4),
        "s_per_instance_total": round(t_serial, 4),
        "max_abs_dE0_batched_vs_per_instance": dE if serial else None,
        "per_basis": per_shape or None,
        "all_steps_full": bool((steps_b == lm).all()),
        "basis_build_s_excluded": round(build_s, 3),
        "cpu_oracle_sample": cpu,
    }


def large_basis(reps: int = 20, lm: int = 200, quick: bool = False):
    """N=16 bosons, M=12 orbitals (nc = 13,037,895), Lieb-Liniger g=50, only
    the single-jump Map stored (7.5 GB; the double-jump maps would be 200 GB)
    and double jumps ranked on the fly.  H.v timed with every term evaluated
    and with zero-skipping; then a 200-step Lanczos ground state from the
    reference's start vector (groundStateLieb.c:578-586) and the adjoint
    (rho1, rho2 / 2) at the ground vector."""
    import mcpi_amd
    N, M, g = 16, 12, 50.0
    t0 = time.time()
    b = mcpi_amd.Basis(N, M, "single")
    build_s = time.time() - t0
    nc = b.nc
    B = 16 * nc + 16 * nc + 4 * M * nc + 4 * M * M * nc
    rng = np.random.default_rng(1612)
    C = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    dC = mcpi_amd.DeviceArray.from_numpy(C / np.linalg.norm(C))
    dO = mcpi_amd.DeviceArray(nc)
    Ho, Hint = mcpi_amd.ll_ho(M), mcpi_amd.ll_hint(M, g)
    out = {"workload": "N=16, M=12, Lieb-Liniger g=50, Map stored, double jumps ranked",
           "nc": nc, "bytes_per_hv": B, "basis_build_s": round(build_s, 3)}
    # every term reading Map (schedule 2, the byte model's path), every term
    # through the pair space (schedule 6: no map read, so no byte model), and
    # the autotuned zero-skipping path
    legs = (("skip_zero", dict(skip_zero=True)),)
    if not quick:
        legs = (("every_term", dict(skip_zero=False, schedule=2)),
                ("every_term_pair_space", dict(skip_zero=False, schedule=6))) + legs
    for label, kw in legs:
        H = mcpi_amd.Hamiltonian(b, Ho, Hint, **kw)
        H.apply(dC, dO)                      # schedule tuning
        ms = H.apply_timed(dC, dO, reps)
        if label == "every_term_pair_space":
            out[label] = {"ms_per_hv": round(ms, 4), "schedule": 6,
                          "pair_product_flops": pair_product_flops(N, M)}
        else:
            out[label] = {"ms_per_hv": round(ms, 4), "gb_s_model": round(B / ms / 1e6, 1),
                          "hbm_frac_model": round(B / ms / 1e6 / HBM_PEAK_GBS, 4), "schedule": H.schedule,
                          "pattern_table_bytes": b.pattern_bytes()}
        H.close()
    # the same H.v on the sector-ordered basis (rows sorted by total
    # momentum, which the Lieb-Liniger Hamiltonian conserves): the kernel on
    # row-order vectors (what every Lanczos step runs), and the reference-
    # order entry point, which moves both vectors through the row order
    t0 = time.time()
    bs = mcpi_amd.Basis(N, M, "single", charge=mcpi_amd.conserved_charge(Ho, Hint))
    sbuild_s = time.time() - t0
    Hs = mcpi_amd.Hamiltonian(bs, Ho, Hint)
    dR, dRo = mcpi_amd.DeviceArray(nc), mcpi_amd.DeviceArray(nc)
    bs.to_rows(dC, dR)
    Hs.apply_rows(dR, dRo)                   # pattern tables built here
    ms_rows = Hs.apply_timed(dR, dRo, reps)
    mcpi_amd.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        Hs.apply(dC, dO)
    mcpi_amd.synchronize()
    ms_ref = (time.perf_counter() - t) / reps * 1e3
    ref_out = None
    if not quick:
        Hp = mcpi_amd.Hamiltonian(b, Ho, Hint)
        dP = mcpi_amd.DeviceArray(nc)
        Hp.apply(dC, dP)
        mcpi_amd.synchronize()
        ref_out = dP.numpy()
        Hp.close()
        del dP
    out["sector_order"] = {
        "nsectors": bs.nsectors, "largest_sector_rows": bs.largest_sector, "basis_build_s": round(sbuild_s, 3),
        "ms_per_hv": round(ms_rows, 4), "gb_s_model": round(B / ms_rows / 1e6, 1),
        "hbm_frac_model": round(B / ms_rows / 1e6 / HBM_PEAK_GBS, 4), "schedule": Hs.schedule,
        "pattern_table_bytes": bs.pattern_bytes(),
        "ms_per_hv_reference_order_api": round(ms_ref, 4),
        "max_rel_diff_vs_plain_basis": (float(np.abs(dO.numpy() - ref_out).max() / np.abs(ref_out).max())
                                        if ref_out is not None else None)}
    del dR, dRo
    if quick:
        return out
    dS = mcpi_amd.DeviceArray.from_numpy(mcpi_amd.start_vector(nc))
    for label, basis, ham in (("ground", b, mcpi_amd.Hamiltonian(b, Ho, Hint)), ("ground_sector_order", bs, Hs)):
        t = time.perf_counter()
        E0, psi, steps = ham.ground(lm, dS)
        t_ground = time.perf_counter() - t
        t = time.perf_counter()
        g1, g2 = basis.vjp(psi)
        t_vjp = time.perf_counter() - t
        out[label] = {"lm": lm, "steps": steps, "E0_per_N": E0 / N, "s": round(t_ground, 3),
                      "vjp_s": round(t_vjp, 3),
                      "energy_identity_residual": float(abs(np.real(np.sum(g1 * Ho) + np.sum(g2 * Hint)) - E0))}
        del psi
    out["ground_sector_order"]["abs_dE0_vs_plain"] = abs(out["ground_sector_order"]["E0_per_N"]
                                                         - out["ground"]["E0_per_N"]) * N
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--profile", action="store_true",
                    help="time only the generic H.v kernel (no CPU baseline, no zero-skipping run)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--schedule", type=int, default=5,
                    help="H.v schedule timed as the headline (5: every term, every stored map "
                         "entry read, double jumps compacted into full wavefronts)")
    ap.add_argument("--skip-zero", action="store_true",
                    help="time the zero-skipping path instead (not the headline)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched Lanczos comparison")
    ap.add_argument("--no-large", action="store_true", help="skip the N=16, M=12 measurements")
    ap.add_argument("--large-only", action="store_true", help="only run the N=16, M=12 measurements")
    ap.add_argument("--large-quick", action="store_true",
                    help="with --large-only: only the zero-skipping H.v (profiling runs)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched Lanczos call, without the per-instance solves "
                         "(profiling runs)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the N=10, M=12 ground state and adjoint of a dense random Hamiltonian")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device-to-device copies with the library's copy kernel "
                         "(reference reading for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (H.v would shard by "
              "output rows with an all-gather of C per step; not built). Run with --gpus 1.",
              file=sys.stderr)
        return 2

    import mcpi_amd

    if args.copy_calibration:
        n = 1 << 30
        src, dst = mcpi_amd.DeviceArray(n // 16), mcpi_amd.DeviceArray(n // 16)
        times = []
        for _ in range(3):
            mcpi_amd.synchronize()
            t = time.perf_counter()
            mcpi_amd._check(mcpi_amd.lib().mcpi_hip_memcpy_d2d(dst.ptr, src.ptr, n), "copy")
            mcpi_amd.synchronize()
            times.append(time.perf_counter() - t)
        print(json.dumps({"copy_calibration": {"bytes_per_copy": n, "copies": 3,
                                               "read_plus_write_GB_s_best": round(2 * n / min(times) / 1e9, 1)}}))
        return 0

    if args.large_only:
        print(json.dumps({"large_basis": large_basis(quick=args.large_quick)}))
        return 0
    if args.batched_only:   # the batched call alone (no per-instance solves)
        print(json.dumps({"batched_lanczos": batched_lanczos(serial=False)}))
        return 0
    if args.adjoint_only:
        print(json.dumps({"adjoint_n10_m12": adjoint_n10_m12()}))
        return 0

    name, cus, mem = mcpi_amd.device_info()
    t0 = time.time()
    b = mcpi_amd.Basis(N_, M_, "full")
    build_s = time.time() - t0
    rng = np.random.default_rng(20240612)
    C = rng.standard_normal(b.nc) + 1j * rng.standard_normal(b.nc)
    C /= np.linalg.norm(C)
    Ho, Hint = mcpi_amd.ll_ho(M_), mcpi_amd.ll_hint(M_, G_)
    dC = mcpi_amd.DeviceArray.from_numpy(C)
    dO = mcpi_amd.DeviceArray(b.nc)
    B = algorithmic_bytes(b.nc, M_, b.ot_len, b.tt_len)

    # the timed path: every term evaluated, every map entry read (schedule 5:
    # kernel specialised on M; the double jumps of each block of 256 rows are
    # compacted into full wavefronts before their packed targets are read)
    H = mcpi_amd.Hamiltonian(b, Ho, Hint, skip_zero=args.skip_zero, schedule=args.schedule)
    for _ in range(args.warmup):
        H.apply(dC, dO)
    mcpi_amd.synchronize()
    t0 = time.perf_counter()
    each = H.apply_timed_each(dC, dO, args.steps)   # exactly K launches, events on their stream
    mcpi_amd.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(each.sum() / args.steps)
    ms_step = wall / args.steps * 1e3
    value = B * args.steps / wall / 1e9
    achieved = B / (kern_ms * 1e-3) / 1e9
    gpu_out = dO.numpy()

    # other schedules of the same H.v (same result), for reference:
    # generic thread-per-row kernel, the thread-per-row specialised kernel
    # reading the same tables, double jumps ranked instead of read, and the
    # autotuned Lieb-Liniger-aware path that skips exact-zero terms
    others = {}
    agree = 0.0
    if not args.profile:
        reps = max(10, args.steps // 10)
        for label, kw in (("generic_rows_kernel", dict(skip_zero=False, schedule=1)),
                          ("packed_rows_per_lane", dict(skip_zero=False, schedule=2)),
                          ("ranked_doubles", dict(skip_zero=False, schedule=3)),
                          ("pair_space_every_term", dict(skip_zero=False, schedule=6)),
                          ("pair_space_factorized", dict(skip_zero=True, schedule=7)),
                          ("skip_zero_autotuned", dict(skip_zero=True, schedule=0))):
            Hx = mcpi_amd.Hamiltonian(b, Ho, Hint, **kw)
            for _ in range(2):
                Hx.apply(dC, dO)
            others[label] = {"ms_per_hv": round(Hx.apply_timed(dC, dO, reps), 4), "schedule": Hx.schedule}
            if Hx.schedule == 4:
                others[label]["pattern_table_bytes"] = b.pattern_bytes()
            if label == "pair_space_every_term":
                # no map is read: the pair product's FP64 flops (Gauss form,
                # three real products per complex one) are its model
                others[label]["pair_product_flops"] = pair_product_flops(N_, M_)
            agree = max(agree, float(np.linalg.norm(dO.numpy() - gpu_out) / np.linalg.norm(gpu_out)))

    # one- and two-body densities of the same vector on the same basis
    # (per call from the host: the Gram kernels, the block sums, the
    # expansion into the full matrices and their copy to the host)
    dens = gpu_rho = None
    if not args.profile:
        gpu_rho = (b.rho1(dC), b.rho2(dC))
        mcpi_amd.synchronize()
        dens = {}
        for label, fn in (("rho1_ms", b.rho1), ("rho2_ms", b.rho2)):
            t = time.perf_counter()
            for _ in range(20):
                fn(dC)
            mcpi_amd.synchronize()
            dens[label] = round((time.perf_counter() - t) / 20 * 1e3, 4)
        dens.update(density_model(N_, M_, dens["rho2_ms"]))

    refcalls = None
    if not args.profile:
        refcalls = reference_shaped_calls()
    batched = large = None
    if not (args.no_batched or args.profile):
        del dO
        batched = batched_lanczos(cpu_sample=not args.no_cpu)
    adjoint = None
    if not args.profile:
        adjoint = adjoint_n10_m12()

    cpu = None
    if not (args.no_cpu or args.profile):
        cpu = cpu_baseline(b, C, Ho, Hint, gpu_out, args.cpu_seconds, gpu_rho)
        cpu["reference_build"] = reference_build_baseline()
    if not (args.no_large or args.profile):
        large = large_basis()

    traffic, traffic_info = traffic_from_profiles()
    line = {
        "metric": "H·v GB/s (% HBM peak)",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded random normalized complex128 vector; analytic Lieb-Liniger Ho/Hint)",
        "config": {
            "workload": "H.v, N=12 bosons, M=12 orbitals, Lieb-Liniger g=10, full jump maps on device",
            "N": N_, "M": M_, "nc": b.nc, "g": G_, "maps": "full",
            "map_ints": {"Map": b.map_len, "MapOT": b.ot_len, "MapTT": b.tt_len},
            "bytes_per_hv": B,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "extra": {
            "device": name, "compute_units": cus, "traffic_record": traffic_info,
            "hbm_fraction_wall": round(value / HBM_PEAK_GBS, 4),
            "kernel_ms_per_hv": round(kern_ms, 4),
            "kernel_ms_p10_p50_p90": [round(float(np.percentile(each, q)), 4) for q in (10, 50, 90)],
            "kernel_ms_total": round(float(each.sum()), 3),
            "basis_build_s": round(build_s, 3),
            "other_schedules": others,
            "densities_n12_m12": dens,
            "reference_shaped_applyHconf_omp_n10_m12": refcalls,
            "batched_lanczos": batched,
            "adjoint_n10_m12": adjoint,
            "large_basis_n16_m12": large,
            "max_rel_diff_between_schedules": agree,
            "speedup_vs_cpu": (round(cpu["ms_per_hv"] / ms_step, 1) if cpu else None),
        },
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
