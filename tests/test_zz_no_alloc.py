"""The allocation-free adjoint (mcpi_hip_vjp_noalloc) against mcpi_hip_vjp.

Called through the C ABI: two calls in a row on one GPU, the second on a
smaller basis, each compared with the allocating entry point's host outputs.
"""
import ctypes as ct

import numpy as np
import pytest

import mcpi_amd
import oracle
from tests.gpu_util import assert_close, note

pytestmark = pytest.mark.gpu


def _noalloc_fn():
    L = mcpi_amd.lib()
    f = L.mcpi_hip_vjp_noalloc
    f.restype = ct.c_int
    f.argtypes = [ct.c_void_p] * 5
    return f


def _run(N, M, maps, seed):
    note(f"vjp_noalloc N={N} M={M} maps={maps} seed={seed}")
    b = mcpi_amd.Basis(N, M, maps)
    C, _, _ = oracle.random_inputs(N, M, seed=seed)
    dC = mcpi_amd.DeviceArray.from_numpy(C)
    g1_ref, g2_ref = b.vjp(dC)
    d1 = mcpi_amd.DeviceArray(M * M)
    d2 = mcpi_amd.DeviceArray(M ** 4)
    # poison the outputs: every entry must be written
    d1.upload(np.full(M * M, np.nan + 1j * np.nan))
    d2.upload(np.full(M ** 4, np.nan + 1j * np.nan))
    mcpi_amd._check(_noalloc_fn()(b.handle, dC.ptr, d1.ptr, d2.ptr, None), "vjp_noalloc")
    mcpi_amd.synchronize()
    g1 = mcpi_amd.unflat(d1.numpy(), M, 2)
    g2 = mcpi_amd.unflat(d2.numpy(), M, 4)
    assert np.isfinite(g1).all() and np.isfinite(g2).all()
    assert_close(g1, g1_ref, 1e-12, "dE/dHo")
    assert_close(g2, g2_ref, 1e-12, "dE/dHint")
    # and the oracle's densities (rho1, rho2 / 2)
    o = oracle.Basis(N, M)
    assert_close(g1, o.rho1(C), 1e-12, "rho1 vs oracle")
    assert_close(g2, 0.5 * o.rho2(C), 1e-12, "rho2/2 vs oracle")
    for x in (d1, d2, dC):
        x.free()
    b.close()


def test_vjp_noalloc_two_calls():
    """N=8, M=11 then N=4, M=6 (smaller basis and fewer orbitals): the second
    call reuses the scratch array the first one filled."""
    _run(8, 11, "full", 5)
    _run(4, 6, "single", 6)


def test_vjp_noalloc_many_orbitals():
    """M = 25 (325 pair columns: four Gram tiles a side, ten tile pairs) with
    N = 2, then M = 32, the library's largest, with N = 2: the 16 MiB array
    holds the Gram tiles and a reduced number of block partials."""
    _run(2, 25, "none", 7)
    _run(2, 32, "none", 8)
