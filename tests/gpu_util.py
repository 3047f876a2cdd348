"""Helpers shared by the GPU parity tests."""
import sys

import numpy as np


def note(msg: str) -> None:
    """Name the case before a GPU step (flushed, so it survives a crash)."""
    sys.stdout.write(f"[gpu] {msg}\n")
    sys.stdout.flush()


def assert_close(x, ref, tol=1e-12, what=""):
    """|x - ref| / |ref| <= tol (normwise) and max |x - ref| <= tol max |ref|."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    nr = np.linalg.norm(ref)
    scale = max(np.abs(ref).max() if ref.size else 0.0, 1e-300)
    err = np.linalg.norm(x - ref)
    assert err <= tol * max(nr, 1e-300) or (nr == 0 and err == 0), f"{what}: normwise {err / max(nr, 1e-300):.3e}"
    assert np.abs(x - ref).max() <= tol * scale or nr == 0, f"{what}: elementwise {np.abs(x - ref).max() / scale:.3e}"


def oracle_on_device_tables(b):
    """An oracle.Basis holding the tables the GPU built for basis ``b``
    (copied to the host), so that the oracle's H.v, densities and Lanczos
    run on exactly the device's tables at sizes where the oracle's own
    builders would take minutes.  The device tables themselves are checked
    entry by entry against the oracle's builders at smaller sizes."""
    import oracle

    o = oracle.Basis.__new__(oracle.Basis)
    o.N, o.M, o.nc = b.N, b.M, b.nc
    o.NCmat = b.ncmat()
    o.IF = b.copy_range("IF", 0, b.nc * b.M)
    o.Map = b.copy_range("Map", 0, b.map_len) if b.map_len else None
    o.MapOT = o.MapTT = o.strideOT = o.strideTT = None
    if b.tt_len:
        o.strideOT = b.copy_range("strideOT", 0, b.nc)
        o.strideTT = b.copy_range("strideTT", 0, b.nc)
        o.MapOT = b.copy_range("MapOT", 0, b.ot_len)
        o.MapTT = b.copy_range("MapTT", 0, b.tt_len)
        o.n_ot, o.n_tt = b.ot_len // (b.M * b.M), b.tt_len // (b.M * b.M)
    return o
