"""CPU oracle for the mcpi Fock-space hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker.  The product
(``mcpi_amd``) never imports it.

It wraps ``oracle/libmcpi_oracle.so`` (the C restatement in
``mcpi_oracle.c``; every function there cites the reference lines it follows)
with numpy conventions:

* ``Ho`` is an ``(M, M)`` complex array, ``Ho[k, l]`` = coefficient of
  a*_k a_l; ``Hint`` is ``(M, M, M, M)``, ``Hint[k, s, q, l]`` = coefficient
  of a*_k a*_s a_q a_l.  Both are passed to C in Fortran order, which is the
  flat layout ``k + M*l`` / ``k + s*M + q*M^2 + l*M^3``.
* ``rho1[k, l] = <a*_k a_l>``, ``rho2[k, s, q, l] = <a*_k a*_s a_q a_l>``.

``oracle.ref`` (module ``oracle/ref.py``) wraps the reference's own code
compiled by ``oracle/Makefile`` into ``oracle/_ref/`` when it is present.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcpi_oracle.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_c128p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64 = ct.c_int64
_int = ct.c_int

_lib = None


def build(force: bool = False) -> None:
    """Compile the C restatement (and, where the reference sources are
    present, the reference build under ``oracle/_ref``)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ct.CDLL(LIB_PATH)
    sig = {
        "orc_nc": (_i64, [_int, _int]),
        "orc_ncmat": (None, [_int, _int, _i32p]),
        "orc_rank": (_i64, [_int, _int, _i32p, _i32p]),
        "orc_unrank": (None, [_i64, _int, _int, _i32p]),
        "orc_setup_focks": (None, [_int, _int, _i32p]),
        "orc_one_one_map": (None, [_int, _int, _i32p, _i32p, _i32p]),
        "orc_count_ot_chunks": (_i64, [_int, _int, _i32p]),
        "orc_count_tt_chunks": (_i64, [_int, _int, _i32p]),
        "orc_stride_ot": (None, [_int, _int, _i32p, _i64p]),
        "orc_stride_tt": (None, [_int, _int, _i32p, _i64p]),
        "orc_one_two_map": (None, [_int, _int, _i32p, _i32p, _i64p, _i32p]),
        "orc_two_two_map": (None, [_int, _int, _i32p, _i32p, _i64p, _i32p]),
        "orc_acth": (None, [_int, _int, _i32p, _i32p, _i32p, _i64p, _i64p, _i32p,
                            _c128p, _c128p, _c128p, _c128p]),
        "orc_acth_range": (None, [_int, _int, _i64, _i64, _i32p, _i32p, _i32p, _i64p, _i64p,
                                  _i32p, _c128p, _c128p, _c128p, _c128p]),
        "orc_acth_hash_rows": (None, [_int, _int, _i32p, _i64, _i64p, _c128p,
                                      _c128p, _c128p, _c128p]),
        "orc_rho1": (None, [_int, _int, _i32p, _i32p, _c128p, _c128p]),
        "orc_rho2": (None, [_int, _int, _i32p, _i32p, _i32p, _i64p, _i64p, _i32p,
                            _c128p, _c128p]),
        "orc_lanczos": (_int, [_int, _int, _int, _i32p, _i32p, _i32p, _i64p, _i64p,
                               _i32p, _c128p, _c128p, _c128p, _c128p, _c128p]),
        "orc_ground": (ct.c_double, [_int, _int, _int, _i32p, _i32p, _i32p, _i64p,
                                     _i64p, _i32p, _c128p, _c128p, _c128p,
                                     ct.POINTER(_int)]),
        "orc_tridiag_min": (ct.c_double, [_int, _f64p, _f64p, _f64p]),
        "orc_ll_ho": (None, [_int, ct.c_double, _c128p]),
        "orc_ll_hint": (None, [_int, ct.c_double, _c128p]),
        "orc_start_vector": (None, [_i64, _c128p]),
        "orc_perf_ho": (None, [_int, _c128p]),
        "orc_perf_hint": (None, [_int, _c128p]),
        "orc_num_threads": (_int, []),
        "orc_set_num_threads": (None, [_int]),
        "orc_wtime": (ct.c_double, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# --------------------------------------------------------------------- L1 --

def nc(N: int, M: int) -> int:
    return int(lib().orc_nc(N, M))


def ncmat(N: int, M: int) -> np.ndarray:
    out = np.zeros((M + 1) * (N + 1), np.int32)
    lib().orc_ncmat(N, M, out)
    return out


def rank(N: int, M: int, NCmat: np.ndarray, v) -> int:
    return int(lib().orc_rank(N, M, NCmat, np.ascontiguousarray(v, np.int32)))


def unrank(k: int, N: int, M: int) -> np.ndarray:
    v = np.zeros(M, np.int32)
    lib().orc_unrank(k, N, M, v)
    return v


def setup_focks(N: int, M: int) -> np.ndarray:
    IF = np.zeros(nc(N, M) * M, np.int32)
    lib().orc_setup_focks(N, M, IF)
    return IF


class Basis:
    """All hashing tables and jump maps of (N, M), built by the oracle.

    ``maps`` = "full" (Map, MapOT, MapTT), "single" (Map only) or "none".
    """

    def __init__(self, N: int, M: int, maps: str = "full"):
        L = lib()
        self.N, self.M = N, M
        self.nc = nc(N, M)
        self.NCmat = ncmat(N, M)
        self.IF = setup_focks(N, M)
        self.Map = self.MapOT = self.MapTT = None
        self.strideOT = self.strideTT = None
        if maps in ("full", "single"):
            self.Map = np.zeros(self.nc * M * M, np.int32)
            L.orc_one_one_map(N, M, self.NCmat, self.IF, self.Map)
        if maps == "full":
            self.n_ot = int(L.orc_count_ot_chunks(N, M, self.IF))
            self.n_tt = int(L.orc_count_tt_chunks(N, M, self.IF))
            self.strideOT = np.zeros(self.nc, np.int64)
            self.strideTT = np.zeros(self.nc, np.int64)
            L.orc_stride_ot(N, M, self.IF, self.strideOT)
            L.orc_stride_tt(N, M, self.IF, self.strideTT)
            self.MapOT = np.zeros(max(self.n_ot, 1) * M * M, np.int32)
            self.MapTT = np.zeros(max(self.n_tt, 1) * M * M, np.int32)
            L.orc_one_two_map(N, M, self.NCmat, self.IF, self.strideOT, self.MapOT)
            L.orc_two_two_map(N, M, self.NCmat, self.IF, self.strideTT, self.MapTT)
            self.MapOT = self.MapOT[: self.n_ot * M * M]
            self.MapTT = self.MapTT[: self.n_tt * M * M]

    # C wants a non-empty buffer even when a map has no chunks
    def _nz(self, a):
        return a if a.size else np.zeros(1, a.dtype)

    def acth(self, C, Ho, Hint) -> np.ndarray:
        out = np.zeros(self.nc, np.complex128)
        lib().orc_acth(self.N, self.M, self.Map, self._nz(self.MapOT), self._nz(self.MapTT),
                       self.strideOT, self.strideTT, self.IF,
                       np.ascontiguousarray(C, np.complex128), flat(Ho), flat(Hint), out)
        return out

    def rho1(self, C) -> np.ndarray:
        out = np.zeros(self.M * self.M, np.complex128)
        lib().orc_rho1(self.N, self.M, self.Map, self.IF,
                       np.ascontiguousarray(C, np.complex128), out)
        return unflat(out, self.M, 2)

    def rho2(self, C) -> np.ndarray:
        out = np.zeros(self.M ** 4, np.complex128)
        lib().orc_rho2(self.N, self.M, self.Map, self._nz(self.MapOT), self._nz(self.MapTT),
                       self.strideOT, self.strideTT, self.IF,
                       np.ascontiguousarray(C, np.complex128), out)
        return unflat(out, self.M, 4)

    def lanczos(self, lm: int, C0, Ho, Hint):
        """Returns (iterations, diag, offdiag, lvec[lm, nc])."""
        diag = np.zeros(lm, np.complex128)
        off = np.zeros(lm, np.complex128)
        lvec = np.zeros((lm, self.nc), np.complex128)
        lvec[0] = C0
        n = lib().orc_lanczos(lm, self.N, self.M, self.Map, self._nz(self.MapOT),
                              self._nz(self.MapTT), self.strideOT, self.strideTT, self.IF,
                              flat(Ho), flat(Hint), diag, off, lvec.reshape(-1))
        return int(n), diag, off, lvec

    def ground(self, lm: int, C0, Ho, Hint):
        """Returns (E0, ground vector, iterations)."""
        C = np.ascontiguousarray(C0, np.complex128).copy()
        it = _int(0)
        E0 = lib().orc_ground(lm, self.N, self.M, self.Map, self._nz(self.MapOT),
                              self._nz(self.MapTT), self.strideOT, self.strideTT, self.IF,
                              flat(Ho), flat(Hint), C, ct.byref(it))
        return float(E0), C, int(it.value)


def acth_hash_rows(N: int, M: int, rows, C, Ho, Hint) -> np.ndarray:
    """H.v restricted to ``rows``, every jump ranked on the fly (no maps)."""
    rows = np.ascontiguousarray(rows, np.int64)
    out = np.zeros(rows.size, np.complex128)
    lib().orc_acth_hash_rows(N, M, ncmat(N, M), rows.size, rows,
                             np.ascontiguousarray(C, np.complex128), flat(Ho), flat(Hint), out)
    return out


def tridiag_min(d, e):
    d = np.ascontiguousarray(d, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    if e.size < d.size:
        e = np.concatenate([e, np.zeros(d.size - e.size)])
    z = np.zeros(d.size, np.float64)
    lam = lib().orc_tridiag_min(d.size, d, e, z)
    return float(lam), z


# ------------------------------------------------------------------ inputs --

def flat(a) -> np.ndarray:
    """(M,)*r complex array -> flat first-index-fastest layout."""
    return np.ascontiguousarray(np.asarray(a, np.complex128).reshape(-1, order="F")
                                if np.ndim(a) > 1 else np.asarray(a, np.complex128))


def unflat(a, M: int, r: int) -> np.ndarray:
    return np.asarray(a).reshape((M,) * r, order="F")


def ll_ho(M: int, L: float = 1.0) -> np.ndarray:
    out = np.zeros(M * M, np.complex128)
    lib().orc_ll_ho(M, L, out)
    return unflat(out, M, 2)


def ll_hint(M: int, g: float) -> np.ndarray:
    out = np.zeros(M ** 4, np.complex128)
    lib().orc_ll_hint(M, g, out)
    return unflat(out, M, 4)


def start_vector(n: int) -> np.ndarray:
    out = np.zeros(n, np.complex128)
    lib().orc_start_vector(n, out)
    return out


def perf_ho(M: int) -> np.ndarray:
    out = np.zeros(M * M, np.complex128)
    lib().orc_perf_ho(M, out)
    return unflat(out, M, 2)


def perf_hint(M: int) -> np.ndarray:
    out = np.zeros(M ** 4, np.complex128)
    lib().orc_perf_hint(M, out)
    return unflat(out, M, 4)


def symmetrize_hint(H: np.ndarray) -> np.ndarray:
    """Average over the 8-element group generated by the creator swap
    (k<->s), the annihilator swap (q<->l) and the conjugate pair exchange
    Hint[k,s,q,l] -> conj(Hint[q,l,k,s]): the full bosonic symmetry the
    reference's folded rules assume (hamiltonianMatrix.h:1543-1547)."""
    a = 0.5 * (H + H.transpose(1, 0, 2, 3))
    a = 0.5 * (a + a.transpose(0, 1, 3, 2))
    a = 0.5 * (a + np.conj(a.transpose(2, 3, 0, 1)))
    return a


def random_inputs(N: int, M: int, seed: int, sigma: float = 1.0):
    """Seeded random normalized C, Hermitian Ho and fully symmetric Hint.

    Entries ~ N(0, sigma) (re and im independent) before symmetrization."""
    rng = np.random.default_rng(seed)
    n = nc(N, M)
    C = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    C /= np.linalg.norm(C)
    A = sigma * (rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M)))
    Ho = 0.5 * (A + A.conj().T)
    B = sigma * (rng.standard_normal((M,) * 4) + 1j * rng.standard_normal((M,) * 4))
    Hint = symmetrize_hint(B)
    return C, Ho, Hint


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(n)
