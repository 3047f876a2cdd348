"""ctypes access to the reference's own C code -- TEST INFRASTRUCTURE ONLY.

``oracle/Makefile`` target ``ref`` compiles the reference headers
(configurationsMap.h, onebodyMatrix.h, twobodyMatrix.h, hamiltonianMatrix.h)
unmodified into ``oracle/_ref/libmcpi_ref.so``.  This module calls those
functions with their own signatures (int32 offsets, ``double complex **``
row-pointer matrices) and converts to/from the oracle's numpy conventions.

The reference functions ``exit()`` the process on INT32 overflow guards
(configurationsMap.h:162-168, 372-377, 462-467, 652-657), so callers keep to
sizes the guards admit.  ``available()`` is False where the build is absent
(e.g. on a machine without the reference sources).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import HERE, flat, unflat

REF_PATH = os.path.join(HERE, "_ref", "libmcpi_ref.so")

_lib = None
_libc = ct.CDLL(None)
_libc.free.argtypes = [ct.c_void_p]
_libc.free.restype = None

_ip = ct.POINTER(ct.c_int)
_vp = ct.c_void_p


def available() -> bool:
    return os.path.exists(REF_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ct.CDLL(REF_PATH)
        for name in ("setupNCmat", "setupFocks", "OneOneMap", "TwoTwoMap", "OneTwoMap"):
            getattr(L, name).restype = _ip
        L.NC.restype = ct.c_int
        L.NC.argtypes = [ct.c_int, ct.c_int]
        L.FockToIndex.restype = ct.c_int
        _lib = L
    return _lib


def _take(ptr, n: int) -> np.ndarray:
    """Copy a malloc'ed int array returned by the reference, then free it."""
    a = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].copy()
    _libc.free(ct.cast(ptr, ct.c_void_p))
    return a.astype(np.int32)


def _ip_of(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class _Cmatrix:
    """double complex ** view of an (M, M) array: row k = Ho[k][.]"""

    def __init__(self, A: np.ndarray):
        self.rows = np.ascontiguousarray(A, np.complex128)
        M = self.rows.shape[0]
        self.ptrs = (ct.c_void_p * M)(*[self.rows[k].ctypes.data for k in range(M)])

    @property
    def ptr(self):
        return ct.cast(self.ptrs, ct.c_void_p)


def NC(N: int, M: int) -> int:
    return int(lib().NC(N, M))


class RefBasis:
    """Tables and maps built by the reference's own builders (int32 strides)."""

    def __init__(self, N: int, M: int):
        L = lib()
        self.N, self.M = N, M
        self.nc = NC(N, M)
        self.NCmat = _take(L.setupNCmat(N, M), (N + 1) * (M + 1))
        self.IF = _take(L.setupFocks(N, M), self.nc * M)
        self.Map = _take(L.OneOneMap(N, M, _ip_of(self.NCmat), _ip_of(self.IF)), self.nc * M * M)
        IFm = self.IF.reshape(self.nc, M)
        n_ot = int((IFm >= 2).sum())
        occ = (IFm >= 1).astype(np.int64)
        n_tt = int(((occ.sum(1) * (occ.sum(1) - 1)) // 2).sum())
        self.strideTT = np.zeros(self.nc, np.int32)
        self.strideOT = np.zeros(self.nc, np.int32)
        self.MapTT = _take(L.TwoTwoMap(N, M, _ip_of(self.NCmat), _ip_of(self.IF),
                                       _ip_of(self.strideTT)), n_tt * M * M)
        self.MapOT = _take(L.OneTwoMap(N, M, _ip_of(self.NCmat), _ip_of(self.IF),
                                       _ip_of(self.strideOT)), n_ot * M * M)

    def _nz(self, a):
        return a if a.size else np.zeros(1, np.int32)

    def acth(self, C, Ho, Hint, variant: str = "omp") -> np.ndarray:
        """applyHconf / _X / _XX / _omp (hamiltonianMatrix.h:31, 431, 891, 1431)."""
        L = lib()
        C = np.ascontiguousarray(C, np.complex128)
        Hf = flat(Hint)
        out = np.zeros(self.nc, np.complex128)
        Hm = _Cmatrix(Ho)
        N, M = self.N, self.M
        MapOT, MapTT = self._nz(self.MapOT), self._nz(self.MapTT)
        if variant == "hash":
            L.applyHconf(N, M, _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data),
                         Hm.ptr, _vp(Hf.ctypes.data), _vp(out.ctypes.data))
        elif variant == "X":
            L.applyHconf_X(N, M, _ip_of(self.Map), _ip_of(MapOT), _ip_of(self.strideOT),
                           _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data), Hm.ptr,
                           _vp(Hf.ctypes.data), _vp(out.ctypes.data))
        else:
            f = L.applyHconf_XX if variant == "XX" else L.applyHconf_omp
            f(N, M, _ip_of(self.Map), _ip_of(MapOT), _ip_of(MapTT), _ip_of(self.strideOT),
              _ip_of(self.strideTT), _ip_of(self.IF), _vp(C.ctypes.data), Hm.ptr,
              _vp(Hf.ctypes.data), _vp(out.ctypes.data))
        return out

    def rho1(self, C, variant: str = "XM") -> np.ndarray:
        """OBrho / _X / _XM (onebodyMatrix.h:34, 116, 190); returns rho[k, l]."""
        L = lib()
        C = np.ascontiguousarray(C, np.complex128)
        R = _Cmatrix(np.zeros((self.M, self.M), np.complex128))
        N, M = self.N, self.M
        if variant == "naive":
            L.OBrho(N, M, _ip_of(self.NCmat), _vp(C.ctypes.data), R.ptr)
        elif variant == "X":
            IF = self.IF.copy()  # OBrho_X writes into IF and restores it
            L.OBrho_X(N, M, _ip_of(self.NCmat), _ip_of(IF), _vp(C.ctypes.data), R.ptr)
        else:
            L.OBrho_XM(N, M, _ip_of(self.Map), _ip_of(self.NCmat), _ip_of(self.IF),
                       _vp(C.ctypes.data), R.ptr)
        return R.rows.copy()

    def rho2(self, C, variant: str = "XX") -> np.ndarray:
        """TBrho / _X / _XX (twobodyMatrix.h:38, 502, 991); returns rho[k,s,q,l]."""
        L = lib()
        C = np.ascontiguousarray(C, np.complex128)
        out = np.zeros(self.M ** 4, np.complex128)
        N, M = self.N, self.M
        MapOT, MapTT = self._nz(self.MapOT), self._nz(self.MapTT)
        if variant == "hash":
            L.TBrho(N, M, _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data),
                    _vp(out.ctypes.data))
        elif variant == "X":
            L.TBrho_X(N, M, _ip_of(self.Map), _ip_of(MapOT), _ip_of(self.strideOT),
                      _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data),
                      _vp(out.ctypes.data))
        else:
            L.TBrho_XX(N, M, _ip_of(self.Map), _ip_of(MapOT), _ip_of(MapTT),
                       _ip_of(self.strideOT), _ip_of(self.strideTT), _ip_of(self.NCmat),
                       _ip_of(self.IF), _vp(C.ctypes.data), _vp(out.ctypes.data))
        return unflat(out, M, 4)


def acth_omp_tables(N: int, M: int, Map, MapOT, MapTT, strideOT, strideTT, IF, C, Ho, Hint) -> np.ndarray:
    """The reference's applyHconf_omp (hamiltonianMatrix.h:1431) on tables the
    caller provides (int32, reference layout; strides int32 with nc entries),
    e.g. tables copied from a device basis at sizes where the reference's
    serial builders take minutes.  Uses the reference's OpenMP threads."""
    L = lib()
    i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
    Map, MapOT, MapTT, strideOT, strideTT, IF = map(i32, (Map, MapOT, MapTT, strideOT, strideTT, IF))
    C = np.ascontiguousarray(C, np.complex128)
    Hf = flat(Hint)
    Hm = _Cmatrix(Ho)
    out = np.zeros(C.size, np.complex128)
    L.applyHconf_omp(N, M, _ip_of(Map), _ip_of(MapOT), _ip_of(MapTT), _ip_of(strideOT), _ip_of(strideTT),
                     _ip_of(IF), _vp(C.ctypes.data), Hm.ptr, _vp(Hf.ctypes.data), _vp(out.ctypes.data))
    return out
