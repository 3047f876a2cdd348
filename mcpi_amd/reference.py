"""Reference-shaped entry points of libmcpi_hip.so, called with the
reference's own conventions (int32 tables returned by malloc, int32 strides,
``double complex **`` matrices).

These are the drop-in functions a program written against mcpi's headers
links to (include/mcpi_hip.h, layer 1): ``setupFocks``, ``OneOneMap``,
``TwoTwoMap``, ``OneTwoMap``, ``applyHconf_omp``, ``applyHconf``,
``OBrho_XM``, ``TBrho_XX``, ``lanczos``, ``ground``, ... .  Like the
reference they exit the process on its INT32 guards, so callers keep to
sizes the guards admit.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import flat, lib, unflat

_libc = ct.CDLL(None)
_libc.free.argtypes = [ct.c_void_p]
_libc.free.restype = None
_ip = ct.POINTER(ct.c_int)
_vp = ct.c_void_p

_ready = False


def _L():
    global _ready
    L = lib()
    if not _ready:
        for name in ("setupNCmat", "setupFocks", "OneOneMap", "TwoTwoMap", "OneTwoMap"):
            getattr(L, name).restype = _ip
        L.NC.restype = ct.c_int
        L.NC.argtypes = [ct.c_int, ct.c_int]
        L.FockToIndex.restype = ct.c_int
        L.lanczos.restype = ct.c_int
        L.ground.restype = ct.c_double
        L.carrMod.restype = ct.c_double
        L.carrMod.argtypes = [ct.c_int, _vp]
        _ready = True
    return L


def _take(ptr, n: int) -> np.ndarray:
    a = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].copy()
    _libc.free(ct.cast(ptr, ct.c_void_p))
    return a.astype(np.int32)


def _ip_of(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class _Cmatrix:
    def __init__(self, A: np.ndarray):
        self.rows = np.ascontiguousarray(A, np.complex128)
        M = self.rows.shape[0]
        self.ptrs = (ct.c_void_p * M)(*[self.rows[k].ctypes.data for k in range(M)])

    @property
    def ptr(self):
        return ct.cast(self.ptrs, ct.c_void_p)


def NC(N: int, M: int) -> int:
    return int(_L().NC(N, M))


def FockToIndex(N: int, M: int, NCmat: np.ndarray, v) -> int:
    v = np.ascontiguousarray(v, np.int32)
    return int(_L().FockToIndex(N, M, _ip_of(NCmat), _ip_of(v)))


def IndexToFock(k: int, N: int, M: int) -> np.ndarray:
    v = np.zeros(M, np.int32)
    _L().IndexToFock(k, N, M, _ip_of(v))
    return v


class RefBasis:
    """Tables and maps through the reference-shaped builders (int32)."""

    def __init__(self, N: int, M: int):
        L = _L()
        self.N, self.M = N, M
        self.nc = NC(N, M)
        self.NCmat = _take(L.setupNCmat(N, M), (N + 1) * (M + 1))
        self.IF = _take(L.setupFocks(N, M), self.nc * M)
        self.Map = _take(L.OneOneMap(N, M, _ip_of(self.NCmat), _ip_of(self.IF)), self.nc * M * M)
        IFm = self.IF.reshape(self.nc, M)
        n_ot = int((IFm >= 2).sum())
        occ = (IFm >= 1).sum(1).astype(np.int64)
        n_tt = int((occ * (occ - 1) // 2).sum())
        self.strideTT = np.zeros(self.nc, np.int32)
        self.strideOT = np.zeros(self.nc, np.int32)
        self.MapTT = _take(L.TwoTwoMap(N, M, _ip_of(self.NCmat), _ip_of(self.IF),
                                       _ip_of(self.strideTT)), n_tt * M * M)
        self.MapOT = _take(L.OneTwoMap(N, M, _ip_of(self.NCmat), _ip_of(self.IF),
                                       _ip_of(self.strideOT)), n_ot * M * M)

    def _nz(self, a):
        return a if a.size else np.zeros(1, np.int32)

    def applyHconf_omp(self, C, Ho, Hint) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        Hf = flat(Hint)
        out = np.zeros(self.nc, np.complex128)
        Hm = _Cmatrix(Ho)
        _L().applyHconf_omp(self.N, self.M, _ip_of(self.Map), _ip_of(self._nz(self.MapOT)),
                            _ip_of(self._nz(self.MapTT)), _ip_of(self.strideOT),
                            _ip_of(self.strideTT), _ip_of(self.IF), _vp(C.ctypes.data), Hm.ptr,
                            _vp(Hf.ctypes.data), _vp(out.ctypes.data))
        return out

    def applyHconf(self, C, Ho, Hint) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        Hf = flat(Hint)
        out = np.zeros(self.nc, np.complex128)
        Hm = _Cmatrix(Ho)
        _L().applyHconf(self.N, self.M, _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data),
                        Hm.ptr, _vp(Hf.ctypes.data), _vp(out.ctypes.data))
        return out

    def applyHconf_X(self, C, Ho, Hint) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        Hf = flat(Hint)
        out = np.zeros(self.nc, np.complex128)
        Hm = _Cmatrix(Ho)
        _L().applyHconf_X(self.N, self.M, _ip_of(self.Map), _ip_of(self._nz(self.MapOT)),
                          _ip_of(self.strideOT), _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data),
                          Hm.ptr, _vp(Hf.ctypes.data), _vp(out.ctypes.data))
        return out

    def OBrho(self, C) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        R = _Cmatrix(np.zeros((self.M, self.M), np.complex128))
        _L().OBrho(self.N, self.M, _ip_of(self.NCmat), _vp(C.ctypes.data), R.ptr)
        return R.rows.copy()

    def OBrho_X(self, C) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        R = _Cmatrix(np.zeros((self.M, self.M), np.complex128))
        _L().OBrho_X(self.N, self.M, _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data), R.ptr)
        return R.rows.copy()

    def TBrho(self, C) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        out = np.zeros(self.M ** 4, np.complex128)
        _L().TBrho(self.N, self.M, _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data),
                   _vp(out.ctypes.data))
        return unflat(out, self.M, 4)

    def TBrho_X(self, C) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        out = np.zeros(self.M ** 4, np.complex128)
        _L().TBrho_X(self.N, self.M, _ip_of(self.Map), _ip_of(self._nz(self.MapOT)), _ip_of(self.strideOT),
                     _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data), _vp(out.ctypes.data))
        return unflat(out, self.M, 4)

    def OBrho_XM(self, C) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        R = _Cmatrix(np.zeros((self.M, self.M), np.complex128))
        _L().OBrho_XM(self.N, self.M, _ip_of(self.Map), _ip_of(self.NCmat), _ip_of(self.IF),
                      _vp(C.ctypes.data), R.ptr)
        return R.rows.copy()

    def TBrho_XX(self, C) -> np.ndarray:
        C = np.ascontiguousarray(C, np.complex128)
        out = np.zeros(self.M ** 4, np.complex128)
        _L().TBrho_XX(self.N, self.M, _ip_of(self.Map), _ip_of(self._nz(self.MapOT)),
                      _ip_of(self._nz(self.MapTT)), _ip_of(self.strideOT), _ip_of(self.strideTT),
                      _ip_of(self.NCmat), _ip_of(self.IF), _vp(C.ctypes.data), _vp(out.ctypes.data))
        return unflat(out, self.M, 4)

    def lanczos(self, lm: int, C0, Ho, Hint):
        """groundStateLieb.c:217 shape; returns (steps, diag, offdiag, lvec)."""
        lvec = np.zeros((lm, self.nc), np.complex128)
        lvec[0] = C0
        ptrs = (ct.c_void_p * lm)(*[lvec[t].ctypes.data for t in range(lm)])
        diag = np.zeros(lm, np.complex128)
        off = np.zeros(lm, np.complex128)
        Hm = _Cmatrix(Ho)
        Hf = flat(Hint)
        n = _L().lanczos(lm, self.N, self.M, _ip_of(self.IF), _ip_of(self.strideOT),
                         _ip_of(self.strideTT), _ip_of(self.Map), _ip_of(self._nz(self.MapOT)),
                         _ip_of(self._nz(self.MapTT)), Hm.ptr, _vp(Hf.ctypes.data),
                         _vp(diag.ctypes.data), _vp(off.ctypes.data), ct.cast(ptrs, ct.c_void_p))
        return int(n), diag, off, lvec

    def ground(self, Niter: int, C0, Ho, Hint):
        """groundStateLieb.c:348 shape; returns (E0, Ritz vector)."""
        C = np.ascontiguousarray(C0, np.complex128).copy()
        Hm = _Cmatrix(Ho)
        Hf = flat(Hint)
        E0 = _L().ground(Niter, self.N, self.M, _ip_of(self.IF), _ip_of(self.strideOT),
                         _ip_of(self.strideTT), _ip_of(self.Map), _ip_of(self._nz(self.MapOT)),
                         _ip_of(self._nz(self.MapTT)), Hm.ptr, _vp(Hf.ctypes.data),
                         _vp(C.ctypes.data))
        return float(E0), C


def setupHo(M: int, L: float = 1.0) -> np.ndarray:
    R = _Cmatrix(np.zeros((M, M), np.complex128))
    _L().setupHo(M, ct.c_double(L), R.ptr)
    return R.rows.copy()


def setupHint(M: int, g: float) -> np.ndarray:
    out = np.zeros(M ** 4, np.complex128)
    _L().setupHint(M, ct.c_double(g), _vp(out.ctypes.data))
    return unflat(out, M, 4)


def release_cached() -> None:
    """Drop the device state the reference-shaped entry points keep between
    calls (mcpi_hip_release_cached)."""
    _L().mcpi_hip_release_cached()


def carrMod(v) -> float:
    v = np.ascontiguousarray(v, np.complex128)
    return float(_L().carrMod(v.size, _vp(v.ctypes.data)))
