"""mcpi_amd -- MI355X (gfx950) implementation of mcpi's Fock-space hot path.

Python mirror of the C ABI in ``include/mcpi_hip.h`` (``libmcpi_hip.so``,
built in-tree by ``mcpi_amd.build``).  Every compute call runs HIP kernels;
there is no CPU fallback: if the native library is missing, importing the
bindings raises.

Conventions follow the reference (andriati-alex/mcpi):

* ``Ho[k, l]`` = coefficient of a*_k a_l, ``Hint[k, s, q, l]`` = coefficient
  of a*_k a*_s a_q a_l (numpy arrays; passed to C in Fortran order, i.e. the
  flat ``k + M*l`` / ``k + s*M + q*M^2 + l*M^3`` layout);
* ``rho1[k, l] = <a*_k a_l>``, ``rho2[k, s, q, l] = <a*_k a*_s a_q a_l>``.

Two layers are exposed:

* :class:`Basis`, :class:`Hamiltonian`, :class:`DeviceArray` and the
  functions below wrap the extended API (device-resident tables, int64
  offsets, device vectors);
* :mod:`mcpi_amd.reference` calls the reference-shaped entry points
  (``applyHconf_omp``, ``OBrho_XM``, ``TBrho_XX``, ``lanczos``, ...) with the
  reference's own argument conventions.
"""
from __future__ import annotations

import atexit
import ctypes as ct
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcpi_hip.so")

MAPS_NONE, MAPS_SINGLE, MAPS_FULL = 0, 1, 2
_MAPS = {"none": MAPS_NONE, "single": MAPS_SINGLE, "full": MAPS_FULL}

_vp = ct.c_void_p
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_c128p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class MCPIError(RuntimeError):
    pass


class _BasisInfo(ct.Structure):
    _fields_ = [("N", ct.c_int), ("M", ct.c_int), ("maps", ct.c_int),
                ("nc", ct.c_int64), ("map_len", ct.c_int64), ("ot_len", ct.c_int64),
                ("tt_len", ct.c_int64), ("device_bytes", ct.c_int64),
                ("pot_len", ct.c_int64), ("ptt_len", ct.c_int64), ("pattern_bytes", ct.c_int64),
                ("nsectors", ct.c_int64), ("largest_sector", ct.c_int64)]


_lib = None
# Live handles, released in a fixed order when the interpreter exits (see
# _shutdown); afterwards the handles' __del__ do nothing.
_live: "weakref.WeakSet" = weakref.WeakSet()
_finalized = False


def lib():
    """Load libmcpi_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if _finalized:
        raise MCPIError("mcpi_amd has been shut down (interpreter exit)")
    if not os.path.exists(LIB_PATH):
        raise MCPIError(f"{LIB_PATH} is missing: build it with `python -m mcpi_amd.build` "
                        "(the HIP library is required; there is no CPU fallback)")
    L = ct.CDLL(LIB_PATH)
    st = ct.c_int
    sig = {
        "mcpi_hip_last_error": (ct.c_char_p, []),
        "mcpi_hip_version": (ct.c_char_p, []),
        "mcpi_hip_malloc": (st, [ct.POINTER(_vp), ct.c_size_t]),
        "mcpi_hip_free": (st, [_vp]),
        "mcpi_hip_memcpy_h2d": (st, [_vp, _vp, ct.c_size_t]),
        "mcpi_hip_memcpy_d2h": (st, [_vp, _vp, ct.c_size_t]),
        "mcpi_hip_memcpy_d2d": (st, [_vp, _vp, ct.c_size_t]),
        "mcpi_hip_memset": (st, [_vp, ct.c_int, ct.c_size_t]),
        "mcpi_hip_synchronize": (st, []),
        "mcpi_hip_device_info": (st, [ct.c_char_p, ct.c_size_t, ct.POINTER(ct.c_int),
                                      ct.POINTER(ct.c_size_t)]),
        "mcpi_hip_basis_create": (st, [ct.c_int, ct.c_int, ct.c_int, ct.POINTER(_vp)]),
        "mcpi_hip_basis_create_sectored": (st, [ct.c_int, ct.c_int, ct.c_int, _i32p, ct.POINTER(_vp)]),
        "mcpi_hip_basis_sectors": (st, [_vp, ct.POINTER(ct.c_int64), _vp]),
        "mcpi_hip_basis_reorder": (st, [_vp, ct.c_int, _vp, _vp, _vp]),
        "mcpi_hip_conserved_charge": (st, [ct.c_int, _c128p, _c128p, _i32p]),
        "mcpi_hip_applyH_rows": (st, [_vp, _vp, _vp, _vp]),
        "mcpi_hip_basis_destroy": (st, [_vp]),
        "mcpi_hip_basis_info": (st, [_vp, ct.POINTER(_BasisInfo)]),
        "mcpi_hip_basis_copy_ncmat": (st, [_vp, _i32p]),
        "mcpi_hip_basis_copy_if": (st, [_vp, _i32p]),
        "mcpi_hip_basis_copy_map": (st, [_vp, _i32p]),
        "mcpi_hip_basis_copy_ot": (st, [_vp, _i64p, _i32p]),
        "mcpi_hip_basis_copy_tt": (st, [_vp, _i64p, _i32p]),
        "mcpi_hip_basis_copy_range": (st, [_vp, ct.c_int, ct.c_int64, ct.c_int64, _vp]),
        "mcpi_hip_basis_rank": (st, [_vp, ct.c_int64, _i32p, _i64p]),
        "mcpi_hip_basis_unrank": (st, [_vp, ct.c_int64, _i64p, _i32p]),
        "mcpi_hip_ham_create": (st, [_vp, _c128p, _c128p, ct.POINTER(_vp)]),
        "mcpi_hip_ham_destroy": (st, [_vp]),
        "mcpi_hip_ham_set_skip_zero": (st, [_vp, ct.c_int]),
        "mcpi_hip_ham_set_schedule": (st, [_vp, ct.c_int]),
        "mcpi_hip_ham_get_schedule": (st, [_vp, ct.POINTER(ct.c_int)]),
        "mcpi_hip_applyH": (st, [_vp, _vp, _vp, _vp]),
        "mcpi_hip_applyH_timed": (st, [_vp, _vp, _vp, ct.c_int, _vp, ct.POINTER(ct.c_double)]),
        "mcpi_hip_applyHconf_dev": (st, [ct.c_int, ct.c_int] + [_vp] * 11),
        "mcpi_hip_applyHconf_dev32": (st, [ct.c_int, ct.c_int] + [_vp] * 11),
        "mcpi_hip_release_cached": (st, []),
        "mcpi_hip_rho1": (st, [_vp, _vp, _c128p]),
        "mcpi_hip_rho2": (st, [_vp, _vp, _c128p]),
        "mcpi_hip_vjp": (st, [_vp, _vp, _c128p, _c128p]),
        "mcpi_hip_lanczos": (st, [_vp, ct.c_int, _vp, _vp, _c128p, _c128p, ct.POINTER(ct.c_int)]),
        "mcpi_hip_ground": (st, [_vp, ct.c_int, _vp, _vp, ct.POINTER(ct.c_double),
                                 ct.POINTER(ct.c_int)]),
        "mcpi_hip_tridiag_lowest": (st, [ct.c_int, _f64p, _f64p, ct.POINTER(ct.c_double), _f64p]),
        "mcpi_hip_lanczos_batched": (st, [ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.c_int,
                                          _f64p, _i32p, ct.POINTER(_vp)]),
        "mcpi_hip_set_batched_kernel": (st, [ct.c_int]),
        "mcpi_hip_applyH_timed_each": (st, [_vp, _vp, _vp, ct.c_int, _vp, _f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _shutdown() -> None:
    """Release every live handle while the HIP runtime is still up: wait for
    the device, then Hamiltonians, device arrays and bases in that order
    (a Hamiltonian refers to its basis).  Runs first thing at interpreter
    exit, before module teardown, so no __del__ reaches the library late."""
    global _finalized
    if _finalized:
        return
    try:
        if _lib is not None:
            _lib.mcpi_hip_synchronize()
            _lib.mcpi_hip_release_cached()
            for kind in (Hamiltonian, DeviceArray, Basis):
                for obj in [o for o in list(_live) if isinstance(o, kind)]:
                    try:
                        obj.close()
                    except Exception:
                        pass
            _lib.mcpi_hip_synchronize()
    finally:
        _finalized = True


atexit.register(_shutdown)


def _check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().mcpi_hip_last_error().decode(errors="replace")
        raise MCPIError(f"{what}: status {status}: {msg}")


def flat(a) -> np.ndarray:
    """(M,)*r complex array -> flat first-index-fastest layout."""
    a = np.asarray(a, np.complex128)
    return np.ascontiguousarray(a.reshape(-1, order="F") if a.ndim > 1 else a)


def unflat(a, M: int, r: int) -> np.ndarray:
    return np.asarray(a).reshape((M,) * r, order="F")


# ------------------------------------------------------------ device --

class DeviceArray:
    """A device allocation of ``n`` elements of ``dtype`` (owned)."""

    def __init__(self, n: int, dtype=np.complex128):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = _vp()
        _check(lib().mcpi_hip_malloc(ct.byref(p), max(self.nbytes, 16)), "malloc")
        self.ptr = p.value
        _live.add(self)

    @classmethod
    def from_numpy(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a) -> None:
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        _check(lib().mcpi_hip_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "h2d")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _check(lib().mcpi_hip_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def at(self, offset_elems: int) -> int:
        return self.ptr + offset_elems * self.dtype.itemsize

    def free(self) -> None:
        if getattr(self, "ptr", None) and not _finalized:
            lib().mcpi_hip_free(self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        if _finalized:
            return
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _check(lib().mcpi_hip_synchronize(), "synchronize")


def device_info():
    name = ct.create_string_buffer(256)
    cu = ct.c_int(0)
    mem = ct.c_size_t(0)
    _check(lib().mcpi_hip_device_info(name, 256, ct.byref(cu), ct.byref(mem)), "device_info")
    return name.value.decode(), cu.value, mem.value


# ------------------------------------------------------------- basis --

def conserved_charge(Ho, Hint):
    """Per-orbital charges q_j that Ho and Hint conserve (q_j = j when they
    conserve the orbital-index sum, e.g. Lieb-Liniger momentum), or None."""
    M = np.asarray(Ho).shape[0]
    q = np.zeros(M, np.int32)
    st = lib().mcpi_hip_conserved_charge(M, flat(Ho), flat(Hint), q)
    return q if st == 0 else None


class Basis:
    """Device-resident basis of N bosons in M orbitals (tables built on the GPU).

    ``maps``: "full" (Map, MapOT, MapTT), "single" (Map; double jumps ranked
    on the fly) or "none" (every jump ranked on the fly).  ``charge``
    (per-orbital integers): rows sorted by the conserved charge
    sum_j q_j v_j (sector order; maps must be "single").  Vectors passed to
    the methods are always in the reference order."""

    def __init__(self, N: int, M: int, maps: str = "full", charge=None):
        h = _vp()
        if charge is None:
            _check(lib().mcpi_hip_basis_create(N, M, _MAPS[maps], ct.byref(h)), f"basis({N},{M})")
        else:
            q = np.ascontiguousarray(charge, np.int32)
            _check(lib().mcpi_hip_basis_create_sectored(N, M, _MAPS[maps], q, ct.byref(h)),
                   f"sectored basis({N},{M})")
        self.handle = h.value
        _live.add(self)
        info = _BasisInfo()
        _check(lib().mcpi_hip_basis_info(self.handle, ct.byref(info)), "basis_info")
        self.N, self.M, self.maps = info.N, info.M, info.maps
        self.nc = int(info.nc)
        self.map_len, self.ot_len, self.tt_len = int(info.map_len), int(info.ot_len), int(info.tt_len)
        self.device_bytes = int(info.device_bytes)
        self.pot_len, self.ptt_len = int(info.pot_len), int(info.ptt_len)
        self.nsectors, self.largest_sector = int(info.nsectors), int(info.largest_sector)

    @property
    def sectored(self) -> bool:
        return self.nsectors > 0

    def sector_offsets(self) -> np.ndarray:
        n = ct.c_int64(0)
        _check(lib().mcpi_hip_basis_sectors(self.handle, ct.byref(n), None), "sectors")
        off = np.zeros(n.value + 1, np.int64)
        _check(lib().mcpi_hip_basis_sectors(self.handle, ct.byref(n), off.ctypes.data), "sectors")
        return off

    def to_rows(self, src: "DeviceArray", dst: "DeviceArray", stream=None) -> None:
        """dst = src moved from the reference order into the basis's row order."""
        _check(lib().mcpi_hip_basis_reorder(self.handle, 1, src.ptr, dst.ptr, stream), "reorder")

    def from_rows(self, src: "DeviceArray", dst: "DeviceArray", stream=None) -> None:
        """dst = src moved from the basis's row order into the reference order."""
        _check(lib().mcpi_hip_basis_reorder(self.handle, 0, src.ptr, dst.ptr, stream), "reorder")

    def ncmat(self) -> np.ndarray:
        out = np.zeros((self.N + 1) * (self.M + 1), np.int32)
        _check(lib().mcpi_hip_basis_copy_ncmat(self.handle, out), "copy_ncmat")
        return out

    def IF(self) -> np.ndarray:
        out = np.zeros(self.nc * self.M, np.int32)
        _check(lib().mcpi_hip_basis_copy_if(self.handle, out), "copy_if")
        return out

    def Map(self) -> np.ndarray:
        out = np.zeros(self.map_len, np.int32)
        _check(lib().mcpi_hip_basis_copy_map(self.handle, out), "copy_map")
        return out

    def OT(self):
        s = np.zeros(self.nc, np.int64)
        m = np.zeros(max(self.ot_len, 1), np.int32)
        _check(lib().mcpi_hip_basis_copy_ot(self.handle, s, m), "copy_ot")
        return s, m[: self.ot_len]

    def TT(self):
        s = np.zeros(self.nc, np.int64)
        m = np.zeros(max(self.tt_len, 1), np.int32)
        _check(lib().mcpi_hip_basis_copy_tt(self.handle, s, m), "copy_tt")
        return s, m[: self.tt_len]

    _WHICH = {"Map": (0, np.int32), "MapOT": (1, np.int32), "MapTT": (2, np.int32),
              "IF": (3, np.int32), "strideOT": (4, np.int64), "strideTT": (5, np.int64),
              "packOT": (6, np.int32), "packTT": (7, np.int32), "poffOT": (8, np.int64),
              "poffTT": (9, np.int64), "perm": (10, np.int32), "iperm": (11, np.int32),
              "IF_ref": (12, np.int32)}

    def pattern_bytes(self) -> int:
        """Device memory of the pattern tables built so far (schedule 4)."""
        info = _BasisInfo()
        _check(lib().mcpi_hip_basis_info(self.handle, ct.byref(info)), "basis_info")
        return int(info.pattern_bytes)

    def copy_range(self, table: str, offset: int, count: int) -> np.ndarray:
        """``count`` entries of a device table starting at ``offset``."""
        which, dt = self._WHICH[table]
        out = np.zeros(max(count, 1), dt)
        _check(lib().mcpi_hip_basis_copy_range(self.handle, which, offset, count, out.ctypes.data),
               "copy_range")
        return out[:count]

    def rank(self, states) -> np.ndarray:
        v = np.ascontiguousarray(states, np.int32).reshape(-1, self.M)
        out = np.zeros(v.shape[0], np.int64)
        _check(lib().mcpi_hip_basis_rank(self.handle, v.shape[0], v.reshape(-1), out), "rank")
        return out

    def unrank(self, idx) -> np.ndarray:
        idx = np.ascontiguousarray(idx, np.int64).reshape(-1)
        out = np.zeros(idx.size * self.M, np.int32)
        _check(lib().mcpi_hip_basis_unrank(self.handle, idx.size, idx, out), "unrank")
        return out.reshape(idx.size, self.M)

    def rho1(self, C: DeviceArray) -> np.ndarray:
        out = np.zeros(self.M * self.M, np.complex128)
        _check(lib().mcpi_hip_rho1(self.handle, C.ptr, out), "rho1")
        return unflat(out, self.M, 2)

    def rho2(self, C: DeviceArray) -> np.ndarray:
        out = np.zeros(self.M ** 4, np.complex128)
        _check(lib().mcpi_hip_rho2(self.handle, C.ptr, out), "rho2")
        return unflat(out, self.M, 4)

    def vjp(self, C: DeviceArray):
        """Hellmann-Feynman adjoint at an eigenvector: (dE/dHo, dE/dHint)."""
        g1 = np.zeros(self.M * self.M, np.complex128)
        g2 = np.zeros(self.M ** 4, np.complex128)
        _check(lib().mcpi_hip_vjp(self.handle, C.ptr, g1, g2), "vjp")
        return unflat(g1, self.M, 2), unflat(g2, self.M, 4)

    def close(self) -> None:
        if getattr(self, "handle", None) and not _finalized:
            lib().mcpi_hip_basis_destroy(self.handle)
        self.handle = None

    def __del__(self):
        if _finalized:
            return
        try:
            self.close()
        except Exception:
            pass


class Hamiltonian:
    """H = sum Ho a*a + 1/2 sum Hint a*a*aa on a :class:`Basis`."""

    def __init__(self, basis: Basis, Ho, Hint, skip_zero: bool = True, schedule: int = 0):
        self.basis = basis
        self.M = basis.M
        h = _vp()
        _check(lib().mcpi_hip_ham_create(basis.handle, flat(Ho), flat(Hint), ct.byref(h)), "ham_create")
        self.handle = h.value
        _live.add(self)
        self.set_skip_zero(skip_zero)
        self.set_schedule(schedule)

    def set_skip_zero(self, on: bool) -> None:
        _check(lib().mcpi_hip_ham_set_skip_zero(self.handle, int(bool(on))), "skip_zero")

    def set_schedule(self, schedule: int) -> None:
        """0 auto-tuned, 1 generic, 2 specialised reading all maps, 3
        specialised ranking double jumps on the fly, 4 specialised reading
        the pattern tables of the zero pattern (zero-skipping), 5 every
        term with each block's double jumps compacted into full wavefronts
        (full maps), 6 every term through the pair space (the two-particle
        annihilation amplitudes over the basis with N-2 particles times
        the pair matrix, the one-body part folded into it, gathered back
        per row; 3 <= M <= 12), 7 the same in the factorized pair form
        (diagonal one-body part, pair matrix block diagonal with blocks of
        rank one, as for Lieb-Liniger; 3 <= M <= 24, N >= 2)."""
        _check(lib().mcpi_hip_ham_set_schedule(self.handle, int(schedule)), "schedule")

    @property
    def schedule(self) -> int:
        s = ct.c_int(0)
        _check(lib().mcpi_hip_ham_get_schedule(self.handle, ct.byref(s)), "get_schedule")
        return s.value

    def apply(self, C: DeviceArray, out: DeviceArray, stream=None) -> None:
        _check(lib().mcpi_hip_applyH(self.handle, C.ptr, out.ptr, stream), "applyH")

    def apply_rows(self, C: DeviceArray, out: DeviceArray, stream=None) -> None:
        """H.v with both vectors in the basis's row order."""
        _check(lib().mcpi_hip_applyH_rows(self.handle, C.ptr, out.ptr, stream), "applyH_rows")

    def apply_timed(self, C: DeviceArray, out: DeviceArray, reps: int, stream=None) -> float:
        ms = ct.c_double(0)
        _check(lib().mcpi_hip_applyH_timed(self.handle, C.ptr, out.ptr, reps, stream, ct.byref(ms)),
               "applyH_timed")
        return ms.value

    def apply_timed_each(self, C: DeviceArray, out: DeviceArray, reps: int, stream=None) -> np.ndarray:
        """reps H.v launches; the duration (ms) of each, from HIP events."""
        ms = np.zeros(reps, np.float64)
        _check(lib().mcpi_hip_applyH_timed_each(self.handle, C.ptr, out.ptr, reps, stream, ms),
               "applyH_timed_each")
        return ms

    def __call__(self, C) -> np.ndarray:
        """H.v of a host vector (upload, apply, download)."""
        dC = DeviceArray.from_numpy(np.asarray(C, np.complex128))
        dO = DeviceArray(self.basis.nc)
        self.apply(dC, dO)
        synchronize()
        return dO.numpy()

    def lanczos(self, lm: int, C0: DeviceArray, lvec: DeviceArray):
        diag = np.zeros(lm, np.complex128)
        off = np.zeros(lm, np.complex128)
        steps = ct.c_int(0)
        _check(lib().mcpi_hip_lanczos(self.handle, lm, C0.ptr, lvec.ptr, diag, off, ct.byref(steps)),
               "lanczos")
        return steps.value, diag, off

    def ground(self, lm: int, C0, out: DeviceArray | None = None):
        """Lanczos ground state from C0 (host or device).  Returns
        (E0, ground vector as DeviceArray, steps)."""
        dC = C0 if isinstance(C0, DeviceArray) else DeviceArray.from_numpy(np.asarray(C0, np.complex128))
        dO = out if out is not None else DeviceArray(self.basis.nc)
        E0 = ct.c_double(0)
        steps = ct.c_int(0)
        _check(lib().mcpi_hip_ground(self.handle, lm, dC.ptr, dO.ptr, ct.byref(E0), ct.byref(steps)),
               "ground")
        return E0.value, dO, steps.value

    def close(self) -> None:
        if getattr(self, "handle", None) and not _finalized:
            lib().mcpi_hip_ham_destroy(self.handle)
        self.handle = None

    def __del__(self):
        if _finalized:
            return
        try:
            self.close()
        except Exception:
            pass


def applyHconf_dev(N, M, Map, MapOT, MapTT, strideOT, strideTT, IF, C, Ho, Hint, out) -> None:
    """CUDA-kernel-shaped H.v on DeviceArrays (int64 strides, Ho/Hint flat)."""
    _check(lib().mcpi_hip_applyHconf_dev(N, M, Map.ptr, MapOT.ptr, MapTT.ptr, strideOT.ptr,
                                         strideTT.ptr, IF.ptr, C.ptr, Ho.ptr, Hint.ptr, out.ptr,
                                         None), "applyHconf_dev")


def applyHconf_dev32(N, M, Map, MapOT, MapTT, strideOT, strideTT, IF, C, Ho, Hint, out) -> None:
    """The same with int32 strides (the CUDA kernel's Iarray strides)."""
    _check(lib().mcpi_hip_applyHconf_dev32(N, M, Map.ptr, MapOT.ptr, MapTT.ptr, strideOT.ptr,
                                           strideTT.ptr, IF.ptr, C.ptr, Ho.ptr, Hint.ptr, out.ptr,
                                           None), "applyHconf_dev32")


def tridiag_lowest(d, e):
    d = np.ascontiguousarray(d, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    if e.size < max(d.size - 1, 1):
        e = np.concatenate([e, np.zeros(max(d.size - 1, 1) - e.size)])
    lam = ct.c_double(0)
    z = np.zeros(d.size, np.float64)
    _check(lib().mcpi_hip_tridiag_lowest(d.size, d, e, ct.byref(lam), z), "tridiag")
    return lam.value, z


def lanczos_batched(hams, starts, lm: int, outs=None):
    """Lanczos ground states of many instances advanced together
    (mcpi_hip_lanczos_batched).  hams: Hamiltonian objects (instances may
    share a Basis); starts: normalized DeviceArray start vectors; outs:
    optional DeviceArrays (or None entries) for the Ritz vectors.  Returns
    (E0[n], steps[n])."""
    n = len(hams)
    if len(starts) != n or (outs is not None and len(outs) != n):
        raise ValueError("one start vector (and output) per instance")
    for h, s in zip(hams, starts):
        if s.n != h.basis.nc:
            raise ValueError("start vector length differs from the instance's basis size")
    H = (_vp * max(n, 1))(*[h.handle for h in hams])
    S = (_vp * max(n, 1))(*[s.ptr for s in starts])
    O = None
    if outs is not None:
        for h, o in zip(hams, outs):
            if o is not None and o.n != h.basis.nc:
                raise ValueError("output length differs from the instance's basis size")
        O = (_vp * max(n, 1))(*[(o.ptr if o is not None else None) for o in outs])
    E0 = np.zeros(max(n, 1), np.float64)
    steps = np.zeros(max(n, 1), np.int32)
    _check(lib().mcpi_hip_lanczos_batched(n, H, S, lm, E0, steps, O), "lanczos_batched")
    return E0[:n], steps[:n]


def set_batched_kernel(kind: int) -> None:
    """H.v kernel of batched groups: 0 timed (default), 1 batched specialised
    kernel, 2 multi-vector kernel where allowed, 3 the factorized pair form
    four members at a time where allowed (mcpi_hip_set_batched_kernel)."""
    _check(lib().mcpi_hip_set_batched_kernel(int(kind)), "set_batched_kernel")


# ----------------------------------------------------------- models --

def ll_ho(M: int, L: float = 1.0) -> np.ndarray:
    """Lieb-Liniger one-body matrix (groundStateLieb.c:129-155)."""
    Ho = np.zeros((M, M), np.complex128)
    for n in range(M):
        k = 2 * 3.141592653589793 * (n - M // 2) / L
        Ho[n, n] = 0.5 * k * k
    return Ho


def ll_hint(M: int, g: float) -> np.ndarray:
    """Contact interaction in plane waves: g where momentum is conserved
    (groundStateLieb.c:159-213)."""
    j = np.arange(M) - M // 2
    cons = (j[:, None, None, None] + j[None, :, None, None]
            - j[None, None, :, None] - j[None, None, None, :]) == 0
    return np.where(cons, g, 0.0).astype(np.complex128)


def start_vector(nc: int) -> np.ndarray:
    """C_i = sin(20 i/nc)(i%13) + i((i%8)-(i%3)), normalized
    (groundStateLieb.c:578-586)."""
    i = np.arange(nc, dtype=np.float64)
    ii = np.arange(nc, dtype=np.int64)
    C = np.sin(20 * i / nc) * (ii % 13) + 1j * ((ii % 8) - (ii % 3))
    return C / np.sqrt(np.sum(C.real ** 2 + C.imag ** 2))
