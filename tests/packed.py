"""Test-side restatement of the wave-packed double-jump layout (basis.hip,
k_packed_counts / k_packed_fill): built from the reference-layout maps
(configurationsMap.h:421-762 chunk order) so that the GPU tables can be
compared entry by entry.  A slot of jump (k, s) (s == k for a same-orbital
jump) holds the entries q <= l of the chunk with neither q nor l in {k, s}
(the others are no double-jump term: the rules fold them into zero
coefficients), in triangular order over the orbitals left when k and s are
dropped, padded to a multiple of four with the row's own index."""
import numpy as np


def kept_orbitals(M: int, k: int, s: int):
    return [o for o in range(M) if o != k and o != s]


def real_entries(M: int, tt: bool) -> int:
    mp = max(M - (2 if tt else 1), 0)
    return mp * (mp + 1) // 2


def entries(M: int, tt: bool) -> int:
    return (real_entries(M, tt) + 3) // 4 * 4


def slot_pairs(M: int, k: int, s: int):
    """(q, l) of each real slot entry, in slot order."""
    kept = kept_orbitals(M, k, s)
    return [(kept[qp], kept[lp]) for lp in range(len(kept)) for qp in range(lp + 1)]


def jumps(M: int, tt: bool):
    if tt:
        return [(k, s) for k in range(M) for s in range(k + 1, M)]
    return [(k, k) for k in range(M)]


def expected_wave(IF, M, rows, Map, stride, stride_base, tt: bool):
    """Packed ints of one wavefront (rows: consecutive row indices, at most
    64).  Map/stride: reference-layout chunk map and its per-row offsets
    (stride[r - stride_base] is row r's first entry in Map)."""
    E = entries(M, tt)
    V = IF[rows]
    ordinal = np.zeros(len(rows), np.int64)
    out = []
    for (k, s) in jumps(M, tt):
        act = (V[:, k] >= 1) & (V[:, s] >= 1) if tt else V[:, k] >= 2
        if not act.any():
            continue
        idx = np.nonzero(act)[0]
        pairs = slot_pairs(M, k, s)
        slot = np.zeros((E // 4, len(idx), 4), np.int64)
        for a, ri in enumerate(idx):
            r = rows[ri]
            beg = stride[r - stride_base] + ordinal[ri] * M * M
            vals = [Map[beg + l + q * M] for (q, l) in pairs] + [r] * (E - len(pairs))
            slot[:, a, :] = np.asarray(vals).reshape(E // 4, 4)
            ordinal[ri] += 1
        out.append(slot.reshape(-1))
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def expected_packed(IF, M, Map, stride, tt: bool):
    nc = IF.shape[0]
    nw = (nc + 63) // 64
    offs = [0]
    parts = []
    for w in range(nw):
        rows = list(range(64 * w, min(64 * w + 64, nc)))
        p = expected_wave(IF, M, rows, Map, stride, 0, tt)
        parts.append(p)
        offs.append(offs[-1] + p.size)
    pack = np.concatenate(parts) if parts else np.zeros(0, np.int64)
    return pack.astype(np.int32), np.asarray(offs, np.int64)
