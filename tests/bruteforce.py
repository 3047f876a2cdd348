"""Independent second-quantization brute force for tiny bases (test pin).

Builds dense matrices of H = sum Ho[k,l] a*_k a_l + 1/2 sum Hint[k,s,q,l]
a*_k a*_s a_q a_l and of the operators behind rho1 / rho2 by applying
bosonic ladder operators to occupation tuples one term at a time.  It shares
nothing with the oracle except the ordering of basis states, which is taken
from an explicit enumeration sorted by the closed-form rank
    rank(v) = nc - 1 - sum_{j=1}^{M-1} C(P_j + j - 1, j),  P_j = sum_{o<j} v_o
(checked separately against the oracle and the reference build).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def states(N: int, M: int):
    out = [v for v in itertools.product(range(N + 1), repeat=M) if sum(v) == N]
    nc = math.comb(N + M - 1, N)
    assert len(out) == nc
    return sorted(out, key=lambda v: closed_rank(v))


def closed_rank(v) -> int:
    N, M = sum(v), len(v)
    nc = math.comb(N + M - 1, N)
    s, P = 0, 0
    for j in range(1, M):
        P += v[j - 1]
        s += math.comb(P + j - 1, j)
    return nc - 1 - s


def apply_ops(ops, v):
    """ops: list of (kind, orbital) applied right to left; kind 'c' create,
    'a' annihilate.  Returns (coefficient, new state) or (0, None)."""
    v = list(v)
    c = 1.0
    for kind, o in reversed(ops):
        if kind == "a":
            if v[o] == 0:
                return 0.0, None
            c *= math.sqrt(v[o])
            v[o] -= 1
        else:
            c *= math.sqrt(v[o] + 1)
            v[o] += 1
    return c, tuple(v)


def dense_h(N: int, M: int, Ho, Hint):
    S = states(N, M)
    idx = {v: i for i, v in enumerate(S)}
    H = np.zeros((len(S), len(S)), np.complex128)
    for j, v in enumerate(S):
        for k in range(M):
            for l in range(M):
                if Ho[k, l] == 0:
                    continue
                c, w = apply_ops([("c", k), ("a", l)], v)
                if w is not None:
                    H[idx[w], j] += Ho[k, l] * c
        for k, s, q, l in itertools.product(range(M), repeat=4):
            h = Hint[k, s, q, l]
            if h == 0:
                continue
            c, w = apply_ops([("c", k), ("c", s), ("a", q), ("a", l)], v)
            if w is not None:
                H[idx[w], j] += 0.5 * h * c
    return H


def rho1(N: int, M: int, C):
    S = states(N, M)
    idx = {v: i for i, v in enumerate(S)}
    R = np.zeros((M, M), np.complex128)
    for j, v in enumerate(S):
        for k in range(M):
            for l in range(M):
                c, w = apply_ops([("c", k), ("a", l)], v)
                if w is not None:
                    R[k, l] += np.conj(C[idx[w]]) * C[j] * c
    return R


def rho2(N: int, M: int, C):
    S = states(N, M)
    idx = {v: i for i, v in enumerate(S)}
    R = np.zeros((M,) * 4, np.complex128)
    for j, v in enumerate(S):
        for k, s, q, l in itertools.product(range(M), repeat=4):
            c, w = apply_ops([("c", k), ("c", s), ("a", q), ("a", l)], v)
            if w is not None:
                R[k, s, q, l] += np.conj(C[idx[w]]) * C[j] * c
    return R
