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
