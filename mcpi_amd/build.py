"""Build libmcpi_hip.so in-tree for gfx950 with hipcc.

    python -m mcpi_amd.build           # or mcpi_amd.build.build()

Objects go to mcpi_amd/_build/, the library to mcpi_amd/libmcpi_hip.so, and
the Lieb-Liniger ground-state program (apps/groundStateLieb.c, linked against
the library) to mcpi_amd/bin/groundStateLieb.
Every source is compiled on every call (the build check must always
compile); translation units are built in parallel.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmcpi_hip.so")
BIN = os.path.join(HERE, "bin")
APPS = ["groundStateLieb"]
ARCH = os.environ.get("MCPI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["basis.hip", "acth.hip", "acth_fast_a.hip", "acth_fast_b.hip", "acth_fast_c.hip", "acth_comp.hip", "acth_pair.hip", "acth_multi_a.hip", "acth_multi_b.hip",
           "acth_multi_c.hip", "rho.hip", "lanczos.hip", "capi.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # hardware FP64 atomics (all device buffers are coarse-grained hipMalloc)
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libmcpi_hip.so")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compilation of {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # surface warnings other than the host pass's note about the FP32 feature
    noise = ("packed-fp32-ops", "warning generated", "warnings generated")
    lines = [ln for ln in r.stderr.splitlines() if ln.strip() and not any(n in ln for n in noise)]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_apps(verbose)
    return LIB


def build_apps(verbose: bool = True) -> None:
    """C programs of the package, linked against libmcpi_hip.so (rpath to it)."""
    os.makedirs(BIN, exist_ok=True)
    for app in APPS:
        exe = os.path.join(BIN, app)
        cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", os.path.join(HERE, "apps", app + ".c"),
               "-I", os.path.join(os.path.dirname(HERE), "include"), "-L", HERE, "-lmcpi_hip",
               "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build of {app} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {exe}")


if __name__ == "__main__":
    build()
