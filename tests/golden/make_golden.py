"""Write the golden fixtures of tests/golden/ from the reference's own code.

Run in a checkout where the reference sources are present (they are compiled
unmodified into oracle/_ref by `make -C oracle ref`):

    python tests/golden/make_golden.py

Fixtures (small .npz files: inputs and expected outputs only):
  cfg1_perf.npz   N=5, M=5: tables and maps from the reference builders, and
                  applyHconf_omp / OBrho_XM / TBrho_XX on the timing harness's
                  deterministic inputs (performanceTest.c:169-219)
  cfg1_random.npz N=5, M=5 with seeded random inputs: normalized C, Hermitian
                  Ho, fully symmetric Hint (entries ~ N(0,1) before averaging)
  small_<N>_<M>.npz  the same random-input outputs for a few more bases
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from oracle import ref  # noqa: E402


def dump(path, N, M, C, Ho, Hint, with_tables):
    r = ref.RefBasis(N, M)
    d = dict(N=N, M=M, C=C, Ho=Ho, Hint=Hint,
             out=r.acth(C, Ho, Hint, "omp"), out_hash=r.acth(C, Ho, Hint, "hash"),
             rho1=r.rho1(C, "XM"), rho2=r.rho2(C, "XX"))
    if with_tables:
        d.update(NCmat=r.NCmat, IF=r.IF, Map=r.Map, MapOT=r.MapOT, MapTT=r.MapTT,
                 strideOT=r.strideOT, strideTT=r.strideTT)
    np.savez_compressed(path, **d)
    print("wrote", path)


def main():
    oracle.build()
    if not ref.available():
        raise SystemExit("reference build missing (make -C oracle ref needs the reference sources)")
    N, M = 5, 5
    dump(os.path.join(HERE, "cfg1_perf.npz"), N, M, oracle.start_vector(oracle.nc(N, M)),
         oracle.perf_ho(M), oracle.perf_hint(M), True)
    C, Ho, Hint = oracle.random_inputs(N, M, seed=2024)
    dump(os.path.join(HERE, "cfg1_random.npz"), N, M, C, Ho, Hint, True)
    for (N, M) in [(2, 3), (3, 4), (6, 4), (4, 6), (1, 5)]:
        C, Ho, Hint = oracle.random_inputs(N, M, seed=100 * N + M)
        dump(os.path.join(HERE, f"small_{N}_{M}.npz"), N, M, C, Ho, Hint, False)


if __name__ == "__main__":
    main()
