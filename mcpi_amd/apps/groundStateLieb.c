/*
 * groundStateLieb -- Lieb-Liniger ground state on the GPU, with the command
 * line and output of the reference's program (groundStateLieb.c:475-645):
 *
 *     groundStateLieb Npar Norb g
 *
 * N bosons on a ring of length L = 1 in Norb plane-wave orbitals
 * (j = -Norb/2 ... ), contact coupling g.  Lanczos with full
 * reorthogonalization for nc/8 steps from the reference's deterministic start
 * vector, then the lowest Ritz value; prints E0 per particle as
 * "g = %.5lf | Eo = %.5lf" (groundStateLieb.c:616).
 *
 * Unlike the reference, the tables are built on the GPU with 64-bit offsets
 * (no INT32 guard on the double-jump maps): every map is stored while the
 * basis is small, above that only the single-jump map (double jumps ranked
 * on the fly), so the program runs up to N = 16, M = 12 and beyond.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "mcpi_hip.h"

static void die(const char *what)
{
    printf("\n\nERROR: %s: %s\n\n", what, mcpi_hip_last_error());
    exit(EXIT_FAILURE);
}

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char *argv[])
{
    int Npar, Morb;
    double g;
    if (argc != 4 || sscanf(argv[1], "%d", &Npar) != 1 || sscanf(argv[2], "%d", &Morb) != 1 ||
        sscanf(argv[3], "%lf", &g) != 1) {
        printf("\n\nERROR: Need two integer numbers from command line ");
        printf("the first number of particles and second the number of ");
        printf("IPS and the interaction strength(float) : \n\n");
        printf("./exe_name Npar Norb g\n\n");
        exit(EXIT_FAILURE);
    }
    if (Npar < 1 || Morb < 1) {
        printf("\n\nERROR: Npar and Norb must be positive\n\n");
        exit(EXIT_FAILURE);
    }

    /* number of configurations C(N+M-1, N), exact in 64 bits here */
    int64_t nc = 1;
    for (int64_t i = 1; i <= Npar; i++) nc = nc * (Morb - 1 + i) / i;
    printf("\nNumber of particles : %3d", Npar);
    printf("\nNumber of orbitals  : %3d", Morb);
    printf("\nNumber of configurations : %lld", (long long)nc);

    const int maps = nc <= 2000000 ? MCPI_MAPS_FULL : MCPI_MAPS_SINGLE;
    mcpi_basis b = NULL;
    if (mcpi_hip_basis_create(Npar, Morb, maps, &b) != MCPI_OK) die("basis");

    /* Lieb-Liniger one- and two-body matrices (groundStateLieb.c:129-213) */
    const double L = 1.0;
    const size_t M2 = (size_t)Morb * Morb, M4 = M2 * M2;
    mcpi_cplx *Ho = calloc(M2, sizeof(mcpi_cplx)), *Hint = calloc(M4, sizeof(mcpi_cplx));
    mcpi_cplx **Horows = malloc(sizeof(mcpi_cplx *) * Morb);
    mcpi_cplx *Hrm = calloc(M2, sizeof(mcpi_cplx));
    if (!Ho || !Hint || !Horows || !Hrm) die("host memory");
    for (int k = 0; k < Morb; k++) Horows[k] = Hrm + (size_t)k * Morb;
    setupHo(Morb, L, Horows);
    setupHint(Morb, g, Hint);
    for (int k = 0; k < Morb; k++)
        for (int l = 0; l < Morb; l++) Ho[k + (size_t)Morb * l] = Horows[k][l];   /* flat k + M*l */
    mcpi_ham h = NULL;
    if (mcpi_hip_ham_create(b, Ho, Hint, &h) != MCPI_OK) die("Hamiltonian");

    /* start vector (groundStateLieb.c:578-586) */
    mcpi_cplx *C = malloc(sizeof(mcpi_cplx) * nc);
    if (!C) die("host memory");
    double sum = 0.0;
    for (int64_t i = 0; i < nc; i++) {
        double re = sin(20 * ((double)i) / nc) * (i % 13), im = (double)((i % 8) - (i % 3));
        C[i] = re + im * I;
        sum += re * re + im * im;
    }
    for (int64_t i = 0; i < nc; i++) C[i] /= sqrt(sum);
    void *dC = NULL;
    if (mcpi_hip_malloc(&dC, sizeof(mcpi_cplx) * nc) != MCPI_OK) die("device memory");
    if (mcpi_hip_memcpy_h2d(dC, C, sizeof(mcpi_cplx) * nc) != MCPI_OK) die("upload");

    printf("\n\n======================================");
    const int lm = (int)(nc / 8 > 0 ? nc / 8 : 1);
    double E0 = 0;
    int steps = 0;
    const double t0 = now();
    if (mcpi_hip_ground(h, lm, dC, dC, &E0, &steps) != MCPI_OK) die("ground state");
    const double t1 = now();
    printf("\n\nTime to find ground state(GPU, %d Lanczos steps) : ", steps);
    printf("%.3lfs", t1 - t0);
    printf("\n\ng = %.5lf | Eo = %.5lf", g, E0 / Npar);
    printf("\n\n======================================\n\n");

    mcpi_hip_free(dC);
    mcpi_hip_ham_destroy(h);
    mcpi_hip_basis_destroy(b);
    free(C);
    free(Ho);
    free(Hint);
    free(Hrm);
    free(Horows);
    printf("\n\nDone.\n\n");
    return 0;
}
