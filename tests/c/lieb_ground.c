/* Lieb-Liniger ground state through the reference-shaped entry points of
 * libmcpi_hip.so, driven the way a reference program drives the reference's
 * headers (groundStateLieb.c main: hash table, configurations, the three
 * jump maps, the model, ground), with only the header swapped for
 * mcpi_hip.h.  Prints E0/N.
 *
 *     lieb_ground N M g [steps]      (steps default to NC(N, M) / 8)
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mcpi_hip.h"

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s N M g [steps]\n", argv[0]);
        return 2;
    }
    const int N = atoi(argv[1]), M = atoi(argv[2]);
    const double g = atof(argv[3]);
    const int nc = NC(N, M);
    const int steps = argc > 4 ? atoi(argv[4]) : nc / 8;

    Iarray NCmat = setupNCmat(N, M);
    Iarray IF = setupFocks(N, M);
    Iarray Map = OneOneMap(N, M, NCmat, IF);
    Iarray strideTT = malloc(nc * sizeof(int));
    Iarray strideOT = malloc(nc * sizeof(int));
    Iarray MapTT = TwoTwoMap(N, M, NCmat, IF, strideTT);
    Iarray MapOT = OneTwoMap(N, M, NCmat, IF, strideOT);

    Cmatrix Ho = malloc(M * sizeof(Carray));
    for (int k = 0; k < M; k++) Ho[k] = malloc(M * sizeof(mcpi_cplx));
    Carray Hint = malloc((size_t)M * M * M * M * sizeof(mcpi_cplx));
    setupHo(M, 1.0, Ho);
    setupHint(M, g, Hint);

    /* deterministic start vector of groundStateLieb.c:578-586 */
    Carray C = malloc(nc * sizeof(mcpi_cplx));
    for (int i = 0; i < nc; i++) C[i] = sin(20.0 * i / nc) * (i % 13) + I * ((i % 8) - (i % 3));
    const double norm = carrMod(nc, C);
    for (int i = 0; i < nc; i++) C[i] /= norm;

    const double E0 = ground(steps, N, M, IF, strideOT, strideTT, Map, MapOT, MapTT, Ho, Hint, C);
    printf("%.10f\n", E0 / N);

    for (int k = 0; k < M; k++) free(Ho[k]);
    free(Ho);
    free(Hint);
    free(C);
    free(NCmat);
    free(IF);
    free(Map);
    free(MapOT);
    free(MapTT);
    free(strideOT);
    free(strideTT);
    return 0;
}
