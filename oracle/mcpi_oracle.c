/*
 * mcpi_oracle.c -- CPU restatement of the mcpi Fock-space hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
 * CPU baseline ("port" kind) for the HIP library in mcpi_amd/.  Only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load it.  The
 * product library never links, loads or calls anything here.
 *
 * Every function restates the behaviour of a reference function (cited
 * file:line, reference = andriati-alex/mcpi).  It is written from the
 * semantics, not copied: indices and offsets are int64 so that the
 * N=12/M=12 and N=16/M=12 bases (whose double-jump maps exceed INT32) can be
 * built; values (configuration indices) stay int32 like the reference.
 *
 * Layout conventions (identical to the HIP library):
 *   IF     [i*M + j]                  occupation of orbital j in state i
 *   NCmat  [n + (N+1)*m]              NC(n, m)   (column m = 0 holds 0)
 *   Map    [i + k*nc + l*M*nc]        rank(v_i - e_k + e_l), -1 if v_k = 0
 *   MapOT  [strideOT[i] + c*M*M + l + q*M]   same-orbital double jump
 *   MapTT  [strideTT[i] + c*M*M + l + q*M]   two-orbital double jump
 *   Ho     [k + M*l]                  coefficient of a*_k a_l
 *   Hint   [k + s*M + q*M^2 + l*M^3]  coefficient of a*_k a*_s a_q a_l
 *   rho1   [k + M*l]                  <a*_k a_l>
 *   rho2   [k + s*M + q*M^2 + l*M^3]  <a*_k a*_s a_q a_l>
 * Complex numbers are C99 double complex (interleaved re, im).
 *
 * Parity status of each function is stated next to it ("pinned by").
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define IDX4(a, b, c, d, M) \
    ((int64_t)(a) + (int64_t)(M) * ((b) + (int64_t)(M) * ((c) + (int64_t)(M) * (d))))

/* ------------------------------------------------------------------------ */
/* L1: counting and hashing                                                 */
/* ------------------------------------------------------------------------ */

/* Number of configurations of N bosons in M orbitals, C(N+M-1, N).
 * Follows configurationsMap.h:145-202 (the value, not the INT32 guard; the
 * guard lives in the reference-shaped wrappers of the product).  M = 0
 * returns 1 like the reference function.  Returns -1 on int64 overflow.
 * Pinned by: math.comb closed form (tests/test_oracle_hash.py). */
int64_t orc_nc(int N, int M)
{
    if (N < 0 || M < 0) return -1;
    if (M == 0) return 1;
    /* C(N+M-1, min(N, M-1)) by the multiplicative formula, exact division */
    int64_t n = N + M - 1;
    int64_t r = (N < M - 1) ? N : M - 1;
    unsigned __int128 acc = 1;
    for (int64_t t = 1; t <= r; t++) {
        acc = acc * (unsigned __int128)(n - r + t) / (unsigned __int128)t;
        if (acc > (unsigned __int128)INT64_MAX) return -1;
    }
    return (int64_t)acc;
}

/* NCmat[n + (N+1)*m] = NC(n, m) for m >= 2, column 1 = 1, column 0 = 0.
 * Follows configurationsMap.h:206-234.  Values fit int32 whenever NC(N, M)
 * does (NC is increasing in both arguments).
 * Pinned by: math.comb closed form + the reference build (oracle/_ref). */
void orc_ncmat(int N, int M, int32_t *NCmat)
{
    for (int n = 0; n <= N; n++) {
        NCmat[n + (N + 1) * 0] = 0;
        if (M >= 1) NCmat[n + (N + 1) * 1] = 1;
        for (int m = 2; m <= M; m++) NCmat[n + (N + 1) * m] = (int32_t)orc_nc(n, m);
    }
}

/* Perfect-hash rank of an occupation vector.  Follows the loop of
 * FockToIndex, configurationsMap.h:273-305: orbitals are emptied from the
 * last one down to orbital 1; every particle removed from orbital i while n
 * particles remain adds NC(n, i).
 * Pinned by: exhaustive bijection onto [0, nc), the prefix-count closed form
 * rank = nc-1 - sum_j C(P_j + j - 1, j), and the reference build. */
int64_t orc_rank(int N, int M, const int32_t *NCmat, const int32_t *v)
{
    const int col = N + 1;
    int64_t k = 0;
    int n_left = N;
    for (int i = M - 1; i > 0; i--) {
        for (int p = v[i]; p > 0; p--) {
            k += NCmat[n_left + col * i];
            n_left--;
        }
    }
    return k;
}

/* Unrank: greedy inverse of orc_rank.  Follows IndexToFock,
 * configurationsMap.h:238-269: from the highest orbital down, place a
 * particle in orbital m whenever the remaining index can pay NC(n, m);
 * whatever is left goes to orbital 0.
 * Pinned by: rank(unrank(k)) == k for all k (exhaustive, small N, M). */
void orc_unrank(int64_t k, int N, int M, int32_t *v)
{
    int m = M - 1;
    int n = N;
    for (int i = 0; i < M; i++) v[i] = 0;
    while (k > 0) {
        while (k - orc_nc(n, m) < 0) m--;
        k -= orc_nc(n, m);
        v[m]++;
        n--;
    }
    if (n > 0) v[0] += n;
}

/* IF table.  Follows setupFocks, configurationsMap.h:309-344. */
void orc_setup_focks(int N, int M, int32_t *IF)
{
    int64_t nc = orc_nc(N, M);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nc; k++) orc_unrank(k, N, M, &IF[k * M]);
}

/* ------------------------------------------------------------------------ */
/* L2: jump maps                                                            */
/* ------------------------------------------------------------------------ */

/* Single-jump map.  Follows OneOneMap, configurationsMap.h:348-417:
 * Map[i + k*nc + l*M*nc] = rank(v - e_k + e_l) when v_k >= 1, else -1
 * (l = k gives i itself).
 * Pinned by: brute force rank(unrank(i) - e_k + e_l) and the reference build. */
void orc_one_one_map(int N, int M, const int32_t *NCmat, const int32_t *IF,
                     int32_t *Map)
{
    int64_t nc = orc_nc(N, M);
#pragma omp parallel
    {
        int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; i++) {
            for (int q = 0; q < M; q++) v[q] = IF[i * M + q];
            for (int k = 0; k < M; k++) {
                for (int l = 0; l < M; l++) {
                    int64_t at = i + (int64_t)k * nc + (int64_t)l * M * nc;
                    if (v[k] < 1) { Map[at] = -1; continue; }
                    v[k]--; v[l]++;
                    Map[at] = (int32_t)orc_rank(N, M, NCmat, v);
                    v[k]++; v[l]--;
                }
            }
        }
        free(v);
    }
}

/* Number of same-orbital double-jump chunks (rows x orbitals with v >= 2).
 * Follows the counting loop of allocOneTwoMap, configurationsMap.h:617-650. */
int64_t orc_count_ot_chunks(int N, int M, const int32_t *IF)
{
    int64_t nc = orc_nc(N, M), chunks = 0;
#pragma omp parallel for reduction(+ : chunks) schedule(static)
    for (int64_t i = 0; i < nc; i++)
        for (int k = 0; k < M; k++)
            if (IF[i * M + k] >= 2) chunks++;
    return chunks;
}

/* Number of two-orbital double-jump chunks (rows x pairs k<s occupied).
 * Follows the counting loop of allocTwoTwoMap, configurationsMap.h:421-460. */
int64_t orc_count_tt_chunks(int N, int M, const int32_t *IF)
{
    int64_t nc = orc_nc(N, M), chunks = 0;
#pragma omp parallel for reduction(+ : chunks) schedule(static)
    for (int64_t i = 0; i < nc; i++)
        for (int k = 0; k < M; k++) {
            if (IF[i * M + k] < 1) continue;
            for (int s = k + 1; s < M; s++)
                if (IF[i * M + s] >= 1) chunks++;
        }
    return chunks;
}

/* Row offsets of the OT map: strideOT[i] = M^2 * (chunks of rows < i).
 * Follows OneTwoMap, configurationsMap.h:713-717 (int64 here). */
void orc_stride_ot(int N, int M, const int32_t *IF, int64_t *strideOT)
{
    int64_t nc = orc_nc(N, M), chunks = 0;
    for (int64_t i = 0; i < nc; i++) {
        strideOT[i] = chunks * M * M;
        for (int k = 0; k < M; k++)
            if (IF[i * M + k] >= 2) chunks++;
    }
}

/* Row offsets of the TT map.  Follows TwoTwoMap, configurationsMap.h:551-557. */
void orc_stride_tt(int N, int M, const int32_t *IF, int64_t *strideTT)
{
    int64_t nc = orc_nc(N, M), chunks = 0;
    for (int64_t i = 0; i < nc; i++) {
        strideTT[i] = chunks * M * M;
        for (int k = 0; k < M; k++) {
            if (IF[i * M + k] < 1) continue;
            for (int s = k + 1; s < M; s++)
                if (IF[i * M + s] >= 1) chunks++;
        }
    }
}

/* Same-orbital double-jump map.  Follows OneTwoMap, configurationsMap.h:
 * 668-762: for every row and every k with v_k >= 2 (ascending k) a chunk of
 * M^2 entries, entry l + q*M = rank(v - 2e_k + e_l + e_q).  strideOT must
 * already hold orc_stride_ot.
 * Pinned by: brute force on tiny bases and the reference build. */
void orc_one_two_map(int N, int M, const int32_t *NCmat, const int32_t *IF,
                     const int64_t *strideOT, int32_t *MapOT)
{
    int64_t nc = orc_nc(N, M);
    int64_t M2 = (int64_t)M * M;
#pragma omp parallel
    {
        int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; i++) {
            for (int q = 0; q < M; q++) v[q] = IF[i * M + q];
            int64_t c = 0;
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                int64_t base = strideOT[i] + c * M2;
                for (int l = 0; l < M; l++)
                    for (int q = 0; q < M; q++) {
                        v[k] -= 2; v[l]++; v[q]++;
                        MapOT[base + l + (int64_t)q * M] = (int32_t)orc_rank(N, M, NCmat, v);
                        v[k] += 2; v[l]--; v[q]--;
                    }
                c++;
            }
        }
        free(v);
    }
}

/* Two-orbital double-jump map.  Follows TwoTwoMap, configurationsMap.h:
 * 478-613: for every row and every pair k<s of occupied orbitals, in
 * lexicographic order, a chunk with entry l + q*M = rank(v - e_k - e_s + e_l
 * + e_q).  strideTT must already hold orc_stride_tt.
 * Pinned by: brute force on tiny bases and the reference build. */
void orc_two_two_map(int N, int M, const int32_t *NCmat, const int32_t *IF,
                     const int64_t *strideTT, int32_t *MapTT)
{
    int64_t nc = orc_nc(N, M);
    int64_t M2 = (int64_t)M * M;
#pragma omp parallel
    {
        int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; i++) {
            for (int q = 0; q < M; q++) v[q] = IF[i * M + q];
            int64_t c = 0;
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++) {
                    if (v[s] < 1) continue;
                    int64_t base = strideTT[i] + c * M2;
                    for (int l = 0; l < M; l++)
                        for (int q = 0; q < M; q++) {
                            v[k]--; v[s]--; v[l]++; v[q]++;
                            MapTT[base + l + (int64_t)q * M] =
                                (int32_t)orc_rank(N, M, NCmat, v);
                            v[k]++; v[s]++; v[l]--; v[q]--;
                        }
                    c++;
                }
            }
        }
        free(v);
    }
}

/* ------------------------------------------------------------------------ */
/* L3: H.v                                                                  */
/* ------------------------------------------------------------------------ */

/* chunk ordinal of orbital k in the OT section of a row: number of orbitals
 * h < k with v_h >= 2 (hamiltonianMatrix.h:1564-1568) */
static inline int64_t ot_ordinal(const int32_t *v, int k)
{
    int64_t c = 0;
    for (int h = 0; h < k; h++)
        if (v[h] > 1) c++;
    return c;
}

/* chunk ordinal of pair (k, s) in the TT section of a row: pairs (h<g) with
 * h < k both occupied, plus occupied g with k < g < s
 * (hamiltonianMatrix.h:1927-1941) */
static inline int64_t tt_ordinal(const int32_t *v, int M, int k, int s)
{
    int64_t c = 0;
    for (int h = 0; h < k; h++)
        for (int g = h + 1; g < M; g++)
            if (v[h] > 0 && v[g] > 0) c++;
    for (int g = k + 1; g < s; g++)
        if (v[g] > 0) c++;
    return c;
}

/* Action of H on C with all three maps for rows [row0, row1) (out_rows[i -
 * row0]).  Restates applyHconf_omp,
 * hamiltonianMatrix.h:1431-1970 (same as applyHconf_XX 891-1425): the same
 * rules, multipliers and order, out = w + 0.5*z.  The chunk ordinals are
 * hoisted out of the innermost loops (same values).
 * Pinned by: dense brute-force second quantization (tiny bases), Hermiticity,
 * agreement with the hash-only restatement, and the reference build. */
void orc_acth_range(int N, int M, int64_t row0, int64_t row1, const int32_t *Map,
                    const int32_t *MapOT, const int32_t *MapTT, const int64_t *strideOT,
                    const int64_t *strideTT, const int32_t *IF, const cplx *C,
                    const cplx *Ho, const cplx *Hint, cplx *out_rows)
{
    int64_t nc = orc_nc(N, M);
    int64_t M2 = (int64_t)M * M;
    cplx *out = out_rows - row0;   /* out[i] for i in [row0, row1) */
#pragma omp parallel
    {
        int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
#pragma omp for schedule(static)
        for (int64_t i = row0; i < row1; i++) {
            cplx w = 0, z = 0;
            for (int k = 0; k < M; k++) v[k] = IF[i * M + k];

            /* one-body, hamiltonianMatrix.h:1494-1512 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                w += Ho[k + M * k] * v[k] * C[i];
                for (int l = 0; l < M; l++) {
                    if (l == k) continue;
                    double f = sqrt((double)v[k] * (v[l] + 1));
                    int64_t j = Map[i + k * nc + l * M * nc];
                    w += Ho[k + M * l] * f * C[j];
                }
            }
            /* rule 1, 1525-1529 */
            for (int k = 0; k < M; k++) {
                double f = (double)v[k] * (v[k] - 1);
                z += Hint[IDX4(k, k, k, k, M)] * C[i] * f;
            }
            /* rule 2, 1536-1549 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++) {
                    double f = (double)v[k] * v[s];
                    z += 4 * Hint[IDX4(k, s, k, s, M)] * f * C[i];
                }
            }
            /* rule 3, 1556-1579 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                int64_t base = strideOT[i] + ot_ordinal(v, k) * M2;
                for (int q = 0; q < M; q++) {
                    if (q == k) continue;
                    double f = sqrt((double)(v[k] - 1) * v[k] * (v[q] + 1) * (v[q] + 2));
                    int64_t j = MapOT[base + q + q * M];
                    z += Hint[IDX4(k, k, q, q, M)] * C[j] * f;
                }
            }
            /* rule 4, 1586-1602 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                for (int l = 0; l < M; l++) {
                    if (l == k) continue;
                    double f = (v[k] - 1) * sqrt((double)v[k] * (v[l] + 1));
                    int64_t j = Map[i + k * nc + l * M * nc];
                    z += 2 * Hint[IDX4(k, k, k, l, M)] * C[j] * f;
                }
            }
            /* rule 5, 1609-1625 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = 0; s < M; s++) {
                    if (s == k || v[s] < 1) continue;
                    double f = v[s] * sqrt((double)v[k] * (v[s] + 1));
                    int64_t j = Map[i + k * nc + s * M * nc];
                    z += 2 * Hint[IDX4(k, s, s, s, M)] * C[j] * f;
                }
            }
            /* rules 6.0 / 6.1 / 6.2, 1632-1732: v_k >= 2, q < l, k not in {q,l}.
             * The three orderings are visited in the reference's order. */
            for (int part = 0; part < 3; part++) {
                for (int a = 0; a < M; a++)
                    for (int b = a + 1; b < M; b++)
                        for (int c = b + 1; c < M; c++) {
                            int k, q, l;
                            if (part == 0) { k = a; q = b; l = c; }      /* k<q<l */
                            else if (part == 1) { q = a; k = b; l = c; } /* q<k<l */
                            else { q = a; l = b; k = c; }                /* q<l<k */
                            if (v[k] < 2) continue;
                            double f = sqrt((double)v[k] * (v[k] - 1) * (v[q] + 1) * (v[l] + 1));
                            int64_t base = strideOT[i] + ot_ordinal(v, k) * M2;
                            int64_t j = MapOT[base + q + l * M];
                            z += 2 * Hint[IDX4(k, k, q, l, M)] * C[j] * f;
                        }
            }
            /* rules 7.0 / 7.1 / 7.2, 1739-1876: k < s occupied, q not in {k,s} */
            for (int part = 0; part < 3; part++) {
                for (int a = 0; a < M; a++)
                    for (int b = a + 1; b < M; b++)
                        for (int c = b + 1; c < M; c++) {
                            int k, s, q;
                            if (part == 0) { q = a; k = b; s = c; }      /* q<k<s */
                            else if (part == 1) { k = a; q = b; s = c; } /* k<q<s */
                            else { k = a; s = b; q = c; }                /* k<s<q */
                            if (v[k] < 1 || v[s] < 1) continue;
                            double f = sqrt((double)v[k] * v[s] * (v[q] + 1) * (v[q] + 2));
                            int64_t base = strideTT[i] + tt_ordinal(v, M, k, s) * M2;
                            int64_t j = MapTT[base + q + q * M];
                            z += 2 * Hint[IDX4(k, s, q, q, M)] * C[j] * f;
                        }
            }
            /* rule 8, 1883-1907 */
            for (int s = 0; s < M; s++) {
                if (v[s] < 1) continue;
                for (int k = 0; k < M; k++) {
                    if (v[k] < 1 || k == s) continue;
                    for (int l = 0; l < M; l++) {
                        if (l == k || l == s) continue;
                        double f = v[s] * sqrt((double)v[k] * (v[l] + 1));
                        int64_t j = Map[i + k * nc + l * M * nc];
                        z += 4 * Hint[IDX4(k, s, s, l, M)] * C[j] * f;
                    }
                }
            }
            /* rule 9, 1913-1961 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++) {
                    if (v[s] < 1) continue;
                    int64_t base = strideTT[i] + tt_ordinal(v, M, k, s) * M2;
                    for (int q = 0; q < M; q++) {
                        if (q == s || q == k) continue;
                        for (int l = q + 1; l < M; l++) {
                            if (l == k || l == s) continue;
                            double f = sqrt((double)v[k] * v[s] * (v[q] + 1) * (v[l] + 1));
                            int64_t j = MapTT[base + q + l * M];
                            z += 4 * Hint[IDX4(k, s, q, l, M)] * C[j] * f;
                        }
                    }
                }
            }
            out[i] = w + 0.5 * z;
        }
        free(v);
    }
}

/* all rows: restates applyHconf_omp over the whole basis */
void orc_acth(int N, int M, const int32_t *Map, const int32_t *MapOT,
              const int32_t *MapTT, const int64_t *strideOT,
              const int64_t *strideTT, const int32_t *IF, const cplx *C,
              const cplx *Ho, const cplx *Hint, cplx *out)
{
    orc_acth_range(N, M, 0, orc_nc(N, M), Map, MapOT, MapTT, strideOT, strideTT, IF, C, Ho,
                   Hint, out);
}

/* Action of H on C for a list of rows, ranking every jump on the fly (no
 * maps).  Restates the hash-only applyHconf, hamiltonianMatrix.h:31-425
 * (rule 9 over all ordered distinct (k,s,q,l) with multiplier 1), using
 * IndexToFock for the row state.  Used for row samples of bases too large
 * for CPU maps.
 * Pinned by: agreement with orc_acth on every row of small bases. */
void orc_acth_hash_rows(int N, int M, const int32_t *NCmat, int64_t nrows,
                        const int64_t *rows, const cplx *C, const cplx *Ho,
                        const cplx *Hint, cplx *out)
{
#pragma omp parallel
    {
        int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrows; r++) {
            int64_t i = rows[r];
            cplx w = 0, z = 0;
            orc_unrank(i, N, M, v);
#define RANK() orc_rank(N, M, NCmat, v)
            /* one-body, 85-102 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                w += Ho[k + M * k] * v[k] * C[i];
                for (int l = 0; l < M; l++) {
                    if (l == k) continue;
                    double f = sqrt((double)v[k] * (v[l] + 1));
                    v[k]--; v[l]++;
                    int64_t j = RANK();
                    v[k]++; v[l]--;
                    w += Ho[k + M * l] * f * C[j];
                }
            }
            /* rule 1, 115-119 */
            for (int k = 0; k < M; k++)
                z += Hint[IDX4(k, k, k, k, M)] * C[i] * ((double)v[k] * (v[k] - 1));
            /* rule 2, 126-139 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++)
                    z += 4 * Hint[IDX4(k, s, k, s, M)] * ((double)v[k] * v[s]) * C[i];
            }
            /* rule 3, 145-159 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                for (int q = 0; q < M; q++) {
                    if (q == k) continue;
                    double f = sqrt((double)(v[k] - 1) * v[k] * (v[q] + 1) * (v[q] + 2));
                    v[k] -= 2; v[q] += 2;
                    int64_t j = RANK();
                    v[k] += 2; v[q] -= 2;
                    z += Hint[IDX4(k, k, q, q, M)] * C[j] * f;
                }
            }
            /* rule 4, 164-179 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                for (int l = 0; l < M; l++) {
                    if (l == k) continue;
                    double f = (v[k] - 1) * sqrt((double)v[k] * (v[l] + 1));
                    v[k]--; v[l]++;
                    int64_t j = RANK();
                    v[k]++; v[l]--;
                    z += 2 * Hint[IDX4(k, k, k, l, M)] * C[j] * f;
                }
            }
            /* rule 5, 185-200 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = 0; s < M; s++) {
                    if (s == k || v[s] < 1) continue;
                    double f = v[s] * sqrt((double)v[k] * (v[s] + 1));
                    v[k]--; v[s]++;
                    int64_t j = RANK();
                    v[k]++; v[s]--;
                    z += 2 * Hint[IDX4(k, s, s, s, M)] * C[j] * f;
                }
            }
            /* rule 6, 206-275: v_k >= 2, q < l, k not in {q, l} */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                for (int q = 0; q < M; q++) {
                    if (q == k) continue;
                    for (int l = q + 1; l < M; l++) {
                        if (l == k) continue;
                        double f = sqrt((double)v[k] * (v[k] - 1) * (v[q] + 1) * (v[l] + 1));
                        v[k] -= 2; v[l]++; v[q]++;
                        int64_t j = RANK();
                        v[k] += 2; v[l]--; v[q]--;
                        z += 2 * Hint[IDX4(k, k, q, l, M)] * C[j] * f;
                    }
                }
            }
            /* rule 7, 282-354: k < s occupied, q not in {k, s} */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++) {
                    if (v[s] < 1) continue;
                    for (int q = 0; q < M; q++) {
                        if (q == k || q == s) continue;
                        double f = sqrt((double)v[k] * v[s] * (v[q] + 1) * (v[q] + 2));
                        v[k]--; v[s]--; v[q] += 2;
                        int64_t j = RANK();
                        v[k]++; v[s]++; v[q] -= 2;
                        z += 2 * Hint[IDX4(k, s, q, q, M)] * C[j] * f;
                    }
                }
            }
            /* rule 8, 361-384 */
            for (int s = 0; s < M; s++) {
                if (v[s] < 1) continue;
                for (int k = 0; k < M; k++) {
                    if (v[k] < 1 || k == s) continue;
                    for (int l = 0; l < M; l++) {
                        if (l == k || l == s) continue;
                        double f = v[s] * sqrt((double)v[k] * (v[l] + 1));
                        v[k]--; v[l]++;
                        int64_t j = RANK();
                        v[k]++; v[l]--;
                        z += 4 * Hint[IDX4(k, s, s, l, M)] * C[j] * f;
                    }
                }
            }
            /* rule 9, 390-416: all ordered distinct (k, s, q, l), multiplier 1 */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = 0; s < M; s++) {
                    if (v[s] < 1 || s == k) continue;
                    for (int q = 0; q < M; q++) {
                        if (q == s || q == k) continue;
                        for (int l = 0; l < M; l++) {
                            if (l == k || l == s || l == q) continue;
                            double f = sqrt((double)v[k] * v[s] * (v[q] + 1) * (v[l] + 1));
                            v[k]--; v[s]--; v[q]++; v[l]++;
                            int64_t j = RANK();
                            v[k]++; v[s]++; v[q]--; v[l]--;
                            z += Hint[IDX4(k, s, q, l, M)] * C[j] * f;
                        }
                    }
                }
            }
#undef RANK
            out[r] = w + 0.5 * z;
        }
        free(v);
    }
}

/* ------------------------------------------------------------------------ */
/* L3: density matrices                                                     */
/* ------------------------------------------------------------------------ */

/* One-body density matrix rho1[k + M*l] = <a*_k a_l>.  Restates OBrho_XM,
 * onebodyMatrix.h:190-251: diagonal sum |C_i|^2 v_k; off-diagonal k < l
 * sum conj(C_i) C_Map[i,k,l] sqrt((v_l+1) v_k), lower triangle = conjugate.
 * Pinned by: brute force, trace = N, Hermiticity, reference build. */
void orc_rho1(int N, int M, const int32_t *Map, const int32_t *IF,
              const cplx *C, cplx *rho)
{
    int64_t nc = orc_nc(N, M);
    for (int k = 0; k < M; k++) {
        double d = 0;
#pragma omp parallel for reduction(+ : d) schedule(static)
        for (int64_t i = 0; i < nc; i++) {
            double m2 = creal(C[i]) * creal(C[i]) + cimag(C[i]) * cimag(C[i]);
            d += m2 * IF[i * M + k];
        }
        rho[k + M * k] = d;
        for (int l = k + 1; l < M; l++) {
            double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
            for (int64_t i = 0; i < nc; i++) {
                int vk = IF[i * M + k];
                if (vk < 1) continue;
                int vl = IF[i * M + l];
                int64_t j = Map[i + k * nc + l * M * nc];
                cplx t = conj(C[i]) * C[j] * sqrt((double)(vl + 1) * vk);
                re += creal(t);
                im += cimag(t);
            }
            rho[k + M * l] = re + I * im;
            rho[l + M * k] = re - I * im;
        }
    }
}

/* Two-body density matrix rho2[k + sM + qM^2 + lM^3] = <a*_k a*_s a_q a_l>.
 * Restates TBrho_XX, twobodyMatrix.h:991-1487.  The reference loops
 * orbital-outer / row-inner; here every row adds its contributions to a
 * per-thread table of the reference's primary entries, which are then
 * reduced and expanded by the reference's symmetry copies in its rule order.
 * Pinned by: brute force, partial-trace identities, energy identity,
 * reference build. */
void orc_rho2(int N, int M, const int32_t *Map, const int32_t *MapOT,
              const int32_t *MapTT, const int64_t *strideOT,
              const int64_t *strideTT, const int32_t *IF, const cplx *C,
              cplx *rho)
{
    int64_t nc = orc_nc(N, M);
    int64_t M2 = (int64_t)M * M, M4 = M2 * M2;
    cplx *acc = (cplx *)calloc(M4, sizeof(cplx));
#pragma omp parallel
    {
        cplx *loc = (cplx *)calloc(M4, sizeof(cplx));
        int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; i++) {
            for (int k = 0; k < M; k++) v[k] = IF[i * M + k];
            cplx ci = C[i], cci = conj(C[i]);
            double m2 = creal(ci) * creal(ci) + cimag(ci) * cimag(ci);
            /* rule 1, 1038-1051 */
            for (int k = 0; k < M; k++)
                if (v[k] >= 2) loc[IDX4(k, k, k, k, M)] += m2 * v[k] * (v[k] - 1);
            /* rule 2, 1058-1078 */
            for (int k = 0; k < M; k++)
                for (int s = k + 1; s < M; s++)
                    loc[IDX4(k, s, k, s, M)] += m2 * v[k] * v[s];
            /* rule 3, 1085-1116 (k < q, v_k >= 2) */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                int64_t base = strideOT[i] + ot_ordinal(v, k) * M2;
                for (int q = k + 1; q < M; q++) {
                    int64_t j = MapOT[base + q + q * M];
                    double f = sqrt((double)(v[k] - 1) * v[k] * (v[q] + 1) * (v[q] + 2));
                    loc[IDX4(k, k, q, q, M)] += cci * C[j] * f;
                }
            }
            /* rule 4, 1123-1146 (k < l, v_k >= 2) */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                for (int l = k + 1; l < M; l++) {
                    int64_t j = Map[i + k * nc + l * M * nc];
                    double f = (v[k] - 1) * sqrt((double)v[k] * (v[l] + 1));
                    loc[IDX4(k, k, k, l, M)] += cci * C[j] * f;
                }
            }
            /* rule 5, 1153-1176 (k < s, both occupied) */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++) {
                    if (v[s] < 1) continue;
                    int64_t j = Map[i + k * nc + s * M * nc];
                    double f = v[s] * sqrt((double)v[k] * (v[s] + 1));
                    loc[IDX4(k, s, s, s, M)] += cci * C[j] * f;
                }
            }
            /* rule 6, 1183-1308 (v_k >= 2, q < l, k not in {q, l}) */
            for (int k = 0; k < M; k++) {
                if (v[k] < 2) continue;
                int64_t base = strideOT[i] + ot_ordinal(v, k) * M2;
                for (int q = 0; q < M; q++) {
                    if (q == k) continue;
                    for (int l = q + 1; l < M; l++) {
                        if (l == k) continue;
                        int64_t j = MapOT[base + l + q * M];
                        double f = sqrt((double)v[k] * (v[k] - 1) * (v[q] + 1) * (v[l] + 1));
                        loc[IDX4(k, k, q, l, M)] += cci * C[j] * f;
                    }
                }
            }
            /* rule 7, 1315-1422 (k < l, s not in {k, l}, v_k, v_s >= 1) */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int l = k + 1; l < M; l++) {
                    int64_t j = Map[i + k * nc + l * M * nc];
                    for (int s = 0; s < M; s++) {
                        if (s == k || s == l || v[s] < 1) continue;
                        double f = v[s] * sqrt((double)v[k] * (v[l] + 1));
                        loc[IDX4(k, s, s, l, M)] += cci * C[j] * f;
                    }
                }
            }
            /* rule 8, 1429-1484 (k < s occupied, q < l, disjoint) */
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++) {
                    if (v[s] < 1) continue;
                    int64_t base = strideTT[i] + tt_ordinal(v, M, k, s) * M2;
                    for (int q = 0; q < M; q++) {
                        if (q == s || q == k) continue;
                        for (int l = q + 1; l < M; l++) {
                            if (l == k || l == s) continue;
                            int64_t j = MapTT[base + q + l * M];
                            double f = sqrt((double)v[k] * v[s] * (v[q] + 1) * (v[l] + 1));
                            loc[IDX4(k, s, q, l, M)] += cci * C[j] * f;
                        }
                    }
                }
            }
        }
#pragma omp critical
        for (int64_t t = 0; t < M4; t++) acc[t] += loc[t];
        free(loc);
        free(v);
    }

    /* expansion by the reference's symmetry copies, in its rule order */
    for (int64_t t = 0; t < M4; t++) rho[t] = 0;
    for (int k = 0; k < M; k++) rho[IDX4(k, k, k, k, M)] = acc[IDX4(k, k, k, k, M)];
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            cplx R = acc[IDX4(k, s, k, s, M)];
            rho[IDX4(k, s, k, s, M)] = R; rho[IDX4(s, k, k, s, M)] = R;
            rho[IDX4(s, k, s, k, M)] = R; rho[IDX4(k, s, s, k, M)] = R;
        }
    for (int k = 0; k < M; k++)
        for (int q = k + 1; q < M; q++) {
            cplx R = acc[IDX4(k, k, q, q, M)];
            rho[IDX4(k, k, q, q, M)] = R; rho[IDX4(q, q, k, k, M)] = conj(R);
        }
    for (int k = 0; k < M; k++)
        for (int l = k + 1; l < M; l++) {
            cplx R = acc[IDX4(k, k, k, l, M)];
            rho[IDX4(k, k, k, l, M)] = R; rho[IDX4(k, k, l, k, M)] = R;
            rho[IDX4(l, k, k, k, M)] = conj(R); rho[IDX4(k, l, k, k, M)] = conj(R);
        }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            cplx R = acc[IDX4(k, s, s, s, M)];
            rho[IDX4(k, s, s, s, M)] = R; rho[IDX4(s, k, s, s, M)] = R;
            rho[IDX4(s, s, s, k, M)] = conj(R); rho[IDX4(s, s, k, s, M)] = conj(R);
        }
    for (int k = 0; k < M; k++)
        for (int q = 0; q < M; q++) {
            if (q == k) continue;
            for (int l = q + 1; l < M; l++) {
                if (l == k) continue;
                cplx R = acc[IDX4(k, k, q, l, M)];
                rho[IDX4(k, k, q, l, M)] = R; rho[IDX4(k, k, l, q, M)] = R;
                rho[IDX4(l, q, k, k, M)] = conj(R); rho[IDX4(q, l, k, k, M)] = conj(R);
            }
        }
    for (int k = 0; k < M; k++)
        for (int l = k + 1; l < M; l++)
            for (int s = 0; s < M; s++) {
                if (s == k || s == l) continue;
                cplx R = acc[IDX4(k, s, s, l, M)];
                rho[IDX4(k, s, s, l, M)] = R; rho[IDX4(s, k, s, l, M)] = R;
                rho[IDX4(s, k, l, s, M)] = R; rho[IDX4(k, s, l, s, M)] = R;
                rho[IDX4(l, s, s, k, M)] = conj(R); rho[IDX4(s, l, s, k, M)] = conj(R);
                rho[IDX4(s, l, k, s, M)] = conj(R); rho[IDX4(l, s, k, s, M)] = conj(R);
            }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
            for (int q = 0; q < M; q++) {
                if (q == s || q == k) continue;
                for (int l = q + 1; l < M; l++) {
                    if (l == k || l == s) continue;
                    cplx R = acc[IDX4(k, s, q, l, M)];
                    rho[IDX4(k, s, q, l, M)] = R; rho[IDX4(s, k, q, l, M)] = R;
                    rho[IDX4(k, s, l, q, M)] = R; rho[IDX4(s, k, l, q, M)] = R;
                }
            }
    free(acc);
}

/* ------------------------------------------------------------------------ */
/* L4: Lanczos and the ground state                                         */
/* ------------------------------------------------------------------------ */

/* <v1, v2> with the first argument conjugated.  groundStateLieb.c:73-85 */
cplx orc_dot(int64_t n, const cplx *v1, const cplx *v2)
{
    double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        cplx t = conj(v1[i]) * v2[i];
        re += creal(t);
        im += cimag(t);
    }
    return re + I * im;
}

/* 2-norm.  groundStateLieb.c:89-101 */
double orc_norm(int64_t n, const cplx *v)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; i++) s += creal(v[i]) * creal(v[i]) + cimag(v[i]) * cimag(v[i]);
    return sqrt(s);
}

/* Lanczos with one-pass classical Gram-Schmidt full reorthogonalization.
 * Restates lanczos, groundStateLieb.c:217-344.  lvec is lm rows of nc
 * (row-major, lvec[0] = normalized start vector on entry).
 * Breakdown (beta_i / max beta < 1e-15, or beta_i == 0) returns i + 1, the
 * size of the invariant Krylov space; the reference returns i (dropping the
 * last alpha) and divides 0/0 when beta_0 = 0 (documented deviation).
 * Pinned by: dense eigvalsh of the brute-force H on small bases, liebGS.dat,
 * and orthonormality of the Lanczos basis. */
int orc_lanczos(int lm, int N, int M, const int32_t *Map, const int32_t *MapOT,
                const int32_t *MapTT, const int64_t *strideOT,
                const int64_t *strideTT, const int32_t *IF, const cplx *Ho,
                const cplx *Hint, cplx *diag, cplx *offdiag, cplx *lvec)
{
    int64_t nc = orc_nc(N, M);
    cplx *HC = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *ortho = (cplx *)malloc(sizeof(cplx) * (lm > 0 ? lm : 1));
    double tol = 1e-15, maxb = 0;
    int done = lm;

    orc_acth(N, M, Map, MapOT, MapTT, strideOT, strideTT, IF, lvec, Ho, Hint, HC);
    diag[0] = orc_dot(nc, lvec, HC);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nc; j++) HC[j] -= diag[0] * lvec[j];

    for (int i = 0; i < lm - 1; i++) {
        double b = orc_norm(nc, HC);
        offdiag[i] = b;
        if (b > maxb) maxb = b;
        if (b == 0.0 || b / maxb < tol) { done = i + 1; break; }
        cplx *vi = lvec + (int64_t)i * nc, *vn = lvec + (int64_t)(i + 1) * nc;
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < nc; j++) vn[j] = HC[j] / b;
        orc_acth(N, M, Map, MapOT, MapTT, strideOT, strideTT, IF, vn, Ho, Hint, HC);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < nc; j++) HC[j] -= b * vi[j];
        diag[i + 1] = orc_dot(nc, vn, HC);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < nc; j++) HC[j] -= diag[i + 1] * vn[j];
        for (int t = 0; t < i + 2; t++) ortho[t] = orc_dot(nc, lvec + (int64_t)t * nc, HC);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < nc; j++) {
            cplx h = HC[j];
            for (int t = 0; t < i + 2; t++) h -= lvec[(int64_t)t * nc + j] * ortho[t];
            HC[j] = h;
        }
    }
    if (done == lm && lm >= 1) offdiag[lm - 1] = 0;
    free(ortho);
    free(HC);
    return done;
}

/* Sturm count: number of eigenvalues of the symmetric tridiagonal (d, e)
 * strictly below x (LDL^T inertia).  e[i] couples i and i+1. */
static int sturm_below(int n, const double *d, const double *e, double x)
{
    int cnt = 0;
    double q = 1.0;
    for (int i = 0; i < n; i++) {
        double e2 = (i > 0) ? e[i - 1] * e[i - 1] : 0.0;
        q = d[i] - x - ((i > 0) ? e2 / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) cnt++;
    }
    return cnt;
}

/* Smallest eigenvalue of a symmetric tridiagonal matrix and its unit
 * eigenvector: bisection on Sturm counts to machine precision, then inverse
 * iteration with partial pivoting.  This stands in for LAPACK dstev at
 * groundStateLieb.c:435 (MKL, unpinned version); parity at that boundary is
 * pinned against numpy.linalg.eigh instead. */
double orc_tridiag_min(int n, const double *d, const double *e, double *z)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; i++) {
        double r = (i > 0 ? fabs(e[i - 1]) : 0) + (i < n - 1 ? fabs(e[i]) : 0);
        if (d[i] - r < lo) lo = d[i] - r;
        if (d[i] + r > hi) hi = d[i] + r;
    }
    double span = hi - lo + 1e-300;
    lo -= 1e-12 * span; hi += 1e-12 * span;
    for (int it = 0; it < 400; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_below(n, d, e, mid) >= 1) hi = mid; else lo = mid;
    }
    double lam = 0.5 * (lo + hi);
    if (z) {
        /* inverse iteration: (T - lam' I) x = b, tridiagonal LU with pivoting */
        double shift = lam - 1e-14 * (fabs(lam) + span);
        double *a = (double *)malloc(sizeof(double) * n * 4);
        double *dl = a, *dd = a + n, *du = a + 2 * n, *du2 = a + 3 * n;
        int *piv = (int *)malloc(sizeof(int) * n);
        for (int i = 0; i < n; i++) z[i] = 1.0 / sqrt((double)n) * (1.0 + 0.01 * i);
        for (int sweep = 0; sweep < 3; sweep++) {
            for (int i = 0; i < n; i++) {
                dd[i] = d[i] - shift;
                du[i] = (i < n - 1) ? e[i] : 0;
                dl[i] = (i < n - 1) ? e[i] : 0;
                du2[i] = 0;
            }
            for (int i = 0; i < n - 1; i++) {
                if (fabs(dd[i]) >= fabs(dl[i])) {
                    piv[i] = i;
                    double f = (dd[i] != 0) ? dl[i] / dd[i] : 0;
                    dl[i] = f;
                    dd[i + 1] -= f * du[i];
                } else {
                    piv[i] = i + 1;
                    double f = dd[i] / dl[i];
                    dd[i] = dl[i];
                    dl[i] = f;
                    double t = du[i];
                    du[i] = dd[i + 1];
                    dd[i + 1] = t - f * dd[i + 1];
                    if (i < n - 2) { du2[i] = du[i + 1]; du[i + 1] = -f * du[i + 1]; }
                }
            }
            for (int i = 0; i < n; i++)
                if (dd[i] == 0) dd[i] = 1e-300;
            for (int i = 0; i < n - 1; i++) {
                if (piv[i] == i) z[i + 1] -= dl[i] * z[i];
                else { double t = z[i]; z[i] = z[i + 1]; z[i + 1] = t - dl[i] * z[i + 1]; }
            }
            z[n - 1] /= dd[n - 1];
            if (n > 1) z[n - 2] = (z[n - 2] - du[n - 2] * z[n - 1]) / dd[n - 2];
            for (int i = n - 3; i >= 0; i--)
                z[i] = (z[i] - du[i] * z[i + 1] - du2[i] * z[i + 2]) / dd[i];
            double s = 0;
            for (int i = 0; i < n; i++) s += z[i] * z[i];
            s = sqrt(s);
            for (int i = 0; i < n; i++) z[i] /= s;
        }
        free(a);
        free(piv);
    }
    return lam;
}

/* Ground state.  Restates ground, groundStateLieb.c:348-471: Lanczos of
 * length lm from C, lowest eigenpair of the real tridiagonal, Ritz vector
 * C = sum_k lvec[k] z[k] written back into C.  Returns E0. */
double orc_ground(int lm, int N, int M, const int32_t *Map, const int32_t *MapOT,
                  const int32_t *MapTT, const int64_t *strideOT,
                  const int64_t *strideTT, const int32_t *IF, const cplx *Ho,
                  const cplx *Hint, cplx *C, int *iters_out)
{
    int64_t nc = orc_nc(N, M);
    cplx *diag = (cplx *)calloc(lm, sizeof(cplx));
    cplx *offd = (cplx *)calloc(lm, sizeof(cplx));
    cplx *lvec = (cplx *)malloc(sizeof(cplx) * (int64_t)lm * nc);
    double *d = (double *)malloc(sizeof(double) * lm);
    double *e = (double *)malloc(sizeof(double) * lm);
    double *z = (double *)malloc(sizeof(double) * lm);
    memcpy(lvec, C, sizeof(cplx) * nc);
    int n = orc_lanczos(lm, N, M, Map, MapOT, MapTT, strideOT, strideTT, IF, Ho,
                        Hint, diag, offd, lvec);
    for (int k = 0; k < n; k++) { d[k] = creal(diag[k]); e[k] = creal(offd[k]); }
    double E0 = orc_tridiag_min(n, d, e, z);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; i++) {
        cplx s = 0;
        for (int k = 0; k < n; k++) s += lvec[(int64_t)k * nc + i] * z[k];
        C[i] = s;
    }
    if (iters_out) *iters_out = n;
    free(diag); free(offd); free(lvec); free(d); free(e); free(z);
    return E0;
}

/* ------------------------------------------------------------------------ */
/* Model inputs                                                             */
/* ------------------------------------------------------------------------ */

/* Lieb-Liniger one-body matrix: diagonal (2 pi (n - M/2) / L)^2 / 2.
 * groundStateLieb.c:129-155 (M/2 is integer division). */
void orc_ll_ho(int M, double L, cplx *Ho)
{
    const double PI = 3.141592653589793;
    for (int n = 0; n < M; n++)
        for (int j = 0; j < M; j++) Ho[n + M * j] = 0;
    for (int n = 0; n < M; n++) {
        double k = 2 * PI * (n - M / 2) / L;
        Ho[n + M * n] = 0.5 * k * k;
    }
}

/* Lieb-Liniger contact interaction: Hint = g when momentum is conserved
 * (n + s == q + l), else 0.  groundStateLieb.c:159-213. */
void orc_ll_hint(int M, double g, cplx *Hint)
{
    for (int n = 0; n < M; n++)
        for (int s = 0; s < M; s++)
            for (int q = 0; q < M; q++)
                for (int l = 0; l < M; l++)
                    Hint[IDX4(n, s, q, l, M)] = ((l - M / 2) + (q - M / 2) - (s - M / 2) - (n - M / 2) == 0) ? g : 0.0;
}

/* Deterministic start vector C_i = sin(20 i/nc)(i%13) + i((i%8)-(i%3)),
 * normalized.  groundStateLieb.c:578-586 and performanceTest.c:169-177. */
void orc_start_vector(int64_t nc, cplx *C)
{
    double sum = 0;
    for (int64_t i = 0; i < nc; i++) {
        double re = sin(20 * ((double)i) / nc) * (i % 13);
        double im = (double)((i % 8) - (i % 3));
        C[i] = re + I * im;
        sum += re * re + im * im;
    }
    double nrm = sqrt(sum);
    for (int64_t i = 0; i < nc; i++) C[i] /= nrm;
}

/* Deterministic Hermitian one-body matrix of the timing harness.
 * performanceTest.c:181-189 (stored as Ho[k + M*l] = Ho_ref[k][l]). */
void orc_perf_ho(int M, cplx *Ho)
{
    for (int i = 0; i < M; i++) {
        Ho[i + M * i] = (i % 4) - 1;
        for (int j = i + 1; j < M; j++) {
            cplx h = (double)(i * (j % 3) - (i % 4) + 5 * (j % 2)) - I * (4.123 * i / (j + 1));
            Ho[i + M * j] = h;
            Ho[j + M * i] = conj(h);
        }
    }
}

/* Deterministic fully symmetric two-body matrix of the timing harness.
 * performanceTest.c:192-219 (all 1.234, then the 8-entry orbit of every
 * disjoint (i<j, q<l) set to z and conj(z), later loops overwriting). */
void orc_perf_hint(int M, cplx *Hint)
{
    int64_t M4 = (int64_t)M * M * M * M;
    for (int64_t t = 0; t < M4; t++) Hint[t] = 1.234;
    for (int i = 0; i < M; i++)
        for (int j = i + 1; j < M; j++)
            for (int q = 0; q < M; q++) {
                if (q == i || q == j) continue;
                for (int l = q + 1; l < M; l++) {
                    if (l == i || l == j) continue;
                    cplx z = (double)(i - 2 + 10 * (j % (i + 1)) - q * l) +
                             I * (((double)i * q - j * l) / M);
                    Hint[IDX4(i, j, q, l, M)] = z;
                    Hint[IDX4(i, j, l, q, M)] = z;
                    Hint[IDX4(j, i, l, q, M)] = z;
                    Hint[IDX4(j, i, q, l, M)] = z;
                    Hint[IDX4(q, l, i, j, M)] = conj(z);
                    Hint[IDX4(q, l, j, i, M)] = conj(z);
                    Hint[IDX4(l, q, i, j, M)] = conj(z);
                    Hint[IDX4(l, q, j, i, M)] = conj(z);
                }
            }
}

/* number of OpenMP threads the oracle uses (CPU-baseline reporting) */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

double orc_wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}
