/*
 * mcpi_hip.h -- C ABI of libmcpi_hip.so, an MI355X (gfx950) implementation of
 * the mcpi matrix-free Fock-space hot path: perfect hashing of bosonic
 * occupation vectors, single/double jump maps, H.v, one- and two-body
 * density matrices, and the Lanczos ground state.
 *
 * Two layers:
 *
 *  1. Reference-shaped entry points (same names, argument order and meaning
 *     as andriati-alex/mcpi's headers; host pointers; synchronous; results in
 *     host memory).  A program written against the reference includes this
 *     header instead of configurationsMap.h / hamiltonianMatrix.h /
 *     onebodyMatrix.h / twobodyMatrix.h, links -lmcpi_hip, and runs the hot
 *     loops on the GPU.  Like the reference they print a message and exit()
 *     on the reference's INT32 overflow guards.
 *
 *  2. The extended API (prefix mcpi_hip_): device-resident bases and
 *     Hamiltonians as opaque handles, int64 offsets (bases beyond the
 *     reference's INT32 limits), device pointers and streams, batched
 *     Lanczos and the Hellmann-Feynman adjoint.  Every call returns an
 *     mcpi_status (0 = success) and never exits; mcpi_hip_last_error() gives
 *     the message of the last failure on the calling thread.
 *
 * Layouts (shared by both layers; "first index fastest"):
 *   IF[i*M + j]                       occupation of orbital j in state i
 *   NCmat[n + (N+1)*m]                NC(n, m) = C(n+m-1, n); column 0 is 0
 *   Map[i + k*nc + l*M*nc]            index of v_i - e_k + e_l, -1 if v_k = 0
 *   MapOT[strideOT[i] + c*M*M + l + q*M]  v_i - 2e_k + e_l + e_q, c = ordinal
 *                                     of k among orbitals with v >= 2
 *   MapTT[strideTT[i] + c*M*M + l + q*M]  v_i - e_k - e_s + e_l + e_q, c =
 *                                     lexicographic ordinal of the pair k<s
 *                                     among occupied pairs
 *   Ho (flat) [k + M*l]               coefficient of a*_k a_l
 *   Hint      [k + s*M + q*M^2 + l*M^3]  coefficient of a*_k a*_s a_q a_l
 *   rho1 (flat) [k + M*l] = <a*_k a_l>;  rho2 [k + sM + qM^2 + lM^3] =
 *                                     <a*_k a*_s a_q a_l>
 * Complex numbers are two doubles (re, im): C99 double _Complex, C++
 * std::complex<double> and hipDoubleComplex all have this layout.
 * H = sum Ho[k,l] a*_k a_l + 1/2 sum Hint[k,s,q,l] a*_k a*_s a_q a_l, with the
 * reference's folded two-body rules (exact for a Hint with full bosonic
 * symmetry Hint[k,s,q,l] = Hint[s,k,q,l] = Hint[k,s,l,q] = conj(Hint[q,l,k,s])).
 */
#ifndef MCPI_HIP_H
#define MCPI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> mcpi_cplx;
extern "C" {
#else
#include <complex.h>
typedef double _Complex mcpi_cplx;
#endif

/* Reference typedefs, configurationsMap.h:27-40 */
typedef double *Rarray;
typedef mcpi_cplx *Carray;
typedef mcpi_cplx **Cmatrix;
typedef int *Iarray;

/* ======================================================================
 * 1. Reference-shaped entry points
 * ====================================================================== */

/* configurationsMap.h:145  NC(N, M) = C(N+M-1, N); exits on INT32 overflow */
int NC(int N, int M);
/* configurationsMap.h:206  malloc'ed (N+1)*(M+1) table, free() it */
Iarray setupNCmat(int N, int M);
/* configurationsMap.h:238  unrank k into v[M] */
void IndexToFock(int k, int N, int M, Iarray v);
/* configurationsMap.h:273  rank of v */
int FockToIndex(int N, int M, Iarray NCmat, Iarray v);
/* configurationsMap.h:309  malloc'ed nc*M IF table (built on the GPU) */
Iarray setupFocks(int N, int M);
/* configurationsMap.h:348  malloc'ed nc*M*M single-jump map (built on the GPU) */
Iarray OneOneMap(int N, int M, Iarray NCmat, Iarray IF);
/* configurationsMap.h:478  malloc'ed two-orbital double-jump map; fills strideC[nc] */
Iarray TwoTwoMap(int N, int M, Iarray NCmat, Iarray IF, Iarray strideC);
/* configurationsMap.h:668  malloc'ed same-orbital double-jump map; fills strideC[nc] */
Iarray OneTwoMap(int N, int M, Iarray NCmat, Iarray IF, Iarray strideC);

/* hamiltonianMatrix.h:1431  out = H C with all maps (host arrays; H.v on the GPU) */
void applyHconf_omp(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT,
                    Iarray strideOT, Iarray strideTT, Iarray IF, Carray C,
                    Cmatrix Ho, Carray Hint, Carray out);
/* hamiltonianMatrix.h:891  same computation as applyHconf_omp */
void applyHconf_XX(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT,
                   Iarray strideOT, Iarray strideTT, Iarray IF, Carray C,
                   Cmatrix Ho, Carray Hint, Carray out);
/* hamiltonianMatrix.h:431  same computation, Map and MapOT given (double
 * jumps are ranked on the fly; MapOT holds the same targets) */
void applyHconf_X(int N, int M, Iarray Map, Iarray MapOT, Iarray strideOT,
                  Iarray NCmat, Iarray IF, Carray C, Cmatrix Ho, Carray Hint,
                  Carray out);
/* hamiltonianMatrix.h:31  hash-only H.v (double jumps ranked on the fly) */
void applyHconf(int N, int M, Iarray NCmat, Iarray IF, Carray C, Cmatrix Ho,
                Carray Hint, Carray out);
/* onebodyMatrix.h:34  rho[k][l] = <a*_k a_l>, no table passed (the basis is
 * generated on the GPU from N, M) */
void OBrho(int N, int M, Iarray NCmat, Carray C, Cmatrix rho);
/* onebodyMatrix.h:116  same, IF given (the reference edits IF in place and
 * restores it; this one only reads it) */
void OBrho_X(int N, int M, Iarray NCmat, Iarray IF, Carray C, Cmatrix rho);
/* onebodyMatrix.h:190  same, IF and Map given */
void OBrho_XM(int N, int M, Iarray Map, Iarray NCmat, Iarray IF, Carray C,
              Cmatrix rho);
/* twobodyMatrix.h:38  rho[k + sM + qM^2 + lM^3] = <a*_k a*_s a_q a_l>, IF given */
void TBrho(int N, int M, Iarray NCmat, Iarray IF, Carray C, Carray rho);
/* twobodyMatrix.h:502  same, IF, Map and MapOT given (double jumps ranked) */
void TBrho_X(int N, int M, Iarray Map, Iarray MapOT, Iarray strideOT,
             Iarray NCmat, Iarray IF, Carray C, Carray rho);
/* twobodyMatrix.h:991  same, every map given */
void TBrho_XX(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT,
              Iarray strideOT, Iarray strideTT, Iarray NCmat, Iarray IF,
              Carray C, Carray rho);
/* groundStateLieb.c:217  Lanczos with full reorthogonalization; lvec[lm][nc],
 * lvec[0] = normalized start vector.  Returns the number of steps kept. */
int lanczos(int lm, int Npar, int Morb, Iarray IF, Iarray strideOT,
            Iarray strideTT, Iarray Map, Iarray MapOT, Iarray MapTT,
            Cmatrix Ho, Carray Hint, Carray diag, Carray offdiag, Cmatrix lvec);
/* groundStateLieb.c:348  lowest eigenpair; C in: start vector, out: Ritz vector */
double ground(int Niter, int Npar, int Morb, Iarray IF, Iarray strideOT,
              Iarray strideTT, Iarray Map, Iarray MapOT, Iarray MapTT,
              Cmatrix Ho, Carray Hint, Carray C);
/* groundStateLieb.c:73, :89  (the C++ view of the complex return value has
 * the same x86-64 ABI as C's double _Complex: two doubles in xmm0/xmm1) */
#if defined(__cplusplus) && defined(__clang__)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wreturn-type-c-linkage"
#endif
mcpi_cplx carrDot(int n, Carray v1, Carray v2);
#if defined(__cplusplus) && defined(__clang__)
#pragma clang diagnostic pop
#endif
double carrMod(int n, Carray v);
/* groundStateLieb.c:129, :159  Lieb-Liniger model on a ring of length L */
void setupHo(int Morb, double L, Cmatrix Ho);
void setupHint(int Morb, double g, Carray Hint);

/* ======================================================================
 * 2. Extended API
 * ====================================================================== */

typedef int mcpi_status;
#define MCPI_OK 0
#define MCPI_ERR_ARG 1       /* invalid argument or size beyond a limit   */
#define MCPI_ERR_HIP 2       /* HIP runtime error (message has details)   */
#define MCPI_ERR_ALLOC 3     /* device or host allocation failed          */
#define MCPI_ERR_STATE 4     /* handle in the wrong state for the call    */

/* storage modes of a basis (the reference's memory-for-time ladder,
 * hamiltonianMatrix.h:31 / :431 / :891) */
#define MCPI_MAPS_NONE 0     /* IF only: every jump ranked on the fly       */
#define MCPI_MAPS_SINGLE 1   /* IF + Map: double jumps ranked on the fly    */
#define MCPI_MAPS_FULL 2     /* IF + Map + MapOT + MapTT                    */

typedef struct mcpi_basis_s *mcpi_basis;
typedef struct mcpi_ham_s *mcpi_ham;

typedef struct {
    int N, M, maps;
    int64_t nc;              /* number of configurations                    */
    int64_t map_len;         /* nc*M*M or 0                                 */
    int64_t ot_len, tt_len;  /* |MapOT|, |MapTT| in ints (0 if not stored)  */
    int64_t device_bytes;    /* device memory held by the basis             */
    int64_t pot_len, ptt_len; /* ints of the wave-packed OT / TT tables that
                                 the H.v kernels read (full maps only; the
                                 layout is given at mcpi_hip_basis_copy_range) */
    int64_t pattern_bytes;    /* device memory of the pattern tables built so
                                 far for H.v schedule 4 (all zero patterns) */
    int64_t nsectors;         /* charge sectors of a sector-ordered basis (0:
                                 reference order) */
    int64_t largest_sector;   /* rows of its largest sector */
} mcpi_basis_info;

const char *mcpi_hip_last_error(void);
const char *mcpi_hip_version(void);
/* The reference-shaped entry points keep device state between calls: the
 * tables uploaded from the caller's arrays (reused while the same arrays are
 * passed again, which are taken to be unchanged), the Hamiltonian of the
 * last Ho/Hint with its tuned H.v schedule (reused while Ho and Hint are
 * equal entry for entry) and two device vectors.  This frees all of it, e.g.
 * after the caller edited tables in place or before it ends. */
mcpi_status mcpi_hip_release_cached(void);

/* device memory and synchronization helpers (HIP runtime of the library) */
mcpi_status mcpi_hip_malloc(void **dptr, size_t bytes);
mcpi_status mcpi_hip_free(void *dptr);
mcpi_status mcpi_hip_memcpy_h2d(void *dst, const void *src, size_t bytes);
mcpi_status mcpi_hip_memcpy_d2h(void *dst, const void *src, size_t bytes);
mcpi_status mcpi_hip_memcpy_d2d(void *dst, const void *src, size_t bytes);
mcpi_status mcpi_hip_memset(void *dptr, int value, size_t bytes);
mcpi_status mcpi_hip_synchronize(void);
mcpi_status mcpi_hip_device_info(char *name, size_t name_len, int *cu_count,
                                 size_t *total_mem);

/* Build a device-resident basis (tables built on the GPU, int64 offsets). */
mcpi_status mcpi_hip_basis_create(int N, int M, int maps, mcpi_basis *out);
mcpi_status mcpi_hip_basis_destroy(mcpi_basis b);
mcpi_status mcpi_hip_basis_info(mcpi_basis b, mcpi_basis_info *info);
/* copies of the tables to host memory (sizes from mcpi_basis_info); IF in the
 * reference order (also for a sector-ordered basis, whose map is in row order
 * and is copied with mcpi_hip_basis_copy_range only) */
mcpi_status mcpi_hip_basis_copy_ncmat(mcpi_basis b, int32_t *NCmat);
mcpi_status mcpi_hip_basis_copy_if(mcpi_basis b, int32_t *IF);
mcpi_status mcpi_hip_basis_copy_map(mcpi_basis b, int32_t *Map);
mcpi_status mcpi_hip_basis_copy_ot(mcpi_basis b, int64_t *strideOT, int32_t *MapOT);
mcpi_status mcpi_hip_basis_copy_tt(mcpi_basis b, int64_t *strideTT, int32_t *MapTT);
/* count entries of one table from offset (which: 0 Map, 1 MapOT, 2 MapTT,
 * 3 IF, 4 strideOT, 5 strideTT, 6 packOT, 7 packTT, 8 poffOT, 9 poffTT,
 * and for sector-ordered bases 10 perm (reference index -> row), 11 iperm
 * (row -> reference index), 12 IF in the reference order; strides and
 * packed offsets are int64, the others int32).  Tables are in the basis's
 * row order and hold row indices.
 * Wave-packed tables (full maps): wavefront w = rows 64w .. 64w+63; for each
 * jump active in at least one of its rows (OT: k with v_k >= 2, ascending;
 * TT: pairs k < s both occupied, lexicographic) one slot of E entries per
 * active row, starting at int offset poff[w] and following each other:
 * slot[g][a][u] = target of entry t = 4g + u of the a-th active row.  A slot
 * holds the entries q <= l of the chunk with neither q nor l a removed
 * orbital of the jump (k, or k and s; the others are no double-jump term),
 * numbered over the MP = M-1 (OT) or M-2 (TT) remaining orbitals
 * o' = o - [k < o] - [s < o]: t = l'(l'+1)/2 + q'; E = MP(MP+1)/2 rounded
 * up to a multiple of four, the padding holding the row's own index.  The
 * target of (q, l) is the reference map's entry l + q*M of that chunk. */
mcpi_status mcpi_hip_basis_copy_range(mcpi_basis b, int which, int64_t offset,
                                      int64_t count, void *dst);
/* Sector order.  A Hamiltonian that conserves an additive charge
 * Q(v) = sum_j q_j v_j (Ho[k,l] != 0 only where q_k = q_l, Hint[k,s,q,l] != 0
 * only where q_k + q_s = q_q + q_l) is block-diagonal over the sectors of
 * equal Q; the Lieb-Liniger model conserves the total momentum, q_j = j
 * (groundStateLieb.c:143-153, 200-207).  A sector-ordered basis stores its
 * rows sorted by Q (reference order kept within a sector), so the H.v of a
 * sector gathers only from that sector's slice of the vector.  Every entry
 * point that takes or returns vectors (mcpi_hip_applyH, the Lanczos and
 * ground-state solvers, the densities) keeps the reference order: the
 * solvers move the start vector into row order once and the Ritz vector
 * back once; mcpi_hip_applyH moves both vectors on every call, and
 * mcpi_hip_applyH_rows / mcpi_hip_applyH_timed* work on row-order vectors.
 * Only the single-jump map is stored (maps must be MCPI_MAPS_SINGLE) and
 * H.v runs the zero-skipping kernel on the pattern tables (schedule 4);
 * 3 <= M <= 13.  Hamiltonians that couple two sectors are refused. */
/* charge[j] = j if Ho and Hint conserve the orbital-index sum (MCPI_OK),
 * MCPI_ERR_STATE otherwise */
mcpi_status mcpi_hip_conserved_charge(int M, const mcpi_cplx *Ho, const mcpi_cplx *Hint,
                                      int32_t *charge);
mcpi_status mcpi_hip_basis_create_sectored(int N, int M, int maps, const int32_t *charge,
                                           mcpi_basis *out);
/* number of sectors and their row offsets (offsets: nsectors + 1 entries,
 * may be NULL); a basis in the reference order reports one sector */
mcpi_status mcpi_hip_basis_sectors(mcpi_basis b, int64_t *nsectors, int64_t *offsets);
/* vector (nc complex, device) from the reference order into row order
 * (to_rows = 1) or back (0), on stream; a plain copy for other bases */
mcpi_status mcpi_hip_basis_reorder(mcpi_basis b, int to_rows, const mcpi_cplx *in_dev,
                                   mcpi_cplx *out_dev, void *stream);

/* rank / unrank of many states on the GPU: v is n*M ints (host) */
mcpi_status mcpi_hip_basis_rank(mcpi_basis b, int64_t n, const int32_t *v, int64_t *idx);
mcpi_status mcpi_hip_basis_unrank(mcpi_basis b, int64_t n, const int64_t *idx, int32_t *v);

/* A Hamiltonian on a basis: Ho (M*M, flat k + M*l) and Hint (M^4), host
 * arrays, copied and pre-folded into the per-rule coefficient tables. */
mcpi_status mcpi_hip_ham_create(mcpi_basis b, const mcpi_cplx *Ho,
                                const mcpi_cplx *Hint, mcpi_ham *out);
mcpi_status mcpi_hip_ham_destroy(mcpi_ham h);
/* Terms whose folded coefficient is exactly zero are skipped (result-neutral
 * for finite inputs; default on).  Off = every term of the storage mode is
 * evaluated and every map entry read, as the reference does. */
mcpi_status mcpi_hip_ham_set_skip_zero(mcpi_ham h, int on);
/* H.v kernel schedule: 0 = auto (the first H.v times the applicable
 * schedules and keeps the fastest), 1 = generic thread-per-row kernel,
 * 2 = kernel specialised on M reading every stored map, 3 = specialised
 * kernel ranking double jumps on the fly even when the maps are stored,
 * 4 = specialised kernel with zero-skipping reading the pattern tables: the
 * targets of the non-zero double-jump terms only, built once per basis and
 * zero pattern on first use and kept with the basis (sparse Hamiltonians
 * with Map stored; falls back to 2 otherwise, or when device memory is
 * short, or on a stream being captured before the tables exist),
 * 5 = every term, with the double jumps of each block of 256 rows
 * compacted into full wavefronts before their targets are read from the
 * wave-packed tables (full maps; falls back to 2 otherwise),
 * 6 = every term through the pair space: the pair amplitudes <u|a_q a_l|C>
 * over the basis with N-2 particles are gathered and multiplied by the
 * Hamiltonian's pair matrix (its one-body part folded in as
 * a*_k a_l = sum_s a*_k a*_s a_s a_l / (N-1)), and each row gathers its
 * M(M+1)/2 products back (any storage mode, reads no
 * map; 3 <= M <= 12, bases in the reference order; a workspace of about
 * 2 M(M+1)/2 NC(N-2, M) complex entries per stream that runs
 * the schedule, allocated on that stream's first use outside stream capture,
 * is kept with the basis; falls back to 2 otherwise or when device memory
 * is short), 7 = 6 in the factorized pair form, for Hamiltonians whose
 * one-body part is diagonal and whose pair matrix is block diagonal with
 * blocks of rank one, W[P][Q] = a[P] b[Q] (contact interactions that conserve
 * a pair charge, such as Lieb-Liniger's): per block r, Z_r = sum_Q b[Q] a_Q C
 * is gathered over the states of N-2 particles and each row gathers its
 * a[P] Z_r back, with no matrix product (3 <= M <= 24, N >= 2; a
 * workspace of R NC(N-2, M) complex entries per stream, R the number of
 * blocks, kept likewise; a smaller workspace stays allocated beside a larger
 * one until the basis is freed; falls back to 2 otherwise).  The autotuner
 * takes 6 and 7 among its candidates only with zero-skipping on, since they
 * read no map, and 7 in place of 6 where it applies.  Where
 * the kernels specialised on M do not exist (M > 13), every fallback is the
 * generic kernel 1 */
mcpi_status mcpi_hip_ham_set_schedule(mcpi_ham h, int schedule);
/* schedule in use (after autotuning; 0 before the first H.v) */
mcpi_status mcpi_hip_ham_get_schedule(mcpi_ham h, int *schedule);

/* out = H C on device pointers (nc complex each), on stream (NULL = default) */
mcpi_status mcpi_hip_applyH(mcpi_ham h, const mcpi_cplx *C_dev, mcpi_cplx *out_dev,
                            void *stream);
/* out = H C with both vectors in the basis's row order (the reference order
 * except for sector-ordered bases), on stream */
mcpi_status mcpi_hip_applyH_rows(mcpi_ham h, const mcpi_cplx *C_dev, mcpi_cplx *out_dev,
                                 void *stream);
/* CUDA-kernel-shaped H.v on raw device tables (cuda/hamiltonianMatrix.cuh:61):
 * every pointer is device memory, Ho flat k + M*l, int64 strides (nc
 * entries).  Every term is evaluated.  Enqueued on stream (NULL = default)
 * after a stream-ordered read of Ho and Hint; the folded coefficients of the
 * last Ho/Hint and the rank table of the last (N, M) are kept and reused. */
mcpi_status mcpi_hip_applyHconf_dev(int N, int M, const int32_t *Map,
                                    const int32_t *MapOT, const int32_t *MapTT,
                                    const int64_t *strideOT, const int64_t *strideTT,
                                    const int32_t *IF, const mcpi_cplx *C,
                                    const mcpi_cplx *Ho, const mcpi_cplx *Hint,
                                    mcpi_cplx *out, void *stream);
/* the same with the CUDA kernel's int32 strides (Iarray strideOT/strideTT,
 * cuda/hamiltonianMatrix.cuh:61-63), widened on the device when read */
mcpi_status mcpi_hip_applyHconf_dev32(int N, int M, const int32_t *Map,
                                      const int32_t *MapOT, const int32_t *MapTT,
                                      const int32_t *strideOT, const int32_t *strideTT,
                                      const int32_t *IF, const mcpi_cplx *C,
                                      const mcpi_cplx *Ho, const mcpi_cplx *Hint,
                                      mcpi_cplx *out, void *stream);
/* average duration (ms) of reps H.v kernels (mcpi_hip_applyH_rows),
 * measured with HIP events on the launch stream */
mcpi_status mcpi_hip_applyH_timed(mcpi_ham h, const mcpi_cplx *C_dev,
                                  mcpi_cplx *out_dev, int reps, void *stream,
                                  double *avg_ms);
/* the same, one duration per call: ms_each[reps] (events recorded between
 * consecutive launches on the stream) */
mcpi_status mcpi_hip_applyH_timed_each(mcpi_ham h, const mcpi_cplx *C_dev,
                                       mcpi_cplx *out_dev, int reps, void *stream,
                                       double *ms_each);

/* density matrices of C (device vector): rho1 M*M, rho2 M^4, host outputs */
mcpi_status mcpi_hip_rho1(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *rho1);
mcpi_status mcpi_hip_rho2(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *rho2);
/* Hellmann-Feynman adjoint at a normalized eigenvector C: dE/dHo = rho1,
 * dE/dHint = rho2 / 2 (host outputs, flat layouts above).  These are the
 * gradients over Hermitian Ho and bosonic Hint (symmetric under k <-> l and
 * under q <-> s, Hint[k,l,q,s] = conj(Hint[q,s,k,l])), the inputs for which
 * the reference's rules apply 1/2 sum Hint a*_k a*_l a_q a_s.  The rules
 * read particular entries of Hint and weight them as if it had that
 * symmetry, so for a Hint with only the weaker symmetry Hint[k,l,q,s] =
 * Hint[l,k,s,q] they apply another operator (the H.v here follows them) and
 * E != sum Ho rho1 + 1/2 sum Hint rho2 in general. */
mcpi_status mcpi_hip_vjp(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *dE_dHo,
                         mcpi_cplx *dE_dHint);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops): the adjoint
 * of mcpi_hip_vjp with device outputs (dE_dHo_dev M*M, dE_dHint_dev M^4),
 * enqueued on stream (NULL = default) without host synchronization.  Its
 * partial sums live in a 16 MiB device array of the library (one per GPU,
 * held until the process ends); it calls no hipMalloc/hipFree.  Every basis
 * the library builds (M <= 32) fits: the Gram tiles take at most 2.9 MiB and
 * the number of block partials is cut to what remains; a basis that would
 * not fit gives MCPI_ERR_ARG.  Calls that overlap on one GPU share that
 * array and race: serialize them. */
mcpi_status mcpi_hip_vjp_noalloc(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *dE_dHo_dev,
                                 mcpi_cplx *dE_dHint_dev, void *stream);

/* Lanczos on the device.  C0_dev: normalized start vector (nc).  lvec_dev:
 * lm*nc device buffer for the basis (row t = v_t).  diag/offdiag: host, lm
 * each.  *steps = size of the tridiagonal actually built. */
mcpi_status mcpi_hip_lanczos(mcpi_ham h, int lm, const mcpi_cplx *C0_dev,
                             mcpi_cplx *lvec_dev, mcpi_cplx *diag, mcpi_cplx *offdiag,
                             int *steps);
/* Ground state: Lanczos of length lm from C0_dev, host tridiagonal solve,
 * Ritz vector into C_out_dev (may alias C0_dev).  *E0 = lowest eigenvalue. */
mcpi_status mcpi_hip_ground(mcpi_ham h, int lm, const mcpi_cplx *C0_dev,
                            mcpi_cplx *C_out_dev, double *E0, int *steps);
/* lowest eigenpair of the real symmetric tridiagonal (d[n], e[n-1]) on the
 * host: implicit QL with Wilkinson shifts; z (n) may be NULL */
mcpi_status mcpi_hip_tridiag_lowest(int n, const double *d, const double *e,
                                    double *lambda, double *z);

/* Batched Lanczos ground states (groundStateLieb.c:348-471 for n instances
 * at once): n Hamiltonian handles (instances may share a basis and differ in
 * Ho/Hint, N, M) and normalized device start vectors; lm steps each, all
 * instances advanced together, one iteration replayed as one hipGraph.
 * Out: E0[n], steps[n] (host); C_out_dev (may be NULL, or hold NULL
 * entries): device Ritz vectors, nc each (may alias the start vectors).
 * The H.v kernel of each group of instances sharing a basis is chosen by
 * timing on the first call and kept for later calls with the same handles
 * and settings. */
mcpi_status mcpi_hip_lanczos_batched(int n, const mcpi_ham *hams,
                                     const mcpi_cplx *const *C0_dev, int lm,
                                     double *E0, int *steps,
                                     mcpi_cplx *const *C_out_dev);
/* H.v kernel of batched groups (instances sharing a basis): 0 = chosen by
 * timing the candidates on the first step (default); 1 = one launch of the
 * specialised kernel over the group; 2 = the multi-vector kernel (one pass
 * over the rows for up to four instances that share the zero pattern of
 * their coefficients), where the group allows it; 3 = the factorized pair
 * form (schedule 7) four members at a time, where every member has it with
 * the same blocks and b factors (Lieb-Liniger couplings of one basis), the
 * timed choice otherwise.  The timed choice (0) includes that candidate.
 * Process-wide. */
mcpi_status mcpi_hip_set_batched_kernel(int kind);

#ifdef __cplusplus
}
#endif
#endif /* MCPI_HIP_H */
