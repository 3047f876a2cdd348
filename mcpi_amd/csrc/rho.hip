// One- and two-body density matrices and the Hellmann-Feynman adjoint.
//
// Reference: OBrho_XM, onebodyMatrix.h:190-251 (rho1), and TBrho_XX,
// twobodyMatrix.h:991-1487 (rho2, rules 1-8 plus symmetry copies).
//
// The reference walks the rows of the basis and adds, for every term its
// maps reach, conj(C_v') C_v times the bosonic factors into the matrix entry
// of the term: about M^4/4 gathered terms per row for rho2.  Here both
// densities are Gram matrices over the basis with fewer particles.  With the
// identity sum_u |u><u| over N-1 or N-2 particles inserted between the
// creators and the annihilators,
//
//   rho1[k + M l]                 = <a*_k a_l>         = sum_u conj(f_u(k)) f_u(l),
//       f_u(k)   = <u| a_k |C>     = sqrt(u_k + 1) C[rank(u + e_k)],          u in B(N-1, M)
//   rho2[a + bM + cM^2 + dM^3]    = <a*_a a*_b a_c a_d> = sum_u conj(g_u(ab)) g_u(cd),
//       g_u(ql)  = <u| a_q a_l |C> = sqrt((u_q + 1)(u_l + 1 + [q = l]))
//                                     C[rank(u + e_q + e_l)],                  u in B(N-2, M)
//
// (bosonic operators commute, so g_u depends on the unordered pair {q, l}).
// Every u + e_q + e_l is a state of the basis, so a row of the matrix
// Phi = [g_u(p)] needs M(M+1)/2 gathers of C and no occupation test, and
// rho2 is the Hermitian product Phi^H Phi: at N = M = 12, 27.5 M gathers and
// 2.1e9 complex multiply-adds (half of them computed, by Hermitian symmetry)
// in place of 2.2e9 gathered terms.  The result is the reference's sum
// regrouped, equal to it up to rounding.
//
// Targets: rank(v) = nc - 1 - sum_{j=1}^{M-1} B[j][P_j] with P_j the
// particles in orbitals below j (closed form of FockToIndex,
// configurationsMap.h:273-305).  Adding e_q + e_l (q <= l) to u raises P_j
// by 1 for q < j <= l and by 2 for j > l, so with the running sums
// S_d(x) = sum_{j=1}^{x} B[j][P_j(u) + d]:
//   rank(u + e_q + e_l) = nc - 1 - S_2(M-1) - (S_0(q) - S_1(q)) - (S_1(l) - S_2(l))
//   rank(u + e_k)       = nc - 1 - S_1(M-1) - (S_0(k) - S_1(k))
// Each row u of a tile is unranked once (greedy over the same B table: the
// combinatorial number system) and its two difference tables are kept in
// LDS, so a gather costs two LDS reads and two subtractions.
//
// Kernel: a block owns one tile pair (TI, TJ) of 80 x 80 entries of the
// Gram matrix and loops over tiles of U rows of Phi (32 for rho2, 128 for
// rho1, whose Phi has at most 32 columns): it unranks the rows (one thread
// per row), gathers the U x 16 NG (x 2 off the diagonal) entries of Phi
// into LDS (lanes over consecutive rows of one column: neighbouring
// targets; GB loads in flight per thread), then every thread adds its
// entry of each 16 x 16 sub-block (the sub-blocks i <= j of a diagonal
// tile pair, all of them off it) with FP64 FMAs.  Each block writes its
// tile once; two fixed-order stages sum the blocks (deterministic result),
// and a last kernel expands the Gram matrix into the full rho1 / rho2
// arrays on the device.  No map is read, so the densities run on every
// storage mode, and the vectors are in the reference order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mcpi_internal.h"

namespace mcpi {

namespace {

constexpr int GT = 80;            // Gram tile side (columns of Phi per tile)
constexpr int GG = GT / 16;       // 16 x 16 sub-blocks per tile side
constexpr int GU2 = 32;           // rows of Phi per LDS tile, rho2
constexpr int GU1 = 128;          // rows of Phi per LDS tile, rho1 (16 or 32 columns)
constexpr int GTHREADS = 256;     // 16 x 16 threads: one entry of each sub-block
constexpr int GTILE = GT * GT;    // entries of one tile
constexpr int GB = 5;             // gathers in flight per thread
constexpr int GSEG = 32;          // segments of the first stage of the block sums

struct GramArgs {
    int N, M, ord;       // particles and orbitals of the basis; annihilators (1: rho1, 2: rho2)
    int np;              // columns of Phi: M (ord 1) or M(M+1)/2 (ord 2)
    int ntile;           // tiles of GT columns
    int rs;              // LDS row stride of a Phi tile in complex entries (16 * groups + 1)
    int ng;              // 16-column groups of a full tile (<= GG)
    int64_t nc;          // states of the basis
    int64_t nu;          // rows of Phi: states of B(N - ord, M)
    int64_t nut;         // tiles of U rows
    const int32_t *B;    // rank table B[j*(N+1) + p]
    const cplx *C;       // vector, reference order
    cplx *partial;       // [tile pair][block][GTILE]
};

__device__ inline cplx ldc(const cplx *p)
{
    double2 t = *reinterpret_cast<const double2 *>(p);
    return cmake(t.x, t.y);
}

// tile pair of an off-diagonal index: TI < TJ, lexicographic
__host__ __device__ inline void off_pair(int idx, int T, int &TI, int &TJ)
{
    TI = 0;
    while (idx >= T - 1 - TI) {
        idx -= T - 1 - TI;
        TI++;
    }
    TJ = TI + 1 + idx;
}

__host__ __device__ inline int off_index(int TI, int TJ, int T)
{
    return TI * (2 * T - TI - 1) / 2 + (TJ - TI - 1);
}

// LDS layout of a block (dynamic shared memory, in this order)
struct GramLds {
    size_t phiA, phiB, tt, dq, dl, sq, bt, ql, pp, bytes;
};

__host__ __device__ inline GramLds gram_lds(int N, int M, int np, int rs, bool diag, int GU, int ord)
{
    GramLds L;
    size_t o = 0;
    auto take = [&o](size_t n, size_t al) {
        o = (o + al - 1) / al * al;
        size_t r = o;
        o += n;
        return r;
    };
    L.phiA = take(sizeof(cplx) * GU * rs, 16);
    L.phiB = diag ? L.phiA : take(sizeof(cplx) * GU * rs, 16);
    L.tt = take(sizeof(int64_t) * GU, 8);
    L.dq = take(sizeof(int64_t) * GU * M, 8);
    L.dl = take(sizeof(int64_t) * GU * M * (ord == 2), 8);
    L.sq = take(sizeof(double) * (N + 3), 8);
    L.bt = take(sizeof(int32_t) * M * (N + 1), 4);
    L.ql = take(sizeof(uint16_t) * np, 2);
    L.pp = take((size_t)GU * (M + 1), 1);
    L.bytes = (o + 15) / 16 * 16;
    return L;
}

// Block partials of one tile pair of Phi^H Phi.  DIAG: tile pair (TI, TI)
// from blockIdx.y, sub-blocks i <= j; else the off-diagonal pair
// off_pair(blockIdx.y), all sub-blocks.
template <bool DIAG, int GU, int NG>
__global__ __launch_bounds__(GTHREADS) __attribute__((amdgpu_waves_per_eu(DIAG ? 3 : 2))) void k_gram(GramArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const GramLds L = gram_lds(a.N, a.M, a.np, a.rs, DIAG, GU, a.ord);
    cplx *phiA = (cplx *)(lds + L.phiA);
    cplx *phiB = (cplx *)(lds + L.phiB);
    int64_t *tt = (int64_t *)(lds + L.tt);
    int64_t *dq = (int64_t *)(lds + L.dq);
    int64_t *dl = (int64_t *)(lds + L.dl);
    double *sq = (double *)(lds + L.sq);
    int32_t *Bt = (int32_t *)(lds + L.bt);
    uint16_t *ql = (uint16_t *)(lds + L.ql);
    uint8_t *pp = lds + L.pp;

    const int t = threadIdx.x;
    const int N = a.N, M = a.M, NB = N + 1, Np = N - a.ord;
    int TI, TJ;
    if (DIAG) TI = TJ = blockIdx.y;
    else off_pair(blockIdx.y, a.ntile, TI, TJ);
    const int tp = DIAG ? TI : a.ntile + (int)blockIdx.y;

    for (int x = t; x < M * NB; x += GTHREADS) Bt[x] = a.B[x];
    for (int x = t; x < N + 3; x += GTHREADS) sq[x] = sqrt((double)x);
    for (int x = t; x < a.np; x += GTHREADS) {
        if (a.ord == 2) {
            int l = 0;
            while ((l + 1) * (l + 2) / 2 <= x) l++;
            ql[x] = (uint16_t)((x - l * (l + 1) / 2) | (l << 8));
        } else {
            ql[x] = (uint16_t)(x | (x << 8));
        }
    }
    // columns of the two sides (NG groups of 16, a.ng == NG; the columns of
    // a narrower last tile past wA / wB hold zeros, so every sub-block is
    // computed without a test)
    const int wA = min(GT, a.np - TI * GT), wB = min(GT, a.np - TJ * GT);
    const int r = t >> 4, c = t & 15;

    cplx acc[NG][NG];
#pragma unroll
    for (int i = 0; i < NG; i++)
#pragma unroll
        for (int j = 0; j < NG; j++) acc[i][j] = cmake(0, 0);

    constexpr int ncolA = 16 * NG, ncolB = DIAG ? 0 : 16 * NG;
    const int nent = GU * (ncolA + ncolB);
    for (int64_t tile = blockIdx.x; tile < a.nut; tile += gridDim.x) {
        __syncthreads();   // the previous tile's Gram reads are done
        const int64_t u0 = tile * GU;
        // rows of the tile: unrank (greedy over B), difference tables
        if (t < GU) {
            const int64_t u = u0 + t;
            uint8_t *P = pp + t * (M + 1);
            if (u < a.nu) {
                // one step per iteration (lower p, or settle orbital j and
                // move down), so the lanes of a wavefront stay together:
                // at most Np + M - 1 iterations
                int64_t R = a.nu - 1 - u;
                int pj = Np, j = M - 1;
                P[M] = (uint8_t)Np;
                while (j >= 1) {
                    const int64_t bj = Bt[j * NB + pj];
                    if (bj > R) {
                        pj--;
                    } else {
                        P[j] = (uint8_t)pj;
                        R -= bj;
                        j--;
                    }
                }
                P[0] = 0;
                int64_t s0 = 0, s1 = 0, s2 = 0;
                int64_t *Dq = dq + t * M;
                Dq[0] = 0;
                if (a.ord == 2) {   // the second table exists for rho2 only
                    int64_t *Dl = dl + t * M;
                    Dl[0] = 0;
                    for (int j = 1; j < M; j++) {
                        const int b = j * NB + P[j];
                        s0 += Bt[b];
                        s1 += Bt[b + 1];
                        s2 += Bt[b + 2];
                        Dq[j] = s0 - s1;
                        Dl[j] = s1 - s2;
                    }
                } else {
                    for (int j = 1; j < M; j++) {
                        const int b = j * NB + P[j];
                        s0 += Bt[b];
                        s1 += Bt[b + 1];
                        Dq[j] = s0 - s1;
                    }
                }
                tt[t] = a.nc - 1 - (a.ord == 2 ? s2 : s1);
            } else {
                tt[t] = -1;
            }
        }
        __syncthreads();
        // gather the tile of Phi: lanes over consecutive rows of one column,
        // GB entries per thread at a time (their targets first, then all
        // their loads in flight together, then the LDS stores: a store to
        // the tile may alias the tables as far as the compiler knows, so
        // one entry at a time would serialize the loads)
        for (int e0 = t; e0 < nent; e0 += GB * GTHREADS) {
            int64_t tg[GB];
            double f[GB];
#pragma unroll
            for (int k = 0; k < GB; k++) {
                const int e = e0 + k * GTHREADS;
                tg[k] = -1;
                f[k] = 0;
                const bool sideB = e >= GU * ncolA;
                const int e2 = sideB ? e - GU * ncolA : e;
                const int col = e2 / GU, ul = e2 % GU;
                const int64_t base = e < nent ? tt[ul] : -1;
                if (base >= 0 && col < (sideB ? wB : wA)) {
                    const int p = (sideB ? TJ : TI) * GT + col;
                    const int q = ql[p] & 255, l = ql[p] >> 8;
                    const uint8_t *P = pp + ul * (M + 1);
                    const int uq = P[q + 1] - P[q];
                    if (a.ord == 2) {
                        const int ulo = P[l + 1] - P[l];
                        tg[k] = base - dq[ul * M + q] - dl[ul * M + l];
                        f[k] = sq[uq + 1] * sq[ulo + 1 + (q == l)];
                    } else {
                        tg[k] = base - dq[ul * M + q];
                        f[k] = sq[uq + 1];
                    }
                }
            }
            cplx v[GB];
#pragma unroll
            for (int k = 0; k < GB; k++) v[k] = tg[k] >= 0 ? ldc(a.C + tg[k]) : cmake(0, 0);
#pragma unroll
            for (int k = 0; k < GB; k++) {
                const int e = e0 + k * GTHREADS;
                if (e < nent) {
                    const bool sideB = e >= GU * ncolA;
                    const int e2 = sideB ? e - GU * ncolA : e;
                    (sideB ? phiB : phiA)[(e2 % GU) * a.rs + e2 / GU] = cscale(v[k], f[k]);
                }
            }
        }
        __syncthreads();
        // Gram: acc[i][j] += conj(Phi[u][16i + r]) Phi[u][16j + c]
        // rho1 tiles (one or two column groups) keep several rows' loads in flight
        constexpr int UN = GU == GU1 ? 4 : 1;
#pragma unroll UN
        for (int ul = 0; ul < GU; ul++) {
            const cplx *ra = phiA + ul * a.rs, *rb = phiB + ul * a.rs;
            cplx av[NG], bv[NG];
#pragma unroll
            for (int i = 0; i < NG; i++) {
                av[i] = ra[16 * i + r];
                bv[i] = rb[16 * i + c];
            }
#pragma unroll
            for (int i = 0; i < NG; i++)
#pragma unroll
                for (int j = 0; j < NG; j++) {
                    if (DIAG && j < i) continue;
                    acc[i][j] = cfma_conj(av[i], bv[j], acc[i][j]);
                }
        }
    }
    cplx *out = a.partial + ((size_t)tp * gridDim.x + blockIdx.x) * GTILE;
#pragma unroll
    for (int i = 0; i < GG; i++)
#pragma unroll
        for (int j = 0; j < GG; j++)
            out[(16 * i + r) * GT + 16 * j + c] = (i < NG && j < NG) ? acc[i < NG ? i : 0][j < NG ? j : 0] : cmake(0, 0);
}

// Block sums in two fixed-order stages.  Stage 1: segment s of the blocks
// (blocks s*seg .. s*seg+seg-1, in order) summed into the partial of its
// first block; stage 2: the segment sums, in order, into G[tp][e].
__global__ void k_gram_sum1(int ntp, int nblocks, int seg, cplx *__restrict__ partial)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)ntp * GTILE) return;
    const int64_t tp = x / GTILE, e = x % GTILE;
    const int b0 = blockIdx.y * seg, b1 = min(nblocks, b0 + seg);
    if (b0 >= b1) return;
    cplx *p = partial + tp * nblocks * GTILE + e;
    double re = 0, im = 0;
    for (int b = b0; b < b1; b++) {
        const cplx z = p[(int64_t)b * GTILE];
        re += z.re;
        im += z.im;
    }
    p[(int64_t)b0 * GTILE] = cmake(re, im);
}

__global__ void k_gram_sum2(int ntp, int nblocks, int seg, const cplx *__restrict__ partial, cplx *__restrict__ G)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)ntp * GTILE) return;
    const int64_t tp = x / GTILE, e = x % GTILE;
    const cplx *p = partial + tp * nblocks * GTILE + e;
    double re = 0, im = 0;
    for (int b = 0; b < nblocks; b += seg) {
        const cplx z = p[(int64_t)b * GTILE];
        re += z.re;
        im += z.im;
    }
    G[x] = cmake(re, im);
}

// entry (p1, p2) of the Hermitian Gram matrix held as the tiles of its upper
// triangle (T tiles a side); a diagonal entry is real
__device__ inline cplx gram_entry(const cplx *G, int T, int p1, int p2)
{
    const bool lower = p1 > p2;
    const int a = lower ? p2 : p1, b = lower ? p1 : p2;
    const int TI = a / GT, TJ = b / GT;
    const int tp = TI == TJ ? TI : T + off_index(TI, TJ, T);
    const cplx z = G[(size_t)tp * GTILE + (a % GT) * GT + b % GT];
    if (a == b) return cmake(z.re, 0);
    return lower ? cmake(z.re, -z.im) : z;
}

__device__ inline int pair_of(int x, int y)
{
    const int lo = x < y ? x : y, hi = x < y ? y : x;
    return hi * (hi + 1) / 2 + lo;
}

// rho1[k + M l] = G1(k, l) and rho2[a + bM + cM^2 + dM^3] = s2 * G2({a,b}, {c,d});
// a null G gives zeros (fewer particles than annihilators)
__global__ void k_gram_expand(int M, int T1, const cplx *__restrict__ G1, cplx *__restrict__ rho1, int T2,
                              const cplx *__restrict__ G2, cplx *__restrict__ rho2, double s2)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t M2 = (int64_t)M * M;
    if (rho1 && x < M2) rho1[x] = G1 ? gram_entry(G1, T1, (int)(x % M), (int)(x / M)) : cmake(0, 0);
    if (!rho2 || x >= M2 * M2) return;
    const int a = (int)(x % M), b = (int)(x / M % M), c = (int)(x / M2 % M), d = (int)(x / (M2 * M));
    rho2[x] = G2 ? cscale(gram_entry(G2, T2, pair_of(a, b), pair_of(c, d)), s2) : cmake(0, 0);
}

}  // namespace

// launch geometry of one Gram product
struct GramPlan {
    GramArgs a;
    int U = GU2;         // rows of Phi per LDS tile
    int ntp = 0;         // tile pairs: ntile diagonal, then the off-diagonal ones
    int nblocks = 0;     // blocks per tile pair
    size_t sm_diag = 0, sm_off = 0;
    size_t partial_entries() const { return (size_t)ntp * nblocks * GTILE; }
    size_t gram_entries() const { return (size_t)ntp * GTILE; }
};

// plan Phi^H Phi for ord annihilators on basis b; nblocks at most max_blocks
static mcpi_status plan_gram(const BasisDev &b, int ord, int max_blocks, GramPlan *pl)
{
    GramArgs &a = pl->a;
    a.N = b.N;
    a.M = b.M;
    a.ord = ord;
    a.np = ord == 2 ? b.M * (b.M + 1) / 2 : b.M;
    a.ntile = (a.np + GT - 1) / GT;
    a.ng = std::min(GG, (a.np + 15) / 16);
    a.rs = 16 * a.ng + 1;   // odd stride: the lanes of a gather write rows on distinct banks
    a.nc = b.nc;
    a.nu = b.N >= ord ? nc_of(b.N - ord, b.M) : 0;
    pl->U = ord == 2 ? GU2 : GU1;
    a.nut = (a.nu + pl->U - 1) / pl->U;
    a.B = b.B;
    a.C = nullptr;
    a.partial = nullptr;
    pl->ntp = a.ntile * (a.ntile + 1) / 2;
    int dev = 0, ncu = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // about three blocks per CU over all tile pairs, at most one per row tile
    const int64_t want = std::max<int64_t>(1, (int64_t)3 * ncu / pl->ntp);
    pl->nblocks = (int)std::max<int64_t>(1, std::min<int64_t>({want, a.nut, (int64_t)max_blocks}));
    pl->sm_diag = gram_lds(a.N, a.M, a.np, a.rs, true, pl->U, ord).bytes;
    pl->sm_off = gram_lds(a.N, a.M, a.np, a.rs, false, pl->U, ord).bytes;
    MCPI_ARG((a.ntile > 1 ? pl->sm_off : pl->sm_diag) <= 160 * 1024, "density-matrix tiles exceed LDS");
    return MCPI_OK;
}

template <bool DIAG, int U, int NG>
static mcpi_status launch_gram(const GramPlan &pl, int ny, size_t sm, hipStream_t s)
{
    (void)hipFuncSetAttribute((const void *)k_gram<DIAG, U, NG>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    hipLaunchKernelGGL((k_gram<DIAG, U, NG>), dim3(pl.nblocks, ny), dim3(GTHREADS), sm, s, pl.a);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// the block-partial kernels of a plan: rho1 (U = GU1, one tile of at most
// 32 columns) or rho2 (U = GU2: the diagonal tile pairs, then the others)
static mcpi_status launch_gram_kernels(const GramPlan &pl, hipStream_t s)
{
    const int T = pl.a.ntile;
    if (pl.U == GU1) {
        MCPI_ARG(T == 1 && pl.a.ng <= 2, "one-body Gram matrix wider than two column groups");
        return pl.a.ng == 1 ? launch_gram<true, GU1, 1>(pl, 1, pl.sm_diag, s)
                            : launch_gram<true, GU1, 2>(pl, 1, pl.sm_diag, s);
    }
    switch (pl.a.ng) {
    case 1: MCPI_CHECK((launch_gram<true, GU2, 1>(pl, T, pl.sm_diag, s))); break;
    case 2: MCPI_CHECK((launch_gram<true, GU2, 2>(pl, T, pl.sm_diag, s))); break;
    case 3: MCPI_CHECK((launch_gram<true, GU2, 3>(pl, T, pl.sm_diag, s))); break;
    case 4: MCPI_CHECK((launch_gram<true, GU2, 4>(pl, T, pl.sm_diag, s))); break;
    default: MCPI_CHECK((launch_gram<true, GU2, GG>(pl, T, pl.sm_diag, s))); break;
    }
    // off-diagonal tile pairs exist only when the columns fill more than one tile
    if (T > 1) MCPI_CHECK((launch_gram<false, GU2, GG>(pl, T * (T - 1) / 2, pl.sm_off, s)));
    return MCPI_OK;
}

// Phi^H Phi into G (pl.gram_entries()) through partial, on stream s
static mcpi_status run_gram(GramPlan pl, const cplx *C, cplx *partial, cplx *G, hipStream_t s)
{
    pl.a.C = C;
    pl.a.partial = partial;
    MCPI_CHECK(launch_gram_kernels(pl, s));
    const int64_t n = (int64_t)pl.ntp * GTILE;
    const int seg = (pl.nblocks + GSEG - 1) / GSEG, nseg = (pl.nblocks + seg - 1) / seg;
    hipLaunchKernelGGL(k_gram_sum1, dim3((unsigned)((n + 255) / 256), (unsigned)nseg), dim3(256), 0, s, pl.ntp,
                       pl.nblocks, seg, partial);
    MCPI_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gram_sum2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pl.ntp, pl.nblocks, seg,
                       (const cplx *)partial, G);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// Densities of C on the device: rho1 (M*M, if want1) and / or rho2 (M^4,
// scaled by s2, if want2) into device memory, on stream s.  ws: device
// workspace of ws_entries complex entries (null: only report the need in
// *need); the number of blocks shrinks to fit a given workspace.
static mcpi_status densities_dev(const BasisDev &b, const cplx *C, bool want1, bool want2, cplx *rho1, cplx *rho2,
                                 double s2, cplx *ws, size_t ws_entries, size_t *need, hipStream_t s)
{
    const int M = b.M;
    GramPlan p1, p2;
    const bool do1 = want1 && b.N >= 1, do2 = want2 && b.N >= 2;
    int maxb = 1 << 20;
    MCPI_CHECK(plan_gram(b, 1, maxb, &p1));
    MCPI_CHECK(plan_gram(b, 2, maxb, &p2));
    const size_t g1 = do1 ? p1.gram_entries() : 0, g2 = do2 ? p2.gram_entries() : 0;
    auto partials = [&](const GramPlan &a, const GramPlan &c) {
        return std::max(do1 ? a.partial_entries() : 0, do2 ? c.partial_entries() : 0);
    };
    if (!ws) {
        *need = g1 + g2 + partials(p1, p2);
        return MCPI_OK;
    }
    if (g1 + g2 + partials(p1, p2) > ws_entries) {
        // fewer blocks: the partials take what the Gram tiles leave
        MCPI_ARG(ws_entries > g1 + g2, "density workspace too small");
        const size_t room = ws_entries - g1 - g2;
        const size_t per1 = (size_t)p1.ntp * GTILE, per2 = (size_t)p2.ntp * GTILE;
        const size_t per = std::max(do1 ? per1 : 0, do2 ? per2 : 0);
        MCPI_ARG(per > 0 && room >= per, "density workspace too small");
        maxb = (int)std::min<size_t>(room / per, 1 << 20);
        MCPI_CHECK(plan_gram(b, 1, maxb, &p1));
        MCPI_CHECK(plan_gram(b, 2, maxb, &p2));
    }
    cplx *G1 = ws, *G2 = ws + g1, *partial = ws + g1 + g2;
    if (do1) MCPI_CHECK(run_gram(p1, C, partial, G1, s));
    if (do2) MCPI_CHECK(run_gram(p2, C, partial, G2, s));
    const int64_t M2 = (int64_t)M * M, n = want2 ? M2 * M2 : M2;
    hipLaunchKernelGGL(k_gram_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, p1.a.ntile,
                       do1 ? (const cplx *)G1 : nullptr, want1 ? rho1 : nullptr, p2.a.ntile,
                       do2 ? (const cplx *)G2 : nullptr, want2 ? rho2 : nullptr, s2);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// The tables the density kernels read: rank table and sizes only; vectors
// are in the reference order also for a sector-ordered basis.
static BasisDev density_view(mcpi_basis bs)
{
    BasisDev b = bs->d;
    b.perm = nullptr;
    return b;
}

// rho1 / rho2 (rho2 scaled by s2) into host arrays; either may be null
static mcpi_status densities(mcpi_basis bs, const cplx *C, cplx *rho1, cplx *rho2, double s2)
{
    std::lock_guard<std::mutex> lock(bs->rho_mu);
    const BasisDev b = density_view(bs);
    const int64_t M2 = (int64_t)b.M * b.M, M4 = M2 * M2;
    size_t need = 0;
    MCPI_CHECK(densities_dev(b, C, rho1, rho2, nullptr, nullptr, s2, nullptr, 0, &need, 0));
    const size_t total = need + (rho1 ? M2 : 0) + (rho2 ? M4 : 0);
    if (bs->rho_ws_bytes < total * sizeof(cplx)) {
        if (bs->rho_ws) MCPI_HIP(hipFree(bs->rho_ws));
        bs->rho_ws = nullptr;
        bs->rho_ws_bytes = 0;
        MCPI_HIP(hipMalloc(&bs->rho_ws, total * sizeof(cplx)));
        bs->rho_ws_bytes = total * sizeof(cplx);
    }
    cplx *ws = (cplx *)bs->rho_ws;
    cplx *d1 = rho1 ? ws + need : nullptr;
    cplx *d2 = rho2 ? ws + need + (rho1 ? M2 : 0) : nullptr;
    MCPI_CHECK(densities_dev(b, C, rho1, rho2, d1, d2, s2, ws, need, &need, 0));
    if (rho1) MCPI_HIP(hipMemcpy(rho1, d1, sizeof(cplx) * M2, hipMemcpyDeviceToHost));
    if (rho2) MCPI_HIP(hipMemcpy(rho2, d2, sizeof(cplx) * M4, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

// ------------------------------------------------ allocation-free adjoint --
// For callers that cannot allocate device memory (code under stream capture,
// serving loops): the block partials and the Gram tiles live in a file-scope
// device array instead of the basis's workspace (fewer blocks when the
// partials would not fit), and the outputs are written on the device, so
// the call allocates nothing and copies nothing to the host.  Every call on
// a GPU shares that GPU's array: calls that overlap on one GPU race.
#define MCPI_NOALLOC_SCRATCH_BYTES (16u << 20)
static __device__ __align__(256) cplx scratch[MCPI_NOALLOC_SCRATCH_BYTES / sizeof(cplx)];

#define MCPI_NOALLOC_MAXDEV 64
static std::mutex g_scratch_mu;
static cplx *g_scratch_dev[MCPI_NOALLOC_MAXDEV];

// device address of this device's copy of `scratch` (looked up once per
// device, with that device current)
static mcpi_status noalloc_scratch(cplx **p)
{
    int dev = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_ARG(dev >= 0 && dev < MCPI_NOALLOC_MAXDEV, "device ordinal beyond the supported range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *a = nullptr;
        MCPI_HIP(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (cplx *)a;
    }
    *p = g_scratch_dev[dev];
    return MCPI_OK;
}

static mcpi_status vjp_noalloc(mcpi_basis bs, const cplx *C, cplx *g1, cplx *g2, hipStream_t s)
{
    cplx *scr = nullptr;
    MCPI_CHECK(noalloc_scratch(&scr));
    const size_t cap = MCPI_NOALLOC_SCRATCH_BYTES / sizeof(cplx);
    const mcpi_status st = densities_dev(density_view(bs), C, true, true, g1, g2, 0.5, scr, cap, nullptr, s);
    if (st == MCPI_ERR_ARG)
        set_error("basis has too many orbitals for the allocation-free adjoint (scratch holds 16 MiB)");
    return st;
}

}  // namespace mcpi

using namespace mcpi;

extern "C" {

mcpi_status mcpi_hip_rho1(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *rho1)
{
    MCPI_ARG(b && C_dev && rho1, "null argument");
    return densities(b, (const cplx *)C_dev, (cplx *)rho1, nullptr, 1.0);
}

mcpi_status mcpi_hip_rho2(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *rho2)
{
    MCPI_ARG(b && C_dev && rho2, "null argument");
    return densities(b, (const cplx *)C_dev, nullptr, (cplx *)rho2, 1.0);
}

mcpi_status mcpi_hip_vjp(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *dE_dHo, mcpi_cplx *dE_dHint)
{
    MCPI_ARG(b && C_dev && dE_dHo && dE_dHint, "null argument");
    // Hellmann-Feynman: E = sum Ho . rho1 + 1/2 sum Hint . rho2
    return densities(b, (const cplx *)C_dev, (cplx *)dE_dHo, (cplx *)dE_dHint, 0.5);
}

mcpi_status mcpi_hip_vjp_noalloc(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *dE_dHo_dev,
                                 mcpi_cplx *dE_dHint_dev, void *stream)
{
    MCPI_ARG(b && C_dev && dE_dHo_dev && dE_dHint_dev, "null argument");
    return vjp_noalloc(b, (const cplx *)C_dev, (cplx *)dE_dHo_dev, (cplx *)dE_dHint_dev, (hipStream_t)stream);
}

}  // extern "C"
