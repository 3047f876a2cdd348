// H.v through the pair space (schedules 6 and 7): every term of the
// reference's operator (hamiltonianMatrix.h:1431-1970, the folded rules of
// acth.hip), evaluated as two gathers around a pair-space step instead of
// one gather per term.
//
// In pair form (build_coefficients) the operator is
//   H = sum_{k,l} h1[k][l] a*_k a_l + sum_{P,Q} W[P][Q] a*_{P1} a*_{P2} a_{Q1} a_{Q2}
// over unordered orbital pairs.  On N >= 2 particles the one-body part is a
// pair term too (a*_k a_l = sum_s a*_k a*_s a_s a_l / (N - 1)).  Inserting
// the identity over the basis with N-2 particles after the creators,
//   (H C)(v) = sum_P <v|a*_{P1} a*_{P2}|u> Y2_u(P),  u = v - e_{P1} - e_{P2}
//   Y2_u(P)  = sum_Q W[P][Q] X[Q][u],  X[Q][u] = <u|a_{Q1} a_{Q2}|C>
// and <u|a_q a_l|C> = sqrt((u_q + 1)(u_l + 1 + [q = l])) C[rank(u + e_q + e_l)]
// is the pair amplitude of the densities' Gram matrix (rho.hip).  At N = M =
// 12 a row has up to 1,636 double-jump terms; here it costs 78 output
// gathers, and the 352,716 rows of 10 particles cost 78 gathers and a 78 x
// 78 complex matrix-vector product each, on the FP64 MFMA (schedule 6).
// When W is block diagonal with rank-one blocks, W = a b^T on each block r
// (a contact interaction that conserves a pair charge), and h1 is diagonal,
// Y2_u(P) = a[P] Z[r(P)][u] with Z[r][u] = sum_{Q in r} b[Q] X[Q][u]: one
// sum per block and no product (schedule 7).
//
// k_pair_amp<M, false>: X, gathered.  k_pair_mul<M>: Y2 = W X on the MFMA.
// k_pair_amp<M, true>: Z, gathered and summed per block.
// k_pair_out<M, FAC>: one thread per row v of the basis ranks the v - e_P
// in the smaller basis from its prefix-count tables and sums the factors
// times Y2 (or a[P] Z).  k_pair_y1<M, U>: for N = 1 only, the one-body
// part over the empty basis of N - 1 particles.  Y2, Z and X are stored
// column-major (Y[col * ld + u]) so that every pass reads and writes them
// coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mcpi_internal.h"

namespace mcpi {

namespace {

constexpr int PT = 256;   // threads per block

struct PairArgs {
    int N;                // particles of the basis
    int64_t nc;           // its states
    int64_t nu;           // states of the basis with N - 2 (or N - 1) particles
    int64_t nut;          // tiles of U rows
    const int32_t *B;     // rank table B[j*(N+1) + p]
    const cplx *C;
    const cplx *W;        // pair matrix pw[Q * PWS + P], or h1 as ph1[l * 16 + k] (M <= 12)
    cplx *Y;              // Y[col * ld + u]
    int64_t ld;           // row stride of Y (and of the pair amplitudes X), >= nu
    const cplx *H1;       // pair product: h1 (ph1[l * 16 + k]) folded into W with weight h1s
    double h1s;           // 1 / (N - 1), or 0
    const int32_t *zord;  // factorized form: pairs block by block (W = zb of the pairs)
    const int32_t *zoff;  // block r at zord[zoff[r] .. zoff[r+1])
    int zR;               // blocks
    int64_t mC, mY;       // factorized form, several members: entries between
                          // the members' C vectors and between their Z tables
};

__device__ inline cplx ldcp(const cplx *p)
{
    double2 t = *reinterpret_cast<const double2 *>(p);
    return cmake(t.x, t.y);
}

// 16-byte load at byte offset off of the buffer [p, p + bytes) (bytes <
// 2^31); an offset at or past the end reads zero without a memory access
__device__ inline cplx ldbuf(const cplx *p, int bytes, uint32_t off)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, bytes, 0x00020000);
    const v4u w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return cmake(__builtin_bit_cast(double, (unsigned long long)w.x | ((unsigned long long)w.y << 32)),
                 __builtin_bit_cast(double, (unsigned long long)w.z | ((unsigned long long)w.w << 32)));
}

template <int M, int U>
struct PairLds {
    static constexpr int RS = M + 1;   // LDS row stride of the tile (odd)
    static size_t bytes(int N)
    {
        size_t o = sizeof(cplx) * ((size_t)M * 16 + (size_t)U * RS);
        o += sizeof(int64_t) * ((size_t)U + (size_t)U * M);
        o += sizeof(double) * (N + 3) + sizeof(int32_t) * M * (N + 1) + (size_t)U * (M + 1);
        return (o + 15) / 16 * 16;
    }
};

// Y1[k][u] = sum_l h1[k][l] <u|a_l|C> over the basis with N-1 particles
template <int M, int U>
__global__ __launch_bounds__(PT) void k_pair_y1(PairArgs a)
{
    using L = PairLds<M, U>;
    constexpr int RS = L::RS, UR = U / 16;
    extern __shared__ __align__(16) unsigned char lds[];
    cplx *Wl = (cplx *)lds;                                   // [M][16]: h1[k][l] at l * 16 + k
    cplx *phi = Wl + M * 16;                                   // [U][RS]
    int64_t *tt = (int64_t *)(phi + U * RS);                  // [U]
    int64_t *dq = tt + U;                                      // [U][M]
    double *sq = (double *)(dq + U * M);                       // [N + 3]
    int32_t *Bt = (int32_t *)(sq + a.N + 3);                   // [M][N + 1]
    uint8_t *pp = (uint8_t *)(Bt + M * (a.N + 1));            // [U][M + 1]

    const int t = threadIdx.x;
    const int N = a.N, NB = N + 1, Np = N - 1;
    for (int x = t; x < M * 16; x += PT) Wl[x] = ldcp(a.W + x);
    for (int x = t; x < M * NB; x += PT) Bt[x] = a.B[x];
    for (int x = t; x < N + 3; x += PT) sq[x] = sqrt((double)x);
    const int pg = t & 15, ug = t >> 4;

    for (int64_t tile = blockIdx.x; tile < a.nut; tile += gridDim.x) {
        __syncthreads();   // the previous tile's product has read phi and the row tables
        const int64_t u0 = tile * U;
        if (t < U) {
            const int64_t u = u0 + t;
            uint8_t *P = pp + t * (M + 1);
            if (u < a.nu) {
                // greedy unranking, one step per iteration (lanes stay together)
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
                int64_t s0 = 0, s1 = 0;
                int64_t *Dq = dq + t * M;
                Dq[0] = 0;
                for (int j2 = 1; j2 < M; j2++) {
                    const int b = j2 * NB + P[j2];
                    s0 += Bt[b];
                    s1 += Bt[b + 1];
                    Dq[j2] = s0 - s1;
                }
                tt[t] = a.nc - 1 - s1;
            } else {
                tt[t] = -1;
            }
        }
        __syncthreads();
        // amplitudes of the tile: lanes over consecutive rows of one column,
        // five loads in flight per thread
        constexpr int NE = U * M, GBP = 5;
        for (int e0 = t; e0 < NE; e0 += GBP * PT) {
            int64_t tg[GBP];
            double f[GBP];
#pragma unroll
            for (int k = 0; k < GBP; k++) {
                const int e = e0 + k * PT;
                const int col = e / U, ul = e % U;
                const int64_t base = e < NE ? tt[ul] : -1;
                tg[k] = -1;
                f[k] = 0;
                if (base >= 0) {
                    const uint8_t *P = pp + ul * (M + 1);
                    tg[k] = base - dq[ul * M + col];
                    f[k] = sq[P[col + 1] - P[col] + 1];
                }
            }
            cplx v[GBP];
#pragma unroll
            for (int k = 0; k < GBP; k++) v[k] = tg[k] >= 0 ? ldcp(a.C + tg[k]) : cmake(0, 0);
#pragma unroll
            for (int k = 0; k < GBP; k++) {
                const int e = e0 + k * PT;
                if (e < NE) phi[(e % U) * RS + e / U] = cscale(v[k], f[k]);
            }
        }
        __syncthreads();
        // thread (pg, ug) forms column pg (an orbital k < M) of the rows
        // ug + 16 j
        cplx acc[UR];
#pragma unroll
        for (int j = 0; j < UR; j++) acc[j] = cmake(0, 0);
#pragma unroll 2
        for (int l = 0; l < M; l++) {
            const cplx w = Wl[l * 16 + pg];
            cplx x[UR];
#pragma unroll
            for (int j = 0; j < UR; j++) x[j] = phi[(ug + 16 * j) * RS + l];
#pragma unroll
            for (int j = 0; j < UR; j++) acc[j] = cfma(w, x[j], acc[j]);
        }
        if (pg < M) {
#pragma unroll
            for (int j = 0; j < UR; j++) {
                const int64_t u = u0 + ug + 16 * j;
                if (u < a.nu) {
                    double2 o;
                    o.x = acc[j].re;
                    o.y = acc[j].im;
                    *reinterpret_cast<double2 *>(a.Y + (int64_t)pg * a.ld + u) = o;
                }
            }
        }
    }
}

// The pair amplitudes X[Q][u] = <u|a_{Q1} a_{Q2}|C> of the basis with N-2
// particles, column-major (X[Q * nu + u]): one thread per row of a 64-row
// tile unranks it, then the block's lanes run over consecutive rows of one
// column at a time (neighbouring targets, coalesced stores), five loads in
// flight per thread.  Small LDS, so many blocks share a CU and hide each
// other's unranking.
template <int M>
struct PairAmpLds {
    static constexpr int U = 64;
    static size_t bytes(int N)
    {
        size_t o = sizeof(int64_t) * ((size_t)U + (size_t)U * M * 2);
        o += sizeof(double) * (N + 3) + sizeof(int32_t) * M * (N + 1) + (size_t)U * (M + 1);
        return (o + 15) / 16 * 16;
    }
};

// FAC: instead of X, Z[r][u] = sum_{Q in block r} zb[Q] X[Q][u] for the
// R blocks of the factorized pair form (W = zb; a wavefront per block and
// 64 rows); Z has R rows of ld entries.  MV > 1 (FAC only): the same for MV
// members that share zb and the blocks, their C vectors mC and their Z
// tables mY entries apart (the unranking and the targets are shared); grid
// row y serves members y MV .. y MV + MV - 1.
template <int M, bool FAC, int MV = 1>
__global__ __launch_bounds__(PT) void k_pair_amp(PairArgs a)
{
    static_assert(MV == 1 || FAC, "several members: factorized form only");
    const int64_t m0 = MV > 1 ? (int64_t)blockIdx.y * MV : 0;   // first member of this grid row
    constexpr int NP = pair_count(M), U = PairAmpLds<M>::U;
    extern __shared__ __align__(16) unsigned char lds[];
    int64_t *tt = (int64_t *)lds;                              // [U]
    int64_t *dq = tt + U;                                      // [U][M]
    int64_t *dl = dq + U * M;                                  // [U][M]
    double *sq = (double *)(dl + U * M);                       // [N + 3]
    int32_t *Bt = (int32_t *)(sq + a.N + 3);                   // [M][N + 1]
    uint8_t *pp = (uint8_t *)(Bt + M * (a.N + 1));            // [U][M + 1]
    const int t = threadIdx.x;
    const int N = a.N, NB = N + 1, Np = N - 2;
    for (int x = t; x < M * NB; x += PT) Bt[x] = a.B[x];
    for (int x = t; x < N + 3; x += PT) sq[x] = sqrt((double)x);
    for (int64_t tile = blockIdx.x; tile < a.nut; tile += gridDim.x) {
        __syncthreads();
        const int64_t u0 = tile * U;
        if (t < U) {
            const int64_t u = u0 + t;
            uint8_t *P = pp + t * (M + 1);
            if (u < a.nu) {
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
                int64_t *Dq = dq + t * M, *Dl = dl + t * M;
                Dq[0] = 0;
                Dl[0] = 0;
                for (int j2 = 1; j2 < M; j2++) {
                    const int b = j2 * NB + P[j2];
                    s0 += Bt[b];
                    s1 += Bt[b + 1];
                    s2 += Bt[b + 2];
                    Dq[j2] = s0 - s1;
                    Dl[j2] = s1 - s2;
                }
                tt[t] = a.nc - 1 - s2;
            } else {
                tt[t] = -1;
            }
        }
        __syncthreads();
        if (FAC) {
            // the pair (q <= l) of index col and its amplitude's target and factor
            auto amp = [&](int col, int ul, int64_t base, int64_t &tg, double &f) {
                const uint8_t *P = pp + ul * (M + 1);
                const int l = pair_hi(col), q = col - l * (l + 1) / 2;
                const int uq = P[q + 1] - P[q], ulo = P[l + 1] - P[l];
                tg = base - dq[ul * M + q] - dl[ul * M + l];
                f = sq[uq + 1] * sq[ulo + 1 + (q == l)];
            };
            for (int e = t; e < a.zR * U; e += PT) {
                const int r = e / U, ul = e % U;          // r: one per wavefront
                const int64_t base = tt[ul];
                if (base < 0) continue;
                const int j0 = a.zoff[r], j1 = a.zoff[r + 1];
                cplx acc[MV];
#pragma unroll
                for (int m = 0; m < MV; m++) acc[m] = cmake(0, 0);
                // the block's pairs two at a time: a small register count
                // (many wavefronts per SIMD) hides the gathers better than
                // more loads per thread (four: +4% at N=16/M=12, eight and
                // two outputs per thread: +20%)
                constexpr int GBK = 2;
                for (int j = j0; j < j1; j += GBK) {
                    int64_t tg[GBK];
                    double f[GBK];
                    cplx w[GBK], v[MV][GBK];
#pragma unroll
                    for (int k = 0; k < GBK; k++) {
                        tg[k] = -1;
                        f[k] = 0;
                        w[k] = cmake(0, 0);
                        if (j + k < j1) {
                            const int col = a.zord[j + k];
                            amp(col, ul, base, tg[k], f[k]);
                            w[k] = ldcp(a.W + col);
                        }
                    }
#pragma unroll
                    for (int m = 0; m < MV; m++)
#pragma unroll
                        for (int k = 0; k < GBK; k++)
                            v[m][k] = tg[k] >= 0 ? ldcp(a.C + (m0 + m) * a.mC + tg[k]) : cmake(0, 0);
#pragma unroll
                    for (int m = 0; m < MV; m++)
#pragma unroll
                        for (int k = 0; k < GBK; k++) acc[m] = cfma(cscale(w[k], f[k]), v[m][k], acc[m]);
                }
#pragma unroll
                for (int m = 0; m < MV; m++) {
                    double2 o;
                    o.x = acc[m].re;
                    o.y = acc[m].im;
                    *reinterpret_cast<double2 *>(a.Y + (m0 + m) * a.mY + (int64_t)r * a.ld + u0 + ul) = o;
                }
            }
            continue;
        }
        constexpr int NE = U * NP, GBP = 5;
        for (int e0 = t; e0 < NE; e0 += GBP * PT) {
            int64_t tg[GBP];
            double f[GBP];
#pragma unroll
            for (int k = 0; k < GBP; k++) {
                const int e = e0 + k * PT;
                const int col = e / U, ul = e % U;
                const int64_t base = e < NE ? tt[ul] : -1;
                tg[k] = -1;
                f[k] = 0;
                if (base >= 0) {
                    const uint8_t *P = pp + ul * (M + 1);
                    int l = 0;
                    while ((l + 1) * (l + 2) / 2 <= col) l++;
                    const int q = col - l * (l + 1) / 2;
                    const int uq = P[q + 1] - P[q], ulo = P[l + 1] - P[l];
                    tg[k] = base - dq[ul * M + q] - dl[ul * M + l];
                    f[k] = sq[uq + 1] * sq[ulo + 1 + (q == l)];
                }
            }
            cplx v[GBP];
#pragma unroll
            for (int k = 0; k < GBP; k++) v[k] = tg[k] >= 0 ? ldcp(a.C + tg[k]) : cmake(0, 0);
#pragma unroll
            for (int k = 0; k < GBP; k++) {
                if (tg[k] < 0) continue;
                const int e = e0 + k * PT;
                const cplx z = cscale(v[k], f[k]);
                double2 o;
                o.x = z.re;
                o.y = z.im;
                *reinterpret_cast<double2 *>(a.Y + (int64_t)(e / U) * a.ld + u0 + e % U) = o;
            }
        }
    }
}

// Y2 = W X through FP64 MFMA (v_mfma_f64_16x16x4f64: lane i + 16k holds
// A[i][k] and B[k][i]; the result D[g + 4r][j] sits in register r of lane
// j + 16g), three real products per complex one (Gauss): T1 = (Xr + Xi)
// Wr, T2 = Xr (Wi - Wr), T3 = Xi (Wr + Wi), then Re = T1 - T3 and Im = T1 +
// T2.  The block stages W once in LDS as B fragments in the order the lanes
// read them: Wa[s][g][lane] = (Wr, Wi - Wr) and Wb[s][g][lane] = Wr + Wi of
// W[4s + k][16g + j], KS = ceil(NP/4) steps of four Q times NG groups of
// 16 pair columns P, zero past NP; conflict-free reads.
// Each wavefront takes 16-row blocks of X on its own and keeps the 16 x NP
// result block in accumulators: per step one A fragment (four Q rows of 16
// consecutive u, coalesced; loaded four steps ahead, across row blocks)
// meets the NG fragments of W, three MFMAs each.
typedef double mcpi_v4d __attribute__((ext_vector_type(4)));

constexpr int PAIR_MUL_WAVES = 8;

// The one-body operator in pair form: on states of N >= 2 particles
//   a*_k a_l = (1 / (N - 1)) sum_s a*_k a*_s a_s a_l
// (the number operator is N - 1 on a_l |v>), so h1 adds
// h1[k][l] / (N - 1) to W[{k, s}][{s, l}] for every s.  This is the sum of
// h1[k][l] over the orderings (k, s) of the pair P and (s, l) of the pair Q
// that share s.
__device__ inline cplx h1_pair(const cplx *h1, int P, int Q)
{
    const int p2 = pair_hi(P), p1 = P - p2 * (p2 + 1) / 2;
    const int q2 = pair_hi(Q), q1 = Q - q2 * (q2 + 1) / 2;
    cplx z = cmake(0, 0);
    for (int o = 0; o < (p1 == p2 ? 1 : 2); o++) {
        const int k = o ? p1 : p2, s = o ? p2 : p1;
        for (int o2 = 0; o2 < (q1 == q2 ? 1 : 2); o2++) {
            const int s2 = o2 ? q2 : q1, l = o2 ? q1 : q2;
            if (s2 == s) z = cadd(z, ldcp(h1 + l * 16 + k));
        }
    }
    return z;
}

template <int M>
struct PairMulLds {
    static constexpr int NP = pair_count(M), KS = (NP + 3) / 4, NG = (NP + 15) / 16;
    static constexpr size_t bytes = (sizeof(double2) + sizeof(double)) * NG * KS * 64;
};

template <int M>
__global__ __launch_bounds__(64 * PAIR_MUL_WAVES) void k_pair_mul(PairArgs a, const cplx *X)
{
    using L = PairMulLds<M>;
    constexpr int NP = L::NP, KS = L::KS, NG = L::NG, PWS = pair_stride(M);
    extern __shared__ __align__(16) unsigned char lds[];
    double2 *wa = (double2 *)lds;                 // [KS][NG][64]
    double *wb = (double *)(wa + NG * KS * 64);   // [KS][NG][64]
    for (int x = threadIdx.x; x < NG * KS * 64; x += blockDim.x) {
        const int s = x / (NG * 64), g = (x / 64) % NG, ln = x % 64;
        const int P = 16 * g + (ln & 15), Q = 4 * s + (ln >> 4);
        double2 v = {0, 0};
        double vb = 0;
        if (P < NP && Q < NP) {
            cplx w = ldcp(a.W + (int64_t)Q * PWS + P);
            if (a.h1s != 0.0) {
                const cplx h = h1_pair(a.H1, P, Q);
                w.re = fma(a.h1s, h.re, w.re);
                w.im = fma(a.h1s, h.im, w.im);
            }
            v.x = w.re;
            v.y = w.im - w.re;
            vb = w.re + w.im;
        }
        wa[x] = v;
        wb[x] = vb;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jl = lane & 15, kl = lane >> 4;
    const int64_t nrb = (a.nu + 15) / 16, stride = (int64_t)gridDim.x * PAIR_MUL_WAVES;
    // (the launch gives ld >= 16 nrb)
    // A fragments: lane (j, k) of step s of row block rb reads X[4s + k][16
    // rb + j].  X has 4 KS rows of ld >= 16 nrb entries, zero past NP rows
    // and nu columns, and Y has 16 NG rows of ld entries, so that neither
    // loads nor stores need a bound; a block past the last is read as the
    // last.
    const int64_t rowstep = 4 * a.ld;
    auto lanep = [&](int64_t rb) { return X + (int64_t)kl * a.ld + (rb < nrb ? rb : nrb - 1) * 16 + jl; };
    // They are loaded four steps before their use, across row blocks: the
    // steps of a block are counted up to KS4 = KS rounded up to four (the
    // steps past KS load step KS - 1 again and skip their MFMAs), so that
    // step t of the ring always sits in register t % 4.
    constexpr int KS4 = (KS + 3) & ~3;
    auto ldx = [&](const cplx *p, int s) { return ldcp(p + (int64_t)(s < KS ? s : KS - 1) * rowstep); };
    int64_t rb = (int64_t)blockIdx.x * PAIR_MUL_WAVES + wave;
    const cplx *pc = lanep(rb);
    cplx x0 = ldx(pc, 0);
    __builtin_amdgcn_sched_barrier(0);
    cplx x1 = ldx(pc, 1);
    __builtin_amdgcn_sched_barrier(0);
    cplx x2 = ldx(pc, 2);
    __builtin_amdgcn_sched_barrier(0);
    cplx x3 = ldx(pc, 3);
    __builtin_amdgcn_sched_barrier(0);
    // one row block: KS4 steps in straight-line code, then the stores
    auto block = [&](int64_t rb, const cplx *pn) {
        auto ldnext = [&](int t) { return t < KS4 ? ldx(pc, t) : ldx(pn, t - KS4); };
        mcpi_v4d acc[3 * NG];
#pragma unroll
        for (int g = 0; g < 3 * NG; g++) acc[g] = mcpi_v4d{0, 0, 0, 0};
        auto step = [&](int s, const cplx xc) {
            if (s >= KS) return;
            const double xs = xc.re + xc.im;
#pragma unroll
            for (int g = 0; g < NG; g++) {
                const double2 w = wa[(s * NG + g) * 64 + lane];
                const double w3 = wb[(s * NG + g) * 64 + lane];
                acc[3 * g] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs, w.x, acc[3 * g], 0, 0, 0);
                acc[3 * g + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(xc.re, w.y, acc[3 * g + 1], 0, 0, 0);
                acc[3 * g + 2] = __builtin_amdgcn_mfma_f64_16x16x4f64(xc.im, w3, acc[3 * g + 2], 0, 0, 0);
            }
            // LDS reads of one group ahead of its three MFMAs, in turn
#pragma unroll
            for (int g = 0; g < NG; g++) {
                __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
            }
        };
#pragma unroll
        for (int s = 0; s < KS4; s += 4) {
            step(s, x0);
            __builtin_amdgcn_sched_barrier(0);
            x0 = ldnext(s + 4);
            step(s + 1, x1);
            __builtin_amdgcn_sched_barrier(0);
            x1 = ldnext(s + 5);
            step(s + 2, x2);
            __builtin_amdgcn_sched_barrier(0);
            x2 = ldnext(s + 6);
            step(s + 3, x3);
            __builtin_amdgcn_sched_barrier(0);
            x3 = ldnext(s + 7);
        }
#pragma unroll
        for (int g = 0; g < NG; g++) {
            cplx *y = a.Y + (int64_t)(16 * g + jl) * a.ld + rb * 16 + kl;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double2 o;
                o.x = acc[3 * g][r] - acc[3 * g + 2][r];
                o.y = acc[3 * g][r] + acc[3 * g + 1][r];
                *reinterpret_cast<double2 *>(y + 4 * r) = o;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    // The first block is peeled, so that the loop is entered with the
    // stores of a block in flight as on every later pass: the first steps of
    // a block then wait only for their own fragments (loaded before those
    // stores), not for the stores to complete.
    if (rb < nrb) {
        const cplx *pn = lanep(rb + stride);
        block(rb, pn);
        pc = pn;
        for (rb += stride; rb < nrb; rb += stride) {
            pn = lanep(rb + stride);
            block(rb, pn);
            pc = pn;
        }
    }
}

// out[v] = sum_k sqrt(v_k) Y1[k][rank'(v - e_k)] + sum_P f_P(v) Y2[P][rank''(v - e_P)];
// FAC (factorized pair form, Y2 = Z with one row per block):
// out[v] = (sum_k h1[k][k] v_k) C[v] + sum_P f_P(v) za[P] Z[zblk[P]][rank''(v - e_P)]
template <int M, bool FAC>
__global__ __launch_bounds__(PT) void k_pair_out(int N, int64_t nc, const int32_t *__restrict__ IF,
                                                 const int32_t *__restrict__ Bg, int64_t nu1, int64_t nu2,
                                                 int64_t ld2, const cplx *__restrict__ Y1,
                                                 const cplx *__restrict__ Y2, const cplx *__restrict__ za,
                                                 const int32_t *__restrict__ zblk, const cplx *__restrict__ h1,
                                                 const cplx *__restrict__ C, cplx *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char lds[];
    double *sq = (double *)lds;                      // [N + 3]
    int32_t *Bt = (int32_t *)(sq + N + 3);          // [M][N + 1]
    const int NB = N + 1;
    for (int x = threadIdx.x; x < N + 3; x += PT) sq[x] = sqrt((double)x);
    for (int x = threadIdx.x; x < M * NB; x += PT) Bt[x] = Bg[x];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= nc) return;
    int v[M];
#pragma unroll
    for (int k = 0; k < M; k++) v[k] = IF[i * M + k];
    // S_d(x) = sum_{j=1..x} B[j][P_j - d] (P_j: particles below orbital j),
    // modulo 2^32 (every rank is below 2^31)
    uint32_t S0[M], S1[M], S2[M];
    {
        int pj = 0;
        uint32_t s0 = 0, s1 = 0, s2 = 0;
        S0[0] = S1[0] = S2[0] = 0;
#pragma unroll
        for (int j = 1; j < M; j++) {
            pj += v[j - 1];
            const int b = j * NB + pj;
            s0 += (uint32_t)Bt[b];
            s1 += (uint32_t)Bt[pj >= 1 ? b - 1 : b];
            s2 += (uint32_t)Bt[pj >= 2 ? b - 2 : b];
            S0[j] = s0;
            S1[j] = s1;
            S2[j] = s2;
        }
    }
    // Every term is read without a branch through a buffer load: an empty
    // orbital (pair) gets an offset past its column, which the buffer's
    // range check turns into a zero without a memory access, so the loads
    // of a batch issue back to back and no lane touches an extra line.
    constexpr uint32_t OUT = 0x80000000u;   // beyond every column (< 2 GiB)
    cplx acc = cmake(0, 0);
    if (FAC) {   // the diagonal one-body part
        cplx d = cmake(0, 0);
#pragma unroll
        for (int k = 0; k < M; k++) {
            const cplx h = ldcp(h1 + k * (pair_h1_stride(M) + 1));
            d.re = fma(h.re, (double)v[k], d.re);
            d.im = fma(h.im, (double)v[k], d.im);
        }
        acc = cfma(d, ldcp(C + i), acc);
    }
    if (nu1 > 0) {
        const uint32_t top1 = (uint32_t)(nu1 - 1) - S1[M - 1];
        const int bytes1 = (int)(nu1 * (int64_t)sizeof(cplx));
        cplx y[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
            const uint32_t off = v[k] >= 1 ? (top1 - (S0[k] - S1[k])) * (uint32_t)sizeof(cplx) : OUT;
            y[k] = ldbuf(Y1 + (int64_t)k * nu1, bytes1, off);
        }
#pragma unroll
        for (int k = 0; k < M; k++) {
            const double f = sq[v[k]];
            acc.re = fma(f, y[k].re, acc.re);
            acc.im = fma(f, y[k].im, acc.im);
        }
    }
    if (nu2 > 0) {
        const uint32_t top2 = (uint32_t)(nu2 - 1) - S2[M - 1];
        const int bytes2 = (int)(nu2 * (int64_t)sizeof(cplx));
        constexpr int NPp = pair_count(M), BT = 13;
        // full unroll: every index below is then a constant (a partial
        // unroll, which plain `unroll` gives from M = 15 on, indexes v and
        // the prefix sums at run time and puts them in scratch)
#pragma clang loop unroll(full)
        for (int b0 = 0; b0 < NPp; b0 += BT) {
            cplx y[BT], cf[BT];
            double f[BT];
#pragma unroll
            for (int j = 0; j < BT; j++) {
                const int P = b0 + j;
                if (P >= NPp) break;
                const int c2 = pair_hi(P), c1 = P - c2 * (c2 + 1) / 2;
                const int blk = FAC ? zblk[P] : P;
                const bool occ = (c1 < c2 ? (v[c1] >= 1 && v[c2] >= 1) : v[c1] >= 2) && blk >= 0;
                const uint32_t off =
                    occ ? (top2 - (S0[c1] - S1[c1]) - (S1[c2] - S2[c2])) * (uint32_t)sizeof(cplx) : OUT;
                f[j] = c1 < c2 ? sq[v[c1]] * sq[v[c2]] : sq[v[c1]] * sq[v[c1] >= 1 ? v[c1] - 1 : 0];
                if (FAC) cf[j] = cscale(ldcp(za + P), f[j]);
                y[j] = ldbuf(Y2 + (int64_t)(blk >= 0 ? blk : 0) * ld2, bytes2, off);
            }
#pragma unroll
            for (int j = 0; j < BT; j++) {
                if (b0 + j >= NPp) break;
                if (FAC) {
                    acc = cfma(cf[j], y[j], acc);
                } else {
                    acc.re = fma(f[j], y[j].re, acc.re);
                    acc.im = fma(f[j], y[j].im, acc.im);
                }
            }
            // an empty statement that takes the sums and fences memory, so
            // that this batch is summed before the next batch's loads issue
            // (all 78 loads up front take every register: one wavefront per
            // SIMD; with the next batch in flight during the sums, two
            // wavefronts per SIMD and no gain)
            asm volatile("" : "+v"(acc.re), "+v"(acc.im) : : "memory");
        }
    }
    double2 o;
    o.x = acc.re;
    o.y = acc.im;
    *reinterpret_cast<double2 *>(out + i) = o;
}

// The output pass of the factorized form for MV members that share the
// basis, zb and the blocks (a group of Lieb-Liniger couplings): the row's
// occupations, prefix sums, pair offsets and factors are computed once and
// serve every member; member m reads its C, Z, za and h1 at C + m mC,
// Z + m mZ and (bytes) za + m mK, h1 + m mK, and writes out + m mC; grid
// row y serves members y MV .. y MV + MV - 1.  Batches of BT terms for all
// members (MV BT loads in flight).
template <int M, int MV>
__global__ __launch_bounds__(PT) void k_pairz_out_mv(int N, int64_t nc, const int32_t *__restrict__ IF,
                                                     const int32_t *__restrict__ Bg, int64_t nu2, int64_t ld2,
                                                     const cplx *__restrict__ Z, int64_t mZ,
                                                     const cplx *__restrict__ za, const int32_t *__restrict__ zblk,
                                                     const cplx *__restrict__ h1, int64_t mK,
                                                     const cplx *__restrict__ C, cplx *__restrict__ out, int64_t mC)
{
    extern __shared__ __align__(16) unsigned char lds[];
    double *sq = (double *)lds;                      // [N + 3]
    int32_t *Bt = (int32_t *)(sq + N + 3);          // [M][N + 1]
    const int NB = N + 1;
    for (int x = threadIdx.x; x < N + 3; x += PT) sq[x] = sqrt((double)x);
    for (int x = threadIdx.x; x < M * NB; x += PT) Bt[x] = Bg[x];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= nc) return;
    {   // this grid row's first member
        const int64_t m0 = (int64_t)blockIdx.y * MV;
        C += m0 * mC;
        out += m0 * mC;
        Z += m0 * mZ;
        za = (const cplx *)((const char *)za + m0 * mK);
        h1 = (const cplx *)((const char *)h1 + m0 * mK);
    }
    int v[M];
#pragma unroll
    for (int k = 0; k < M; k++) v[k] = IF[i * M + k];
    uint32_t S0[M], S1[M], S2[M];
    {
        int pj = 0;
        uint32_t s0 = 0, s1 = 0, s2 = 0;
        S0[0] = S1[0] = S2[0] = 0;
#pragma unroll
        for (int j = 1; j < M; j++) {
            pj += v[j - 1];
            const int b = j * NB + pj;
            s0 += (uint32_t)Bt[b];
            s1 += (uint32_t)Bt[pj >= 1 ? b - 1 : b];
            s2 += (uint32_t)Bt[pj >= 2 ? b - 2 : b];
            S0[j] = s0;
            S1[j] = s1;
            S2[j] = s2;
        }
    }
    constexpr uint32_t OUT = 0x80000000u;
    auto mza = [&](int m) { return (const cplx *)((const char *)za + m * mK); };
    cplx acc[MV];
#pragma unroll
    for (int m = 0; m < MV; m++) {   // the diagonal one-body part
        const cplx *h = (const cplx *)((const char *)h1 + m * mK);
        cplx d = cmake(0, 0);
#pragma unroll
        for (int k = 0; k < M; k++) {
            const cplx hk = ldcp(h + k * (pair_h1_stride(M) + 1));
            d.re = fma(hk.re, (double)v[k], d.re);
            d.im = fma(hk.im, (double)v[k], d.im);
        }
        acc[m] = cfma(d, ldcp(C + m * mC + i), cmake(0, 0));
    }
    const uint32_t top2 = (uint32_t)(nu2 - 1) - S2[M - 1];
    const int bytes2 = (int)(nu2 * (int64_t)sizeof(cplx));
    constexpr int NPp = pair_count(M), BT = MV >= 4 ? 4 : 6;
#pragma clang loop unroll(full)
    for (int b0 = 0; b0 < NPp; b0 += BT) {
        uint32_t off[BT];
        int bk[BT];
        double f[BT];
        cplx y[MV][BT];
#pragma unroll
        for (int j = 0; j < BT; j++) {
            const int P = b0 + j;
            if (P >= NPp) break;
            const int c2 = pair_hi(P), c1 = P - c2 * (c2 + 1) / 2;
            const int blk = zblk[P];
            const bool occ = (c1 < c2 ? (v[c1] >= 1 && v[c2] >= 1) : v[c1] >= 2) && blk >= 0;
            off[j] = occ ? (top2 - (S0[c1] - S1[c1]) - (S1[c2] - S2[c2])) * (uint32_t)sizeof(cplx) : OUT;
            f[j] = c1 < c2 ? sq[v[c1]] * sq[v[c2]] : sq[v[c1]] * sq[v[c1] >= 1 ? v[c1] - 1 : 0];
            bk[j] = blk >= 0 ? blk : 0;
        }
#pragma unroll
        for (int m = 0; m < MV; m++)
#pragma unroll
            for (int j = 0; j < BT; j++) {
                if (b0 + j >= NPp) break;
                y[m][j] = ldbuf(Z + m * mZ + (int64_t)bk[j] * ld2, bytes2, off[j]);
            }
#pragma unroll
        for (int m = 0; m < MV; m++)
#pragma unroll
            for (int j = 0; j < BT; j++) {
                if (b0 + j >= NPp) break;
                acc[m] = cfma(cscale(ldcp(mza(m) + b0 + j), f[j]), y[m][j], acc[m]);
            }
        // each batch summed before the next one's loads issue
#pragma unroll
        for (int m = 0; m < MV; m++) asm volatile("" : "+v"(acc[m].re), "+v"(acc[m].im) : : "memory");
    }
#pragma unroll
    for (int m = 0; m < MV; m++) {
        double2 o;
        o.x = acc[m].re;
        o.y = acc[m].im;
        *reinterpret_cast<double2 *>(out + m * mC + i) = o;
    }
}

template <int M, int U>
mcpi_status launch_y1(const PairArgs &a, int ncu, hipStream_t st)
{
    const size_t sm = PairLds<M, U>::bytes(a.N);
    MCPI_ARG(sm <= 160 * 1024, "pair-space tile exceeds LDS");
    const int per_cu = std::max(1, (int)(160 * 1024 / sm));
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(a.nut, (int64_t)ncu * per_cu));
    (void)hipFuncSetAttribute((const void *)k_pair_y1<M, U>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    hipLaunchKernelGGL((k_pair_y1<M, U>), dim3((unsigned)nb), dim3(PT), sm, st, a);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

template <int M>
mcpi_status launch_pair_m(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, cplx *ws,
                          hipStream_t st)
{
    int dev = 0, ncu = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int N = b.N;
    const int64_t nu1 = N >= 1 ? nc_of(N - 1, M) : 0, nu2 = N >= 2 ? nc_of(N - 2, M) : 0;
    const int64_t ld2 = pair_ld(nu2);
    static_assert(pair_h1_stride(M) == 16, "the pair product reads h1 with row stride 16");
    // with N >= 2 the one-body part rides in the pair product (h1 folded
    // into W); the pass over N - 1 particles runs only for N = 1
    const int64_t nu1p = N == 1 ? nu1 : 0;
    cplx *Y1 = ws, *Y2 = ws + (size_t)M * nu1p;
    if (nu1p > 0) {
        PairArgs a{N, b.nc, nu1, (nu1 + 127) / 128, b.B, C, c.ph1, Y1, nu1, nullptr, 0.0, nullptr, nullptr, 0};
        MCPI_CHECK((launch_y1<M, 128>(a, ncu, st)));
    }
    if (nu2 > 0) {
        // amplitudes X into the workspace after Y2, then Y2 = W X (rows of
        // ld2 entries; Y2 has pair_stride(M) rows, X pair_x_rows(M))
        constexpr int U = PairAmpLds<M>::U;
        static_assert(16 * PairMulLds<M>::NG == pair_stride(M) && 4 * PairMulLds<M>::KS == pair_x_rows(M),
                      "pair workspace rows");
        cplx *X = Y2 + (size_t)pair_stride(M) * ld2;
        PairArgs a{N, b.nc, nu2, (nu2 + U - 1) / U, b.B, C, c.pw, X, ld2, c.ph1, 1.0 / (N - 1),
                   nullptr, nullptr, 0};
        const size_t sm = PairAmpLds<M>::bytes(N);
        int per_cu = 0;
        MCPI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pair_amp<M, false>, PT, sm));
        int64_t nb = std::max<int64_t>(1, std::min<int64_t>(a.nut, (int64_t)ncu * std::max(per_cu, 1)));
        hipLaunchKernelGGL((k_pair_amp<M, false>), dim3((unsigned)nb), dim3(PT), sm, st, a);
        MCPI_HIP(hipGetLastError());
        a.Y = Y2;
        const size_t smw = PairMulLds<M>::bytes;
        MCPI_ARG(smw <= 160 * 1024, "pair matrix exceeds LDS");
        (void)hipFuncSetAttribute((const void *)k_pair_mul<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)smw);
        const int nt = 64 * PAIR_MUL_WAVES;
        MCPI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pair_mul<M>, nt, smw));
        const int64_t nrb = (nu2 + 15) / 16;
        nb = std::max<int64_t>(1, std::min<int64_t>((nrb + PAIR_MUL_WAVES - 1) / PAIR_MUL_WAVES,
                                                    (int64_t)ncu * std::max(per_cu, 1)));
        hipLaunchKernelGGL(k_pair_mul<M>, dim3((unsigned)nb), dim3(nt), smw, st, a, (const cplx *)X);
        MCPI_HIP(hipGetLastError());
    }
    const size_t sm = sizeof(double) * (N + 3) + sizeof(int32_t) * M * (N + 1);
    hipLaunchKernelGGL((k_pair_out<M, false>), dim3((unsigned)((b.nc + PT - 1) / PT)), dim3(PT), sm, st, N,
                       b.nc, (const int32_t *)b.IF, (const int32_t *)b.B, nu1p, nu2, ld2, (const cplx *)Y1,
                       (const cplx *)Y2, (const cplx *)nullptr, (const int32_t *)nullptr,
                       (const cplx *)nullptr, (const cplx *)nullptr, out);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// the factorized pair form (schedule 7): Z (pzR rows of ld2) at the
// start of the workspace, then the output pass
template <int M>
mcpi_status launch_pairz_m(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, cplx *ws,
                           hipStream_t st)
{
    int dev = 0, ncu = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int N = b.N;
    const int64_t nu2 = nc_of(N - 2, M), ld2 = pair_ld(nu2);
    cplx *Z = ws;
    constexpr int U = PairAmpLds<M>::U;
    PairArgs a{N, b.nc, nu2, (nu2 + U - 1) / U, b.B, C, c.pzb, Z, ld2, nullptr, 0.0, c.pzord, c.pzoff, c.pzR};
    const size_t sm = PairAmpLds<M>::bytes(N);
    int per_cu = 0;
    MCPI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pair_amp<M, true>, PT, sm));
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(a.nut, (int64_t)ncu * std::max(per_cu, 1)));
    hipLaunchKernelGGL((k_pair_amp<M, true>), dim3((unsigned)nb), dim3(PT), sm, st, a);
    MCPI_HIP(hipGetLastError());
    const size_t so = sizeof(double) * (N + 3) + sizeof(int32_t) * M * (N + 1);
    hipLaunchKernelGGL((k_pair_out<M, true>), dim3((unsigned)((b.nc + PT - 1) / PT)), dim3(PT), so, st, N, b.nc,
                       (const int32_t *)b.IF, (const int32_t *)b.B, (int64_t)0, nu2, ld2, (const cplx *)nullptr,
                       (const cplx *)Z, (const cplx *)c.pza, (const int32_t *)c.pzblk, (const cplx *)c.ph1, C,
                       out);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// the factorized pair form for MV members at once (member m's Z at
// ws + m * R * ld2; vectors mC entries and coefficient blocks mK bytes
// apart, member 0's tables in c)
template <int M, int MV>
mcpi_status launch_pairz_mv_m(const BasisDev &b, const HamCoefDev &c, int chunks, const cplx *C, cplx *out,
                              int64_t mC, int64_t mK, cplx *ws, hipStream_t st)
{
    int dev = 0, ncu = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int N = b.N;
    const int64_t nu2 = nc_of(N - 2, M), ld2 = pair_ld(nu2), mZ = (int64_t)c.pzR * ld2;
    constexpr int U = PairAmpLds<M>::U;
    PairArgs a{N,   b.nc,    nu2,     (nu2 + U - 1) / U, b.B, C, c.pzb, ws, ld2, nullptr, 0.0,
               c.pzord, c.pzoff, c.pzR, mC,                mZ};
    const size_t sm = PairAmpLds<M>::bytes(N);
    int per_cu = 0;
    MCPI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pair_amp<M, true, MV>, PT, sm));
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(a.nut, (int64_t)ncu * std::max(per_cu, 1)));
    hipLaunchKernelGGL((k_pair_amp<M, true, MV>), dim3((unsigned)nb, (unsigned)chunks), dim3(PT), sm, st, a);
    MCPI_HIP(hipGetLastError());
    const size_t so = sizeof(double) * (N + 3) + sizeof(int32_t) * M * (N + 1);
    hipLaunchKernelGGL((k_pairz_out_mv<M, MV>), dim3((unsigned)((b.nc + PT - 1) / PT), (unsigned)chunks), dim3(PT), so,
                       st, N, b.nc,
                       (const int32_t *)b.IF, (const int32_t *)b.B, nu2, ld2, (const cplx *)ws, mZ,
                       (const cplx *)c.pza, (const int32_t *)c.pzblk, (const cplx *)c.ph1, mK, C, out, mC);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

}  // namespace

// complex entries of the pair-space workspace of a basis: Y1 (M rows of
// NC(N-1, M), for N = 1 only), then Y2 (pair_stride(M) rows) and the pair
// amplitudes X (pair_x_rows(M) rows) with rows of pair_ld(NC(N-2, M))
// entries; the caller zeroes it once (the pad rows and columns of X are
// read, never written)
size_t pair_workspace_entries(int N, int M)
{
    const int64_t nu1 = N == 1 ? nc_of(N - 1, M) : 0, nu2 = N >= 2 ? nc_of(N - 2, M) : 0;
    return (size_t)M * nu1 + (size_t)(pair_stride(M) + pair_x_rows(M)) * pair_ld(nu2);
}

bool pair_applicable(const BasisDev &b, const HamCoefDev &c)
{
    // (the output pass reads Y1 and Y2 columns as buffers of < 2 GiB:
    // NC(N-1, M) >= NC(N-2, M) entries of 16 bytes)
    return c.pw && c.ph1 && b.M >= 3 && b.M <= MCPI_PAIR_MAXM && !b.perm && b.IF && b.nc <= INT32_MAX &&
           b.N <= MCPI_MAXN && (b.N < 1 || nc_of(b.N - 1, b.M) < ((int64_t)1 << 27));
}

mcpi_status launch_pair(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, cplx *ws,
                        hipStream_t st)
{
    MCPI_ARG(pair_applicable(b, c), "the pair-space H.v does not apply to this basis");
    switch (b.M) {
    case 3: return launch_pair_m<3>(b, c, C, out, ws, st);
    case 4: return launch_pair_m<4>(b, c, C, out, ws, st);
    case 5: return launch_pair_m<5>(b, c, C, out, ws, st);
    case 6: return launch_pair_m<6>(b, c, C, out, ws, st);
    case 7: return launch_pair_m<7>(b, c, C, out, ws, st);
    case 8: return launch_pair_m<8>(b, c, C, out, ws, st);
    case 9: return launch_pair_m<9>(b, c, C, out, ws, st);
    case 10: return launch_pair_m<10>(b, c, C, out, ws, st);
    case 11: return launch_pair_m<11>(b, c, C, out, ws, st);
    case 12: return launch_pair_m<12>(b, c, C, out, ws, st);
    default: return MCPI_ERR_ARG;
    }
}

bool pairz_applicable(const BasisDev &b, const HamCoefDev &c)
{
    return c.pza && c.pzR > 0 && b.M >= 3 && b.M <= MCPI_PAIRZ_MAXM && !b.perm && b.IF && b.nc <= INT32_MAX &&
           b.N >= 2 && b.N <= MCPI_MAXN && nc_of(b.N - 1, b.M) < ((int64_t)1 << 27);
}

mcpi_status launch_pairz(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, cplx *ws,
                        hipStream_t st)
{
    MCPI_ARG(pairz_applicable(b, c), "the factorized pair-space H.v does not apply to this basis and Hamiltonian");
    switch (b.M) {
    case 3: return launch_pairz_m<3>(b, c, C, out, ws, st);
    case 4: return launch_pairz_m<4>(b, c, C, out, ws, st);
    case 5: return launch_pairz_m<5>(b, c, C, out, ws, st);
    case 6: return launch_pairz_m<6>(b, c, C, out, ws, st);
    case 7: return launch_pairz_m<7>(b, c, C, out, ws, st);
    case 8: return launch_pairz_m<8>(b, c, C, out, ws, st);
    case 9: return launch_pairz_m<9>(b, c, C, out, ws, st);
    case 10: return launch_pairz_m<10>(b, c, C, out, ws, st);
    case 11: return launch_pairz_m<11>(b, c, C, out, ws, st);
    case 12: return launch_pairz_m<12>(b, c, C, out, ws, st);
    case 13: return launch_pairz_m<13>(b, c, C, out, ws, st);
    case 14: return launch_pairz_m<14>(b, c, C, out, ws, st);
    case 15: return launch_pairz_m<15>(b, c, C, out, ws, st);
    case 16: return launch_pairz_m<16>(b, c, C, out, ws, st);
    case 17: return launch_pairz_m<17>(b, c, C, out, ws, st);
    case 18: return launch_pairz_m<18>(b, c, C, out, ws, st);
    case 19: return launch_pairz_m<19>(b, c, C, out, ws, st);
    case 20: return launch_pairz_m<20>(b, c, C, out, ws, st);
    case 21: return launch_pairz_m<21>(b, c, C, out, ws, st);
    case 22: return launch_pairz_m<22>(b, c, C, out, ws, st);
    case 23: return launch_pairz_m<23>(b, c, C, out, ws, st);
    case 24: return launch_pairz_m<24>(b, c, C, out, ws, st);
    default: return MCPI_ERR_ARG;
    }
}

mcpi_status launch_pairz_multi(const BasisDev &b, const HamCoefDev &c, int mv, int chunks, const cplx *C, cplx *out,
                               int64_t mC, int64_t mK, cplx *ws, hipStream_t st)
{
    MCPI_ARG(pairz_applicable(b, c) && b.M <= MCPI_PAIRZ_MULTI_MAXM && (mv == 2 || mv == 4),
             "the factorized pair form for several members does not apply");
    MCPI_ARG(mK % 16 == 0 && mC >= b.nc && chunks >= 1 && chunks <= 65535, "member strides and chunks");
#define MCPI_PZMV(MM) \
    case MM: return mv == 4 ? launch_pairz_mv_m<MM, 4>(b, c, chunks, C, out, mC, mK, ws, st) \
                            : launch_pairz_mv_m<MM, 2>(b, c, chunks, C, out, mC, mK, ws, st);
    switch (b.M) {
        MCPI_PZMV(3) MCPI_PZMV(4) MCPI_PZMV(5) MCPI_PZMV(6) MCPI_PZMV(7) MCPI_PZMV(8) MCPI_PZMV(9) MCPI_PZMV(10)
        MCPI_PZMV(11) MCPI_PZMV(12) MCPI_PZMV(13)
    default: return MCPI_ERR_ARG;
    }
#undef MCPI_PZMV
}

}  // namespace mcpi
