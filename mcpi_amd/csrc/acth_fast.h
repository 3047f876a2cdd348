// H.v kernels specialised on the number of orbitals M.
//
// Same maths as the generic kernel in acth.hip (reference applyHconf_omp,
// hamiltonianMatrix.h:1431-1970, folded into per-target coefficient tables).
// With M known at compile time a row's occupations, square roots and prefix
// counts live in registers, and the loops over orbitals are uniform across
// the wavefront: coefficients are scalar loads, and the terms of one chunk
// form a single basic block, so the loads of map entries and the gathers
// C[j] of a whole chunk are issued together instead of one term at a time.
//
// Lanes are consecutive rows: the terms a wavefront evaluates together
// gather C[j] from neighbouring targets (about four gathers per 128-byte
// line at N = M = 12).
//
// Template parameters:
//   MAP1  single jumps from the stored Map (else ranked on the fly)
//   DMAP  double jumps from the stored MapOT/MapTT (else ranked on the fly)
//   SKIP  terms whose folded coefficient is exactly zero are not evaluated:
//         the kernel walks per-chunk lists of the non-zero terms instead of
//         the whole chunk (result-neutral for finite inputs)
//
// Reading double-jump targets (DMAP): the wave-packed tables of the basis
// (basis.hip) hold, per wavefront and active jump, the q <= l targets of the
// active lanes interleaved in 16-byte groups, so one load instruction of the
// wavefront reads one contiguous run instead of 64 separate cache lines.
// The slot of a jump is found by a ballot over the lanes' activity: the
// jumps are visited in the table's order and every active one advances the
// wavefront's offset by (active lanes) x (entries per slot).
//
// Ranking on the fly: with prefix counts P_j = sum_{o<j} v_o,
// rank(v) = nc - 1 - sum_{j=1}^{M-1} B[j][P_j] (closed form of the
// FockToIndex loop, configurationsMap.h:273-305).  Removing particles from a
// uniform set of orbitals shifts P_j by -A_j (uniform) and adding them at
// (q, l) by Q_j = [q<j] + [l<j], a step function of j.  With
// U_a[x] = sum_{j=1}^{x} B[j][P_j - A_j + a] (a = 0, 1, 2) the target is
//   rank = nc - 1 - (U0[q] + U1[l] - U1[q] + U2[M-1] - U2[l])   (q < l)
//   rank = nc - 1 - (U0[q] + U2[M-1] - U2[q])                   (q == l)
// and with D01 = U0 - U1, D12 = U1 - U2, T = U2[M-1] both cases read
//   rank = nc - 1 - T - D01[q] - D12[l],
// so per chunk each lane builds two M-entry tables and every term costs two
// adds (a single jump k -> l: rank = nc - 1 - U1[M-1] - D01[l]).  Arithmetic
// is modulo 2^32 (exact: the result lies in [0, nc) and nc < 2^31).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "mcpi_internal.h"

namespace mcpi {

__device__ inline cplx ld_c(const cplx *p)
{
    double2 t = *reinterpret_cast<const double2 *>(p);
    return cmake(t.x, t.y);
}

// 16-byte piece of a chunk; dword alignment suffices for global_load_dwordx4
struct __attribute__((aligned(4))) Piece4 {
    int32_t x[4];
};

// Streamed table reads (packed and pattern slots, Map) are each used once
// per H.v: non-temporal loads keep them from displacing the gathered C
// lines in the caches (1.8% faster at N = M = 12).  Slot groups are 16-byte
// aligned (slot offsets are multiples of four ints).
typedef int32_t mcpi_v4i __attribute__((ext_vector_type(4)));
__device__ inline Piece4 ld_group_nt(const int32_t *p)
{
    const mcpi_v4i v = __builtin_nontemporal_load(reinterpret_cast<const mcpi_v4i *>(p));
    Piece4 g;
    g.x[0] = v.x;
    g.x[1] = v.y;
    g.x[2] = v.z;
    g.x[3] = v.w;
    return g;
}
__device__ inline int32_t ld_nt(const int32_t *p) { return __builtin_nontemporal_load(p); }

// acc += coef * (x * c) with a real scale x; RC: coef is known to be real
template <bool RC>
__device__ inline cplx term(cplx coef, double x, cplx c, cplx acc)
{
    if (RC) {
        const double w = coef.re * x;
        acc.re = fma(w, c.re, acc.re);
        acc.im = fma(w, c.im, acc.im);
    } else {
        const double wr = coef.re * x, wi = coef.im * x;
        acc.re = fma(wr, c.re, acc.re);
        acc.re = fma(-wi, c.im, acc.re);
        acc.im = fma(wr, c.im, acc.im);
        acc.im = fma(wi, c.re, acc.im);
    }
    return acc;
}

// Hardware block b runs on XCD b % 8 as that XCD's (b / 8)-th block.  Chunks
// of G consecutive row blocks dealt round-robin to the XCDs: each XCD gets
// rows from the whole basis (the H.v work per row differs between regions
// of the basis, so contiguous eighths leave XCDs idle at the end) while
// neighbouring blocks still share an L2.  Blocks past the last full round
// of chunks keep their index.
__device__ inline unsigned chunk_xcd(unsigned b, unsigned n, unsigned G)
{
    const unsigned full = n / (8 * G) * (8 * G);
    if (b >= full) return b;
    const unsigned x = b % 8, k = b / 8;
    return ((k / G) * 8 + x) * G + k % G;
}

// Per-row state kept in registers: the occupations packed 8 bits per orbital
// (M <= 16) and sqrt(v + 1).  Everything else is derived on use (prefix
// counts inside the rank tables, sqrt(v) and sqrt(v + 2) from the LDS table),
// which keeps the whole-chunk basic blocks within 128 VGPRs.
template <int M>
struct RowState {
    uint64_t pk[2];
    double r1[M];   // sqrt(v+1)
};

// occupation of orbital k (compile-time or uniform runtime k)
template <int M>
__device__ inline int occ(const RowState<M> &r, int k)
{
    return (int)((k < 8 ? (r.pk[0] >> (8 * k)) : (r.pk[1] >> (8 * (k - 8)))) & 0xff);
}

template <int M>
__device__ inline void load_row(const BasisDev &b, int64_t i, const double *sq, RowState<M> &r)
{
    r.pk[0] = r.pk[1] = 0;
#pragma unroll
    for (int j = 0; j < M; j++) {
        const int v = b.IF[i * M + j];
        r.r1[j] = sq[v + 1];
        if (j < 8) r.pk[0] |= (uint64_t)v << (8 * j);
        else r.pk[1] |= (uint64_t)v << (8 * (j - 8));
    }
}

// Difference tables of the prefix sums U_a (see the header comment) for the
// removal profile A_j = [ra<j] + [rb<j] (ra, rb uniform; rb = -1 removes a
// single particle).  DBL: the move adds two particles (tables D01, D12, total
// U2[M-1]); otherwise one (D01 and total U1[M-1]).  Table indices are clamped:
// out-of-range entries only enter differences that cancel.
template <int M, bool DBL>
__device__ inline void build_d(const RowState<M> &r, const uint32_t *Bt, int N, int ra, int rb,
                               uint32_t (&D0)[M], uint32_t (&D1)[M], uint32_t &tot)
{
    uint32_t s0 = 0, s1 = 0, s2 = 0;
    int P = 0;   // prefix count sum_{o<j} v_o
    D0[0] = D1[0] = 0;
#pragma unroll
    for (int j = 1; j < M; j++) {
        P += occ<M>(r, j - 1);
        const int p = P - (ra < j) - (rb >= 0 && rb < j);
        const uint32_t *row = Bt + j * (N + 1);
        const int p0 = p < 0 ? 0 : (p > N ? N : p);
        const int p1 = p + 1 < 0 ? 0 : (p + 1 > N ? N : p + 1);
        s0 += row[p0];
        s1 += row[p1];
        if (DBL) {
            const int p2 = p + 2 < 0 ? 0 : (p + 2 > N ? N : p + 2);
            s2 += row[p2];
        }
        D0[j] = s0 - s1;
        D1[j] = DBL ? s1 - s2 : 0u;
    }
    tot = DBL ? s2 : s1;
}

// per-lane copy of the difference tables in LDS (layout [a][x][lane]) for
// lookups at runtime orbitals: register arrays indexed at runtime would be
// demoted to scratch memory
template <int M>
__device__ inline void store_d(const uint32_t (&D0)[M], const uint32_t (&D1)[M], uint32_t *Dl, int lane, int T)
{
#pragma unroll
    for (int x = 0; x < M; x++) {
        Dl[x * T + lane] = D0[x];
        Dl[(M + x) * T + lane] = D1[x];
    }
}

template <int M>
__device__ inline uint32_t dl(const uint32_t *Dl, int lane, int T, int a, int x)
{
    return Dl[(a * M + x) * T + lane];
}

// ------------------------------------------------------- chunk bodies --

// Occupations of a row with the removed orbitals k and s (s == k: one) taken
// out and the rest renumbered o' = o - [k < o] - [s < o], 8 bits each: the
// orbital numbering of a packed slot (packed_pos in mcpi_internal.h)
__device__ inline void compress_occ(const uint64_t (&pk)[2], int k, int s, uint64_t (&out)[2])
{
    typedef unsigned __int128 u128;
    u128 x = ((u128)pk[1] << 64) | (u128)pk[0];
    auto drop = [](u128 v, int o) {
        const u128 lo = ((u128)1 << (8 * o)) - 1;
        return (v & lo) | ((v >> 8) & ~lo);
    };
    if (s != k) x = drop(x, s);   // s > k: dropping s first leaves k in place
    x = drop(x, k);
    out[0] = (uint64_t)x;
    out[1] = (uint64_t)(x >> 64);
}

template <int O>
__device__ inline int occ_c(const uint64_t (&pkc)[2])
{
    return (int)((O < 8 ? (pkc[0] >> (8 * O)) : (pkc[1] >> (8 * (O - 8)))) & 0xff);
}

__device__ inline cplx ldcoef(const cplx *p, int x) { return p[x]; }
__device__ inline cplx ldcoef(const __attribute__((address_space(4))) cplx *p, int x)
{
    return cmake(p[x].re, p[x].im);   // field by field (no generic-address copy)
}

// rows l' = L .. MP-1 of a packed chunk (compile-time recursion keeps every
// index a constant)
template <int M, bool RC, class CP, int MP, int L>
__device__ inline void chunk_rows(const int32_t *idx, const cplx *__restrict__ C, const uint64_t (&pkc)[2],
                                  const double *sq, CP coef, cplx &part)
{
    if constexpr (L < MP) {
        cplx pl = cmake(0, 0);
#pragma unroll
        for (int q = 0; q <= L; q++) {
            const int t = L * (L + 1) / 2 + q;
            const cplx cj = ld_c(C + (int64_t)idx[t]);
            const int vq = (int)((q < 8 ? (pkc[0] >> (8 * q)) : (pkc[1] >> (8 * (q - 8)))) & 0xff);
            if (q < L) pl = term<RC>(ldcoef(coef, t), sq[vq + 1], cj, pl);
            else part = term<RC>(ldcoef(coef, t), sq[vq + 1] * sq[vq + 2], cj, part);
        }
        const double rl = sq[occ_c<L>(pkc) + 1];
        part.re = fma(rl, pl.re, part.re);
        part.im = fma(rl, pl.im, part.im);
        chunk_rows<M, RC, CP, MP, L + 1>(idx, C, pkc, sq, coef, part);
    }
}

// The chunk sum of one (row, jump) pair from its packed slot: entry t of the
// lane at tab + off + (t / 4) * nact4 + t % 4 (16-byte groups, off and the
// stride 32-bit lane values from a uniform base), over the triangle of the
// MP orbitals left when the jump's removed orbitals are dropped.  pkc: the
// row's occupations in that numbering (compress_occ); coef: the jump's
// coefficients in slot order (HamCoefDev::otc / ttc).  Row l':
// sum_{q'<l'} coef r1[q'] C[j] is scaled once by r1[l']; q' == l' uses
// sqrt((v+1)(v+2)).  Square roots come from the LDS table sq.
// All groups are loaded first and a scheduling barrier keeps the compiler
// from sinking each group load next to its gathers: left to itself it issued
// a group load, waited for it (vmcnt 0), issued its gathers, and so on, with
// a few loads in flight; with all index loads up front the gathers of a
// chunk pipeline 12-18 deep (N = M = 12: 5.95 ms -> 4.47 ms per H.v).
template <int M, bool RC, bool TT, class CP>
__device__ inline cplx chunk_packed(const int32_t *__restrict__ tab, uint32_t off, uint32_t nact4,
                                    const cplx *__restrict__ C, const uint64_t (&pkc)[2], const double *sq, CP coef)
{
    constexpr int MP = packed_orbitals(M, TT), G = packed_entries(M, TT) / 4;
    int32_t idx[4 * G + 4];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const Piece4 pc = ld_group_nt(tab + (off + (uint32_t)g * nact4));
#pragma unroll
        for (int u = 0; u < 4; u++) idx[4 * g + u] = pc.x[u];
    }
    __builtin_amdgcn_sched_barrier(0);
    cplx part = cmake(0, 0);
    chunk_rows<M, RC, CP, MP, 0>(idx, C, pkc, sq, coef, part);
    return part;
}
// The chunk sum of one jump (k, s) with targets ranked from the row's
// difference tables (D0, D1 over all M orbitals, base = nc - 1 - total):
// only the rule terms, over the triangle of the MP orbitals left when k and
// s are dropped (as in the packed slots), the tables re-indexed to that
// numbering by a select per entry.  coef: the jump's coefficients in slot
// order (HamCoefDev::otc / ttc).
template <int M, bool RC, bool TT>
__device__ inline cplx chunk_ranked(const uint32_t (&D0)[M], const uint32_t (&D1)[M], uint32_t base,
                                    const cplx *__restrict__ C, const RowState<M> &r, const double *sq,
                                    const cplx *coef, int k, int s)
{
    constexpr int MP = packed_orbitals(M, TT), G = packed_entries(M, TT) / 4;
    // orbital of x: x below k, x + 1 between k and s, x + 2 above s (masks on
    // the loaded values: a select between table entries would be folded
    // into a runtime index and send the tables to scratch memory)
    uint32_t d0[MP], d1[MP];
#pragma unroll
    for (int x = 0; x < MP; x++) {
        const uint32_t lo = 0u - (uint32_t)(x < k);                          // x < k
        const uint32_t hi = TT ? 0u - (uint32_t)(x + 1 >= s) : 0u;           // x + 2 (TT)
        const uint32_t mid = ~lo & ~hi;                                       // x + 1
        const uint32_t a0 = D0[x], a1 = D0[x + 1], a2 = TT ? D0[x + 2 < M ? x + 2 : M - 1] : 0u;
        const uint32_t b0 = D1[x], b1 = D1[x + 1], b2 = TT ? D1[x + 2 < M ? x + 2 : M - 1] : 0u;
        d0[x] = (a0 & lo) | (a1 & mid) | (a2 & hi & ~lo);
        d1[x] = (b0 & lo) | (b1 & mid) | (b2 & hi & ~lo);
    }
    int32_t idx[4 * G + 4];
#pragma unroll
    for (int lp = 0; lp < MP; lp++)
#pragma unroll
        for (int qp = 0; qp <= lp; qp++) idx[lp * (lp + 1) / 2 + qp] = (int32_t)(base - d0[qp] - d1[lp]);
    uint64_t pkc[2];
    compress_occ(r.pk, k, s, pkc);
    // all targets first, then the gathers (see chunk_packed)
    __builtin_amdgcn_sched_barrier(0);
    cplx part = cmake(0, 0);
    chunk_rows<M, RC, const cplx *, MP, 0>(idx, C, pkc, sq, coef, part);
    return part;
}

// the non-zero entries of one chunk from a uniform list (four at a time);
// q and l are runtime values here, so occupations come from the packed copy,
// square roots from the LDS table and prefix differences from the LDS copy
// (DMAP: the slot of jump (k, s), entry (q, l) at packed_pos(k, s, q, l))
template <int M, bool DMAP, bool RC>
__device__ inline cplx chunk_list(const int32_t *__restrict__ slot, int nact4, const uint32_t *Dl, int lane, int T,
                                  uint32_t base, const cplx *__restrict__ C, const RowState<M> &r,
                                  const double *sq, const int32_t *ents, const cplx *coefs, int n, int k, int s)
{
    cplx part = cmake(0, 0);
    for (int t = 0; t < n; t += 4) {
        int64_t j[4];
        double x[4];
        cplx cf[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = t + u < n;  // uniform
            const int e = ok ? ents[t + u] : 0;
            const int q = e % M, l = e / M;
            if (DMAP) {
                // listed entries are rule terms: q <= l, neither in {k, s}
                const int t = ok ? packed_pos(k, s, q, l) : 0;
                j[u] = ld_nt(slot + (int64_t)(t >> 2) * nact4 + (t & 3));
            }
            else j[u] = (int64_t)(uint32_t)(base - dl<M>(Dl, lane, T, 0, q) - dl<M>(Dl, lane, T, 1, l));
            const int vq = occ<M>(r, q), vl = occ<M>(r, l);
            x[u] = ok ? ((q == l) ? sq[vq + 1] * sq[vq + 2] : sq[vq + 1] * sq[vl + 1]) : 0.0;
            cf[u] = ok ? coefs[t + u] : cmake(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) part = term<RC>(cf[u], x[u], ld_c(C + j[u]), part);
    }
    return part;
}

// the listed entries of one jump read from the pattern tables: entry t of
// this lane at slot + t * nact (entry-major slots), four entries at a time
// with their target loads ahead of the gathers; past n the lane's own row
// stands in (coefficient zero)
template <int M, bool RC>
__device__ inline cplx chunk_pat(const int32_t *__restrict__ slot, int nact, const cplx *__restrict__ C,
                                 const RowState<M> &r, const double *sq, const int32_t *ents, const cplx *coefs,
                                 int n, int32_t self)
{
    cplx part = cmake(0, 0);
    for (int t = 0; t < n; t += 4) {
        Piece4 g;
#pragma unroll
        for (int u = 0; u < 4; u++) g.x[u] = t + u < n ? ld_nt(slot + (int64_t)(t + u) * nact) : self;
        double x[4];
        cplx cf[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = t + u < n;  // uniform
            const int e = ok ? ents[t + u] : 0;
            const int q = e % M, l = e / M;
            const int vq = occ<M>(r, q), vl = occ<M>(r, l);
            x[u] = ok ? ((q == l) ? sq[vq + 1] * sq[vq + 2] : sq[vq + 1] * sq[vl + 1]) : 0.0;
            cf[u] = ok ? coefs[t + u] : cmake(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) part = term<RC>(cf[u], x[u], ld_c(C + (int64_t)g.x[u]), part);
    }
    return part;
}

// ---------------------------------------------------------------- kernel --

// Occupancy: the zero-skipping kernels fit in 128 VGPRs (4 waves per SIMD);
// the every-term kernels keep a whole chunk's gathers in flight and are
// given up to 256 VGPRs rather than spill (the packed-table variant uses
// about 150, 3 waves per SIMD; forcing 4 waves spills 16-112 bytes per lane
// depending on M and gained 3% at N = M = 12).
// BATCH: nbatch instances share the basis.  Instance x's coefficient tables
// are instance 0's shifted by x * cstride bytes and its vectors sit
// x * vstride elements after C0 / out0, so every pointer derives from a
// kernel argument and keeps its global / scalar-load form.  The 1-D grid
// holds nbatch blocks per row block with the instance index fastest, and
// chunk_xcd deals chunks of 32 row blocks (32 * nbatch blocks) to the XCDs:
// the nbatch blocks of one row range run back to back on the same XCD, so
// the row range's tables are read from HBM once and served to the other
// instances from that XCD's L2.
// PAT (with SKIP): double-jump targets of the listed terms from the
// pattern tables of the Hamiltonian's zero pattern (no ranking, no packed
// table lookups).
template <int M, bool MAP1, bool DMAP, bool SKIP, bool RC, bool BATCH, bool PAT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SKIP ? 4 : 2))) void k_acth_fast(
    BasisDev b, HamCoefDev c0, const cplx *__restrict__ C0, cplx *__restrict__ out0, int nbatch, size_t cstride,
    int64_t vstride)
{
    // chunks of 32 row blocks (of the nbatch blocks of each with BATCH) dealt
    // round-robin over the XCDs (N = 16, M = 12 zero-skipping: 3.35 ms,
    // against 3.50 ms with single blocks dealt round-robin and 3.66 ms with
    // each XCD walking one contiguous eighth of the rows)
    const unsigned vb = chunk_xcd(blockIdx.x, gridDim.x, BATCH ? 32u * (unsigned)nbatch : 32u);
    const unsigned inst = BATCH ? vb % (unsigned)nbatch : 0u;
    const unsigned rb = BATCH ? vb / (unsigned)nbatch : vb;
    const HamCoefDev c = BATCH ? shift_coef(c0, (ptrdiff_t)((size_t)inst * cstride)) : c0;
    const cplx *__restrict__ C = BATCH ? C0 + (int64_t)inst * vstride : C0;
    cplx *__restrict__ out = BATCH ? out0 + (int64_t)inst * vstride : out0;
    constexpr int MM = M * M;
    extern __shared__ double fsm[];
    double *sq = fsm;                                   // N+3
    uint32_t *Bt = (uint32_t *)(sq + b.N + 3);          // M*(N+1)
    uint32_t *Dl = Bt + M * (b.N + 1);                  // 2*M*blockDim (SKIP with ranking only)
    const int lane = threadIdx.x, T = blockDim.x;
    for (int x = threadIdx.x; x < b.N + 3; x += blockDim.x) sq[x] = sqrt((double)x);
    if (!MAP1 || (!DMAP && !PAT))
        for (int x = threadIdx.x; x < M * (b.N + 1); x += blockDim.x) Bt[x] = (uint32_t)b.B[x];
    __syncthreads();
    const int64_t i = (int64_t)rb * blockDim.x + threadIdx.x;
    if (i >= b.nc) return;
    const int N = b.N;
    const int64_t nc = b.nc;
    const uint32_t top = (uint32_t)(nc - 1);
    RowState<M> r;
    load_row<M>(b, i, sq, r);

    // diagonal: 1-body diagonal, rules 1 and 2
    cplx dsum = cmake(0, 0);
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int ik = occ<M>(r, k);
        const double vk = (double)ik;
        dsum = cfma(c.dk[k], cmake(vk, 0), dsum);
        dsum = cfma(c.d1[k], cmake(vk * (ik - 1), 0), dsum);
#pragma unroll
        for (int s = k + 1; s < M; s++) dsum = cfma(c.d2[k + M * s], cmake(vk * occ<M>(r, s), 0), dsum);
    }
    cplx acc = cfma(dsum, ld_c(C + i), cmake(0, 0));

    // single jumps k -> l: 1-body off-diagonal, rules 4, 5, 8
    for (int k = 0; k < M; k++) {
        const uint32_t lmask = c.m1j[k];   // l with a non-zero coefficient
        if (SKIP && lmask == 0) continue;
        const int vk = occ<M>(r, k);
        const bool act = vk >= 1;
        if (!SKIP && __ballot(act) == 0) continue;   // orbital k empty in every row of the wavefront
        uint32_t D0[M], D1[M], tot = 0;
        if (!MAP1) build_d<M, false>(r, Bt, N, k, -1, D0, D1, tot);
        if (!MAP1 && SKIP) store_d<M>(D0, D1, Dl, lane, T);
        const uint32_t base = top - tot;
        const double fk = act ? sq[vk] : 0.0;
        if (!SKIP) {
#pragma unroll
            for (int l = 0; l < M; l++) {
                const int kl = k + M * l;   // the l == k entry has zero coefficients
                cplx h = c.A0[kl];
                const cplx *Ak = c.A + (int64_t)kl * M;
#pragma unroll
                for (int s = 0; s < M; s++) {
                    h.re = fma((double)occ<M>(r, s), Ak[s].re, h.re);
                    h.im = fma((double)occ<M>(r, s), Ak[s].im, h.im);
                }
                int64_t j;
                if (MAP1) j = ld_nt(b.Map + i + (int64_t)kl * nc);
                else j = (int64_t)(uint32_t)(base - D0[l]);
                j = act ? j : i;
                acc = term<RC>(h, fk * r.r1[l], ld_c(C + j), acc);
            }
        } else {
            uint32_t m = lmask;
            while (m) {
                const int l = __ffs(m) - 1;
                m &= m - 1;
                const int kl = k + M * l;
                cplx h = c.A0[kl];
                const cplx *Ak = c.A + (int64_t)kl * M;
#pragma unroll
                for (int s = 0; s < M; s++) {
                    h.re = fma((double)occ<M>(r, s), Ak[s].re, h.re);
                    h.im = fma((double)occ<M>(r, s), Ak[s].im, h.im);
                }
                int64_t j;
                if (MAP1) j = act ? ld_nt(b.Map + i + (int64_t)kl * nc) : i;
                else j = act ? (int64_t)(uint32_t)(base - dl<M>(Dl, lane, T, 0, l)) : i;
                acc = term<RC>(h, fk * sq[occ<M>(r, l) + 1], ld_c(C + j), acc);
            }
        }
    }

    // double jumps: slots of the packed tables (DMAP) follow the jumps
    // active in any lane, in the order visited below
    constexpr int EO = packed_entries(M, false), ET = packed_entries(M, true);
    const int lw = threadIdx.x & 63;
    const uint64_t below = lw ? (~0ull >> (64 - lw)) : 0ull;
    const int64_t wv = i >> 6;
    if (PAT) {
        // pattern tables: slots only for jumps with a non-empty list
        int64_t pb = b.ppoffOT[wv];
        for (int k = 0; k < M; k++) {
            const int beg = c.ot_off[k], n_nz = c.ot_off[k + 1] - beg;
            if (n_nz == 0) continue;
            const int vk = occ<M>(r, k);
            const bool act = vk >= 2;
            const uint64_t m = __ballot(act);
            const int nact = __popcll(m);
            const int32_t *slot = b.patOT + pb + __popcll(m & below);
            pb += (int64_t)nact * n_nz;
            if (!act) continue;
            const cplx part = chunk_pat<M, RC>(slot, nact, C, r, sq, c.nz_e + beg, c.nz_c + beg, n_nz, (int32_t)i);
            const double f = sq[vk] * sq[vk - 1];
            acc.re = fma(f, part.re, acc.re);
            acc.im = fma(f, part.im, acc.im);
        }
        pb = b.ppoffTT[wv];
        int p = 0;
        for (int k = 0; k < M; k++) {
            const int vk = occ<M>(r, k);
            for (int s = k + 1; s < M; s++, p++) {
                const int beg = c.tt_off[p], n_nz = c.tt_off[p + 1] - beg;
                if (n_nz == 0) continue;
                const int vs = occ<M>(r, s);
                const bool act = vk >= 1 && vs >= 1;
                const uint64_t m = __ballot(act);
                const int nact = __popcll(m);
                const int32_t *slot = b.patTT + pb + __popcll(m & below);
                pb += (int64_t)nact * n_nz;
                if (!act) continue;
                const cplx part =
                    chunk_pat<M, RC>(slot, nact, C, r, sq, c.nz_e + beg, c.nz_c + beg, n_nz, (int32_t)i);
                const double f = sq[vk] * sq[vs];
                acc.re = fma(f, part.re, acc.re);
                acc.im = fma(f, part.im, acc.im);
            }
        }
    }
    // same-orbital double jumps: rules 3, 6
    if (!PAT) {
        int64_t pb = DMAP ? b.poffOT[wv] : 0;
        for (int k = 0; k < M; k++) {
            const int n_nz = SKIP ? c.ot_off[k + 1] - c.ot_off[k] : 1;
            const int vk = occ<M>(r, k);
            const bool act = vk >= 2;
            const uint64_t m = DMAP ? __ballot(act) : 0ull;
            const int nact = __popcll(m);
            const int32_t *slot0 = DMAP ? b.packOT + pb : nullptr;   // the jump's slot, uniform
            const uint32_t soff = 4u * (uint32_t)__popcll(m & below);
            const int32_t *slot = DMAP ? slot0 + soff : nullptr;
            pb += (int64_t)nact * EO;
            if (SKIP && n_nz == 0) continue;
            if (!act) continue;
            uint32_t D0[M], D1[M], tot = 0;
            if (!DMAP) build_d<M, true>(r, Bt, N, k, k, D0, D1, tot);
            if (!DMAP && SKIP) store_d<M>(D0, D1, Dl, lane, T);
            const uint32_t base = top - tot;
            const double f = sq[vk] * sq[vk - 1];
            cplx part;
            if (!SKIP) {
                if (DMAP) {
                    uint64_t pkc[2];
                    compress_occ(r.pk, k, k, pkc);
                    part = chunk_packed<M, RC, false>(slot0, soff, 4u * (uint32_t)nact, C, pkc, sq,
                                                      (const cplx *)(c.otc + (int64_t)k * EO));
                }
                else part = chunk_ranked<M, RC, false>(D0, D1, base, C, r, sq, c.otc + (int64_t)k * EO, k, k);
            } else {
                part = chunk_list<M, DMAP, RC>(slot, 4 * nact, Dl, lane, T, base, C, r, sq, c.nz_e + c.ot_off[k],
                                               c.nz_c + c.ot_off[k], n_nz, k, k);
            }
            acc.re = fma(f, part.re, acc.re);
            acc.im = fma(f, part.im, acc.im);
        }
    }
    // two-orbital double jumps: rules 7, 9
    if (!PAT) {
        int64_t pb = DMAP ? b.poffTT[wv] : 0;
        int p = 0;
        for (int k = 0; k < M; k++) {
            const int vk = occ<M>(r, k);
            for (int s = k + 1; s < M; s++, p++) {
                const int n_nz = SKIP ? c.tt_off[p + 1] - c.tt_off[p] : 1;
                const int vs = occ<M>(r, s);
                const bool act = vk >= 1 && vs >= 1;
                const uint64_t m = DMAP ? __ballot(act) : 0ull;
                const int nact = __popcll(m);
                const int32_t *slot0 = DMAP ? b.packTT + pb : nullptr;
                const uint32_t soff = 4u * (uint32_t)__popcll(m & below);
                const int32_t *slot = DMAP ? slot0 + soff : nullptr;
                pb += (int64_t)nact * ET;
                if (SKIP && n_nz == 0) continue;
                if (!act) continue;
                uint32_t D0[M], D1[M], tot = 0;
                if (!DMAP) build_d<M, true>(r, Bt, N, k, s, D0, D1, tot);
                if (!DMAP && SKIP) store_d<M>(D0, D1, Dl, lane, T);
                const uint32_t base = top - tot;
                const double f = sq[vk] * sq[vs];
                cplx part;
                if (!SKIP) {
                    if (DMAP) {
                        uint64_t pkc[2];
                        compress_occ(r.pk, k, s, pkc);
                        part = chunk_packed<M, RC, true>(slot0, soff, 4u * (uint32_t)nact, C, pkc, sq,
                                                         (const cplx *)(c.ttc + (int64_t)p * ET));
                    }
                    else part = chunk_ranked<M, RC, true>(D0, D1, base, C, r, sq, c.ttc + (int64_t)p * ET, k, s);
                } else {
                    part = chunk_list<M, DMAP, RC>(slot, 4 * nact, Dl, lane, T, base, C, r, sq,
                                                   c.nz_e + c.tt_off[p], c.nz_c + c.tt_off[p], n_nz, k, s);
                }
                acc.re = fma(f, part.re, acc.re);
                acc.im = fma(f, part.im, acc.im);
            }
        }
    }
    double2 o;
    o.x = acc.re;
    o.y = acc.im;
    *reinterpret_cast<double2 *>(out + i) = o;
}

// launch one specialisation; rank_doubles: rank double jumps even when the
// maps are stored; real: every folded coefficient has a zero imaginary part.
// nbatch > 0: batched launch over nbatch instances at fixed strides.
template <int M>
inline mcpi_status launch_fast_m(const BasisDev &b, int maps, const HamCoefDev &c, int skip, int rank_doubles,
                                 const cplx *C, cplx *out, hipStream_t st, int nbatch = 0, size_t cstride = 0,
                                 int64_t vstride = 0)
{
    dim3 grid((unsigned)((b.nc + 255) / 256) * (unsigned)(nbatch > 0 ? nbatch : 1)), blk(256);
    size_t sm = sizeof(double) * (b.N + 3) + sizeof(uint32_t) * M * (b.N + 1) + 16;
    const bool map1 = maps >= MCPI_MAPS_SINGLE;
    if (skip && !(map1 && b.patTT && b.patOT)) sm += sizeof(uint32_t) * 2 * M * 256;
    const bool dmap = maps == MCPI_MAPS_FULL && !rank_doubles && b.packTT && b.packOT;
    const bool real = c.real;
    const bool pat = skip && map1 && b.patTT && b.patOT;
#define MCPI_LP(A, B, S, R, P)                                                                              \
    {                                                                                                       \
        if (nbatch > 0)                                                                                     \
            hipLaunchKernelGGL((k_acth_fast<M, A, B, S, R, true, P>), grid, blk, sm, st, b, c, C, out,      \
                               nbatch, cstride, vstride);                                                   \
        else                                                                                                \
            hipLaunchKernelGGL((k_acth_fast<M, A, B, S, R, false, P>), grid, blk, sm, st, b, c, C, out,     \
                               nbatch, cstride, vstride);                                                   \
    }
#define MCPI_L(A, B, S, R) MCPI_LP(A, B, S, R, false)
#define MCPI_L2(A, B, S) { if (real) MCPI_L(A, B, S, true) else MCPI_L(A, B, S, false) }
    if (pat) { if (real) MCPI_LP(true, false, true, true, true) else MCPI_LP(true, false, true, false, true) }
    else if (map1 && dmap) { if (skip) MCPI_L2(true, true, true) else MCPI_L2(true, true, false) }
    else if (map1) { if (skip) MCPI_L2(true, false, true) else MCPI_L2(true, false, false) }
    else { if (skip) MCPI_L2(false, false, true) else MCPI_L2(false, false, false) }
#undef MCPI_L2
#undef MCPI_L
#undef MCPI_LP
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

}  // namespace mcpi
