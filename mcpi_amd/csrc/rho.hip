// One- and two-body density matrices and the Hellmann-Feynman adjoint.
//
// Reference: OBrho_XM, onebodyMatrix.h:190-251 (rho1), and TBrho_XX,
// twobodyMatrix.h:991-1487 (rho2, rules 1-8 plus symmetry copies).
//
// The reference loops orbital-outer / row-inner.  Here every class of
// representative terms is summed into LDS accumulators of a block
// (atomically: the wavefronts of a block add in the order they arrive, so
// the last bits of a sum can differ from run to run); each block writes its
// accumulators once and a second kernel sums the blocks in a fixed order.
// The host expands the representative terms into full matrices with the
// reference's symmetry copies.  Single jumps and the diagonal (class 0):
// lanes are rows and each term is a wavefront reduction.  Double jumps
// (classes 1, 2): lanes are terms and a wavefront steps through the rows
// where a jump is active, so no reduction is needed.
//
// Single-jump targets come from Map when the basis has it and are ranked on
// the fly otherwise (closed-form FockToIndex over the row's prefix counts,
// configurationsMap.h:273-305); double-jump targets are ranked from per-row
// difference tables (M <= 16) or read from the maps, so the adjoint also
// runs on bases that store only Map, or no map at all.
//
// Term classes (compact layouts, P = M(M-1)/2 pairs, pq = pair index):
//   class 0 (single jumps and diagonal; class 3: r1d and r1o only):
//     r1d[k] | r1o[kl] | R1[k] | R2[ks] | R4[kl] | R5[ks] | R7[kl][s]
//   class 1 (same-orbital double jumps):  R3[kq] | R6[k][ql]
//   class 2 (two-orbital double jumps):   R8[ks - p0][ql]   (pairs p0..p1-1)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mcpi_internal.h"
#include "rho_expand.h"

namespace mcpi {

__device__ inline cplx ldcr(const cplx *p)
{
    double2 t = *reinterpret_cast<const double2 *>(p);
    return cmake(t.x, t.y);
}

__device__ inline double wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// add the wave total of z to acc[idx] (one LDS atomic pair per wave)
__device__ inline void wave_add(cplx z, cplx *acc, int idx, bool any)
{
    if (!any) return;  // uniform: no lane contributes
    double re = wave_sum(z.re), im = wave_sum(z.im);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[idx].re, re);
        atomicAdd(&acc[idx].im, im);
    }
}

__device__ inline bool wave_any(bool p) { return __ballot(p) != 0; }

// One exchange step of the batched reduction: of its 2H values the lane keeps
// the half selected by lane bit `mask` (upper half if set), adds its
// partner's copy of that half and passes the other half to the partner.
template <int H>
__device__ inline void xchg_step(cplx (&z)[8], int mask)
{
    const bool hi = (threadIdx.x & mask) != 0;
#pragma unroll
    for (int u = 0; u < H; u++) {
        const cplx keep = hi ? z[u + H] : z[u];
        const cplx send = hi ? z[u] : z[u + H];
        z[u].re = keep.re + __shfl_xor(send.re, mask, 64);
        z[u].im = keep.im + __shfl_xor(send.im, mask, 64);
    }
}

// Wavefront totals of 8 per-lane values z[0..7], added to acc[slot[u]] for
// the u in the uniform bit mask `use`.  Three exchange steps over lane bits
// 0, 1, 2 leave lane L with a partial of term u(L) = 4 b0 + 2 b1 + b2 over
// the 8 lanes L ^ x, x < 8; three butterfly steps over lane bits 3, 4, 5
// complete it in every lane, and lanes 0..7 (one per term) add the totals:
// 10 complex shuffles and one LDS atomic pair for the 8 terms instead of 48
// shuffles and 8 atomic pairs one after another.
__device__ inline void wave_add8(cplx (&z)[8], cplx *acc, const int (&slot)[8], uint32_t use)
{
    xchg_step<4>(z, 1);
    xchg_step<2>(z, 2);
    xchg_step<1>(z, 4);
    double re = z[0].re, im = z[0].im;
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
    }
    const int lane = threadIdx.x & 63;
    const int u = ((lane & 1) << 2) | (lane & 2) | ((lane >> 2) & 1);
    // slot[u] by masks: a select chain would be folded into a runtime index
    // and send the slot array to scratch memory
    int sl = 0;
#pragma unroll
    for (int x = 0; x < 8; x++) sl |= slot[x] & (0 - (int)(u == x));
    if (lane < 8 && ((use >> u) & 1u)) {
        atomicAdd(&acc[sl].re, re);
        atomicAdd(&acc[sl].im, im);
    }
}

// rank of this lane's row with particles removed from a, b2 (b2 = -1: none)
// and added to c, d (d = -1: none); occupations from the LDS tile
// ([orbital][row of tile], T rows)
__device__ inline int64_t rank_lane(const uint8_t *Vs, int t, int T, const int32_t *Bt, int N, int M, int64_t nc,
                                    int a, int b2, int c, int d)
{
    int P = 0;
    int64_t s = 0;
    for (int j = 1; j < M; j++) {
        P += Vs[(j - 1) * T + t];
        const int p = P - (a < j) - (b2 >= 0 && b2 < j) + (c < j) + (d >= 0 && d < j);
        s += Bt[j * (N + 1) + p];
    }
    return nc - 1 - s;
}

// Terms of a double jump in lane order (lane-per-term kernels below).
// Class 2, jump (k, s), k < s: the pairs q < l of the M - 2 orbitals other
// than k and s, term u = l'(l'-1)/2 + q' in the renumbering of packed_pos;
// output (p - p0) P + pair(q, l).  Class 1, jump k: first the M - 1 - k
// pairs (q, q), q = k + 1 + u, output pair(k, q); then the pairs q < l of
// the M - 1 orbitals other than k, output P + k P + pair(q, l).  e: entry of
// the packed slot; tix: entry of the row's M x M block of the
// reference-layout map (q + l M for class 2, l + q M for class 1).
struct JumpTerm {
    int q, l, slot, e, tix;
    bool diag, ok;
};

// (q', l'), q' < l', of the strict-triangle index u = l'(l'-1)/2 + q'
__device__ inline void strict_tri(int u, int &qp, int &lp)
{
    int L = (int)((1.0 + sqrt(1.0 + 8.0 * (double)u)) * 0.5);
    while (L * (L - 1) / 2 > u) L--;
    while ((L + 1) * L / 2 <= u) L++;
    lp = L;
    qp = u - L * (L - 1) / 2;
}

__device__ inline int jump_terms(int cls, int M, int k)
{
    const int MP = cls == 2 ? M - 2 : M - 1;
    return (cls == 2 ? 0 : M - 1 - k) + (MP > 1 ? MP * (MP - 1) / 2 : 0);
}

template <int CLASS>
__device__ inline JumpTerm jump_term(int M, int P, int k, int s, int p, int p0, int u)
{
    JumpTerm J;
    J.ok = u < jump_terms(CLASS, M, k);
    int qp, lp;
    if (CLASS == 2) {
        strict_tri(u, qp, lp);
        J.q = packed_orbital(k, s, qp);
        J.l = packed_orbital(k, s, lp);
        J.diag = false;
        J.slot = (p - p0) * P + (J.ok ? pair_index(J.q, J.l, M) : 0);
        J.tix = J.q + J.l * M;
    } else {
        const int n3 = M - 1 - k;
        if (u < n3) {
            J.q = J.l = k + 1 + u;
            J.diag = true;
            J.slot = pair_index(k, J.q, M);
        } else {
            strict_tri(u - n3, qp, lp);
            J.q = packed_orbital(k, k, qp);
            J.l = packed_orbital(k, k, lp);
            J.diag = false;
            J.slot = J.ok ? P + k * P + pair_index(J.q, J.l, M) : 0;
        }
        J.tix = J.l + J.q * M;
    }
    if (!J.ok) J.q = J.l = J.slot = J.tix = 0;
    J.e = J.ok ? packed_pos(k, s, J.q, J.l) : 0;
    return J;
}

// rank of a row (occupations row[0..M-1]) with particles removed from a and
// b2 and added to c and d (configurationsMap.h:273-305, closed form)
__device__ inline int64_t rank_row(const uint8_t *row, const int32_t *Bt, int N, int M, int64_t nc, int a, int b2,
                                   int c, int d)
{
    int P = 0;
    int64_t s = 0;
    for (int j = 1; j < M; j++) {
        P += row[j - 1];
        const int p = P - (a < j) - (b2 < j) + (c < j) + (d < j);
        s += Bt[j * (N + 1) + p];
    }
    return nc - 1 - s;
}

// inclusive prefix sum over each row of 16 lanes (DPP row shifts; lanes
// shifted in from outside the row read 0)
__device__ inline uint32_t row_scan16(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    return v;
}

__device__ inline double readlane_d(double x, int lane)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ inline int64_t readlane_l(int64_t x, int lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)x, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Class 0.  A block of T = 1024 threads walks tiles of T*R rows (lanes =
// rows); lane t holds rows tile + t + T*r, r < R, and sums its R rows'
// contributions to a term before the wavefront reduction, which divides the
// number of reductions and LDS atomics by R.  ALL: every class-0 term (the
// two-body density); else r1d and r1o only (the one-body density).  With
// ALL the terms of a pair (k, l) (R2, r1o, R4, R5 and the M - 2 terms of
// R7) are reduced 8 at a time (wave_add8); without, the r1o terms of eight
// pairs.
template <bool ALL, int R>
__global__ __launch_bounds__(1024) void k_rho(BasisDev b, const cplx *__restrict__ C, int nterms,
                                             cplx *__restrict__ partial)
{
    extern __shared__ double smem_r[];
    const int M = b.M, N = b.N, T = blockDim.x, t = threadIdx.x, TR = T * R;
    const int P = M * (M - 1) / 2;
    const int64_t nc = b.nc;
    cplx *acc = (cplx *)smem_r;                       // nterms
    double *sq = (double *)(acc + nterms);            // N+3
    int32_t *Bt = (int32_t *)(sq + N + 3);            // M*(N+1) rank table
    uint8_t *Vs = (uint8_t *)(Bt + M * (N + 1));      // M*T*R occupations, [j][row of tile]
    for (int x = t; x < nterms; x += T) acc[x] = cmake(0, 0);
    for (int x = t; x < N + 3; x += T) sq[x] = sqrt((double)x);
    for (int x = t; x < M * (N + 1); x += T) Bt[x] = b.B[x];
    __syncthreads();

    for (int64_t tile = (int64_t)blockIdx.x * TR; tile < nc; tile += (int64_t)gridDim.x * TR) {
        int64_t i[R];
        bool valid[R];
        cplx ci[R];
        double m2[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            i[r] = tile + t + (int64_t)T * r;
            valid[r] = i[r] < nc;
            for (int j = 0; j < M; j++) Vs[j * TR + t + T * r] = (uint8_t)(valid[r] ? b.IF[i[r] * M + j] : 0);
            ci[r] = valid[r] ? ldcr(C + i[r]) : cmake(0, 0);
            m2[r] = ci[r].re * ci[r].re + ci[r].im * ci[r].im;
        }
        __syncthreads();
#define V(j, r) ((int)Vs[(j) * TR + t + T * (r)])
        {
            // r1d, R1
            for (int k = 0; k < M; k++) {
                double a = 0, a2 = 0;
                bool any2 = false;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int vk = V(k, r);
                    a += m2[r] * vk;
                    a2 += m2[r] * vk * (vk - 1);
                    any2 = any2 || vk >= 2;
                }
                wave_add(cmake(a, 0), acc, k, true);
                if (ALL) wave_add(cmake(a2, 0), acc, M + P + k, wave_any(any2));
            }
            const int o_r1o = M, o_R2 = M + P + M, o_R4 = o_R2 + P, o_R5 = o_R4 + P, o_R7 = o_R5 + P;
            if (!ALL) {
                // r1o only: eight pairs a batch, their targets and gathers
                // in flight together and one batched reduction
                int k = 0, l = 1;
                for (int pq0 = 0; pq0 < P; pq0 += 8) {
                    cplx zb[8];
                    int sb[8];
                    uint32_t use = 0;
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        zb[u] = cmake(0, 0);
                        sb[u] = 0;
                        if (pq0 + u >= P) continue;   // uniform
                        use |= 1u << u;
                        sb[u] = o_r1o + pq0 + u;
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const int vk = V(k, r), vl = V(l, r);
                            if (valid[r] && vk >= 1) {
                                const int64_t j = b.Map ? (int64_t)b.Map[i[r] + (int64_t)(k + M * l) * nc]
                                                        : rank_lane(Vs, t + T * r, TR, Bt, N, M, nc, k, -1, l, -1);
                                const cplx g = cfma_conj(ci[r], ldcr(C + j), cmake(0, 0));   // conj(C_i) C_j
                                zb[u] = cfma(g, cmake(sq[vk] * sq[vl + 1], 0), zb[u]);
                            }
                        }
                        if (++l == M) {
                            ++k;
                            l = k + 1;
                        }
                    }
                    wave_add8(zb, acc, sb, use);
                }
            }
            int pq = 0;
            for (int k = 0; ALL && k < M; k++)
                for (int l = k + 1; l < M; l++, pq++) {
                    // R2: |C|^2 v_k v_l over all rows
                    double a2 = 0;
                    bool anyact = false;
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int vk = V(k, r), vl = V(l, r);
                        a2 += m2[r] * vk * vl;
                        anyact = anyact || (valid[r] && vk >= 1);
                    }
                    if (!wave_any(anyact)) continue;   // no row with v_k >= 1: every term is zero
                    cplx g[R];
                    double root[R];
                    cplx z1 = cmake(0, 0), z4 = cmake(0, 0), z5 = cmake(0, 0);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int vk = V(k, r), vl = V(l, r);
                        const bool act = valid[r] && vk >= 1;
                        g[r] = cmake(0, 0);
                        if (act) {
                            const int64_t j = b.Map ? (int64_t)b.Map[i[r] + (int64_t)(k + M * l) * nc]
                                                    : rank_lane(Vs, t + T * r, TR, Bt, N, M, nc, k, -1, l, -1);
                            g[r] = cfma_conj(ci[r], ldcr(C + j), cmake(0, 0));   // conj(C_i) C_j
                        }
                        root[r] = act ? sq[vk] * sq[vl + 1] : 0.0;
                        z1 = cfma(g[r], cmake(root[r], 0), z1);
                        z4 = cfma(g[r], cmake(act && vk >= 2 ? (vk - 1) * root[r] : 0.0, 0), z4);
                        z5 = cfma(g[r], cmake(act && vl >= 1 ? vl * root[r] : 0.0, 0), z5);
                    }
                    // terms x of the pair: R2, r1o, R4, R5, then R7 for the
                    // x - 4'th orbital s other than k and l
                    const int nx = 4 + (M - 2);
                    for (int x0 = 0; x0 < nx; x0 += 8) {
                        cplx zb[8];
                        int sb[8];
                        uint32_t use = 0;
#pragma unroll
                        for (int u = 0; u < 8; u++) {
                            const int x = x0 + u;   // uniform
                            zb[u] = cmake(0, 0);
                            sb[u] = 0;
                            if (x >= nx) continue;
                            use |= 1u << u;
                            if (x == 0) { zb[u] = cmake(a2, 0); sb[u] = o_R2 + pq; }
                            else if (x == 1) { zb[u] = z1; sb[u] = o_r1o + pq; }
                            else if (x == 2) { zb[u] = z4; sb[u] = o_R4 + pq; }
                            else if (x == 3) { zb[u] = z5; sb[u] = o_R5 + pq; }
                            else {
                                int so = x - 4;
                                so += (so >= k);
                                so += (so >= l);
#pragma unroll
                                for (int r = 0; r < R; r++) {
                                    const int vs = V(so, r);
                                    const bool act = valid[r] && V(k, r) >= 1;
                                    zb[u] = cfma(g[r], cmake(act && vs >= 1 ? vs * root[r] : 0.0, 0), zb[u]);
                                }
                                sb[u] = o_R7 + pq * M + so;
                            }
                        }
                        wave_add8(zb, acc, sb, use);
                    }
                }
        }
#undef V
        __syncthreads();
    }
    for (int x = t; x < nterms; x += T) partial[(int64_t)blockIdx.x * nterms + x] = acc[x];
}

// Classes 1 and 2, lane per term.  A wavefront owns 64 consecutive rows (one
// wavefront of the packed tables, whose slot order it follows) and walks the
// jumps; for each jump it steps through the rows where the jump is active,
// RHO_G at a time, and lane u evaluates term u of the jump (jump_term) on
// each of them, summing in registers.  One LDS atomic pair per lane and jump
// adds the wavefront's sums to the block's accumulators: no reductions
// across lanes, and each load instruction carries up to 64 useful lanes
// (with rows as lanes only the rows where the jump is active would be
// useful, about a quarter of them at N = M = 12).  Targets ranked on the
// fly (mode 2), or for M > 16 from the packed tables (mode 0) or the
// reference-layout maps (1) when the basis has them.
// The row's amplitude is broadcast with readlane; its occupations are read
// from the wavefront's LDS rows ([row][orbital], MS bytes a row).
#define RHO_G 8

template <int CLASS>
__global__ __launch_bounds__(1024) void k_rho_jumps(BasisDev b, const cplx *__restrict__ C, int p0, int p1,
                                                   int nterms, cplx *__restrict__ partial)
{
    static_assert(CLASS == 1 || CLASS == 2, "double-jump classes");
    constexpr bool TT = CLASS == 2;
    extern __shared__ double smem_r[];
    const int M = b.M, N = b.N, T = blockDim.x, t = threadIdx.x;
    const int MS = (M + 3) & ~3;
    const int P = M * (M - 1) / 2;
    const int64_t nc = b.nc, MM = (int64_t)M * M, nwv = (nc + 63) >> 6;
    const int lane = t & 63, wl = t >> 6, nwb = T >> 6;
    cplx *acc = (cplx *)smem_r;                                              // nterms
    double *sq = (double *)(acc + nterms);                                   // N+3
    int32_t *Bt = (int32_t *)(sq + N + 3);                                   // M*(N+1) rank table
    uint8_t *Vs = (uint8_t *)(Bt + M * (N + 1)) + (size_t)wl * 64 * MS;     // this wavefront's rows
    uint8_t *PF = (uint8_t *)(Bt + M * (N + 1)) + (size_t)(nwb + wl) * 64 * MS;  // their prefix counts
    for (int x = t; x < nterms; x += T) acc[x] = cmake(0, 0);
    for (int x = t; x < N + 3; x += T) sq[x] = sqrt((double)x);
    for (int x = t; x < M * (N + 1); x += T) Bt[x] = b.B[x];
    __syncthreads();

    const int32_t *pack = TT ? b.packTT : b.packOT;
    const int64_t *poff = TT ? b.poffTT : b.poffOT;
    const int32_t *tab = TT ? b.MapTT : b.MapOT;
    const int64_t *stride = TT ? b.strideTT : b.strideOT;
    // M <= 16: targets ranked from difference tables built once per row and
    // jump (one row of 16 lanes per table), which is cheaper than reading
    // them from the maps (N = M = 12: 7.6 against 10.7 ms for rho2); else
    // from the packed tables, the reference-layout maps or rank_row per term
    const bool dtab = M <= 16;
    const int mode = dtab ? 2 : (pack ? 0 : (tab ? 1 : 2));
    const int E = packed_entries(M, TT);
    const uint32_t top = (uint32_t)(nc - 1);

    for (int64_t w = (int64_t)blockIdx.x * nwb + wl; w < nwv; w += (int64_t)gridDim.x * nwb) {
        const int64_t i = w * 64 + lane;
        const bool valid = i < nc;
        uint8_t *own = Vs + lane * MS;
        {
            int pc = 0;
            for (int j = 0; j < M; j++) {
                const int v = valid ? b.IF[i * M + j] : 0;
                own[j] = (uint8_t)v;
                PF[lane * MS + j] = (uint8_t)pc;   // particles in orbitals below j
                pc += v;
            }
        }
        __builtin_amdgcn_wave_barrier();
        const cplx ci = valid ? ldcr(C + i) : cmake(0, 0);
        int64_t pb = mode == 0 ? poff[w] : 0;                       // the wavefront's slots
        int64_t rbase = (mode == 1 && valid) ? stride[i] : 0;       // the row's map blocks
        int p = 0;
        for (int k = 0; k < M; k++) {
            for (int s = TT ? k + 1 : k; s < (TT ? M : k + 1); s++, p++) {
                const int vk = own[k], vs = own[s];
                const bool act = valid && (TT ? (vk >= 1 && vs >= 1) : vk >= 2);
                const uint64_t m = __ballot(act);
                const int nact = __popcll(m);
                const int64_t jpb = pb;
                pb += (int64_t)nact * E;
                const int64_t mybase = rbase;
                if (mode == 1 && act) rbase += MM;
                if (TT && (p < p0 || p >= p1)) continue;
                if (m == 0) continue;
                const int nt = jump_terms(CLASS, M, k);
                for (int c0 = 0; c0 < nt; c0 += 64) {
                    const JumpTerm J = jump_term<CLASS>(M, P, k, s, p, p0, c0 + lane);
                    double zre = 0, zim = 0;
                    uint64_t rem = m;
                    int r0 = 0;   // active rows before the group (slot order)
                    while (rem) {
                        int a[RHO_G], ng = 0;
#pragma unroll
                        for (int g = 0; g < RHO_G; g++) {
                            a[g] = 0;
                            if (rem) {
                                a[g] = __builtin_ctzll(rem);
                                rem &= rem - 1;
                                ng++;
                            }
                        }
                        int64_t j[RHO_G];
                        if (dtab) {
                            // rows 4h .. 4h+3 of the group: lane 16 x + o builds entry o of
                            // the tables of row 4h + x (removal profile k, s); the target
                            // of added q <= l is top - tot - D0[q] - D1[l]
#pragma unroll
                            for (int h = 0; h < RHO_G / 4; h++) {
                                const int x = lane >> 4, o = lane & 15;
                                int ar = a[4 * h];
#pragma unroll
                                for (int y = 1; y < 4; y++) ar = (x == y) ? a[4 * h + y] : ar;
                                uint32_t d0 = 0, d1 = 0, c2 = 0;
                                if (4 * h + x < ng && o >= 1 && o < M) {
                                    const int pp = PF[ar * MS + o] - (k < o) - (s < o);
                                    const int32_t *bt = Bt + o * (N + 1);
                                    const uint32_t c0 = (uint32_t)bt[min(max(pp, 0), N)];
                                    const uint32_t c1 = (uint32_t)bt[min(max(pp + 1, 0), N)];
                                    c2 = (uint32_t)bt[min(max(pp + 2, 0), N)];
                                    d0 = c0 - c1;
                                    d1 = c1 - c2;
                                }
                                d0 = row_scan16(d0);
                                d1 = row_scan16(d1);
                                c2 = row_scan16(c2);
#pragma unroll
                                for (int y = 0; y < 4; y++) {
                                    const int g = 4 * h + y;
                                    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)c2, 16 * y + M - 1);
                                    const uint32_t D0 = (uint32_t)__shfl((int)d0, 16 * y + J.q, 64);
                                    const uint32_t D1 = (uint32_t)__shfl((int)d1, 16 * y + J.l, 64);
                                    j[g] = (g < ng && J.ok) ? (int64_t)(top - tot - D0 - D1) : 0;
                                }
                            }
                        } else {
#pragma unroll
                            for (int g = 0; g < RHO_G; g++) {
                                j[g] = 0;
                                if (g < ng && J.ok) {
                                    if (mode == 0)
                                        j[g] = pack[jpb + (int64_t)(J.e >> 2) * 4 * nact + 4 * (r0 + g) + (J.e & 3)];
                                    else if (mode == 1)
                                        j[g] = tab[readlane_l(mybase, a[g]) + J.tix];
                                    else
                                        j[g] = rank_row(Vs + a[g] * MS, Bt, N, M, nc, k, s, J.q, J.l);
                                }
                            }
                        }
                        cplx cj[RHO_G];
#pragma unroll
                        for (int g = 0; g < RHO_G; g++) cj[g] = (g < ng && J.ok) ? ldcr(C + j[g]) : cmake(0, 0);
#pragma unroll
                        for (int g = 0; g < RHO_G; g++) {
                            if (g >= ng) break;   // uniform
                            const cplx cr = cmake(readlane_d(ci.re, a[g]), readlane_d(ci.im, a[g]));
                            const uint8_t *row = Vs + a[g] * MS;
                            const int rk = row[k], rs = row[s];
                            const double f = TT ? sq[rk] * sq[rs] : sq[rk] * sq[rk - 1];
                            const int vq = row[J.q], vl = row[J.l];
                            const double x = f * (J.diag ? sq[vq + 1] * sq[vq + 2] : sq[vq + 1] * sq[vl + 1]);
                            const cplx z = cfma_conj(cr, cj[g], cmake(0, 0));   // conj(C_i) C_j
                            zre = fma(z.re, x, zre);
                            zim = fma(z.im, x, zim);
                        }
                        r0 += ng;
                    }
                    if (J.ok) {
                        atomicAdd(&acc[J.slot].re, zre);
                        atomicAdd(&acc[J.slot].im, zim);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();   // the rows are rewritten for the next wavefront
    }
    __syncthreads();
    for (int x = t; x < nterms; x += T) partial[(int64_t)blockIdx.x * nterms + x] = acc[x];
}

__global__ void k_sum_blocks(int nblocks, int nterms, const cplx *__restrict__ partial, cplx *__restrict__ out)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nterms) return;
    double re = 0, im = 0;
    for (int bl = 0; bl < nblocks; bl++) {
        cplx z = partial[(int64_t)bl * nterms + x];
        re += z.re;
        im += z.im;
    }
    out[x] = cmake(re, im);
}

// launch geometry of one class
struct ClassPlan {
    int T = 1024;       // threads per block
    int R = 1;          // rows per lane (classes 0 and 3)
    int nblocks = 1;    // blocks (grid-stride over the rows)
    size_t sm = 0;      // dynamic LDS bytes
};

static mcpi_status plan_class(const BasisDev &b, int cls, int nterms, ClassPlan *pl)
{
    int dev = 0, ncu = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t max_sm = 160 * 1024;
    const size_t base_sm = sizeof(cplx) * nterms + sizeof(double) * (b.N + 3) + sizeof(int32_t) * b.M * (b.N + 1) + 16;
    int T = 1024;
    if (cls == 1 || cls == 2) {
        // lane per term: up to 16 wavefronts a block, one LDS row tile
        // (occupations and prefix counts) each; fewer when the accumulators
        // leave too little LDS (large M)
        const size_t ms = (size_t)((b.M + 3) & ~3);
        while (T > 64 && base_sm + 2 * (size_t)T * ms > max_sm) T /= 2;
        const size_t sm = base_sm + 2 * (size_t)T * ms;
        MCPI_ARG(sm <= max_sm, "density-matrix accumulators exceed LDS");
        const int64_t nwv = (b.nc + 63) / 64;
        pl->T = T;
        pl->R = 1;
        pl->sm = sm;
        pl->nblocks = (int)std::min<int64_t>(std::max<int64_t>((nwv + T / 64 - 1) / (T / 64), 1), (int64_t)ncu);
        return MCPI_OK;
    }
    // classes 0 and 3, rows per lane: as many as the occupation tile leaves
    // room for, at most 2 (class 0; 4 spills) or 4 (class 3)
    int R = cls == 0 ? 2 : 4;
    while (R > 1 && base_sm + (size_t)b.M * T * R > max_sm) R /= 2;
    // and few enough that every CU gets a tile
    while (R > 1 && (b.nc + (int64_t)T * R - 1) / ((int64_t)T * R) < ncu) R /= 2;
    const size_t sm = base_sm + (size_t)b.M * T * R;
    MCPI_ARG(sm <= max_sm, "density-matrix accumulators exceed LDS");
    const int64_t tiles = (b.nc + (int64_t)T * R - 1) / ((int64_t)T * R);
    pl->T = T;
    pl->R = R;
    pl->sm = sm;
    pl->nblocks = (int)std::min<int64_t>(std::max<int64_t>(tiles, 1), (int64_t)ncu);
    return MCPI_OK;
}

template <bool ALL, int R>
static void launch_rho(const BasisDev &b, const cplx *C, int nterms, cplx *partial, const ClassPlan &pl,
                       hipStream_t s)
{
    (void)hipFuncSetAttribute((const void *)k_rho<ALL, R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.sm);
    hipLaunchKernelGGL((k_rho<ALL, R>), dim3(pl.nblocks), dim3(pl.T), pl.sm, s, b, C, nterms, partial);
}

template <int CLASS>
static void launch_jumps(const BasisDev &b, const cplx *C, int p0, int p1, int nterms, cplx *partial,
                         const ClassPlan &pl, hipStream_t s)
{
    (void)hipFuncSetAttribute((const void *)k_rho_jumps<CLASS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)pl.sm);
    hipLaunchKernelGGL((k_rho_jumps<CLASS>), dim3(pl.nblocks), dim3(pl.T), pl.sm, s, b, C, p0, p1, nterms, partial);
}

// one class on stream s: block partials into partial (pl.nblocks * nterms),
// their sums in a fixed order into out (nterms)
static mcpi_status launch_class(const BasisDev &b, int cls, const cplx *C, int p0, int p1, int nterms,
                                const ClassPlan &pl, cplx *partial, cplx *out, hipStream_t s)
{
    if (cls == 1) launch_jumps<1>(b, C, p0, p1, nterms, partial, pl, s);
    else if (cls == 2) launch_jumps<2>(b, C, p0, p1, nterms, partial, pl, s);
    else if (cls == 0) {
        if (pl.R == 2) launch_rho<true, 2>(b, C, nterms, partial, pl, s);
        else launch_rho<true, 1>(b, C, nterms, partial, pl, s);
    } else {
        if (pl.R == 4) launch_rho<false, 4>(b, C, nterms, partial, pl, s);
        else if (pl.R == 2) launch_rho<false, 2>(b, C, nterms, partial, pl, s);
        else launch_rho<false, 1>(b, C, nterms, partial, pl, s);
    }
    MCPI_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_blocks, dim3((nterms + 255) / 256), dim3(256), 0, s, pl.nblocks, nterms, partial, out);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// the basis's density scratch: nblocks * nterms block partials, then
// nterms sums (grown on demand; the caller holds rho_mu)
static mcpi_status density_scratch(mcpi_basis bs, int nblocks, int nterms, cplx **partial, cplx **out)
{
    const size_t need = sizeof(cplx) * ((size_t)nblocks * nterms + (size_t)nterms);
    if (bs->rho_ws_bytes < need) {
        if (bs->rho_ws) MCPI_HIP(hipFree(bs->rho_ws));
        bs->rho_ws = nullptr;
        bs->rho_ws_bytes = 0;
        MCPI_HIP(hipMalloc(&bs->rho_ws, need));
        bs->rho_ws_bytes = need;
    }
    *partial = (cplx *)bs->rho_ws;
    *out = *partial + (size_t)nblocks * nterms;
    return MCPI_OK;
}

// The tables the density kernels read: the basis's own, or for a
// sector-ordered basis the reference-order IF with no map (targets ranked
// on the fly), so that densities take vectors in the reference order.
static BasisDev density_view(mcpi_basis bs)
{
    BasisDev b = bs->d;
    if (bs->sectored()) {
        b.IF = bs->IF_ref;
        b.Map = nullptr;
        b.perm = nullptr;
    }
    return b;
}

// run one class and bring its compact terms to the host
static mcpi_status run_class(mcpi_basis bs, int cls, const cplx *C, int p0, int p1, int nterms,
                             std::vector<cplx> &host)
{
    const BasisDev b = density_view(bs);
    ClassPlan pl;
    MCPI_CHECK(plan_class(b, cls, nterms, &pl));
    cplx *partial = nullptr, *out = nullptr;
    MCPI_CHECK(density_scratch(bs, pl.nblocks, nterms, &partial, &out));
    MCPI_CHECK(launch_class(b, cls, C, p0, p1, nterms, pl, partial, out, 0));
    host.resize(nterms);
    MCPI_HIP(hipMemcpy(host.data(), out, sizeof(cplx) * nterms, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

static inline int64_t I4(int a, int b, int c, int d, int M)
{
    return a + (int64_t)M * (b + (int64_t)M * (c + (int64_t)M * d));
}
static inline cplx conjc(cplx z) { return cmake(z.re, -z.im); }

// rho1 / rho2 from the device; either output may be null
static mcpi_status densities(mcpi_basis bs, const cplx *C, cplx *rho1, cplx *rho2)
{
    std::lock_guard<std::mutex> lock(bs->rho_mu);
    const int M = bs->d.M;
    const int P = M * (M - 1) / 2;
    std::vector<cplx> a;
    const int nA = M + P + M + P + P + P + P * M;
    // class 0 in full for the two-body density, else its r1d and r1o terms
    // only (class 3)
    MCPI_CHECK(run_class(bs, rho2 ? 0 : 3, C, 0, 0, nA, a));
    const cplx *r1d = a.data(), *r1o = r1d + M, *R1 = r1o + P, *R2 = R1 + M, *R4 = R2 + P,
               *R5 = R4 + P, *R7 = R5 + P;
    if (rho1) {
        for (int k = 0; k < M; k++) {
            rho1[k + M * k] = r1d[k];
            for (int l = k + 1; l < M; l++) {
                cplx r = r1o[pair_index(k, l, M)];
                rho1[k + M * l] = r;        // <a*_k a_l>
                rho1[l + M * k] = conjc(r);
            }
        }
    }
    if (!rho2) return MCPI_OK;
    std::vector<cplx> ob, tt((size_t)P * P);
    MCPI_CHECK(run_class(bs, 1, C, 0, 0, P + M * P, ob));
    // class 2 in pair ranges that fit in LDS (about 120 KB of accumulators)
    int chunk = std::max(1, (int)(120 * 1024 / (16 * std::max(P, 1))));
    for (int p0 = 0; p0 < P; p0 += chunk) {
        int p1 = std::min(P, p0 + chunk);
        std::vector<cplx> part;
        MCPI_CHECK(run_class(bs, 2, C, p0, p1, (p1 - p0) * P, part));
        memcpy(tt.data() + (size_t)p0 * P, part.data(), sizeof(cplx) * part.size());
    }
    const cplx *R3 = ob.data(), *R6 = R3 + P;
    const int64_t M4 = (int64_t)M * M * M * M;
    for (int64_t x = 0; x < M4; x++) rho2[x] = cmake(0, 0);
    // expansion by the reference's symmetry copies, in its rule order
    // (twobodyMatrix.h:1050, 1073-1076, 1113-1114, 1141-1144, 1171-1174,
    //  1214-1217 (x3), 1335-1343 (x3), 1477-1480)
    for (int k = 0; k < M; k++) rho2[I4(k, k, k, k, M)] = R1[k];
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            cplx r = R2[pair_index(k, s, M)];
            rho2[I4(k, s, k, s, M)] = r; rho2[I4(s, k, k, s, M)] = r;
            rho2[I4(s, k, s, k, M)] = r; rho2[I4(k, s, s, k, M)] = r;
        }
    for (int k = 0; k < M; k++)
        for (int q = k + 1; q < M; q++) {
            cplx r = R3[pair_index(k, q, M)];
            rho2[I4(k, k, q, q, M)] = r; rho2[I4(q, q, k, k, M)] = conjc(r);
        }
    for (int k = 0; k < M; k++)
        for (int l = k + 1; l < M; l++) {
            cplx r = R4[pair_index(k, l, M)];
            rho2[I4(k, k, k, l, M)] = r; rho2[I4(k, k, l, k, M)] = r;
            rho2[I4(l, k, k, k, M)] = conjc(r); rho2[I4(k, l, k, k, M)] = conjc(r);
        }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            cplx r = R5[pair_index(k, s, M)];
            rho2[I4(k, s, s, s, M)] = r; rho2[I4(s, k, s, s, M)] = r;
            rho2[I4(s, s, s, k, M)] = conjc(r); rho2[I4(s, s, k, s, M)] = conjc(r);
        }
    for (int k = 0; k < M; k++)
        for (int q = 0; q < M; q++)
            for (int l = q + 1; l < M; l++) {
                if (q == k || l == k) continue;
                cplx r = R6[k * P + pair_index(q, l, M)];
                rho2[I4(k, k, q, l, M)] = r; rho2[I4(k, k, l, q, M)] = r;
                rho2[I4(l, q, k, k, M)] = conjc(r); rho2[I4(q, l, k, k, M)] = conjc(r);
            }
    for (int k = 0; k < M; k++)
        for (int l = k + 1; l < M; l++)
            for (int s = 0; s < M; s++) {
                if (s == k || s == l) continue;
                cplx r = R7[pair_index(k, l, M) * M + s];
                rho2[I4(k, s, s, l, M)] = r; rho2[I4(s, k, s, l, M)] = r;
                rho2[I4(s, k, l, s, M)] = r; rho2[I4(k, s, l, s, M)] = r;
                rho2[I4(l, s, s, k, M)] = conjc(r); rho2[I4(s, l, s, k, M)] = conjc(r);
                rho2[I4(s, l, k, s, M)] = conjc(r); rho2[I4(l, s, k, s, M)] = conjc(r);
            }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
            for (int q = 0; q < M; q++)
                for (int l = q + 1; l < M; l++) {
                    if (q == k || q == s || l == k || l == s) continue;
                    cplx r = tt[(size_t)pair_index(k, s, M) * P + pair_index(q, l, M)];
                    rho2[I4(k, s, q, l, M)] = r; rho2[I4(s, k, q, l, M)] = r;
                    rho2[I4(k, s, l, q, M)] = r; rho2[I4(s, k, l, q, M)] = r;
                }
    return MCPI_OK;
}

// ------------------------------------------------ allocation-free adjoint --
// For callers that cannot allocate device memory (code under stream capture,
// serving loops): the block partials and the representative sums live in a
// file-scope device array instead of the basis's scratch, and the full
// matrices are expanded on the device (rho2_entry), so the call allocates
// nothing and copies nothing to the host.  Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU race.
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

// rho1 (M*M) and rho2 (M^4) scaled by s2, one thread per entry
__global__ void k_expand_rho(int M, const cplx *__restrict__ A, const cplx *__restrict__ OB,
                             const cplx *__restrict__ TT, cplx *__restrict__ rho1, cplx *__restrict__ rho2,
                             double s2)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    RepLayout L{M, M * (M - 1) / 2};
    const int64_t M2 = (int64_t)M * M;
    if (x < M2) rho1[x] = rho1_entry((int)(x % M), (int)(x / M), L, A);
    if (x >= M2 * M2) return;
    const int a = (int)(x % M), b = (int)(x / M % M), c = (int)(x / M2 % M), d = (int)(x / (M2 * M));
    rho2[x] = cscale(rho2_entry(a, b, c, d, L, A, OB, TT), s2);
}

static mcpi_status vjp_noalloc(mcpi_basis bs, const cplx *C, cplx *g1, cplx *g2, hipStream_t s)
{
    const BasisDev b = density_view(bs);
    const int M = b.M, P = M * (M - 1) / 2;
    const RepLayout L{M, P};
    const size_t cap = MCPI_NOALLOC_SCRATCH_BYTES / sizeof(cplx);
    // representative sums: class 0 | class 1 | class 2 (P x P), then partials
    const size_t nA = (size_t)L.n_class0(), nOB = (size_t)L.n_class1(), nTT = (size_t)P * P;
    const size_t reps = (nA + nOB + nTT + 15) & ~(size_t)15;
    const int chunk = std::max(1, (int)(120 * 1024 / (16 * std::max(P, 1))));
    const size_t max_terms = std::max({nA, nOB, (size_t)std::min(chunk, std::max(P, 1)) * P});
    MCPI_ARG(reps + max_terms <= cap,
             "basis has too many orbitals for the allocation-free adjoint (scratch holds 16 MiB)");
    cplx *scr = nullptr;
    MCPI_CHECK(noalloc_scratch(&scr));
    cplx *A = scr, *OB = A + nA, *TT = OB + nOB, *partial = scr + reps;
    const size_t room = cap - reps;
    auto run = [&](int cls, int p0, int p1, int nterms, cplx *out) -> mcpi_status {
        ClassPlan pl;
        MCPI_CHECK(plan_class(b, cls, nterms, &pl));
        pl.nblocks = (int)std::max<size_t>(1, std::min<size_t>((size_t)pl.nblocks, room / (size_t)nterms));
        return launch_class(b, cls, C, p0, p1, nterms, pl, partial, out, s);
    };
    MCPI_CHECK(run(0, 0, 0, (int)nA, A));
    MCPI_CHECK(run(1, 0, 0, (int)nOB, OB));
    for (int p0 = 0; p0 < P; p0 += chunk) {
        const int p1 = std::min(P, p0 + chunk);
        MCPI_CHECK(run(2, p0, p1, (p1 - p0) * P, TT + (size_t)p0 * P));
    }
    const int64_t M4 = (int64_t)M * M * M * M;
    hipLaunchKernelGGL(k_expand_rho, dim3((unsigned)((M4 + 255) / 256)), dim3(256), 0, s, M, (const cplx *)A,
                       (const cplx *)OB, (const cplx *)TT, g1, g2, 0.5);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

}  // namespace mcpi

using namespace mcpi;

extern "C" {

mcpi_status mcpi_hip_rho1(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *rho1)
{
    MCPI_ARG(b && C_dev && rho1, "null argument");
    return densities(b, (const cplx *)C_dev, (cplx *)rho1, nullptr);
}

mcpi_status mcpi_hip_rho2(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *rho2)
{
    MCPI_ARG(b && C_dev && rho2, "null argument");
    return densities(b, (const cplx *)C_dev, nullptr, (cplx *)rho2);
}

mcpi_status mcpi_hip_vjp(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *dE_dHo, mcpi_cplx *dE_dHint)
{
    MCPI_ARG(b && C_dev && dE_dHo && dE_dHint, "null argument");
    MCPI_CHECK(densities(b, (const cplx *)C_dev, (cplx *)dE_dHo, (cplx *)dE_dHint));
    // Hellmann-Feynman: E = sum Ho . rho1 + 1/2 sum Hint . rho2
    const int64_t M4 = (int64_t)b->d.M * b->d.M * b->d.M * b->d.M;
    cplx *g = (cplx *)dE_dHint;
    for (int64_t x = 0; x < M4; x++) g[x] = cscale(g[x], 0.5);
    return MCPI_OK;
}

mcpi_status mcpi_hip_vjp_noalloc(mcpi_basis b, const mcpi_cplx *C_dev, mcpi_cplx *dE_dHo_dev,
                                 mcpi_cplx *dE_dHint_dev, void *stream)
{
    MCPI_ARG(b && C_dev && dE_dHo_dev && dE_dHint_dev, "null argument");
    return vjp_noalloc(b, (const cplx *)C_dev, (cplx *)dE_dHo_dev, (cplx *)dE_dHint_dev, (hipStream_t)stream);
}

}  // extern "C"
