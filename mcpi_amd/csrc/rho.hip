// One- and two-body density matrices and the Hellmann-Feynman adjoint.
//
// Reference: OBrho_XM, onebodyMatrix.h:190-251 (rho1), and TBrho_XX,
// twobodyMatrix.h:991-1487 (rho2, rules 1-8 plus symmetry copies).
//
// The reference loops orbital-outer / row-inner.  Here a block walks row
// tiles (lanes = rows) and, for every representative term of a class, sums
// the lanes' contributions with a wavefront reduction and adds the result to
// an LDS accumulator (atomically: the wavefronts of a block add in the order
// they arrive, so the last bits of a sum can differ from run to run).  Each
// block writes its accumulators once; a second kernel sums the blocks in a
// fixed order.  The host expands the representative terms into full
// matrices with the reference's symmetry copies.
//
// Targets come from the stored maps when the basis has them and are ranked
// on the fly otherwise (closed-form FockToIndex over the lane's prefix
// counts, configurationsMap.h:273-305), so the adjoint also runs on bases
// that store only Map, or no map at all.
//
// Term classes (compact layouts, P = M(M-1)/2 pairs, pq = pair index):
//   class 0 (single jumps and diagonal):
//     r1d[k] | r1o[kl] | R1[k] | R2[ks] | R4[kl] | R5[ks] | R7[kl][s]
//   class 1 (same-orbital double jumps):  R3[kq] | R6[k][ql]
//   class 2 (two-orbital double jumps):   R8[ks - p0][ql]   (pairs p0..p1-1)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "mcpi_internal.h"

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

// rank of this lane's row with particles removed from a, b2 (b2 = -1: none)
// and added to c, d (d = -1: none); occupations from the LDS tile
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

// U terms (q[u], l[u]) of one double jump (k, s), rows r < R of the lane:
// every target first (map entries or ranks), then every gather, then the
// sums and U wavefront reductions, so that the memory accesses of the batch
// are in flight together instead of one dependent chain per term.
// Targets: packed (pk[r] != nullptr: the row's slot in the wave-packed
// tables, entry packed_pos(k, s, q, l) at group stride n4[r]), else tab
// (MapOT or MapTT at the row's chunk offset base[r]), else ranked on the
// fly.  f[r]: the jump's factor (0 for rows without the jump); diag: q == l
// terms (sqrt((v+1)(v+2)) instead of sqrt(v_q+1) sqrt(v_l+1)).
template <int U, int R>
__device__ inline void rho_batch(const int32_t *const (&pk)[R], const uint32_t (&n4)[R], const int32_t *tab,
                                 const int64_t (&base)[R], const bool (&act)[R],
                                 const double (&f)[R], const cplx (&ci)[R], const cplx *__restrict__ C,
                                 const uint8_t *Vs, int t, int T, int TR, const int32_t *Bt, int N, int M, int64_t nc,
                                 const double *sq, int k, int s, const int (&q)[U], const int (&l)[U], int nu,
                                 bool diag, cplx *acc, const int (&slot)[U])
{
    int64_t j[U][R];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int qq = u < nu ? q[u] : q[0], ll = u < nu ? l[u] : l[0];
            j[u][r] = 0;
            if (act[r]) {
                if (pk[r]) {
                    const int e = packed_pos(k, s, qq < ll ? qq : ll, qq < ll ? ll : qq);
                    j[u][r] = pk[r][(e >> 2) * n4[r] + (e & 3)];
                } else if (tab) {
                    j[u][r] = tab[base[r] + ll + qq * M];
                } else {
                    j[u][r] = rank_lane(Vs, t + T * r, TR, Bt, N, M, nc, k, s, qq, ll);
                }
            }
        }
    cplx g[U][R];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int r = 0; r < R; r++) g[u][r] = act[r] ? ldcr(C + j[u][r]) : cmake(0, 0);
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (u >= nu) break;   // uniform
        cplx z = cmake(0, 0);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int vq = Vs[q[u] * TR + t + T * r], vl = Vs[l[u] * TR + t + T * r];
            const double x = diag ? sq[vq + 1] * sq[vq + 2] : sq[vq + 1] * sq[vl + 1];
            z = cfma(cfma_conj(ci[r], g[u][r], cmake(0, 0)), cmake(f[r] * x, 0), z);
        }
        wave_add(z, acc, slot[u], true);
    }
}

// A block of T = 1024 threads (16 wavefronts: the accumulators, up to
// 120 KB of LDS, allow one block per CU, so the block has to fill the CU)
// walks tiles of T*R rows; lane t holds rows tile + t + T*r, r < R, and sums
// its R rows' contributions to a term before the wavefront reduction, which
// divides the number of reductions and LDS atomics by R.
#define RHO_U 8   // terms per batch in classes 1 and 2

template <int CLASS, int R>
__global__ __launch_bounds__(1024) void k_rho(BasisDev b, const cplx *__restrict__ C, int p0, int p1,
                                             int nterms, cplx *__restrict__ partial)
{
    extern __shared__ double smem_r[];
    const int M = b.M, N = b.N, T = blockDim.x, t = threadIdx.x, TR = T * R;
    const int P = M * (M - 1) / 2;
    const int64_t nc = b.nc, MM = (int64_t)M * M, nwv = (nc + 63) >> 6;
    const int lw = t & 63;
    const uint64_t below = lw ? (~0ull >> (64 - lw)) : 0ull;
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
        if (CLASS == 0) {
            int o = 0;
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
                wave_add(cmake(a, 0), acc, o + k, true);
                wave_add(cmake(a2, 0), acc, o + P + M + k, wave_any(any2));
            }
            o += M;  // r1o at o
            const int o_r1o = o, o_R2 = M + P + M, o_R4 = o_R2 + P, o_R5 = o_R4 + P, o_R7 = o_R5 + P;
            int pq = 0;
            for (int k = 0; k < M; k++)
                for (int l = k + 1; l < M; l++, pq++) {
                    // R2: |C|^2 v_k v_l over all rows
                    double a = 0;
                    bool any = false, anyact = false;
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int vk = V(k, r), vl = V(l, r);
                        a += m2[r] * vk * vl;
                        any = any || (vk && vl);
                        anyact = anyact || (valid[r] && vk >= 1);
                    }
                    wave_add(cmake(a, 0), acc, o_R2 + pq, wave_any(any));
                    if (!wave_any(anyact)) continue;
                    cplx g[R];
                    double root[R];
                    cplx z1 = cmake(0, 0), z4 = cmake(0, 0), z5 = cmake(0, 0);
                    bool any4 = false, any5 = false;
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
                        any4 = any4 || (act && vk >= 2);
                        any5 = any5 || (act && vl >= 1);
                    }
                    wave_add(z1, acc, o_r1o + pq, true);
                    wave_add(z4, acc, o_R4 + pq, wave_any(any4));
                    wave_add(z5, acc, o_R5 + pq, wave_any(any5));
                    for (int s = 0; s < M; s++) {
                        if (s == k || s == l) continue;
                        cplx z7 = cmake(0, 0);
                        bool any7 = false;
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const int vs = V(s, r);
                            const bool act = valid[r] && V(k, r) >= 1;
                            z7 = cfma(g[r], cmake(act && vs >= 1 ? vs * root[r] : 0.0, 0), z7);
                            any7 = any7 || (act && vs >= 1);
                        }
                        wave_add(z7, acc, o_R7 + pq * M + s, wave_any(any7));
                    }
                }
        } else if (CLASS == 1) {
            // R3 (k < q) and R6 (q < l, k not in {q, l}), v_k >= 2
            // packed tables (full maps): the wavefronts of a tile are the
            // tables' wavefronts; slot offsets advance over every jump
            const bool packed = b.packOT != nullptr;
            const int EO = packed_entries(M, false);
            int cot[R];
            int64_t pb[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                cot[r] = 0;
                const int64_t w = i[r] >> 6;
                pb[r] = (packed && w < nwv) ? b.poffOT[w] : 0;
            }
            for (int k = 0; k < M; k++) {
                bool act[R], anyact = false;
                int64_t base[R];
                double fk[R];
                const int32_t *pks[R];
                uint32_t n4[R];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int vk = V(k, r);
                    act[r] = valid[r] && vk >= 2;
                    base[r] = (!packed && act[r] && b.MapOT) ? b.strideOT[i[r]] + (int64_t)cot[r] * MM : 0;
                    if (act[r]) cot[r]++;
                    fk[r] = act[r] ? sq[vk] * sq[vk - 1] : 0.0;
                    anyact = anyact || act[r];
                    const uint64_t m = __ballot(act[r]);
                    pks[r] = packed ? b.packOT + pb[r] + 4 * __popcll(m & below) : nullptr;
                    n4[r] = 4u * (uint32_t)__popcll(m);
                    pb[r] += (int64_t)__popcll(m) * EO;
                }
                if (!wave_any(anyact)) continue;
                constexpr int U = RHO_U;
                int qs[U], ls[U], sl[U], nu = 0;
                // R3: q == l (k < q); the map entry q + q M is read as l + q M with l = q
                for (int q = k + 1; q < M; q++) {
                    qs[nu] = q; ls[nu] = q; sl[nu] = pair_index(k, q, M);
                    if (++nu == U) {
                        rho_batch<U, R>(pks, n4, b.MapOT, base, act, fk, ci, C, Vs, t, T, TR, Bt, N, M, nc, sq, k, k,
                                        qs, ls, nu, true, acc, sl);
                        nu = 0;
                    }
                }
                if (nu) rho_batch<U, R>(pks, n4, b.MapOT, base, act, fk, ci, C, Vs, t, T, TR, Bt, N, M, nc, sq, k, k,
                                        qs, ls, nu, true, acc, sl);
                nu = 0;
                // R6: q < l, k not in {q, l}; map entry l + q M
                int pq = 0;
                for (int q = 0; q < M; q++)
                    for (int l = q + 1; l < M; l++, pq++) {
                        if (q == k || l == k) continue;
                        qs[nu] = q; ls[nu] = l; sl[nu] = P + k * P + pq;
                        if (++nu == U) {
                            rho_batch<U, R>(pks, n4, b.MapOT, base, act, fk, ci, C, Vs, t, T, TR, Bt, N, M, nc, sq, k,
                                            k, qs, ls, nu, false, acc, sl);
                            nu = 0;
                        }
                    }
                if (nu) rho_batch<U, R>(pks, n4, b.MapOT, base, act, fk, ci, C, Vs, t, T, TR, Bt, N, M, nc, sq, k, k,
                                        qs, ls, nu, false, acc, sl);
            }
        } else {
            // R8: k < s occupied, q < l, {q, l} disjoint from {k, s}
            const bool packed = b.packTT != nullptr;
            const int ET = packed_entries(M, true);
            int ctt[R];
            int64_t pb[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                ctt[r] = 0;
                const int64_t w = i[r] >> 6;
                pb[r] = (packed && w < nwv) ? b.poffTT[w] : 0;
            }
            int p = 0;
            for (int k = 0; k < M; k++)
                for (int s = k + 1; s < M; s++, p++) {
                    bool act[R], anyact = false;
                    int64_t base[R];
                    double fks[R];
                    const int32_t *pks[R];
                    uint32_t n4[R];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int vk = V(k, r), vs = V(s, r);
                        act[r] = valid[r] && vk >= 1 && vs >= 1;
                        base[r] = (!packed && act[r] && b.MapTT) ? b.strideTT[i[r]] + (int64_t)ctt[r] * MM : 0;
                        if (act[r]) ctt[r]++;
                        fks[r] = act[r] ? sq[vk] * sq[vs] : 0.0;
                        anyact = anyact || act[r];
                        const uint64_t m = __ballot(act[r]);
                        pks[r] = packed ? b.packTT + pb[r] + 4 * __popcll(m & below) : nullptr;
                        n4[r] = 4u * (uint32_t)__popcll(m);
                        pb[r] += (int64_t)__popcll(m) * ET;
                    }
                    if (p < p0 || p >= p1) continue;
                    if (!wave_any(anyact)) continue;
                    constexpr int U = RHO_U;
                    int qs[U], ls[U], sl[U], nu = 0;
                    int pq = 0;
                    // MapTT entry q + l M (the transposed entry l + q M holds the same target)
                    for (int q = 0; q < M; q++)
                        for (int l = q + 1; l < M; l++, pq++) {
                            if (q == k || q == s || l == k || l == s) continue;
                            qs[nu] = l; ls[nu] = q; sl[nu] = (p - p0) * P + pq;
                            if (++nu == U) {
                                rho_batch<U, R>(pks, n4, b.MapTT, base, act, fks, ci, C, Vs, t, T, TR, Bt, N, M, nc, sq,
                                                k, s, qs, ls, nu, false, acc, sl);
                                nu = 0;
                            }
                        }
                    if (nu) rho_batch<U, R>(pks, n4, b.MapTT, base, act, fks, ci, C, Vs, t, T, TR, Bt, N, M, nc, sq, k, s,
                                            qs, ls, nu, false, acc, sl);
                }
        }
#undef V
        __syncthreads();
    }
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

// run one class and bring its compact terms to the host
template <int CLASS, int R>
static void launch_rho(const BasisDev &b, const cplx *C, int p0, int p1, int nterms, cplx *partial, int nblocks,
                       int T, size_t sm)
{
    (void)hipFuncSetAttribute((const void *)k_rho<CLASS, R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    hipLaunchKernelGGL((k_rho<CLASS, R>), dim3(nblocks), dim3(T), sm, 0, b, C, p0, p1, nterms, partial);
}

static mcpi_status run_class(mcpi_basis bs, int cls, const cplx *C, int p0, int p1, int nterms,
                             std::vector<cplx> &host)
{
    const BasisDev &b = bs->d;
    const int T = 1024;
    int dev = 0, ncu = 0;
    MCPI_HIP(hipGetDevice(&dev));
    MCPI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t max_sm = 160 * 1024;
    const size_t base_sm = sizeof(cplx) * nterms + sizeof(double) * (b.N + 3) + sizeof(int32_t) * b.M * (b.N + 1) + 16;
    // rows per lane: as many as the occupation tile leaves room for, at most
    // 4 for class 0 and 2 for the batched double-jump classes (4 spills)
    int R = cls == 0 ? 4 : 2;
    while (R > 1 && base_sm + (size_t)b.M * T * R > max_sm) R /= 2;
    // and few enough that every CU gets a tile
    while (R > 1 && (b.nc + (int64_t)T * R - 1) / ((int64_t)T * R) < ncu) R /= 2;
    const size_t sm = base_sm + (size_t)b.M * T * R;
    MCPI_ARG(sm <= max_sm, "density-matrix accumulators exceed LDS");
    int64_t tiles = (b.nc + (int64_t)T * R - 1) / ((int64_t)T * R);
    int nblocks = (int)std::min<int64_t>(std::max<int64_t>(tiles, 1), (int64_t)ncu);
    cplx *partial = nullptr, *out = nullptr;
    MCPI_HIP(hipMalloc(&partial, sizeof(cplx) * (size_t)nblocks * nterms));
    MCPI_HIP(hipMalloc(&out, sizeof(cplx) * nterms));
    if (cls == 0) {
        if (R == 4) launch_rho<0, 4>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
        else if (R == 2) launch_rho<0, 2>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
        else launch_rho<0, 1>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
    } else if (cls == 1) {
        if (R == 2) launch_rho<1, 2>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
        else launch_rho<1, 1>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
    } else {
        if (R == 2) launch_rho<2, 2>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
        else launch_rho<2, 1>(b, C, p0, p1, nterms, partial, nblocks, T, sm);
    }
    MCPI_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_blocks, dim3((nterms + 255) / 256), dim3(256), 0, 0, nblocks, nterms, partial, out);
    MCPI_HIP(hipGetLastError());
    host.resize(nterms);
    MCPI_HIP(hipMemcpy(host.data(), out, sizeof(cplx) * nterms, hipMemcpyDeviceToHost));
    MCPI_HIP(hipFree(partial));
    MCPI_HIP(hipFree(out));
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
    const int M = bs->d.M;
    const int P = M * (M - 1) / 2;
    std::vector<cplx> a;
    const int nA = M + P + M + P + P + P + P * M;
    MCPI_CHECK(run_class(bs, 0, C, 0, 0, nA, a));
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

}  // extern "C"
