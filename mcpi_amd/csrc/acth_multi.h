// Multi-vector H.v: several Hamiltonians on one basis in one pass.
//
// Same maths as k_acth_fast (acth_fast.h; reference applyHconf_omp,
// hamiltonianMatrix.h:1431-1970), for a group of instances that share the
// basis and the zero pattern of their folded coefficients (Lieb-Liniger
// Hamiltonians of different couplings g, for instance).  A thread owns one
// row for NV instances: the work that depends only on the row and the jump -
// walking the non-zero lists, the target index (packed map or on-the-fly
// rank), the occupation factors - is done once, and each instance only adds
// its coefficient times its own gathered C[j].
//
// Layout as for the batched kernel: instance v's coefficient block is
// instance 0's shifted by v * cstride bytes, its vectors sit v * vstride
// elements after C0 / out0.  The 1-D grid holds ceil(nbatch / NV) blocks per
// row block, group index fastest; chunks of 32 row blocks go to the XCDs
// round-robin (chunk_xcd), so a row block's groups share an L2 (batched
// Lanczos, 64 instances: 1.31 s per call, 1.35 s with each XCD walking one
// contiguous eighth of the rows).
#pragma once

#include "acth_fast.h"

namespace mcpi {

template <class T>
__device__ inline const T *shifted(const T *p, size_t bytes)
{
    return (const T *)((const char *)p + bytes);
}

// the non-zero entries of one chunk of jump (k, s) for NV instances (lists
// shared; each instance's coefficients at its own block offset)
template <int M, int NV, bool DMAP, bool RC>
__device__ inline void chunk_list_multi(const int32_t *__restrict__ slot, int nact4, const uint32_t *Dl, int lane,
                                        int T, uint32_t base, const RowState<M> &r, const double *sq,
                                        const int32_t *ents, const cplx *coefs0, size_t cstride,
                                        const cplx *__restrict__ C0, int64_t vstride, int nv, int n, int k, int s,
                                        cplx (&part)[NV])
{
    for (int t = 0; t < n; t += 4) {
        int64_t j[4];
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = t + u < n;  // uniform
            const int e = ok ? ents[t + u] : 0;
            const int q = e % M, l = e / M;
            if (DMAP) {
                const int tt = ok ? packed_pos(k, s, q, l) : 0;   // listed: q, l not in {k, s}
                j[u] = ld_nt(slot + (int64_t)(tt >> 2) * nact4 + (tt & 3));
            } else {
                j[u] = (int64_t)(uint32_t)(base - dl<M>(Dl, lane, T, 0, q) - dl<M>(Dl, lane, T, 1, l));
            }
            const int vq = occ<M>(r, q), vl = occ<M>(r, l);
            x[u] = ok ? ((q == l) ? sq[vq + 1] * sq[vq + 2] : sq[vq + 1] * sq[vl + 1]) : 0.0;
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if (v >= nv) break;   // uniform
            const cplx *cf = shifted(coefs0, (size_t)v * cstride);
            const cplx *Cv = C0 + (int64_t)v * vstride;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const cplx c = (t + u < n) ? cf[t + u] : cmake(0, 0);
                part[v] = term<RC>(c, x[u], ld_c(Cv + j[u]), part[v]);
            }
        }
    }
}

// the listed entries of one jump from the pattern tables (entry t of this
// lane at slot + t * nact), for NV instances; past n the lane's own row
// stands in (coefficient zero)
template <int M, int NV, bool RC>
__device__ inline void chunk_pat_multi(const int32_t *__restrict__ slot, int nact, const RowState<M> &r,
                                       const double *sq, const int32_t *ents, const cplx *coefs0, size_t cstride,
                                       const cplx *__restrict__ C0, int64_t vstride, int nv, int n, int32_t self,
                                       cplx (&part)[NV])
{
    for (int t = 0; t < n; t += 4) {
        Piece4 g;
#pragma unroll
        for (int u = 0; u < 4; u++) g.x[u] = t + u < n ? ld_nt(slot + (int64_t)(t + u) * nact) : self;
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = t + u < n;  // uniform
            const int e = ok ? ents[t + u] : 0;
            const int q = e % M, l = e / M;
            const int vq = occ<M>(r, q), vl = occ<M>(r, l);
            x[u] = ok ? ((q == l) ? sq[vq + 1] * sq[vq + 2] : sq[vq + 1] * sq[vl + 1]) : 0.0;
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if (v >= nv) break;   // uniform
            const cplx *cf = shifted(coefs0, (size_t)v * cstride);
            const cplx *Cv = C0 + (int64_t)v * vstride;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const cplx c = (t + u < n) ? cf[t + u] : cmake(0, 0);
                part[v] = term<RC>(c, x[u], ld_c(Cv + (int64_t)g.x[u]), part[v]);
            }
        }
    }
}

// PAT: double-jump targets of the listed terms from the pattern tables
template <int M, int NV, bool DMAP, bool RC, bool PAT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_acth_multi(
    BasisDev b, HamCoefDev c0, const cplx *__restrict__ C0, cplx *__restrict__ out0, int nbatch, size_t cstride,
    int64_t vstride)
{
    const unsigned G = (unsigned)((nbatch + NV - 1) / NV);
    const unsigned vb = chunk_xcd(blockIdx.x, gridDim.x, 32 * G);
    const unsigned grp = vb % G, rb = vb / G;
    const int vfirst = (int)grp * NV;
    const int nv = min(NV, nbatch - vfirst);
    const size_t cbase = (size_t)vfirst * cstride;   // this group's first coefficient block
    const cplx *__restrict__ C = C0 + (int64_t)vfirst * vstride;
    cplx *__restrict__ out = out0 + (int64_t)vfirst * vstride;
    constexpr int MM = M * M;
    extern __shared__ double fsm[];
    double *sq = fsm;                                   // N+3
    uint32_t *Bt = (uint32_t *)(sq + b.N + 3);          // M*(N+1)
    uint32_t *Dl = Bt + M * (b.N + 1);                  // 2*M*blockDim (ranked doubles)
    const int lane = threadIdx.x, T = blockDim.x;
    for (int x = threadIdx.x; x < b.N + 3; x += blockDim.x) sq[x] = sqrt((double)x);
    if (!DMAP && !PAT)
        for (int x = threadIdx.x; x < M * (b.N + 1); x += blockDim.x) Bt[x] = (uint32_t)b.B[x];
    __syncthreads();
    const int64_t i = (int64_t)rb * blockDim.x + threadIdx.x;
    if (i >= b.nc) return;
    const int N = b.N;
    const int64_t nc = b.nc;
    const uint32_t top = (uint32_t)(nc - 1);
    RowState<M> r;
    load_row<M>(b, i, sq, r);

    cplx acc[NV];
    // diagonal: 1-body diagonal, rules 1 and 2
#pragma unroll
    for (int v = 0; v < NV; v++) {
        acc[v] = cmake(0, 0);
        if (v >= nv) continue;
        const size_t off = cbase + (size_t)v * cstride;
        const cplx *dk = shifted(c0.dk, off), *d1 = shifted(c0.d1, off), *d2 = shifted(c0.d2, off);
        cplx dsum = cmake(0, 0);
#pragma unroll
        for (int k = 0; k < M; k++) {
            const int ik = occ<M>(r, k);
            const double vk = (double)ik;
            dsum = cfma(dk[k], cmake(vk, 0), dsum);
            dsum = cfma(d1[k], cmake(vk * (ik - 1), 0), dsum);
#pragma unroll
            for (int s = k + 1; s < M; s++) dsum = cfma(d2[k + M * s], cmake(vk * occ<M>(r, s), 0), dsum);
        }
        acc[v] = cfma(dsum, ld_c(C + (int64_t)v * vstride + i), cmake(0, 0));
    }

    // single jumps k -> l from Map (pattern c0.m1j shared by the group)
    for (int k = 0; k < M; k++) {
        uint32_t m = c0.m1j[k];
        if (m == 0) continue;
        const int vk = occ<M>(r, k);
        const bool act = vk >= 1;
        const double fk = act ? sq[vk] : 0.0;
        while (m) {
            const int l = __ffs(m) - 1;
            m &= m - 1;
            const int kl = k + M * l;
            const int64_t j = act ? (int64_t)ld_nt(b.Map + i + (int64_t)kl * nc) : i;
            const double x = fk * sq[occ<M>(r, l) + 1];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                if (v >= nv) break;
                const size_t off = cbase + (size_t)v * cstride;
                cplx h = shifted(c0.A0, off)[kl];
                const cplx *Ak = shifted(c0.A, off) + (int64_t)kl * M;
#pragma unroll
                for (int s = 0; s < M; s++) {
                    h.re = fma((double)occ<M>(r, s), Ak[s].re, h.re);
                    h.im = fma((double)occ<M>(r, s), Ak[s].im, h.im);
                }
                acc[v] = term<RC>(h, x, ld_c(C + (int64_t)v * vstride + j), acc[v]);
            }
        }
    }

    constexpr int EO = packed_entries(M, false), ET = packed_entries(M, true);
    const int lw = threadIdx.x & 63;
    const uint64_t below = lw ? (~0ull >> (64 - lw)) : 0ull;
    const int64_t wv = i >> 6;
    const cplx *nzc = shifted(c0.nz_c, cbase);
    if (PAT) {
        // pattern tables: slots only for jumps with a non-empty list
        int64_t pb = b.ppoffOT[wv];
        for (int k = 0; k < M; k++) {
            const int beg = c0.ot_off[k], n_nz = c0.ot_off[k + 1] - beg;
            if (n_nz == 0) continue;
            const int vk = occ<M>(r, k);
            const bool act = vk >= 2;
            const uint64_t m = __ballot(act);
            const int nact = __popcll(m);
            const int32_t *slot = b.patOT + pb + __popcll(m & below);
            pb += (int64_t)nact * n_nz;
            if (!act) continue;
            cplx part[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) part[v] = cmake(0, 0);
            chunk_pat_multi<M, NV, RC>(slot, nact, r, sq, c0.nz_e + beg, nzc + beg, cstride, C, vstride, nv,
                                       n_nz, (int32_t)i, part);
            const double f = sq[vk] * sq[vk - 1];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                acc[v].re = fma(f, part[v].re, acc[v].re);
                acc[v].im = fma(f, part[v].im, acc[v].im);
            }
        }
        pb = b.ppoffTT[wv];
        int p = 0;
        for (int k = 0; k < M; k++) {
            const int vk = occ<M>(r, k);
            for (int s = k + 1; s < M; s++, p++) {
                const int beg = c0.tt_off[p], n_nz = c0.tt_off[p + 1] - beg;
                if (n_nz == 0) continue;
                const int vs = occ<M>(r, s);
                const bool act = vk >= 1 && vs >= 1;
                const uint64_t m = __ballot(act);
                const int nact = __popcll(m);
                const int32_t *slot = b.patTT + pb + __popcll(m & below);
                pb += (int64_t)nact * n_nz;
                if (!act) continue;
                cplx part[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) part[v] = cmake(0, 0);
                chunk_pat_multi<M, NV, RC>(slot, nact, r, sq, c0.nz_e + beg, nzc + beg, cstride, C, vstride,
                                           nv, n_nz, (int32_t)i, part);
                const double f = sq[vk] * sq[vs];
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    acc[v].re = fma(f, part[v].re, acc[v].re);
                    acc[v].im = fma(f, part[v].im, acc[v].im);
                }
            }
        }
    }
    // same-orbital double jumps: rules 3, 6
    if (!PAT) {
        int64_t pb = DMAP ? b.poffOT[wv] : 0;
        for (int k = 0; k < M; k++) {
            const int n_nz = c0.ot_off[k + 1] - c0.ot_off[k];
            const int vk = occ<M>(r, k);
            const bool act = vk >= 2;
            const uint64_t m = DMAP ? __ballot(act) : 0ull;
            const int nact = __popcll(m);
            const int32_t *slot = DMAP ? b.packOT + pb + 4 * __popcll(m & below) : nullptr;
            pb += (int64_t)nact * EO;
            if (n_nz == 0 || !act) continue;
            uint32_t D0[M], D1[M], tot = 0;
            if (!DMAP) {
                build_d<M, true>(r, Bt, N, k, k, D0, D1, tot);
                store_d<M>(D0, D1, Dl, lane, T);
            }
            const double f = sq[vk] * sq[vk - 1];
            cplx part[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) part[v] = cmake(0, 0);
            chunk_list_multi<M, NV, DMAP, RC>(slot, 4 * nact, Dl, lane, T, top - tot, r, sq, c0.nz_e + c0.ot_off[k],
                                              nzc + c0.ot_off[k], cstride, C, vstride, nv, n_nz, k, k, part);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                acc[v].re = fma(f, part[v].re, acc[v].re);
                acc[v].im = fma(f, part[v].im, acc[v].im);
            }
        }
    }
    // two-orbital double jumps: rules 7, 9
    if (!PAT) {
        int64_t pb = DMAP ? b.poffTT[wv] : 0;
        int p = 0;
        for (int k = 0; k < M; k++) {
            const int vk = occ<M>(r, k);
            for (int s = k + 1; s < M; s++, p++) {
                const int n_nz = c0.tt_off[p + 1] - c0.tt_off[p];
                const int vs = occ<M>(r, s);
                const bool act = vk >= 1 && vs >= 1;
                const uint64_t m = DMAP ? __ballot(act) : 0ull;
                const int nact = __popcll(m);
                const int32_t *slot = DMAP ? b.packTT + pb + 4 * __popcll(m & below) : nullptr;
                pb += (int64_t)nact * ET;
                if (n_nz == 0 || !act) continue;
                uint32_t D0[M], D1[M], tot = 0;
                if (!DMAP) {
                    build_d<M, true>(r, Bt, N, k, s, D0, D1, tot);
                    store_d<M>(D0, D1, Dl, lane, T);
                }
                const double f = sq[vk] * sq[vs];
                cplx part[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) part[v] = cmake(0, 0);
                chunk_list_multi<M, NV, DMAP, RC>(slot, 4 * nact, Dl, lane, T, top - tot, r, sq,
                                                  c0.nz_e + c0.tt_off[p], nzc + c0.tt_off[p], cstride, C, vstride,
                                                  nv, n_nz, k, s, part);
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    acc[v].re = fma(f, part[v].re, acc[v].re);
                    acc[v].im = fma(f, part[v].im, acc[v].im);
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        if (v >= nv) break;
        double2 o;
        o.x = acc[v].re;
        o.y = acc[v].im;
        *reinterpret_cast<double2 *>(out + (int64_t)v * vstride + i) = o;
    }
    (void)MM;
}

// instances per thread of the multi-vector kernel
#define MCPI_MULTI_NV 4

template <int M>
inline mcpi_status launch_multi_m(const BasisDev &b, const HamCoefDev &c0, int rank_doubles, int nbatch,
                                  size_t cstride, const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st)
{
    constexpr int NV = MCPI_MULTI_NV;
    const unsigned G = (unsigned)((nbatch + NV - 1) / NV);
    dim3 grid((unsigned)((b.nc + 255) / 256) * G), blk(256);
    size_t sm = sizeof(double) * (b.N + 3) + sizeof(uint32_t) * M * (b.N + 1) + 16;
    const bool pat = b.patTT && b.patOT;
    const bool dmap = !pat && !rank_doubles && b.packTT && b.packOT;
    if (!dmap && !pat) sm += sizeof(uint32_t) * 2 * M * 256;
#define MCPI_LM(D, R, P)                                                                                       \
    hipLaunchKernelGGL((k_acth_multi<M, NV, D, R, P>), grid, blk, sm, st, b, c0, C0, out0, nbatch, cstride, vstride)
    if (pat) {
        if (c0.real) MCPI_LM(false, true, true);
        else MCPI_LM(false, false, true);
    } else if (dmap) {
        if (c0.real) MCPI_LM(true, true, false);
        else MCPI_LM(true, false, false);
    } else {
        if (c0.real) MCPI_LM(false, true, false);
        else MCPI_LM(false, false, false);
    }
#undef MCPI_LM
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

}  // namespace mcpi
