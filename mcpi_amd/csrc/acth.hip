// H.v (actH) on the GPU.
//
// Reference: applyHconf_omp, hamiltonianMatrix.h:1431-1970 (rules 1-9 and
// the one-body term, out = w + 0.5 z; identical maths in applyHconf_XX
// 891-1425, applyHconf_X 431-887 and the hash-only applyHconf 31-425).
//
// The rules are folded once per Hamiltonian into coefficient tables indexed
// by the target a row gathers from (build_coefficients below):
//   * target i itself        (1-body diagonal, rules 1 and 2)
//   * single jump k -> l     (1-body off-diagonal, rules 4, 5 and 8 all
//                             gather C[Map[i,k,l]]: one gather instead of up
//                             to 2 + #occupied)
//   * same-orbital double    (rules 3 and 6, OT chunk entry q + l*M, q <= l)
//   * two-orbital double     (rules 7 and 9, TT chunk entry q + l*M, q <= l)
// so every gathered C[j] is multiplied by one complex coefficient times a
// product of square roots of occupations.  Coefficients are uniform across
// a wavefront (they depend only on the orbitals), occupations per lane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mcpi_internal.h"
#include "pair_factor.h"

namespace mcpi {

// --------------------------------------------------------- coefficients --

mcpi_status build_coefficients(int M, const cplx *Ho, const cplx *Hint, HamCoefDev *c, void **block,
                               size_t *block_bytes)
{
    const int64_t M2 = (int64_t)M * M, M3 = M2 * M;
    const int P = M * (M - 1) / 2;
    auto H4 = [&](int k, int s, int q, int l) -> cplx {
        return Hint[k + M * (s + (int64_t)M * (q + (int64_t)M * l))];
    };
    auto H2 = [&](int k, int l) -> cplx { return Ho[k + M * l]; };
    std::vector<cplx> dk(M), d1(M), d2(M2), A0(M2), A(M3), ot(M * M2), tt((size_t)std::max(P, 1) * M2);
    std::vector<uint8_t> nz1(M2), nzo(M * M2), nzt((size_t)std::max(P, 1) * M2);
    auto nzc = [](cplx z) { return z.re != 0.0 || z.im != 0.0; };
    for (int k = 0; k < M; k++) {
        dk[k] = H2(k, k);                                 // w: Ho[k][k] v_k
        d1[k] = cscale(H4(k, k, k, k), 0.5);              // rule 1 (x 1/2)
        for (int s = k + 1; s < M; s++) d2[k + M * s] = cscale(H4(k, s, k, s), 2.0);  // rule 2: 4 x 1/2
    }
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++) {
            if (l == k) continue;
            int kl = k + M * l;
            // 1-body Ho[k][l]; rule 4: 2 Hint[kkkl] (v_k - 1) x 1/2; rule 5 (s = l):
            // 2 Hint[klll] v_l x 1/2; rule 8: 4 Hint[kssl] v_s x 1/2 for s not in {k,l}
            cplx h4 = H4(k, k, k, l), h5 = H4(k, l, l, l);
            A0[kl] = cmake(H2(k, l).re - h4.re, H2(k, l).im - h4.im);
            for (int s = 0; s < M; s++) {
                cplx a;
                if (s == k) a = h4;
                else if (s == l) a = h5;
                else a = cscale(H4(k, s, s, l), 2.0);
                A[(int64_t)kl * M + s] = a;
                if (nzc(a)) nz1[kl] = 1;
            }
            if (nzc(A0[kl])) nz1[kl] = 1;
        }
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++)
            for (int q = 0; q <= l; q++) {
                if (q == k || l == k) continue;
                int64_t e = q + (int64_t)M * l;
                // rule 3 (q == l): Hint[kkqq] x 1/2; rule 6 (q < l): 2 Hint[kkql] x 1/2
                cplx a = (q == l) ? cscale(H4(k, k, q, q), 0.5) : H4(k, k, q, l);
                ot[k * M2 + e] = a;
                nzo[k * M2 + e] = nzc(a);
            }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            int p = pair_index(k, s, M);
            for (int l = 0; l < M; l++)
                for (int q = 0; q <= l; q++) {
                    if (q == k || q == s || l == k || l == s) continue;
                    int64_t e = q + (int64_t)M * l;
                    // rule 7 (q == l): 2 Hint[ksqq] x 1/2; rule 9 (q < l): 4 Hint[ksql] x 1/2
                    cplx a = (q == l) ? H4(k, s, q, q) : cscale(H4(k, s, q, l), 2.0);
                    tt[p * M2 + e] = a;
                    nzt[p * M2 + e] = nzc(a);
                }
        }
    // the same coefficients in packed-slot order (entries (q, l) outside the
    // removed orbitals, renumbered without them; padding zero)
    const int EO = packed_entries(M, false), ET = packed_entries(M, true);
    const int MPO = packed_orbitals(M, false), MPT = packed_orbitals(M, true);
    std::vector<cplx> otc((size_t)M * EO, cmake(0, 0)), ttc((size_t)std::max(P, 1) * ET, cmake(0, 0));
    for (int k = 0; k < M; k++) {
        int t = 0;
        for (int lp = 0; lp < MPO; lp++)
            for (int qp = 0; qp <= lp; qp++, t++)
                otc[(size_t)k * EO + t] =
                    ot[k * M2 + packed_orbital(k, k, qp) + (int64_t)M * packed_orbital(k, k, lp)];
    }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            const int p = pair_index(k, s, M);
            int t = 0;
            for (int lp = 0; lp < MPT; lp++)
                for (int qp = 0; qp <= lp; qp++, t++)
                    ttc[(size_t)p * ET + t] =
                        tt[p * M2 + packed_orbital(k, s, qp) + (int64_t)M * packed_orbital(k, s, lp)];
        }
    // zero-skipping lists: per-k single-jump masks and, per OT / TT chunk,
    // the entries with a non-zero folded coefficient
    std::vector<uint32_t> m1j(M, 0);
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++)
            if (l != k && nz1[k + M * l]) m1j[k] |= 1u << l;
    std::vector<int32_t> ot_off(M + 1), tt_off(std::max(P, 0) + 1), nz_e;
    std::vector<cplx> nz_c;
    for (int k = 0; k < M; k++) {
        ot_off[k] = (int32_t)nz_e.size();
        for (int e = 0; e < M2; e++)
            if (nzo[k * M2 + e]) { nz_e.push_back(e); nz_c.push_back(ot[k * M2 + e]); }
    }
    ot_off[M] = (int32_t)nz_e.size();
    for (int p = 0; p < P; p++) {
        tt_off[p] = (int32_t)nz_e.size();
        for (int e = 0; e < M2; e++)
            if (nzt[p * M2 + e]) { nz_e.push_back(e); nz_c.push_back(tt[p * M2 + e]); }
    }
    tt_off[P] = (int32_t)nz_e.size();
    // the lists are padded to their largest possible length, so that every
    // Hamiltonian with M orbitals has the same block layout and size (batched
    // H.v addresses instance x's tables at a fixed stride from instance 0's)
    nz_e.resize((size_t)(M + std::max(P, 1)) * M2 + 1, 0);
    nz_c.resize(nz_e.size(), cmake(0, 0));
    // The operator in pair form (schedule 6): every folded term is a
    // normal-ordered product a*_x a*_y a_z a_w or a*_x a_z.  In the gather
    // form a row v collects C(w) for
    //   diagonal      d1[k] v_k(v_k-1) = a*_k a*_k a_k a_k,  d2[k,s] v_k v_s = a*_k a*_s a_s a_k
    //   k -> l        A0[k,l] a*_k a_l;  with a spectator s, A[k,l,s] v_s a*_k a_l, which is
    //                 a*_k a*_s a_s a_l (s not k), or a*_k a*_k a_k a_l + a*_k a_l (s = k)
    //   OT, TT        a*_k a*_k a_q a_l,  a*_k a*_s a_q a_l
    // so the pair matrix collects the two-body coefficients and h1 the
    // one-body ones.  Both are uploaded with the other tables.
    const bool pair = M >= 3 && M <= MCPI_PAIRZ_MAXM;
    const int PWS = pair_stride(M);
    const int HS = pair_h1_stride(M);
    std::vector<cplx> pw(pair ? (size_t)PWS * PWS : 0, cmake(0, 0)), ph1(pair ? (size_t)HS * HS : 0, cmake(0, 0));
    if (pair) {
        auto addw = [&](int c1, int c2, int a1, int a2, cplx z) {
            cplx &w = pw[(size_t)pair_of2(a1, a2) * PWS + pair_of2(c1, c2)];
            w = cadd(w, z);
        };
        auto addh = [&](int k, int l, cplx z) { ph1[(size_t)l * HS + k] = cadd(ph1[(size_t)l * HS + k], z); };
        for (int k = 0; k < M; k++) {
            addh(k, k, dk[k]);
            addw(k, k, k, k, d1[k]);
            for (int s = k + 1; s < M; s++) addw(k, s, k, s, d2[k + M * s]);
        }
        for (int k = 0; k < M; k++)
            for (int l = 0; l < M; l++) {
                if (l == k) continue;
                const int kl = k + M * l;
                addh(k, l, A0[kl]);
                for (int s = 0; s < M; s++) {
                    const cplx a = A[(int64_t)kl * M + s];
                    if (s == k) {
                        addw(k, k, k, l, a);
                        addh(k, l, a);
                    } else {
                        addw(k, s, s, l, a);
                    }
                }
            }
        for (int k = 0; k < M; k++)
            for (int l = 0; l < M; l++)
                for (int q = 0; q <= l; q++) {
                    if (q == k || l == k) continue;
                    addw(k, k, q, l, ot[k * M2 + q + (int64_t)M * l]);
                }
        for (int k = 0; k < M; k++)
            for (int s = k + 1; s < M; s++) {
                const int p = pair_index(k, s, M);
                for (int l = 0; l < M; l++)
                    for (int q = 0; q <= l; q++) {
                        if (q == k || q == s || l == k || l == s) continue;
                        addw(k, s, q, l, tt[p * M2 + q + (int64_t)M * l]);
                    }
            }
    }
    // the factorized pair form of schedule 7 (pair_factor.h)
    std::vector<cplx> za, zb;
    std::vector<int32_t> zblk, zord, zoff;
    if (pair) {
        PairFactor<cplx> pf;
        if (pair_factorize(M, pair_count(M), PWS, HS, pw.data(), ph1.data(), pf)) {
            za = std::move(pf.za);
            zb = std::move(pf.zb);
            zblk = std::move(pf.zblk);
            zord = std::move(pf.zord);
            zoff = std::move(pf.zoff);
        }
    }
    // one device block: complex tables first (16-byte aligned), then integers
    size_t nC = dk.size() + d1.size() + d2.size() + A0.size() + A.size() + ot.size() + tt.size() + nz_c.size() +
                otc.size() + ttc.size() + pw.size() + ph1.size() + za.size() + zb.size();
    size_t nB = nz1.size() + nzo.size() + nzt.size() + 4 * (m1j.size() + ot_off.size() + tt_off.size() +
                                                           nz_e.size() + zblk.size() + zord.size() + zoff.size()) + 64;
    size_t bytes = (nC * sizeof(cplx) + nB + 64 + 255) & ~(size_t)255;
    char *dev = nullptr;
    MCPI_HIP(hipMalloc(&dev, bytes));
    std::vector<char> host(bytes, 0);
    size_t off = 0;
    auto put = [&](const void *src, size_t n) { memcpy(host.data() + off, src, n); size_t o = off; off += n; return o; };
    size_t o_dk = put(dk.data(), dk.size() * sizeof(cplx));
    size_t o_d1 = put(d1.data(), d1.size() * sizeof(cplx));
    size_t o_d2 = put(d2.data(), d2.size() * sizeof(cplx));
    size_t o_A0 = put(A0.data(), A0.size() * sizeof(cplx));
    size_t o_A = put(A.data(), A.size() * sizeof(cplx));
    size_t o_ot = put(ot.data(), ot.size() * sizeof(cplx));
    size_t o_tt = put(tt.data(), tt.size() * sizeof(cplx));
    size_t o_nc = put(nz_c.data(), nz_c.size() * sizeof(cplx));
    size_t o_otc = put(otc.data(), otc.size() * sizeof(cplx));
    size_t o_ttc = put(ttc.data(), ttc.size() * sizeof(cplx));
    const size_t o_pw = pw.empty() ? 0 : put(pw.data(), pw.size() * sizeof(cplx));
    const size_t o_ph1 = ph1.empty() ? 0 : put(ph1.data(), ph1.size() * sizeof(cplx));
    const size_t o_za = za.empty() ? 0 : put(za.data(), za.size() * sizeof(cplx));
    const size_t o_zb = zb.empty() ? 0 : put(zb.data(), zb.size() * sizeof(cplx));
    size_t o_m1 = put(m1j.data(), m1j.size() * 4);
    size_t o_oo = put(ot_off.data(), ot_off.size() * 4);
    size_t o_to = put(tt_off.data(), tt_off.size() * 4);
    size_t o_ne = put(nz_e.data(), nz_e.size() * 4);
    const size_t o_zblk = zblk.empty() ? 0 : put(zblk.data(), zblk.size() * 4);
    const size_t o_zord = zblk.empty() ? 0 : put(zord.data(), zord.size() * 4);
    const size_t o_zoff = zblk.empty() ? 0 : put(zoff.data(), zoff.size() * 4);
    size_t o_n1 = put(nz1.data(), nz1.size());
    size_t o_no = put(nzo.data(), nzo.size());
    size_t o_nt = put(nzt.data(), nzt.size());
    hipError_t e = hipMemcpy(dev, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return hip_fail(e, "coefficient upload", __FILE__, __LINE__);
    }
    c->dk = (cplx *)(dev + o_dk);
    c->d1 = (cplx *)(dev + o_d1);
    c->d2 = (cplx *)(dev + o_d2);
    c->A0 = (cplx *)(dev + o_A0);
    c->A = (cplx *)(dev + o_A);
    c->ot = (cplx *)(dev + o_ot);
    c->tt = (cplx *)(dev + o_tt);
    c->otc = (cplx *)(dev + o_otc);
    c->ttc = (cplx *)(dev + o_ttc);
    c->pw = pair ? (cplx *)(dev + o_pw) : nullptr;
    c->ph1 = pair ? (cplx *)(dev + o_ph1) : nullptr;
    const bool fac = !zblk.empty();
    c->pza = fac ? (cplx *)(dev + o_za) : nullptr;
    c->pzb = fac ? (cplx *)(dev + o_zb) : nullptr;
    c->pzblk = fac ? (int32_t *)(dev + o_zblk) : nullptr;
    c->pzord = fac ? (int32_t *)(dev + o_zord) : nullptr;
    c->pzoff = fac ? (int32_t *)(dev + o_zoff) : nullptr;
    c->pzR = fac ? (int)zoff.size() - 1 : 0;
    c->nz_1j = (uint8_t *)(dev + o_n1);
    c->nz_ot = (uint8_t *)(dev + o_no);
    c->nz_tt = (uint8_t *)(dev + o_nt);
    c->nz_c = (cplx *)(dev + o_nc);
    c->m1j = (uint32_t *)(dev + o_m1);
    c->ot_off = (int32_t *)(dev + o_oo);
    c->tt_off = (int32_t *)(dev + o_to);
    c->nz_e = (int32_t *)(dev + o_ne);
    bool real = true;
    for (const std::vector<cplx> *v : {&dk, &d1, &d2, &A0, &A, &ot, &tt})
        for (const cplx &z : *v) real = real && z.im == 0.0;
    c->real = real ? 1 : 0;
    uint64_t sig = 1469598103934665603ull;   // FNV-1a over the zero pattern
    auto mix = [&sig](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t x = 0; x < n; x++) sig = (sig ^ c[x]) * 1099511628211ull;
    };
    mix(m1j.data(), m1j.size() * 4);
    mix(ot_off.data(), ot_off.size() * 4);
    mix(tt_off.data(), tt_off.size() * 4);
    mix(nz_e.data(), (size_t)tt_off[P] * 4);
    c->zsig = sig;
    const int64_t listed = tt_off[P];   // every listed (non-zero) double-jump term
    const int64_t half = (int64_t)(M + P) * (M * (M + 1) / 2);   // q <= l entries of every chunk kind
    c->dense = (2 * listed >= half) ? 1 : 0;
    *block = dev;
    if (block_bytes) *block_bytes = bytes;
    return MCPI_OK;
}

// ------------------------------------------------------------- kernels --

__device__ inline cplx ldc(const cplx *p)
{
    double2 t = *reinterpret_cast<const double2 *>(p);
    return cmake(t.x, t.y);
}

// rank of v - e_a - e_b + e_c + e_d (b, d may be -1) from the row's prefix
// counts held in LDS (column layout Ps[j*T + t]) and the table B.
__device__ inline int64_t rank_lds(const uint8_t *Ps, int T, int t, const int32_t *B, int N, int M,
                                   int64_t nc, int a, int b, int c, int d)
{
    int64_t s = 0;
    for (int j = 1; j < M; j++) {
        int p = (int)Ps[j * T + t] - (a < j) - (b >= 0 && b < j) + (c < j) + (d >= 0 && d < j);
        s += B[j * (N + 1) + p];
    }
    return nc - 1 - s;
}

// Thread-per-row H.v for all storage modes.
//   MODE 2: Map + MapOT + MapTT (applyHconf_omp / _XX)
//   MODE 1: Map, double jumps ranked (applyHconf_X ranks TT only; here both)
//   MODE 0: every jump ranked (applyHconf)
// Lanes are consecutive rows, so a term's map reads are coalesced (Map) and
// its gathers C[j] land close together; the loops over orbitals are uniform.
template <int MODE>
__global__ __launch_bounds__(256) void k_acth_rows(BasisDev b, HamCoefDev c, const cplx *__restrict__ C,
                                                   cplx *__restrict__ out, int skip_zero)
{
    extern __shared__ double smem_d[];
    const int M = b.M, N = b.N, T = blockDim.x, t = threadIdx.x;
    const int64_t nc = b.nc, MM = (int64_t)M * M;
    double *sq = smem_d;                                    // sqrt(n), n = 0..N+2
    int32_t *Bt = (int32_t *)(sq + N + 3);                  // M*(N+1)
    uint8_t *Vs = (uint8_t *)(Bt + M * (N + 1));            // M*T
    uint8_t *Ps = Vs + M * T;                               // M*T (prefix counts)
    for (int x = t; x < N + 3; x += T) sq[x] = sqrt((double)x);
    if (MODE < 2)
        for (int x = t; x < M * (N + 1); x += T) Bt[x] = b.B[x];
    const unsigned per = gridDim.x / 8;   // XCD-contiguous row ranges (see acth_fast.h)
    const unsigned bx = blockIdx.x < per * 8 ? (blockIdx.x % 8) * per + blockIdx.x / 8 : blockIdx.x;
    const int64_t i = (int64_t)bx * T + t;
    const bool valid = i < nc;
    {
        int acc = 0;
        for (int j = 0; j < M; j++) {
            int x = valid ? b.IF[i * M + j] : 0;
            Vs[j * T + t] = (uint8_t)x;
            Ps[j * T + t] = (uint8_t)acc;
            acc += x;
        }
    }
    __syncthreads();
    if (!valid) return;
#define V(j) ((int)Vs[(j) * T + t])
    const cplx ci = ldc(C + i);

    // diagonal: 1-body diagonal, rules 1 and 2
    cplx dsum = cmake(0, 0);
    for (int k = 0; k < M; k++) {
        int vk = V(k);
        if (vk == 0) continue;
        dsum = cfma(c.dk[k], cmake((double)vk, 0), dsum);
        dsum = cfma(c.d1[k], cmake((double)vk * (vk - 1), 0), dsum);
        for (int s = k + 1; s < M; s++) dsum = cfma(c.d2[k + M * s], cmake((double)vk * V(s), 0), dsum);
    }
    cplx acc = cmake(0, 0);
    acc = cfma(dsum, ci, acc);

    // single jumps k -> l: 1-body off-diagonal, rules 4, 5, 8
    for (int k = 0; k < M; k++) {
        const int vk = V(k);
        // l with a non-zero coefficient (every l != k without zero-skipping)
        uint32_t lmask = skip_zero ? c.m1j[k] : (~0u >> (32 - M)) & ~(1u << k);
        while (lmask) {
            const int l = __ffs(lmask) - 1;
            lmask &= lmask - 1;
            const int kl = k + M * l;
            if (vk < 1) continue;
            cplx h = c.A0[kl];
            const cplx *Ak = c.A + (int64_t)kl * M;
            for (int s = 0; s < M; s++) {
                double vs = (double)V(s);
                h.re = fma(vs, Ak[s].re, h.re);
                h.im = fma(vs, Ak[s].im, h.im);
            }
            int64_t j;
            if (MODE >= 1) j = b.Map[i + (int64_t)kl * nc];
            else j = rank_lds(Ps, T, t, Bt, N, M, nc, k, -1, l, -1);
            double f = sq[vk] * sq[V(l) + 1];
            acc = cfma(h, cscale(ldc(C + j), f), acc);
        }
    }

    // same-orbital double jumps: rules 3 and 6.  With zero-skipping the
    // chunk's list of non-zero entries is walked (HamCoefDev::ot_off / nz_e)
    // instead of testing all M(M+1)/2 entries.
    int64_t cot = 0;
    for (int k = 0; k < M; k++) {
        const int vk = V(k);
        if (vk < 2) continue;
        const double fk = sq[vk] * sq[vk - 1];
        const int64_t base = (MODE == 2) ? b.strideOT[i] + cot * MM : 0;
        cot++;
        if (skip_zero) {
            for (int x0 = c.ot_off[k], x1 = c.ot_off[k + 1]; x0 < x1; x0++) {
                const int e = c.nz_e[x0], q = e % M, l = e / M;
                double x = (q == l) ? sq[V(q) + 1] * sq[V(q) + 2] : sq[V(q) + 1] * sq[V(l) + 1];
                int64_t j;
                if (MODE == 2) j = b.MapOT[base + e];
                else j = rank_lds(Ps, T, t, Bt, N, M, nc, k, k, q, l);
                acc = cfma(c.nz_c[x0], cscale(ldc(C + j), fk * x), acc);
            }
            continue;
        }
        for (int l = 0; l < M; l++) {
            if (l == k) continue;
            for (int q = 0; q <= l; q++) {
                if (q == k) continue;
                const int e = q + M * l;
                double x = (q == l) ? sq[V(q) + 1] * sq[V(q) + 2] : sq[V(q) + 1] * sq[V(l) + 1];
                int64_t j;
                if (MODE == 2) j = b.MapOT[base + e];
                else j = rank_lds(Ps, T, t, Bt, N, M, nc, k, k, q, l);
                acc = cfma(c.ot[k * MM + e], cscale(ldc(C + j), fk * x), acc);
            }
        }
    }

    // two-orbital double jumps: rules 7 and 9
    int64_t ctt = 0;
    int p = 0;
    for (int k = 0; k < M; k++) {
        const int vk = V(k);
        for (int s = k + 1; s < M; s++, p++) {
            const int vs = V(s);
            if (vk < 1 || vs < 1) continue;
            const double fks = sq[vk] * sq[vs];
            const int64_t base = (MODE == 2) ? b.strideTT[i] + ctt * MM : 0;
            ctt++;
            if (skip_zero) {
                for (int x0 = c.tt_off[p], x1 = c.tt_off[p + 1]; x0 < x1; x0++) {
                    const int e = c.nz_e[x0], q = e % M, l = e / M;
                    double x = (q == l) ? sq[V(q) + 1] * sq[V(q) + 2] : sq[V(q) + 1] * sq[V(l) + 1];
                    int64_t j;
                    if (MODE == 2) j = b.MapTT[base + e];
                    else j = rank_lds(Ps, T, t, Bt, N, M, nc, k, s, q, l);
                    acc = cfma(c.nz_c[x0], cscale(ldc(C + j), fks * x), acc);
                }
                continue;
            }
            const cplx *cp = c.tt + (int64_t)p * MM;
            for (int l = 0; l < M; l++) {
                if (l == k || l == s) continue;
                for (int q = 0; q <= l; q++) {
                    if (q == k || q == s) continue;
                    const int e = q + M * l;
                    double x = (q == l) ? sq[V(q) + 1] * sq[V(q) + 2] : sq[V(q) + 1] * sq[V(l) + 1];
                    int64_t j;
                    if (MODE == 2) j = b.MapTT[base + e];
                    else j = rank_lds(Ps, T, t, Bt, N, M, nc, k, s, q, l);
                    acc = cfma(cp[e], cscale(ldc(C + j), fks * x), acc);
                }
            }
        }
    }
#undef V
    double2 r;
    r.x = acc.re;
    r.y = acc.im;
    *reinterpret_cast<double2 *>(out + i) = r;
}

static size_t rows_smem(int N, int M, int T)
{
    return sizeof(double) * (N + 3) + sizeof(int32_t) * M * (N + 1) + 2 * (size_t)M * T + 16;
}

#define MCPI_FAST_ARGS const BasisDev &, int, const HamCoefDev &, int, int, const cplx *, cplx *, hipStream_t, \
                       int, size_t, int64_t
mcpi_status launch_fast_a(MCPI_FAST_ARGS);
mcpi_status launch_fast_b(MCPI_FAST_ARGS);
mcpi_status launch_fast_c(MCPI_FAST_ARGS);
#undef MCPI_FAST_ARGS

// specialised kernel for b.M, or MCPI_ERR_STATE when there is none.
// nbatch > 0: one launch over nbatch instances laid out at fixed strides
// (coefficient tables cstride bytes apart, vectors vstride elements apart)
static mcpi_status launch_fast(const BasisDev &b, int maps, const HamCoefDev &c, int skip, int rank_doubles,
                               const cplx *C, cplx *out, hipStream_t st, int nbatch = 0, size_t cstride = 0,
                               int64_t vstride = 0)
{
    if (b.N > MCPI_MAXN) return MCPI_ERR_STATE;
    if (b.M >= 3 && b.M <= 6) return launch_fast_a(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    if (b.M >= 7 && b.M <= 10)
        return launch_fast_b(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    if (b.M >= 11 && b.M <= 13)
        return launch_fast_c(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    return MCPI_ERR_STATE;
}

// One launch of the specialised schedule `sched` (2 or 3) over nbatch
// instances sharing a basis.  c0: instance 0's tables; instance x's are c0's
// shifted by x * cstride bytes (same layout, build_coefficients), its input
// and output vectors C0 + x * vstride, out0 + x * vstride.  c0.real must hold
// for every instance.  MCPI_ERR_STATE if no specialisation exists.
mcpi_status launch_apply_batch(const BasisDev &b, int maps, const HamCoefDev &c0, int skip, int sched, int nbatch,
                               size_t cstride, const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st)
{
    return launch_fast(b, maps, c0, skip, sched == 3, C0, out0, st, nbatch, cstride, vstride);
}

bool has_fast_kernel(int N, int M) { return N <= MCPI_MAXN && M >= 3 && M <= 13; }

#define MCPI_MULTI_ARGS const BasisDev &, const HamCoefDev &, int, int, size_t, const cplx *, cplx *, int64_t, hipStream_t
mcpi_status launch_multi_a(MCPI_MULTI_ARGS);
mcpi_status launch_multi_b(MCPI_MULTI_ARGS);
mcpi_status launch_multi_c(MCPI_MULTI_ARGS);
#undef MCPI_MULTI_ARGS

// One pass of the multi-vector kernel over nbatch instances that share the
// basis and the zero pattern (c0.zsig), in the batched layout.  sched 2 reads
// the packed double-jump tables, 3 ranks the targets.  Needs Map.
mcpi_status launch_apply_multi(const BasisDev &b, const HamCoefDev &c0, int sched, int nbatch, size_t cstride,
                               const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st)
{
    if (b.N > MCPI_MAXN || !b.Map) return MCPI_ERR_STATE;
    const int rd = sched == 3;
    if (b.M >= 3 && b.M <= 6) return launch_multi_a(b, c0, rd, nbatch, cstride, C0, out0, vstride, st);
    if (b.M >= 7 && b.M <= 10) return launch_multi_b(b, c0, rd, nbatch, cstride, C0, out0, vstride, st);
    if (b.M >= 11 && b.M <= 13) return launch_multi_c(b, c0, rd, nbatch, cstride, C0, out0, vstride, st);
    return MCPI_ERR_STATE;
}

mcpi_status launch_apply_raw(const BasisDev &b, int maps, const HamCoefDev &c, int skip_zero,
                             int schedule, const cplx *C, cplx *out, hipStream_t st)
{
    if (schedule == 5) {
        if (maps == MCPI_MAPS_FULL) {
            mcpi_status s = launch_comp(b, c, C, out, st);
            if (s != MCPI_ERR_STATE) return s;
        }
        schedule = 2;
    }
    if (schedule == 2 || schedule == 3) {
        mcpi_status s = launch_fast(b, maps, c, skip_zero && !c.dense, schedule == 3, C, out, st);
        if (s != MCPI_ERR_STATE) return s;
        set_error("no specialised H.v kernel for this number of orbitals");
        return MCPI_ERR_ARG;
    }
    const int T = 256;
    dim3 grid((unsigned)((b.nc + T - 1) / T));
    size_t sm = rows_smem(b.N, b.M, T);
    if (maps == MCPI_MAPS_FULL)
        hipLaunchKernelGGL(k_acth_rows<2>, grid, dim3(T), sm, st, b, c, C, out, skip_zero);
    else if (maps == MCPI_MAPS_SINGLE)
        hipLaunchKernelGGL(k_acth_rows<1>, grid, dim3(T), sm, st, b, c, C, out, skip_zero);
    else
        hipLaunchKernelGGL(k_acth_rows<0>, grid, dim3(T), sm, st, b, c, C, out, skip_zero);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

// Autotuner: the first H.v of a Hamiltonian (per zero-skipping setting, and
// outside graph capture) times every applicable schedule with HIP events on
// the caller's stream and keeps the fastest.  All schedules compute the same
// result; the last one timed leaves its output in `out`.
// schedule 4 applies: zero-skipping on a sparse Hamiltonian with Map stored
static bool pattern_schedule_ok(mcpi_ham h)
{
    const BasisDev &b = h->basis->d;
    return h->skip_zero && !h->c.dense && h->basis->maps >= MCPI_MAPS_SINGLE && has_fast_kernel(b.N, b.M);
}

// complex entries of the pair-space workspace schedule `sched` writes: the
// factorized form (7) only its Z rows, one per block of the pair matrix
static size_t pair_need(mcpi_ham h, int sched)
{
    const BasisDev &b = h->basis->d;
    if (sched == 7) return (size_t)h->c.pzR * pair_ld(nc_of(b.N - 2, b.M));
    return pair_workspace_entries(b.N, b.M);
}

// the basis's pair-space workspace for stream st with at least `entries`
// complex entries, allocated when absent or too small, if allowed (not
// under capture, memory free beyond the reserve); nullptr when not
// available. A smaller one stays allocated beside a larger one until the
// basis is freed: a graph the caller captured may still point at it.
static cplx *pair_workspace(mcpi_basis b, hipStream_t st, bool may_alloc, size_t entries)
{
    std::lock_guard<std::mutex> lock(b->pair_mu);
    const size_t need = sizeof(cplx) * std::max<size_t>(entries, 1);
    for (const auto &w : b->pair_ws) {
        if (w.st != st) continue;
        size_t have = 0;
        if (hipMemPtrGetInfo(w.p, &have) == hipSuccess && have >= need) return (cplx *)w.p;
        (void)hipGetLastError();
    }
    if (!may_alloc) return nullptr;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < need + ((size_t)4 << 30)) {
        (void)hipGetLastError();
        return nullptr;
    }
    void *p = nullptr;
    if (hipMalloc(&p, need) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    // zero once: the rows of the pair amplitudes past the pairs are read
    // (against zero rows of the pair matrix) and never written
    if (hipMemset(p, 0, need) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(p);
        return nullptr;
    }
    b->pair_ws.push_back({st, p});
    return (cplx *)p;
}

static bool pair_schedule_ok(mcpi_ham h) { return pair_applicable(h->basis->d, h->c); }

// one H.v with schedule `sched`; 4 = specialised kernel reading the pattern
// tables of the Hamiltonian's zero pattern (built on first use unless the
// stream is capturing or memory is short; schedule 2 otherwise)
static mcpi_status launch_sched(mcpi_ham h, int sched, const cplx *C, cplx *out, hipStream_t st)
{
    const BasisDev &b = h->basis->d;
    if (h->basis->sectored()) {
        // rows in sector order: only the pattern tables (built with the
        // basis's row indices) give the targets; every other schedule ranks
        // or reads reference-order tables
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        MCPI_HIP(hipStreamIsCapturing(st, &cs));
        const PatternTables *pt = basis_pattern(h->basis, h->c, cs == hipStreamCaptureStatusNone, st);
        if (!pt) {
            set_error("sector-ordered basis: the pattern tables of this Hamiltonian could not be built "
                      "(device memory short, or first H.v under stream capture)");
            return MCPI_ERR_STATE;
        }
        BasisDev bd = b;
        bd.patOT = pt->tabOT;
        bd.patTT = pt->tabTT;
        bd.ppoffOT = pt->offOT;
        bd.ppoffTT = pt->offTT;
        return launch_apply_raw(bd, h->basis->maps, h->c, 1, 2, C, out, st);
    }
    if (sched == 6 || sched == 7) {
        const bool ok = sched == 7 ? pairz_applicable(b, h->c) : pair_schedule_ok(h);
        if (ok) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            MCPI_HIP(hipStreamIsCapturing(st, &cs));
            cplx *ws = pair_workspace(h->basis, st, cs == hipStreamCaptureStatusNone, pair_need(h, sched));
            if (ws) return sched == 7 ? launch_pairz(b, h->c, C, out, ws, st) : launch_pair(b, h->c, C, out, ws, st);
        }
        sched = 2;
    }
    if (sched == 4) {
        const PatternTables *pt = nullptr;
        if (pattern_schedule_ok(h)) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            MCPI_HIP(hipStreamIsCapturing(st, &cs));
            pt = basis_pattern(h->basis, h->c, cs == hipStreamCaptureStatusNone, st);
        }
        if (pt) {
            BasisDev bd = b;
            bd.patOT = pt->tabOT;
            bd.patTT = pt->tabTT;
            bd.ppoffOT = pt->offOT;
            bd.ppoffTT = pt->offTT;
            return launch_apply_raw(bd, h->basis->maps, h->c, 1, 2, C, out, st);
        }
        sched = 2;
    }
    // every fallback lands on 2, which exists for 3 <= M <= 13 only
    if (sched >= 2 && !has_fast_kernel(b.N, b.M)) sched = 1;
    return launch_apply_raw(b, h->basis->maps, h->c, h->skip_zero, sched, C, out, st);
}

static mcpi_status autotune(mcpi_ham h, const cplx *C, cplx *out, hipStream_t st)
{
    const BasisDev &b = h->basis->d;
    const int maps = h->basis->maps;
    if (h->basis->sectored()) {   // one schedule applies
        h->tuned[h->skip_zero ? 1 : 0] = 4;
        return MCPI_OK;
    }
    int cands[8], n = 0;
    cands[n++] = 1;
    if (has_fast_kernel(b.N, b.M)) {
        cands[n++] = 2;
        if (maps == MCPI_MAPS_FULL) cands[n++] = 3;
        if (maps == MCPI_MAPS_FULL && b.packOT && b.packTT) cands[n++] = 5;
        if (pattern_schedule_ok(h) && basis_pattern(h->basis, h->c, true, st)) cands[n++] = 4;
    }
    // the pair-space schedules read no map: candidates only when the caller
    // has not asked for every term of every map (skip_zero off); the
    // factorized form, where it applies, does less work than the dense
    // product in a smaller workspace, and stands in for it
    if (h->skip_zero && pairz_applicable(b, h->c)) {
        if (pair_workspace(h->basis, st, true, pair_need(h, 7))) cands[n++] = 7;
    } else if (h->skip_zero && has_fast_kernel(b.N, b.M) && pair_schedule_ok(h) &&
               pair_workspace(h->basis, st, true, pair_need(h, 6))) {
        cands[n++] = 6;
    }
    int best = cands[n - 1];
    if (n > 1) {
        hipEvent_t e0, e1;
        MCPI_HIP(hipEventCreate(&e0));
        MCPI_HIP(hipEventCreate(&e1));
        float best_ms = 1e30f;
        for (int x = 0; x < n; x++) {
            MCPI_CHECK(launch_sched(h, cands[x], C, out, st));  // warm-up
            MCPI_HIP(hipEventRecord(e0, st));
            MCPI_CHECK(launch_sched(h, cands[x], C, out, st));
            MCPI_HIP(hipEventRecord(e1, st));
            MCPI_HIP(hipEventSynchronize(e1));
            float ms = 0;
            MCPI_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) { best_ms = ms; best = cands[x]; }
        }
        MCPI_HIP(hipEventDestroy(e0));
        MCPI_HIP(hipEventDestroy(e1));
    }
    h->tuned[h->skip_zero ? 1 : 0] = best;
    return MCPI_OK;
}

mcpi_status launch_apply(mcpi_ham h, const cplx *C, cplx *out, hipStream_t st)
{
    int sched = h->schedule;
    if (sched == 0) {
        int &t = h->tuned[h->skip_zero ? 1 : 0];
        if (t == 0) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            MCPI_HIP(hipStreamIsCapturing(st, &cs));
            if (cs == hipStreamCaptureStatusNone) MCPI_CHECK(autotune(h, C, out, st));
        }
        sched = t ? t : (has_fast_kernel(h->basis->d.N, h->basis->d.M) ? 2 : 1);
    }
    return launch_sched(h, sched, C, out, st);
}

int ham_schedule(mcpi_ham h) { return h->schedule ? h->schedule : h->tuned[h->skip_zero ? 1 : 0]; }

}  // namespace mcpi
