// H.v with the double jumps of a whole block compacted into full wavefronts
// (schedule 5: every term, Map and the wave-packed double-jump tables).
//
// Same maths as k_acth_fast (acth_fast.h; reference applyHconf_omp,
// hamiltonianMatrix.h:1431-1970, folded into per-target coefficients).  In
// the thread-per-row kernels a double jump (k, s) runs its chunk body for a
// whole wavefront when any of its 64 rows has the jump, and at N = M = 12
// only 28 of the 64 lanes do on average.  A vector memory instruction costs
// the texture unit about 16 cycles however few lanes it serves, and the
// chunk bodies (20 table loads and 78 gathers per jump) keep that unit
// busy: lanes left idle are the time lost.
//
// Here a block of 256 rows (four wavefronts) first evaluates the diagonal
// and the single jumps row per lane, and lists, per double jump and per
// wavefront, which of its rows have the jump.  Then the jump's active rows
// of all four wavefronts are dealt out 64 at a time to full wavefronts
// (batches of one jump, dealt round robin over the block's wavefronts): a
// lane takes one (row, jump) pair, reads its targets from the row's
// wave-packed slot (the table of the thread-per-row kernels, unchanged) and
// adds the chunk's sum to a per-wavefront copy of the block's accumulators
// in LDS (a batch holds each row at most once, so no atomics).  The copies
// are summed in a fixed order at the end: results do not depend on timing.
#include "acth_fast.h"

namespace mcpi {

namespace {

// Coefficient tables are read through constant (address space 4) pointers:
// with a generic pointer the accumulator stores to LDS inside the batch loop
// count as possible clobbers and the compiler turns the uniform coefficient
// reads into per-lane vector loads (78 more vector memory instructions per
// batch at M = 12)
#define MCPI_CONST __attribute__((address_space(4)))
typedef const MCPI_CONST cplx *coef_ptr;

// wavefronts per block: 4 (43.6 KB of LDS at M = 12, three blocks per CU).
// Larger blocks need fewer batches (6% fewer with 6 wavefronts, 9% with 8)
// but were slower at N = M = 12: 2.83 ms instead of 2.43 with 6 (77 KB, two
// blocks per CU), 3.81 instead of 3.37 with 8 (119 KB, one block per CU)
constexpr int CW = 4;
constexpr int CR = 64 * CW;    // rows per block

// LDS layout (bytes) for one block; accumulators first for 16-byte alignment
template <int M>
struct CompLayout {
    static constexpr int NJ = M + M * (M - 1) / 2;    // OT jumps k, then TT pairs k < s
    static constexpr size_t acc = 0;                                  // cplx [CW][CR]
    static constexpr size_t pk = acc + sizeof(cplx) * CW * CR;        // uint64 [2][CR]
    static constexpr size_t wb = pk + sizeof(uint64_t) * 2 * CR;      // int32 [2][CW] wave bases OT, TT
    static constexpr size_t cnt = wb + sizeof(int32_t) * 2 * CW;      // int32 [NJ][CW] active rows
    static constexpr size_t pb = cnt + sizeof(int32_t) * NJ * CW;     // int32 [NJ][CW] slot offsets
    static constexpr size_t rows = pb + sizeof(int32_t) * NJ * CW;    // uint8 [NJ][CW][64] active lanes
    static constexpr size_t sq = (rows + NJ * CW * 64 + 15) & ~(size_t)15;   // double [N+3]
    static size_t bytes(int N) { return sq + sizeof(double) * (N + 3); }
};

// diagonal and single jumps k -> l of one row (every term; Map stored):
// 1-body diagonal and off-diagonal, rules 1, 2, 4, 5, 8
template <int M, bool RC>
__device__ inline cplx row_part(const BasisDev &b, const HamCoefDev &c, const cplx *__restrict__ C, int64_t i,
                                const RowState<M> &r, const double *sq)
{
    const int64_t nc = b.nc;
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
    for (int k = 0; k < M; k++) {
        const int vk = occ<M>(r, k);
        const bool act = vk >= 1;
        if (__ballot(act) == 0) continue;   // orbital k empty in every row of the wavefront
        const double fk = act ? sq[vk] : 0.0;
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
            int64_t j = ld_nt(b.Map + i + (int64_t)kl * nc);
            j = act ? j : i;
            acc = term<RC>(h, fk * r.r1[l], ld_c(C + j), acc);
        }
    }
    return acc;
}

// the batches of one double jump J (OT when TT is false: k == s) dealt to
// this wavefront; bctr counts the block's batches so far
template <int M, bool RC, bool TT>
__device__ inline void comp_jump(const int32_t *__restrict__ tab, const cplx *__restrict__ C, coef_ptr coef, int J,
                                 int k, int s, unsigned char *sm, const double *sq, int wave, int lane, int &bctr)
{
    using L = CompLayout<M>;
    // counts are the same for every lane: readfirstlane makes that visible to
    // the compiler, so the batch loop stays uniform and the coefficients
    // remain scalar loads
    const int32_t *cnt = reinterpret_cast<const int32_t *>(sm + L::cnt) + J * CW;
    int cw[CW], pref[CW];   // active rows per wavefront and their prefix sums
    int nact = 0;
#pragma unroll
    for (int u = 0; u < CW; u++) {
        cw[u] = __builtin_amdgcn_readfirstlane(cnt[u]);
        pref[u] = nact;
        nact += cw[u];
    }
    if (nact == 0) return;
    const int nb = (nact + 63) >> 6;
    int beta = (wave - bctr % CW + CW) % CW;
    bctr += nb;
    const int32_t *pbl = reinterpret_cast<const int32_t *>(sm + L::pb) + J * CW;
    const unsigned char *rows = sm + L::rows + (size_t)J * CW * 64;
    const int32_t *wb = reinterpret_cast<const int32_t *>(sm + L::wb) + (TT ? CW : 0);
    const uint64_t *pk = reinterpret_cast<const uint64_t *>(sm + L::pk);
    cplx *acc = reinterpret_cast<cplx *>(sm + L::acc) + wave * CR;
    for (; beta < nb; beta += CW) {
        const int a = beta * 64 + lane;
        if (a >= nact) continue;   // tail of the last batch
        // the lane's wavefront w (the last one whose prefix is <= a)
        int w = 0, base = 0, nw = cw[0];
#pragma unroll
        for (int u = 1; u < CW; u++)
            if (a >= pref[u]) {
                w = u;
                base = pref[u];
                nw = cw[u];
            }
        const int aw = a - base;
        const int rho = w * 64 + rows[w * 64 + aw];
        const uint64_t rp[2] = {pk[rho], pk[CR + rho]};
        const uint32_t off = (uint32_t)(wb[w] + pbl[w] + 4 * aw);
        uint64_t pkc[2];
        compress_occ(rp, k, s, pkc);
        const cplx part = chunk_packed<M, RC, TT>(tab, off, 4u * (uint32_t)nw, C, pkc, sq, coef);
        const int vk = (int)((k < 8 ? (rp[0] >> (8 * k)) : (rp[1] >> (8 * (k - 8)))) & 0xff);
        const int vs = (int)((s < 8 ? (rp[0] >> (8 * s)) : (rp[1] >> (8 * (s - 8)))) & 0xff);
        const double f = TT ? sq[vk] * sq[vs] : sq[vk] * sq[vk - 1];
        cplx x = acc[rho];
        x.re = fma(f, part.re, x.re);
        x.im = fma(f, part.im, x.im);
        acc[rho] = x;
    }
}

template <int M, bool RC>
__global__ __launch_bounds__(CR) __attribute__((amdgpu_waves_per_eu(2))) void k_acth_comp(
    BasisDev b, HamCoefDev c, const cplx *__restrict__ C, cplx *__restrict__ out)
{
    using L = CompLayout<M>;
    constexpr int EO = packed_entries(M, false), ET = packed_entries(M, true);
    extern __shared__ __attribute__((aligned(16))) unsigned char sm[];
    double *sq = reinterpret_cast<double *>(sm + L::sq);
    for (int x = threadIdx.x; x < b.N + 3; x += blockDim.x) sq[x] = sqrt((double)x);
    __syncthreads();

    // chunks of 32 consecutive row blocks dealt round-robin over the XCDs:
    // neighbouring blocks share an XCD's L2, and every XCD gets rows from the
    // whole basis.  N = M = 12: 2.27 ms per H.v; 2.37 ms with single blocks
    // dealt round-robin, 2.44 ms with each XCD walking one contiguous eighth
    // of the rows (whose H.v work differs by up to 40% between eighths),
    // 2.53 ms with the blocks sorted costliest first
    const unsigned vb = chunk_xcd(blockIdx.x, gridDim.x, 32);
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);   // uniform per wavefront
    const int64_t i = (int64_t)vb * CR + t;
    const bool valid = i < b.nc;
    const int64_t nwv = (b.nc + 63) >> 6;
    const int64_t wv = i >> 6;

    // row phase: diagonal and single jumps, row per lane
    RowState<M> r;
    r.pk[0] = r.pk[1] = 0;
    cplx acc = cmake(0, 0);
    if (valid) {
        load_row<M>(b, i, sq, r);
        acc = row_part<M, RC>(b, c, C, i, r, sq);
    }
    uint64_t *pk = reinterpret_cast<uint64_t *>(sm + L::pk);
    pk[t] = r.pk[0];
    pk[CR + t] = r.pk[1];
    cplx *accl = reinterpret_cast<cplx *>(sm + L::acc) + wave * CR;
#pragma unroll
    for (int u = 0; u < CW; u++) accl[u * 64 + lane] = u == wave ? acc : cmake(0, 0);

    // per double jump and wavefront: active rows (lanes), slot offsets
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int32_t *cnt = reinterpret_cast<int32_t *>(sm + L::cnt);
    int32_t *pbl = reinterpret_cast<int32_t *>(sm + L::pb);
    unsigned char *rows = sm + L::rows;
    // slots are addressed from the block's first wavefront's offset (a
    // block's tables span less than 2^31 ints: at most 4 x 91 jumps x 64
    // rows x 92 entries)
    const int64_t w0 = (int64_t)vb * CW;   // < nwv: every block has a valid row
    const int64_t b0OT = b.poffOT[w0], b0TT = b.poffTT[w0];
    int32_t *wb = reinterpret_cast<int32_t *>(sm + L::wb);
    if (lane == 0) {
        wb[wave] = wv < nwv ? (int32_t)(b.poffOT[wv] - b0OT) : 0;
        wb[CW + wave] = wv < nwv ? (int32_t)(b.poffTT[wv] - b0TT) : 0;
    }
    int32_t pb = 0;
    for (int k = 0; k < M; k++) {
        const bool act = valid && occ<M>(r, k) >= 2;
        const uint64_t m = __ballot(act);
        const int J = k;
        if (lane == 0) {
            cnt[J * CW + wave] = __popcll(m);
            pbl[J * CW + wave] = pb;
        }
        if (act) rows[(J * CW + wave) * 64 + __popcll(m & below)] = (unsigned char)lane;
        pb += __popcll(m) * EO;
    }
    pb = 0;
    {
        int J = M;
        for (int k = 0; k < M; k++) {
            const bool ak = valid && occ<M>(r, k) >= 1;
            for (int s = k + 1; s < M; s++, J++) {
                const bool act = ak && occ<M>(r, s) >= 1;
                const uint64_t m = __ballot(act);
                if (lane == 0) {
                    cnt[J * CW + wave] = __popcll(m);
                    pbl[J * CW + wave] = pb;
                }
                if (act) rows[(J * CW + wave) * 64 + __popcll(m & below)] = (unsigned char)lane;
                pb += __popcll(m) * ET;
            }
        }
    }
    __syncthreads();

    // jump phase: same-orbital (rules 3, 6), then two-orbital (rules 7, 9)
    int bctr = 0;
    for (int k = 0; k < M; k++)
        comp_jump<M, RC, false>(b.packOT + b0OT, C, (coef_ptr)(c.otc + (int64_t)k * EO), k, k, k, sm, sq, wave, lane, bctr);
    {
        int J = M, p = 0;
        for (int k = 0; k < M; k++)
            for (int s = k + 1; s < M; s++, J++, p++)
                comp_jump<M, RC, true>(b.packTT + b0TT, C, (coef_ptr)(c.ttc + (int64_t)p * ET), J, k, s, sm, sq, wave, lane,
                                       bctr);
    }
    __syncthreads();

    if (valid) {
        const cplx *a0 = reinterpret_cast<const cplx *>(sm + L::acc);
        cplx sum = a0[t];
#pragma unroll
        for (int u = 1; u < CW; u++) {
            const cplx y = a0[u * CR + t];
            sum.re += y.re;
            sum.im += y.im;
        }
        double2 o;
        o.x = sum.re;
        o.y = sum.im;
        *reinterpret_cast<double2 *>(out + i) = o;
    }
}

template <int M>
mcpi_status launch_comp_m(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, hipStream_t st)
{
    const dim3 grid((unsigned)((b.nc + CR - 1) / CR)), blk(CR);
    const size_t sm = CompLayout<M>::bytes(b.N);
    if (c.real) hipLaunchKernelGGL((k_acth_comp<M, true>), grid, blk, sm, st, b, c, C, out);
    else hipLaunchKernelGGL((k_acth_comp<M, false>), grid, blk, sm, st, b, c, C, out);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

}  // namespace

// schedule 5; MCPI_ERR_STATE when the basis has no packed tables or no
// specialisation exists for its M
mcpi_status launch_comp(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, hipStream_t st)
{
    if (b.N > MCPI_MAXN || !b.Map || !b.packOT || !b.packTT) return MCPI_ERR_STATE;
    if (b.nc == 0) return MCPI_OK;
    switch (b.M) {
    case 3: return launch_comp_m<3>(b, c, C, out, st);
    case 4: return launch_comp_m<4>(b, c, C, out, st);
    case 5: return launch_comp_m<5>(b, c, C, out, st);
    case 6: return launch_comp_m<6>(b, c, C, out, st);
    case 7: return launch_comp_m<7>(b, c, C, out, st);
    case 8: return launch_comp_m<8>(b, c, C, out, st);
    case 9: return launch_comp_m<9>(b, c, C, out, st);
    case 10: return launch_comp_m<10>(b, c, C, out, st);
    case 11: return launch_comp_m<11>(b, c, C, out, st);
    case 12: return launch_comp_m<12>(b, c, C, out, st);
    case 13: return launch_comp_m<13>(b, c, C, out, st);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
