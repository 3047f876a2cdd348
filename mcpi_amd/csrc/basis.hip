// Device-resident Fock bases: hashing tables and jump maps built on the GPU.
//
// Reference behaviour (andriati-alex/mcpi, configurationsMap.h):
//   NC / setupNCmat        145-234
//   IndexToFock (unrank)   238-269
//   FockToIndex (rank)     273-305
//   setupFocks (IF)        309-344
//   OneOneMap              348-417
//   allocTwoTwoMap / TwoTwoMap   421-613
//   allocOneTwoMap / OneTwoMap   617-762
// The values are bit-identical to the reference; offsets are int64 so bases
// whose maps exceed INT32 (the reference refuses them) can be built.
//
// Ranking uses the closed form of the FockToIndex loop.  With prefix counts
// P_j = sum_{o<j} v_o the loop adds, for every orbital i >= 1, the sum of
// NC(n, i) over n in (P_i, P_{i+1}]; summing the hockey-stick identity gives
//     rank(v) = nc - 1 - sum_{j=1}^{M-1} C(P_j + j - 1, j),
// so a jump that moves particles only changes the P_j between the orbitals
// involved, and the rank of any target is one pass over j with table
// lookups B[j][p] = C(p + j - 1, j).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "mcpi_internal.h"

namespace mcpi {

int64_t binom(int64_t n, int64_t r)
{
    if (r < 0 || n < 0 || r > n) return 0;
    if (r > n - r) r = n - r;
    unsigned __int128 acc = 1;
    for (int64_t t = 1; t <= r; t++) {
        acc = acc * (unsigned __int128)(n - r + t) / (unsigned __int128)t;
        if (acc > (unsigned __int128)INT64_MAX) return -1;
    }
    return (int64_t)acc;
}

int64_t nc_of(int N, int M)
{
    if (M == 0) return 1;  // reference NC(N, 0) == 1
    return binom((int64_t)N + M - 1, N);
}

// ------------------------------------------------------------- kernels --

// target rank from prefix counts of the source row and a move of particles:
// removed from orbitals a, b (b = -1: none) and added to c, d (d = -1: none).
__device__ inline int64_t rank_moved(const int *P, const int32_t *B, int N, int M, int64_t nc,
                                     int a, int b, int c, int d)
{
    int64_t s = 0;
    for (int j = 1; j < M; j++) {
        int p = P[j] - (a < j) - (b >= 0 && b < j) + (c < j) + (d >= 0 && d < j);
        s += B[j * (N + 1) + p];
    }
    return nc - 1 - s;
}

// IF table: one thread per configuration, greedy unranking
// (IndexToFock, configurationsMap.h:238-269).  NCmat columns m >= 1 hold
// NC(n, m); NC(n, 1) = 1, so the greedy never goes below orbital 1 while k > 0.
__global__ void k_unrank_if(int N, int M, int64_t nc, const int32_t *__restrict__ NCmat,
                            const int64_t *__restrict__ idx_list, int32_t *__restrict__ IF)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc) return;
    int64_t k = idx_list ? idx_list[t] : t;
    int32_t *v = IF + t * M;
    for (int j = 0; j < M; j++) v[j] = 0;
    int n = N, m = M - 1, cnt = 0;
    while (k > 0) {
        int64_t cost = NCmat[n + (N + 1) * m];
        while (k < cost) {
            if (cnt) { v[m] = cnt; cnt = 0; }
            m--;
            cost = NCmat[n + (N + 1) * m];
        }
        k -= cost;
        cnt++;
        n--;
    }
    if (cnt) v[m] += cnt;
    if (n > 0) v[0] += n;
}

// ranks of arbitrary states (FockToIndex loop, configurationsMap.h:273-305)
__global__ void k_rank_states(int N, int M, int64_t n_states, const int32_t *__restrict__ NCmat,
                              const int32_t *__restrict__ v, int64_t *__restrict__ out)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_states) return;
    const int32_t *w = v + t * M;
    int64_t k = 0;
    int n_left = N;
    for (int i = M - 1; i > 0; i--)
        for (int p = w[i]; p > 0; p--) { k += NCmat[n_left + (N + 1) * i]; n_left--; }
    out[t] = k;
}

// Single-jump map (OneOneMap, configurationsMap.h:348-417): one thread per
// row, writes coalesced over rows for every (k, l).
__global__ void k_one_one_map(mcpi::BasisDev b)
{
    extern __shared__ int smem[];
    const int M = b.M, N = b.N;
    int32_t *Bt = smem;                               // M*(N+1)
    int *Ploc = smem + M * (N + 1) + threadIdx.x * (M + 1);
    for (int t = threadIdx.x; t < M * (N + 1); t += blockDim.x) Bt[t] = b.B[t];
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.nc) return;
    const int32_t *v = b.IF + i * M;
    int acc = 0;
    for (int j = 0; j < M; j++) { Ploc[j] = acc; acc += v[j]; }
    for (int k = 0; k < M; k++) {
        bool occ = v[k] >= 1;
        for (int l = 0; l < M; l++) {
            int64_t at = i + (int64_t)(k + l * M) * b.nc;
            int32_t val = -1;
            if (occ) {
                if (l == k) val = (int32_t)i;
                else {
                    const int64_t j = rank_moved(Ploc, Bt, N, M, b.nc, k, -1, l, -1);
                    val = b.perm ? b.perm[j] : (int32_t)j;
                }
            }
            b.Map[at] = val;
        }
    }
}

// per-row chunk counts: OT = #{k: v_k >= 2}, TT = #{k<s: v_k, v_s >= 1}
// (allocOneTwoMap 617-650, allocTwoTwoMap 421-460)
__global__ void k_chunk_counts(int M, int64_t nc, const int32_t *__restrict__ IF,
                               int32_t *__restrict__ cnt_ot, int32_t *__restrict__ cnt_tt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    int n1 = 0, n2 = 0;
    for (int j = 0; j < M; j++) {
        int x = IF[i * M + j];
        n1 += (x >= 1);
        n2 += (x >= 2);
    }
    cnt_ot[i] = n2;
    cnt_tt[i] = n1 * (n1 - 1) / 2;
}

// ---- exclusive scan of int32 counts into int64 strides (times a factor) --
#define SCAN_BLOCK 256
#define SCAN_ITEMS 8
#define SCAN_TILE (SCAN_BLOCK * SCAN_ITEMS)

__device__ inline int64_t block_exclusive_scan(int64_t x, int64_t *total)
{
    __shared__ int64_t wsum[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t inc = x;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int64_t before = 0, tot = 0;
    for (int t = 0; t < SCAN_BLOCK / 64; t++) {
        if (t < w) before += wsum[t];
        tot += wsum[t];
    }
    __syncthreads();
    *total = tot;
    return before + inc - x;
}

__global__ void k_scan_reduce(int64_t n, const int32_t *__restrict__ cnt, int64_t *__restrict__ partial)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
    for (int t = 0; t < SCAN_ITEMS; t++) {
        int64_t i = base + (int64_t)t * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += cnt[i];
    }
    int64_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void k_scan_partials(int64_t nb, int64_t *__restrict__ partial, int64_t *__restrict__ grand)
{
    // single block: sequential chunks of SCAN_BLOCK
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += SCAN_BLOCK) {
        int64_t i = base + threadIdx.x;
        int64_t x = (i < nb) ? partial[i] : 0;
        int64_t tot;
        int64_t ex = block_exclusive_scan(x, &tot);
        if (i < nb) partial[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void k_scan_final(int64_t n, const int32_t *__restrict__ cnt,
                             const int64_t *__restrict__ partial, int64_t factor,
                             int64_t *__restrict__ out)
{
    // each thread owns SCAN_ITEMS consecutive entries
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int32_t loc[SCAN_ITEMS];
    int64_t s = 0;
    for (int t = 0; t < SCAN_ITEMS; t++) {
        int64_t i = base + t;
        loc[t] = (i < n) ? cnt[i] : 0;
        s += loc[t];
    }
    int64_t tot;
    int64_t ex = block_exclusive_scan(s, &tot) + partial[blockIdx.x];
    for (int t = 0; t < SCAN_ITEMS; t++) {
        int64_t i = base + t;
        if (i < n) out[i] = ex * factor;
        ex += loc[t];
    }
}

__global__ void k_scan_last(int64_t n, const int64_t *__restrict__ grand, int64_t factor,
                            int64_t *__restrict__ out)
{
    out[n] = (*grand) * factor;
}

// Double-jump chunk maps, one wavefront per row (lanes over the row's
// contiguous section, so stores are coalesced).
//   TT (TwoTwoMap 478-613): chunk c = c-th occupied pair (k<s) in
//       lexicographic order; entry l + q*M = rank(v - e_k - e_s + e_l + e_q)
//   OT (OneTwoMap 668-762): chunk c = c-th orbital k with v_k >= 2;
//       entry l + q*M = rank(v - 2 e_k + e_l + e_q)
template <bool TT>
__global__ void k_double_map(mcpi::BasisDev b)
{
    extern __shared__ int smem[];
    const int M = b.M, N = b.N;
    const int waves = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int32_t *Bt = smem;                               // M*(N+1)
    int *Pw = smem + M * (N + 1) + w * (MCPI_MAXM + 1 + MCPI_MAXM * (MCPI_MAXM - 1) / 2);
    int *pairs = Pw + MCPI_MAXM + 1;                  // packed k | s << 8
    for (int t = threadIdx.x; t < M * (N + 1); t += blockDim.x) Bt[t] = b.B[t];
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * waves + w;
    const bool valid = i < b.nc;
    const int32_t *v = b.IF + (valid ? i : 0) * M;
    if (valid && lane == 0) {
        int acc = 0, np = 0;
        for (int j = 0; j < M; j++) { Pw[j] = acc; acc += v[j]; }
        if (TT) {
            for (int k = 0; k < M; k++) {
                if (v[k] < 1) continue;
                for (int s = k + 1; s < M; s++)
                    if (v[s] >= 1) pairs[np++] = k | (s << 8);
            }
        } else {
            for (int k = 0; k < M; k++)
                if (v[k] >= 2) pairs[np++] = k | (k << 8);
        }
    }
    __syncthreads();  // every wave of the block reaches this barrier
    if (!valid) return;
    const int64_t *stride = TT ? b.strideTT : b.strideOT;
    int32_t *map = TT ? b.MapTT : b.MapOT;
    int64_t beg = stride[i], end = stride[i + 1];
    const int MM = M * M;
    for (int64_t p = beg + lane; p < end; p += 64) {
        int64_t off = p - beg;
        int c = (int)(off / MM), e = (int)(off % MM);
        int pk = pairs[c];
        int k = pk & 0xff, s = pk >> 8;
        int l = e % M, q = e / M;
        map[p] = (int32_t)rank_moved(Pw, Bt, N, M, b.nc, k, s, l, q);
    }
}

// ---- wave-packed double-jump tables --
//
// The H.v kernel runs one thread per row with lanes = 64 consecutive rows and
// visits the jumps (k, s) of a chunk kind in a fixed order, uniformly across
// the wavefront.  In the reference layout every lane streams its own row's
// chunk, so each 16-byte load of a wavefront touches 64 different cache
// lines.  The packed layout stores, for wavefront w (rows 64w .. 64w+63) and
// each jump active in at least one lane (ballot != 0), one slot
//     slot[g][a][u] = target of entry t = 4g + u for the a-th active lane
// over the entries q <= l of the chunk (the half H.v reads) that are rule
// terms: q and l outside the jump's removed orbitals {k, s}, at t = l'(l'+1)/2
// + q' with the removed orbitals dropped from the numbering
// (packed_pos in mcpi_internal.h); entries past the real ones hold the row's
// own index.  A 16-byte load of all active lanes is then one contiguous
// run.  Slots follow the reference's chunk order (pairs k < s
// lexicographic; OT: k ascending), so a lane's slot sequence is its row's
// chunk sequence with the rows of the wavefront interleaved.

// ints per wavefront (counts fit int32: at most C(32,2) * 64 * 528)
__global__ void k_packed_counts(int M, int64_t nc, const int32_t *__restrict__ IF, int32_t *__restrict__ cnt_ot,
                                int32_t *__restrict__ cnt_tt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = i >> 6;
    const int64_t nw = (nc + 63) >> 6;
    const bool valid = i < nc;
    uint32_t m1 = 0, m2 = 0;
    if (valid)
        for (int j = 0; j < M; j++) {
            const int x = IF[i * M + j];
            m1 |= (uint32_t)(x >= 1) << j;
            m2 |= (uint32_t)(x >= 2) << j;
        }
    int ot = 0, tt = 0;
    for (int k = 0; k < M; k++) {
        ot += __popcll(__ballot(valid && ((m2 >> k) & 1u)));
        for (int s = k + 1; s < M; s++) tt += __popcll(__ballot(valid && ((m1 >> k) & 1u) && ((m1 >> s) & 1u)));
    }
    if ((threadIdx.x & 63) == 0 && w < nw) {
        cnt_ot[w] = ot * packed_entries(M, false);
        cnt_tt[w] = tt * packed_entries(M, true);
    }
}

template <bool TT>
__global__ void k_packed_fill(mcpi::BasisDev b)
{
    extern __shared__ int smem[];
    const int M = b.M, N = b.N;
    int32_t *Bt = smem;                                                 // M*(N+1)
    int *P = smem + M * (N + 1) + threadIdx.x * (MCPI_MAXM + 1);        // prefix counts
    for (int t = threadIdx.x; t < M * (N + 1); t += blockDim.x) Bt[t] = b.B[t];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = i >> 6;
    const int64_t nw = (b.nc + 63) >> 6;
    if (w >= nw) return;   // whole wavefronts past the last row
    const bool valid = i < b.nc;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t m1 = 0, m2 = 0;
    int acc = 0;
    for (int j = 0; j < M; j++) {
        const int x = valid ? b.IF[i * M + j] : 0;
        P[j] = acc;
        acc += x;
        m1 |= (uint32_t)(x >= 1) << j;
        m2 |= (uint32_t)(x >= 2) << j;
    }
    const int E = packed_entries(M, TT), MP = packed_orbitals(M, TT);
    int32_t *pack = TT ? b.packTT : b.packOT;
    int64_t base = (TT ? b.poffTT : b.poffOT)[w];
    for (int k = 0; k < M; k++) {
        for (int s = TT ? k + 1 : k; s < (TT ? M : k + 1); s++) {
            const bool act = valid && (TT ? (((m1 >> k) & 1u) && ((m1 >> s) & 1u)) : ((m2 >> k) & 1u));
            const uint64_t m = __ballot(act);
            if (m == 0) continue;
            const int nact = __popcll(m), my = __popcll(m & below);
            if (act) {
                int t = 0;
                for (int lp = 0; lp < MP; lp++)
                    for (int qp = 0; qp <= lp; qp++, t++) {
                        const int q = packed_orbital(k, s, qp), l = packed_orbital(k, s, lp);
                        pack[base + ((int64_t)(t >> 2) * nact + my) * 4 + (t & 3)] =
                            (int32_t)rank_moved(P, Bt, N, M, b.nc, k, s, l, q);
                    }
                for (; t < E; t++) pack[base + ((int64_t)(t >> 2) * nact + my) * 4 + (t & 3)] = (int32_t)i;
            }
            base += (int64_t)nact * E;
        }
    }
}

// ---- pattern tables (targets of the non-zero terms of one zero pattern) --
//
// For a Hamiltonian whose folded double-jump coefficients are mostly zero
// (Lieb-Liniger keeps about 3 of the 78 entries of a chunk), the zero-
// skipping H.v only needs the targets of the listed entries.  These tables
// hold exactly those: for wavefront w and each jump with a non-empty list of
// n entries that is active in at least one lane, one slot [t][a] = target
// of list entry t of the a-th active lane (entry-major: the wavefront reads
// entry t of all its active lanes as one contiguous run; no padding, so the
// tables hold 4 bytes per listed term and row).

__global__ void k_pattern_counts(mcpi::BasisDev b, HamCoefDev c, int32_t *__restrict__ cnt_ot,
                                 int32_t *__restrict__ cnt_tt)
{
    const int M = b.M;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = i >> 6, nw = (b.nc + 63) >> 6;
    const bool valid = i < b.nc;
    uint32_t m1 = 0, m2 = 0;
    if (valid)
        for (int j = 0; j < M; j++) {
            const int x = b.IF[i * M + j];
            m1 |= (uint32_t)(x >= 1) << j;
            m2 |= (uint32_t)(x >= 2) << j;
        }
    int64_t ot = 0, tt = 0;
    for (int k = 0; k < M; k++) {
        const int n = c.ot_off[k + 1] - c.ot_off[k];
        if (n) ot += (int64_t)__popcll(__ballot(valid && ((m2 >> k) & 1u))) * n;
    }
    int p = 0;
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++, p++) {
            const int n = c.tt_off[p + 1] - c.tt_off[p];
            if (n) tt += (int64_t)__popcll(__ballot(valid && ((m1 >> k) & 1u) && ((m1 >> s) & 1u))) * n;
        }
    if ((threadIdx.x & 63) == 0 && w < nw) {
        cnt_ot[w] = (int32_t)ot;
        cnt_tt[w] = (int32_t)tt;
    }
}

template <bool TT>
__global__ void k_pattern_fill(mcpi::BasisDev b, HamCoefDev c, const int64_t *__restrict__ off,
                               int32_t *__restrict__ tab)
{
    extern __shared__ int smem[];
    const int M = b.M, N = b.N;
    int32_t *Bt = smem;                                                 // M*(N+1)
    int *P = smem + M * (N + 1) + threadIdx.x * (MCPI_MAXM + 1);        // prefix counts
    for (int t = threadIdx.x; t < M * (N + 1); t += blockDim.x) Bt[t] = b.B[t];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = i >> 6, nw = (b.nc + 63) >> 6;
    if (w >= nw) return;   // whole wavefronts past the last row
    const bool valid = i < b.nc;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t m1 = 0, m2 = 0;
    int acc = 0;
    for (int j = 0; j < M; j++) {
        const int x = valid ? b.IF[i * M + j] : 0;
        P[j] = acc;
        acc += x;
        m1 |= (uint32_t)(x >= 1) << j;
        m2 |= (uint32_t)(x >= 2) << j;
    }
    int64_t base = off[w];
    int p = 0;
    for (int k = 0; k < M; k++) {
        for (int s = TT ? k + 1 : k; s < (TT ? M : k + 1); s++, p++) {
            const int32_t *lo = TT ? c.tt_off + p : c.ot_off + k;
            const int beg = lo[0], n = lo[1] - lo[0];
            if (n == 0) continue;
            const bool act = valid && (TT ? (((m1 >> k) & 1u) && ((m1 >> s) & 1u)) : ((m2 >> k) & 1u));
            const uint64_t m = __ballot(act);
            if (m == 0) continue;
            const int nact = __popcll(m), my = __popcll(m & below);
            if (act)
                for (int t = 0; t < n; t++) {
                    const int e = c.nz_e[beg + t];
                    const int64_t j = rank_moved(P, Bt, N, M, b.nc, k, s, e / M, e % M);
                    tab[base + (int64_t)t * nact + my] = b.perm ? b.perm[j] : (int32_t)j;
                }
            base += (int64_t)nact * n;
        }
    }
}

// ---- sector order --
// charge of every state: Q(v) = sum_j q_j v_j
__global__ void k_state_charge(int M, int64_t nc, const int32_t *__restrict__ IF, const int32_t *__restrict__ q,
                               int32_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    int32_t s = 0;
    for (int j = 0; j < M; j++) s += q[j] * IF[i * M + j];
    out[i] = s;
}

__global__ void k_gather_rows(int64_t n, const int32_t *__restrict__ idx, const cplx *__restrict__ in,
                              cplx *__restrict__ out)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256)
        out[r] = in[idx[r]];
}

__global__ void k_scatter_rows(int64_t n, const int32_t *__restrict__ idx, const cplx *__restrict__ in,
                               cplx *__restrict__ out)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256)
        out[idx[r]] = in[r];
}

__global__ void k_widen_rows(int64_t n, const int32_t *__restrict__ a, int64_t *__restrict__ w)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n) w[r] = a[r];
}

static unsigned row_blocks(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 8192); }

mcpi_status gather_rows(int64_t n, const int32_t *idx, const cplx *in, cplx *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_rows, dim3(row_blocks(n)), dim3(256), 0, st, n, idx, in, out);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

mcpi_status scatter_rows(int64_t n, const int32_t *idx, const cplx *in, cplx *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_scatter_rows, dim3(row_blocks(n)), dim3(256), 0, st, n, idx, in, out);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

}  // namespace mcpi

using namespace mcpi;

// ------------------------------------------------------------ host side --

static int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

static mcpi_status scan_to_strides(int64_t n, const int32_t *cnt, int64_t factor, int64_t *out,
                                   int64_t *total_host)
{
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t *partial = nullptr, *grand = nullptr;
    MCPI_HIP(hipMalloc(&partial, sizeof(int64_t) * (nb + 1)));
    MCPI_HIP(hipMalloc(&grand, sizeof(int64_t)));
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLOCK), 0, 0, n, cnt, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_BLOCK), 0, 0, nb, partial, grand);
    hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(SCAN_BLOCK), 0, 0, n, cnt, partial, factor, out);
    hipLaunchKernelGGL(k_scan_last, dim3(1), dim3(1), 0, 0, n, grand, factor, out);
    MCPI_HIP(hipGetLastError());
    int64_t chunks = 0;
    MCPI_HIP(hipMemcpy(&chunks, grand, sizeof(int64_t), hipMemcpyDeviceToHost));
    *total_host = chunks * factor;   // table length, matching out[n]
    MCPI_HIP(hipFree(partial));
    MCPI_HIP(hipFree(grand));
    return MCPI_OK;
}

static void host_tables(int N, int M, std::vector<int32_t> &NCmat, std::vector<int32_t> &B)
{
    NCmat.assign((size_t)(N + 1) * (M + 1), 0);
    for (int n = 0; n <= N; n++) {
        NCmat[n + (N + 1) * 0] = 0;
        if (M >= 1) NCmat[n + (N + 1) * 1] = 1;
        for (int m = 2; m <= M; m++) NCmat[n + (N + 1) * m] = (int32_t)nc_of(n, m);
    }
    B.assign((size_t)M * (N + 1), 0);
    for (int j = 1; j < M; j++)
        for (int p = 0; p <= N; p++) B[j * (N + 1) + p] = (int32_t)binom((int64_t)p + j - 1, j);
}

static void free_basis(mcpi_basis b)
{
    if (!b) return;
    void *ptrs[] = {b->d.IF,     b->d.Map,    b->d.MapOT,  b->d.MapTT,  b->d.strideOT, b->d.strideTT,
                    b->d.B,      b->d.NCmat,  b->d.packOT, b->d.packTT, b->d.poffOT,   b->d.poffTT};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (PatternTables &t : b->patterns) {
        void *q[] = {t.tabOT, t.tabTT, t.offOT, t.offTT};
        for (void *p : q)
            if (p) (void)hipFree(p);
    }
    if (b->rho_ws) (void)hipFree(b->rho_ws);
    for (auto &w : b->pair_ws) (void)hipFree(w.p);
    void *sec[] = {(void *)b->d.perm, b->iperm, b->IF_ref, b->vt[0], b->vt[1]};
    for (void *p : sec)
        if (p) (void)hipFree(p);
    delete b;
}

// Every table gets 64 bytes of padding: the specialised H.v kernels read
// map chunks in 16-byte pieces, and for odd M the last piece of the last
// chunk extends up to 12 bytes past the table.
static mcpi_status alloc_dev(void **p, int64_t bytes, int64_t *acct)
{
    hipError_t e = hipMalloc(p, (size_t)std::max<int64_t>(bytes, 0) + 64);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        char msg[160];
        snprintf(msg, sizeof msg, "device allocation of %lld bytes failed: %s", (long long)bytes,
                 hipGetErrorString(e));
        set_error(msg);
        *p = nullptr;
        return MCPI_ERR_ALLOC;
    }
    *acct += bytes;
    return MCPI_OK;
}

// common part of basis creation: tables, IF (computed or uploaded)
static mcpi_status basis_init(int N, int M, const int32_t *IF_host, const int32_t *NCmat_host,
                              mcpi_basis *out)
{
    MCPI_ARG(N >= 0 && N <= MCPI_MAXN, "N out of range");
    MCPI_ARG(M >= 1 && M <= MCPI_MAXM, "M out of range (1..32)");
    int64_t nc = nc_of(N, M);
    MCPI_ARG(nc > 0 && nc <= INT32_MAX, "number of configurations exceeds int32 indices");
    mcpi_basis b = new mcpi_basis_s();
    memset(&b->d, 0, sizeof b->d);
    b->d.N = N;
    b->d.M = M;
    b->d.nc = nc;
    std::vector<int32_t> NCmat, B;
    host_tables(N, M, NCmat, B);
    if (NCmat_host) memcpy(NCmat.data(), NCmat_host, NCmat.size() * sizeof(int32_t));
    mcpi_status st;
    if ((st = alloc_dev((void **)&b->d.NCmat, NCmat.size() * 4, &b->device_bytes)) ||
        (st = alloc_dev((void **)&b->d.B, B.size() * 4 + 4, &b->device_bytes)) ||
        (st = alloc_dev((void **)&b->d.IF, nc * M * 4, &b->device_bytes))) {
        free_basis(b);
        return st;
    }
    hipError_t e = hipMemcpy(b->d.NCmat, NCmat.data(), NCmat.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d.B, B.data(), B.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && IF_host)
        e = hipMemcpy(b->d.IF, IF_host, (size_t)nc * M * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_basis(b);
        return hip_fail(e, "basis table upload", __FILE__, __LINE__);
    }
    if (!IF_host) {
        hipLaunchKernelGGL(k_unrank_if, dim3((nc + 255) / 256), dim3(256), 0, 0, N, M, nc,
                           b->d.NCmat, (const int64_t *)nullptr, b->d.IF);
        e = hipGetLastError();
        if (e != hipSuccess) {
            free_basis(b);
            return hip_fail(e, "unrank kernel", __FILE__, __LINE__);
        }
    }
    *out = b;
    return MCPI_OK;
}

// wave-packed double-jump tables (built from IF, independent of MapOT/MapTT)
static mcpi_status build_packed(mcpi_basis b)
{
    const int M = b->d.M, N = b->d.N;
    const int64_t nc = b->d.nc, nw = (nc + 63) / 64;
    if (b->d.packTT) return MCPI_OK;
    int32_t *cnt_ot = nullptr, *cnt_tt = nullptr;
    MCPI_HIP(hipMalloc(&cnt_ot, nw * 4));
    MCPI_HIP(hipMalloc(&cnt_tt, nw * 4));
    const int64_t grid = (nw * 64 + 255) / 256;
    hipLaunchKernelGGL(k_packed_counts, dim3(grid), dim3(256), 0, 0, M, nc, b->d.IF, cnt_ot, cnt_tt);
    MCPI_HIP(hipGetLastError());
    MCPI_CHECK(alloc_dev((void **)&b->d.poffOT, (nw + 1) * 8, &b->device_bytes));
    MCPI_CHECK(alloc_dev((void **)&b->d.poffTT, (nw + 1) * 8, &b->device_bytes));
    MCPI_CHECK(scan_to_strides(nw, cnt_ot, 1, b->d.poffOT, &b->pot_len));
    MCPI_CHECK(scan_to_strides(nw, cnt_tt, 1, b->d.poffTT, &b->ptt_len));
    MCPI_CHECK(alloc_dev((void **)&b->d.packOT, b->pot_len * 4, &b->device_bytes));
    MCPI_CHECK(alloc_dev((void **)&b->d.packTT, b->ptt_len * 4, &b->device_bytes));
    const size_t sm = sizeof(int) * (M * (N + 1) + 256 * (MCPI_MAXM + 1));
    if (b->pot_len > 0) hipLaunchKernelGGL(k_packed_fill<false>, dim3(grid), dim3(256), sm, 0, b->d);
    if (b->ptt_len > 0) hipLaunchKernelGGL(k_packed_fill<true>, dim3(grid), dim3(256), sm, 0, b->d);
    MCPI_HIP(hipGetLastError());
    MCPI_HIP(hipFree(cnt_ot));
    MCPI_HIP(hipFree(cnt_tt));
    return MCPI_OK;
}

static mcpi_status build_maps_parts(mcpi_basis b, bool want_map, bool want_ot, bool want_tt)
{
    const int M = b->d.M, N = b->d.N;
    const int64_t nc = b->d.nc;
    if (want_map && !b->d.Map) {
        MCPI_CHECK(alloc_dev((void **)&b->d.Map, nc * M * M * 4, &b->device_bytes));
        b->map_len = nc * M * M;
        size_t sm = sizeof(int) * (M * (N + 1) + 256 * (M + 1));
        hipLaunchKernelGGL(k_one_one_map, dim3((nc + 255) / 256), dim3(256), sm, 0, b->d);
        MCPI_HIP(hipGetLastError());
    }
    if (want_ot || want_tt) {
        int32_t *cnt_ot = nullptr, *cnt_tt = nullptr;
        MCPI_HIP(hipMalloc(&cnt_ot, nc * 4));
        MCPI_HIP(hipMalloc(&cnt_tt, nc * 4));
        hipLaunchKernelGGL(k_chunk_counts, dim3((nc + 255) / 256), dim3(256), 0, 0, M, nc, b->d.IF,
                           cnt_ot, cnt_tt);
        MCPI_HIP(hipGetLastError());
        const int waves = 4;
        size_t sm = sizeof(int) * (M * (N + 1) +
                                   waves * (MCPI_MAXM + 1 + MCPI_MAXM * (MCPI_MAXM - 1) / 2));
        int64_t grid = (nc + waves - 1) / waves;
        if (want_ot && !b->d.MapOT) {
            MCPI_CHECK(alloc_dev((void **)&b->d.strideOT, (nc + 1) * 8, &b->device_bytes));
            MCPI_CHECK(scan_to_strides(nc, cnt_ot, (int64_t)M * M, b->d.strideOT, &b->ot_len));
            MCPI_CHECK(alloc_dev((void **)&b->d.MapOT, b->ot_len * 4, &b->device_bytes));
            if (b->ot_len > 0)
                hipLaunchKernelGGL(k_double_map<false>, dim3(grid), dim3(64 * waves), sm, 0, b->d);
            MCPI_HIP(hipGetLastError());
        }
        if (want_tt && !b->d.MapTT) {
            MCPI_CHECK(alloc_dev((void **)&b->d.strideTT, (nc + 1) * 8, &b->device_bytes));
            MCPI_CHECK(scan_to_strides(nc, cnt_tt, (int64_t)M * M, b->d.strideTT, &b->tt_len));
            MCPI_CHECK(alloc_dev((void **)&b->d.MapTT, b->tt_len * 4, &b->device_bytes));
            if (b->tt_len > 0)
                hipLaunchKernelGGL(k_double_map<true>, dim3(grid), dim3(64 * waves), sm, 0, b->d);
            MCPI_HIP(hipGetLastError());
        }
        MCPI_HIP(hipFree(cnt_ot));
        MCPI_HIP(hipFree(cnt_tt));
    }
    if (b->d.Map && b->d.MapOT && b->d.MapTT) b->maps = MCPI_MAPS_FULL;
    else if (b->d.Map) b->maps = MCPI_MAPS_SINGLE;
    else b->maps = MCPI_MAPS_NONE;
    if (b->maps == MCPI_MAPS_FULL) MCPI_CHECK(build_packed(b));
    MCPI_HIP(hipDeviceSynchronize());
    return MCPI_OK;
}

static mcpi_status basis_build_maps(mcpi_basis b, int maps)
{
    return build_maps_parts(b, maps >= MCPI_MAPS_SINGLE, maps >= MCPI_MAPS_FULL, maps >= MCPI_MAPS_FULL);
}

// Sector order: rows sorted by the conserved charge Q(v) = sum_j q_j v_j,
// the reference order kept inside a sector (stable counting sort on the
// host over the charges computed on the device).  IF is rebuilt in row
// order by unranking the reference indices, and the single-jump map holds
// row indices.  A Hamiltonian that conserves Q only couples rows of one
// sector, so the H.v of a sector gathers from that sector's slice of the
// vector.
static mcpi_status make_sectored(mcpi_basis b, const int32_t *charge)
{
    const int M = b->d.M, N = b->d.N;
    const int64_t nc = b->d.nc;
    b->charge.assign(charge, charge + M);
    int32_t *dq = nullptr, *dQ = nullptr;
    MCPI_HIP(hipMalloc(&dq, sizeof(int32_t) * M));
    MCPI_HIP(hipMalloc(&dQ, sizeof(int32_t) * nc));
    MCPI_HIP(hipMemcpy(dq, charge, sizeof(int32_t) * M, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_state_charge, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, M, nc, b->d.IF, dq, dQ);
    MCPI_HIP(hipGetLastError());
    std::vector<int32_t> Q(nc);
    MCPI_HIP(hipMemcpy(Q.data(), dQ, sizeof(int32_t) * nc, hipMemcpyDeviceToHost));
    MCPI_HIP(hipFree(dq));
    const int32_t qmin = *std::min_element(Q.begin(), Q.end()), qmax = *std::max_element(Q.begin(), Q.end());
    std::vector<int64_t> start((size_t)(qmax - qmin) + 2, 0);
    for (int64_t i = 0; i < nc; i++) start[(size_t)(Q[i] - qmin) + 1]++;
    b->sec_off.assign(1, 0);
    for (size_t c = 1; c < start.size(); c++) {
        if (start[c]) b->sec_off.push_back(b->sec_off.back() + start[c]);
        start[c] += start[c - 1];
    }
    std::vector<int32_t> iperm(nc), perm(nc);
    for (int64_t i = 0; i < nc; i++) {
        const int64_t r = start[(size_t)(Q[i] - qmin)]++;
        iperm[r] = (int32_t)i;
        perm[i] = (int32_t)r;
    }
    int32_t *dperm = nullptr;
    MCPI_CHECK(alloc_dev((void **)&dperm, nc * 4, &b->device_bytes));
    b->d.perm = dperm;
    MCPI_CHECK(alloc_dev((void **)&b->iperm, nc * 4, &b->device_bytes));
    MCPI_HIP(hipMemcpy(dperm, perm.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice));
    MCPI_HIP(hipMemcpy(b->iperm, iperm.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice));
    // IF in row order: unrank the reference index of every row (the
    // reference-order IF is kept for the density matrices and for copies)
    MCPI_HIP(hipFree(dQ));
    int64_t *idx64 = nullptr;
    MCPI_HIP(hipMalloc(&idx64, sizeof(int64_t) * nc));
    hipLaunchKernelGGL(k_widen_rows, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, nc, b->iperm, idx64);
    int32_t *IFrow = nullptr;
    MCPI_CHECK(alloc_dev((void **)&IFrow, nc * M * 4, &b->device_bytes));
    hipLaunchKernelGGL(k_unrank_if, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, N, M, nc, b->d.NCmat,
                       (const int64_t *)idx64, IFrow);
    MCPI_HIP(hipGetLastError());
    MCPI_HIP(hipDeviceSynchronize());
    MCPI_HIP(hipFree(idx64));
    b->IF_ref = b->d.IF;
    b->d.IF = IFrow;
    return MCPI_OK;
}

namespace mcpi {

mcpi_status basis_vectors(mcpi_basis b, cplx **v0, cplx **v1)
{
    for (int k = 0; k < 2; k++)
        if (!b->vt[k]) MCPI_CHECK(alloc_dev((void **)&b->vt[k], sizeof(cplx) * b->d.nc, &b->device_bytes));
    *v0 = b->vt[0];
    if (v1) *v1 = b->vt[1];
    return MCPI_OK;
}

// Pattern tables are kept for at most this many zero patterns per basis
// (least recently used dropped first) and only while they leave this much
// device memory free.
#define MCPI_PATTERNS_PER_BASIS 4
#define MCPI_PATTERN_RESERVE ((int64_t)4 << 30)

static void free_pattern(PatternTables &t)
{
    void *q[] = {t.tabOT, t.tabTT, t.offOT, t.offTT};
    for (void *p : q)
        if (p) (void)hipFree(p);
    t = PatternTables();
}

const PatternTables *basis_pattern(mcpi_basis b, const HamCoefDev &c, bool may_build, hipStream_t st)
{
    for (PatternTables &t : b->patterns)
        if (t.zsig == c.zsig) {
            t.last_use = ++b->use_clock;
            return &t;
        }
    if (!may_build || b->d.M > MCPI_MAXM) return nullptr;
    const int M = b->d.M, N = b->d.N;
    const int64_t nc = b->d.nc, nw = (nc + 63) / 64;
    PatternTables t;
    t.zsig = c.zsig;
    int32_t *cnt = nullptr;
    if (hipMalloc(&cnt, sizeof(int32_t) * 2 * nw) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    const int64_t grid = (nw * 64 + 255) / 256;
    bool ok = true;
    hipLaunchKernelGGL(k_pattern_counts, dim3(grid), dim3(256), 0, st, b->d, c, cnt, cnt + nw);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    ok = ok && hipMalloc(&t.offOT, sizeof(int64_t) * (nw + 1)) == hipSuccess &&
         hipMalloc(&t.offTT, sizeof(int64_t) * (nw + 1)) == hipSuccess;
    ok = ok && scan_to_strides(nw, cnt, 1, t.offOT, &t.lenOT) == MCPI_OK &&
         scan_to_strides(nw, cnt + nw, 1, t.offTT, &t.lenTT) == MCPI_OK;
    (void)hipFree(cnt);
    size_t free_b = 0, total_b = 0;
    ok = ok && hipMemGetInfo(&free_b, &total_b) == hipSuccess;
    t.bytes = 4 * (t.lenOT + t.lenTT) + 16 * (nw + 1);
    ok = ok && (int64_t)free_b > 4 * (t.lenOT + t.lenTT) + MCPI_PATTERN_RESERVE;
    ok = ok && hipMalloc(&t.tabOT, 4 * std::max<int64_t>(t.lenOT, 1)) == hipSuccess &&
         hipMalloc(&t.tabTT, 4 * std::max<int64_t>(t.lenTT, 1)) == hipSuccess;
    if (ok) {
        const size_t sm = sizeof(int) * (M * (N + 1) + 256 * (MCPI_MAXM + 1));
        if (t.lenOT) hipLaunchKernelGGL(k_pattern_fill<false>, dim3(grid), dim3(256), sm, st, b->d, c, t.offOT, t.tabOT);
        if (t.lenTT) hipLaunchKernelGGL(k_pattern_fill<true>, dim3(grid), dim3(256), sm, st, b->d, c, t.offTT, t.tabTT);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        free_pattern(t);
        return nullptr;
    }
    if ((int)b->patterns.size() >= MCPI_PATTERNS_PER_BASIS) {
        size_t old = 0;
        for (size_t x = 1; x < b->patterns.size(); x++)
            if (b->patterns[x].last_use < b->patterns[old].last_use) old = x;
        free_pattern(b->patterns[old]);
        b->patterns.erase(b->patterns.begin() + old);
    }
    t.last_use = ++b->use_clock;
    b->patterns.push_back(t);
    return &b->patterns.back();
}

// Upload of caller-provided tables (reference-shaped entry points).  int32
// strides of the reference become int64 with a closing entry.
mcpi_status basis_upload_raw(int N, int M, const int32_t *IF, const int32_t *NCmat,
                             const int32_t *Map, const int32_t *MapOT, const int32_t *MapTT,
                             const int *strideOT, const int *strideTT, int64_t ot_len,
                             int64_t tt_len, mcpi_basis *out)
{
    mcpi_basis b = nullptr;
    MCPI_CHECK(basis_init(N, M, IF, NCmat, &b));
    const int64_t nc = b->d.nc;
    mcpi_status st = MCPI_OK;
    b->maps = MCPI_MAPS_NONE;
    if (Map) {
        b->maps = MCPI_MAPS_SINGLE;
        b->map_len = nc * M * M;
        st = alloc_dev((void **)&b->d.Map, b->map_len * 4, &b->device_bytes);
        if (!st && hipMemcpy(b->d.Map, Map, b->map_len * 4, hipMemcpyHostToDevice) != hipSuccess)
            st = MCPI_ERR_HIP;
    }
    if (!st && MapOT && MapTT && strideOT && strideTT) {
        b->maps = MCPI_MAPS_FULL;
        b->ot_len = ot_len;
        b->tt_len = tt_len;
        std::vector<int64_t> s64(nc + 1);
        st = alloc_dev((void **)&b->d.strideOT, (nc + 1) * 8, &b->device_bytes);
        if (!st) st = alloc_dev((void **)&b->d.strideTT, (nc + 1) * 8, &b->device_bytes);
        if (!st) st = alloc_dev((void **)&b->d.MapOT, ot_len * 4, &b->device_bytes);
        if (!st) st = alloc_dev((void **)&b->d.MapTT, tt_len * 4, &b->device_bytes);
        if (!st) {
            for (int64_t i = 0; i < nc; i++) s64[i] = strideOT[i];
            s64[nc] = ot_len;
            if (hipMemcpy(b->d.strideOT, s64.data(), (nc + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
                st = MCPI_ERR_HIP;
            for (int64_t i = 0; i < nc; i++) s64[i] = strideTT[i];
            s64[nc] = tt_len;
            if (!st && hipMemcpy(b->d.strideTT, s64.data(), (nc + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
                st = MCPI_ERR_HIP;
            if (!st && ot_len && hipMemcpy(b->d.MapOT, MapOT, ot_len * 4, hipMemcpyHostToDevice) != hipSuccess)
                st = MCPI_ERR_HIP;
            if (!st && tt_len && hipMemcpy(b->d.MapTT, MapTT, tt_len * 4, hipMemcpyHostToDevice) != hipSuccess)
                st = MCPI_ERR_HIP;
        }
    }
    if (!st && b->maps == MCPI_MAPS_FULL) {
        st = build_packed(b);
        if (!st && hipDeviceSynchronize() != hipSuccess) st = MCPI_ERR_HIP;
    }
    if (st) {
        if (st == MCPI_ERR_HIP) set_error("upload of reference tables failed");
        free_basis(b);
        return st;
    }
    *out = b;
    return MCPI_OK;
}

mcpi_status rank_tables_dev(int N, int M, int32_t **B, int32_t **NCmat)
{
    MCPI_ARG(N >= 0 && N <= MCPI_MAXN && M >= 1 && M <= MCPI_MAXM, "N or M out of range");
    std::vector<int32_t> nch, bh;
    host_tables(N, M, nch, bh);
    *B = *NCmat = nullptr;
    int64_t acct = 0;
    MCPI_CHECK(alloc_dev((void **)B, bh.size() * 4 + 4, &acct));
    mcpi_status st = alloc_dev((void **)NCmat, nch.size() * 4, &acct);
    if (st == MCPI_OK && (hipMemcpy(*B, bh.data(), bh.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                          hipMemcpy(*NCmat, nch.data(), nch.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
        st = hip_fail(hipGetLastError(), "rank table upload", __FILE__, __LINE__);
    if (st != MCPI_OK) {
        if (*B) (void)hipFree(*B);
        if (*NCmat) (void)hipFree(*NCmat);
        *B = *NCmat = nullptr;
    }
    return st;
}

__global__ void k_widen32(int64_t n, const int32_t *__restrict__ a, int64_t *__restrict__ w)
{
    for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (int64_t)gridDim.x * 256) w[x] = a[x];
}

mcpi_status widen_strides(int64_t n, const int32_t *a, int64_t *w, hipStream_t st)
{
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_widen32, dim3((unsigned)blocks), dim3(256), 0, st, n, a, w);
    MCPI_HIP(hipGetLastError());
    return MCPI_OK;
}

mcpi_status basis_build_parts(mcpi_basis b, int want_map, int want_ot, int want_tt)
{
    MCPI_ARG(b, "null basis");
    return build_maps_parts(b, want_map != 0, want_ot != 0, want_tt != 0);
}

}  // namespace mcpi

// streaming copy, 16 bytes per lane (also a kernel of exactly known traffic)
__global__ __launch_bounds__(256) void k_copy16(int64_t n16, const double2 *__restrict__ src,
                                                double2 *__restrict__ dst)
{
    for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < n16; x += (int64_t)gridDim.x * 256)
        dst[x] = src[x];
}

extern "C" {

mcpi_status mcpi_hip_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    MCPI_ARG(dst && src, "null pointer");
    const uintptr_t a = (uintptr_t)dst | (uintptr_t)src | (uintptr_t)bytes;
    if (bytes && (a & 15) == 0) {
        int64_t n16 = bytes / 16;
        int64_t grid = std::min<int64_t>((n16 + 255) / 256, 8192);
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, 0, n16, (const double2 *)src,
                           (double2 *)dst);
        MCPI_HIP(hipGetLastError());
        MCPI_HIP(hipDeviceSynchronize());
    } else if (bytes) {
        MCPI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    }
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_create(int N, int M, int maps, mcpi_basis *out)
{
    MCPI_ARG(out, "null output handle");
    MCPI_ARG(maps >= MCPI_MAPS_NONE && maps <= MCPI_MAPS_FULL, "maps must be 0, 1 or 2");
    mcpi_basis b = nullptr;
    MCPI_CHECK(basis_init(N, M, nullptr, nullptr, &b));
    mcpi_status st = basis_build_maps(b, maps);
    if (st != MCPI_OK) {
        free_basis(b);
        return st;
    }
    *out = b;
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_create_sectored(int N, int M, int maps, const int32_t *charge, mcpi_basis *out)
{
    MCPI_ARG(out && charge, "null argument");
    MCPI_ARG(maps == MCPI_MAPS_SINGLE,
             "a sector-ordered basis stores the single-jump map only (maps must be MCPI_MAPS_SINGLE)");
    MCPI_ARG(M >= 1 && M <= MCPI_MAXM, "M out of range (1..32)");
    MCPI_ARG(has_fast_kernel(N, M), "a sector-ordered basis needs 3 <= M <= 13 (its H.v kernel is specialised on M)");
    int64_t span = 0;
    for (int j = 0; j < M; j++) {
        MCPI_ARG(charge[j] >= -65536 && charge[j] <= 65536, "charges must lie in [-65536, 65536]");
        span = std::max<int64_t>(span, std::abs((int64_t)charge[j]));
    }
    MCPI_ARG(span * N <= INT32_MAX / 2, "charge times N out of int32 range");
    mcpi_basis b = nullptr;
    MCPI_CHECK(basis_init(N, M, nullptr, nullptr, &b));
    mcpi_status st = make_sectored(b, charge);
    if (st == MCPI_OK) st = basis_build_maps(b, maps);
    if (st != MCPI_OK) {
        free_basis(b);
        return st;
    }
    *out = b;
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_sectors(mcpi_basis b, int64_t *nsectors, int64_t *offsets)
{
    MCPI_ARG(b && nsectors, "null argument");
    if (!b->sectored()) {
        *nsectors = 1;
        if (offsets) { offsets[0] = 0; offsets[1] = b->d.nc; }
        return MCPI_OK;
    }
    *nsectors = (int64_t)b->sec_off.size() - 1;
    if (offsets) std::copy(b->sec_off.begin(), b->sec_off.end(), offsets);
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_reorder(mcpi_basis b, int to_rows, const mcpi_cplx *in_dev, mcpi_cplx *out_dev,
                                   void *stream)
{
    MCPI_ARG(b && in_dev && out_dev, "null argument");
    MCPI_ARG((const void *)in_dev != (const void *)out_dev, "input and output must not alias");
    const hipStream_t st = (hipStream_t)stream;
    if (!b->sectored()) {
        MCPI_HIP(hipMemcpyAsync(out_dev, in_dev, sizeof(cplx) * b->d.nc, hipMemcpyDeviceToDevice, st));
        return MCPI_OK;
    }
    if (to_rows) return gather_rows(b->d.nc, b->iperm, (const cplx *)in_dev, (cplx *)out_dev, st);
    return scatter_rows(b->d.nc, b->iperm, (const cplx *)in_dev, (cplx *)out_dev, st);
}

mcpi_status mcpi_hip_basis_destroy(mcpi_basis b)
{
    if (b) {
        (void)hipDeviceSynchronize();
        free_basis(b);
    }
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_info(mcpi_basis b, mcpi_basis_info *info)
{
    MCPI_ARG(b && info, "null argument");
    info->N = b->d.N;
    info->M = b->d.M;
    info->maps = b->maps;
    info->nc = b->d.nc;
    info->map_len = b->map_len;
    info->ot_len = b->ot_len;
    info->tt_len = b->tt_len;
    info->device_bytes = b->device_bytes;
    info->pot_len = b->pot_len;
    info->ptt_len = b->ptt_len;
    info->pattern_bytes = 0;
    for (const PatternTables &t : b->patterns) info->pattern_bytes += t.bytes;
    info->nsectors = b->sectored() ? (int64_t)b->sec_off.size() - 1 : 0;
    info->largest_sector = 0;
    for (size_t x = 0; x + 1 < b->sec_off.size(); x++)
        info->largest_sector = std::max(info->largest_sector, b->sec_off[x + 1] - b->sec_off[x]);
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_copy_ncmat(mcpi_basis b, int32_t *NCmat)
{
    MCPI_ARG(b && NCmat, "null argument");
    MCPI_HIP(hipMemcpy(NCmat, b->d.NCmat, sizeof(int32_t) * (b->d.N + 1) * (b->d.M + 1),
                       hipMemcpyDeviceToHost));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_copy_if(mcpi_basis b, int32_t *IF)
{
    MCPI_ARG(b && IF, "null argument");
    const int32_t *src = b->sectored() ? b->IF_ref : b->d.IF;   // reference order
    MCPI_HIP(hipMemcpy(IF, src, sizeof(int32_t) * b->d.nc * b->d.M, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_copy_map(mcpi_basis b, int32_t *Map)
{
    MCPI_ARG(b && Map, "null argument");
    MCPI_ARG(b->d.Map, "basis has no single-jump map");
    MCPI_ARG(!b->sectored(), "the single-jump map of a sector-ordered basis is in row order: "
                             "copy it with mcpi_hip_basis_copy_range (which 0)");
    MCPI_HIP(hipMemcpy(Map, b->d.Map, sizeof(int32_t) * b->map_len, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_copy_ot(mcpi_basis b, int64_t *strideOT, int32_t *MapOT)
{
    MCPI_ARG(b, "null basis");
    MCPI_ARG(b->d.MapOT, "basis has no double-jump maps");
    if (strideOT)
        MCPI_HIP(hipMemcpy(strideOT, b->d.strideOT, sizeof(int64_t) * b->d.nc, hipMemcpyDeviceToHost));
    if (MapOT && b->ot_len)
        MCPI_HIP(hipMemcpy(MapOT, b->d.MapOT, sizeof(int32_t) * b->ot_len, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_copy_tt(mcpi_basis b, int64_t *strideTT, int32_t *MapTT)
{
    MCPI_ARG(b, "null basis");
    MCPI_ARG(b->d.MapTT, "basis has no double-jump maps");
    if (strideTT)
        MCPI_HIP(hipMemcpy(strideTT, b->d.strideTT, sizeof(int64_t) * b->d.nc, hipMemcpyDeviceToHost));
    if (MapTT && b->tt_len)
        MCPI_HIP(hipMemcpy(MapTT, b->d.MapTT, sizeof(int32_t) * b->tt_len, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_copy_range(mcpi_basis b, int which, int64_t offset, int64_t count, void *dst)
{
    MCPI_ARG(b && dst && offset >= 0 && count >= 0, "bad argument");
    const void *src = nullptr;
    int64_t len = 0, esize = 4;
    switch (which) {
    case 0: src = b->d.Map; len = b->map_len; break;
    case 1: src = b->d.MapOT; len = b->ot_len; break;
    case 2: src = b->d.MapTT; len = b->tt_len; break;
    case 3: src = b->d.IF; len = b->d.nc * b->d.M; break;
    case 4: src = b->d.strideOT; len = b->d.strideOT ? b->d.nc + 1 : 0; esize = 8; break;
    case 5: src = b->d.strideTT; len = b->d.strideTT ? b->d.nc + 1 : 0; esize = 8; break;
    case 6: src = b->d.packOT; len = b->pot_len; break;
    case 7: src = b->d.packTT; len = b->ptt_len; break;
    case 8: src = b->d.poffOT; len = b->d.poffOT ? (b->d.nc + 63) / 64 + 1 : 0; esize = 8; break;
    case 9: src = b->d.poffTT; len = b->d.poffTT ? (b->d.nc + 63) / 64 + 1 : 0; esize = 8; break;
    case 10: src = b->d.perm; len = b->d.perm ? b->d.nc : 0; break;
    case 11: src = b->iperm; len = b->iperm ? b->d.nc : 0; break;
    case 12: src = b->IF_ref; len = b->IF_ref ? b->d.nc * b->d.M : 0; break;
    default: MCPI_ARG(false, "which must be 0..12");
    }
    MCPI_ARG(src, "table not stored in this basis");
    MCPI_ARG(offset + count <= len, "range beyond the table");
    if (count)
        MCPI_HIP(hipMemcpy(dst, (const char *)src + offset * esize, count * esize, hipMemcpyDeviceToHost));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_rank(mcpi_basis b, int64_t n, const int32_t *v, int64_t *idx)
{
    MCPI_ARG(b && v && idx && n >= 0, "bad argument");
    if (n == 0) return MCPI_OK;
    const int M = b->d.M, N = b->d.N;
    for (int64_t t = 0; t < n; t++) {
        int64_t s = 0;
        for (int j = 0; j < M; j++) {
            MCPI_ARG(v[t * M + j] >= 0, "negative occupation");
            s += v[t * M + j];
        }
        MCPI_ARG(s == N, "occupation vector does not hold N particles");
    }
    int32_t *dv = nullptr;
    int64_t *di = nullptr;
    MCPI_HIP(hipMalloc(&dv, n * M * 4));
    MCPI_HIP(hipMalloc(&di, n * 8));
    MCPI_HIP(hipMemcpy(dv, v, n * M * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_rank_states, dim3((n + 255) / 256), dim3(256), 0, 0, N, M, n, b->d.NCmat, dv, di);
    MCPI_HIP(hipGetLastError());
    MCPI_HIP(hipMemcpy(idx, di, n * 8, hipMemcpyDeviceToHost));
    MCPI_HIP(hipFree(dv));
    MCPI_HIP(hipFree(di));
    return MCPI_OK;
}

mcpi_status mcpi_hip_basis_unrank(mcpi_basis b, int64_t n, const int64_t *idx, int32_t *v)
{
    MCPI_ARG(b && v && idx && n >= 0, "bad argument");
    if (n == 0) return MCPI_OK;
    for (int64_t t = 0; t < n; t++) MCPI_ARG(idx[t] >= 0 && idx[t] < b->d.nc, "index out of range");
    const int M = b->d.M, N = b->d.N;
    int32_t *dv = nullptr;
    int64_t *di = nullptr;
    MCPI_HIP(hipMalloc(&dv, n * M * 4));
    MCPI_HIP(hipMalloc(&di, n * 8));
    MCPI_HIP(hipMemcpy(di, idx, n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_unrank_if, dim3((n + 255) / 256), dim3(256), 0, 0, N, M, n, b->d.NCmat,
                       (const int64_t *)di, dv);
    MCPI_HIP(hipGetLastError());
    MCPI_HIP(hipMemcpy(v, dv, n * M * 4, hipMemcpyDeviceToHost));
    MCPI_HIP(hipFree(dv));
    MCPI_HIP(hipFree(di));
    return MCPI_OK;
}

}  // extern "C"
