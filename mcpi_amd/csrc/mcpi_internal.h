// Internal definitions shared by the HIP sources of libmcpi_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>
#include <stdint.h>

#include <string>

#include "../../include/mcpi_hip.h"

// ---------------------------------------------------------------- limits --
// Orbitals and particles a basis may have.  The per-row occupation of a
// state is held in 8-bit lanes, and tables indexed by orbital are sized for
// MCPI_MAXM.
#define MCPI_MAXM 32
#define MCPI_MAXN 120

// ----------------------------------------------------------- error state --
namespace mcpi {

void set_error(const std::string &msg);
mcpi_status hip_fail(hipError_t e, const char *what, const char *file, int line);

}  // namespace mcpi

#define MCPI_HIP(call)                                                        \
    do {                                                                      \
        hipError_t _e = (call);                                               \
        if (_e != hipSuccess) return mcpi::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define MCPI_CHECK(call)                                                      \
    do {                                                                      \
        mcpi_status _s = (call);                                              \
        if (_s != MCPI_OK) return _s;                                         \
    } while (0)

#define MCPI_ARG(cond, msg)                                                   \
    do {                                                                      \
        if (!(cond)) { mcpi::set_error(msg); return MCPI_ERR_ARG; }           \
    } while (0)

// --------------------------------------------------------------- complex --
struct cplx {
    double re, im;
};

__host__ __device__ inline cplx cmake(double r, double i) { cplx z; z.re = r; z.im = i; return z; }
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return cmake(a.re + b.re, a.im + b.im); }
__host__ __device__ inline cplx cscale(cplx a, double s) { return cmake(a.re * s, a.im * s); }
// acc += a * b
__host__ __device__ inline cplx cfma(cplx a, cplx b, cplx acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
    return acc;
}
// acc += conj(a) * b
__host__ __device__ inline cplx cfma_conj(cplx a, cplx b, cplx acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
    return acc;
}

namespace mcpi {

// --------------------------------------------------------- combinatorics --
// Exact C(n, r) in int64 (-1 on overflow).
int64_t binom(int64_t n, int64_t r);
// NC(N, M) = C(N+M-1, N), the reference's NC (configurationsMap.h:145-202)
// without the INT32 guard; NC(N, 0) = 1 like the reference.
int64_t nc_of(int N, int M);

// Device-side description of a basis: everything a kernel needs to rank,
// unrank and walk the maps.  Lives in device memory (basis.hip).
//   rank(v) = nc - 1 - sum_{j=1}^{M-1} B[j][P_j],  P_j = sum_{o<j} v_o,
//   B[j][p] = C(p + j - 1, j)
// (closed form of the FockToIndex loop, configurationsMap.h:273-305).
struct BasisDev {
    int N, M;
    int64_t nc;
    int32_t *IF;        // nc*M
    int32_t *Map;       // nc*M*M or null
    int32_t *MapOT;     // or null
    int32_t *MapTT;     // or null
    int64_t *strideOT;  // nc+1 entries (last = |MapOT|) or null
    int64_t *strideTT;  // nc+1 entries (last = |MapTT|) or null
    int32_t *B;         // (M) x (N+1) table B[j*(N+1) + p], row 0 unused
    int32_t *NCmat;     // (N+1)*(M+1), reference layout
    // Wave-packed double-jump targets (full maps only; see basis.hip): for
    // each wavefront of 64 consecutive rows w, the slots of the jumps active
    // in any of its lanes, from int offset poffOT[w] / poffTT[w]
    int32_t *packOT, *packTT;
    int64_t *poffOT, *poffTT;   // nw + 1 entries, nw = ceil(nc / 64)
    // Pattern tables of one zero pattern (set only in the copy a launch
    // passes): per wavefront and jump, entry-major slots holding only the
    // targets of the listed (non-zero) terms of each jump, in list order
    int32_t *patOT, *patTT;
    int64_t *ppoffOT, *ppoffTT;
    // Sector order (null: the reference's order).  perm[i] = row of the
    // state with reference index i; IF, Map and the pattern tables are in
    // row order and hold row indices (builders apply perm to every rank).
    const int32_t *perm;
};

// pattern tables of a basis for one zero pattern (see basis.hip)
struct PatternTables {
    uint64_t zsig = 0;
    int32_t *tabOT = nullptr, *tabTT = nullptr;
    int64_t *offOT = nullptr, *offTT = nullptr;
    int64_t lenOT = 0, lenTT = 0;
    int64_t bytes = 0;
    uint64_t last_use = 0;
};

// Entries of a packed slot.  A double jump removes particles at k and s
// (s == k for a same-orbital jump) and adds them at q <= l; entries with q or
// l in {k, s} are no rule term (their folded coefficients are zero), so a
// slot holds only the others.  With the removed orbitals dropped, the
// remaining MP = M - 1 (same-orbital) or M - 2 (two-orbital) orbitals are
// renumbered o' = o - [k < o] - [s < o] (s counted once when s == k), and
// entry (q, l) sits at t = l'(l'+1)/2 + q' of the triangle over MP
// orbitals; the slot is padded to a multiple of four (padding holds the
// row's own index).
__host__ __device__ constexpr int packed_orbitals(int M, bool tt) { return tt ? (M >= 2 ? M - 2 : 0) : (M >= 1 ? M - 1 : 0); }
__host__ __device__ constexpr int packed_real(int M, bool tt)
{
    return packed_orbitals(M, tt) * (packed_orbitals(M, tt) + 1) / 2;
}
__host__ __device__ constexpr int packed_entries(int M, bool tt) { return (packed_real(M, tt) + 3) / 4 * 4; }
// position of entry (q, l), q <= l, neither in {k, s}, in a slot of jump (k, s)
__host__ __device__ inline int packed_pos(int k, int s, int q, int l)
{
    const int qq = q - (k < q) - (s != k && s < q), ll = l - (k < l) - (s != k && s < l);
    return ll * (ll + 1) / 2 + qq;
}
// orbital of renumbered index o' (inverse of the renumbering above)
__host__ __device__ inline int packed_orbital(int k, int s, int op)
{
    int o = op + (op >= k);
    if (s != k) o += (o >= s);
    return o;
}

}  // namespace mcpi

// internal basis/hamiltonian handles
struct mcpi_basis_s {
    mcpi::BasisDev d;
    int maps;
    int64_t ot_len, tt_len, map_len;
    int64_t pot_len, ptt_len;   // ints in packOT / packTT
    int64_t device_bytes;
    std::vector<mcpi::PatternTables> patterns;   // built on demand, a few zero patterns
    uint64_t use_clock = 0;
    // density-matrix scratch (block partials and their sums), grown on
    // demand and kept; one density call at a time per basis uses it
    void *rho_ws = nullptr;
    size_t rho_ws_bytes = 0;
    std::mutex rho_mu;
    // pair-space H.v workspaces (schedule 6: the pair amplitudes and the Y
    // tables of the bases with N-1 and N-2 particles; schedule 7: the Z
    // rows), per stream that ran a pair-space schedule (so that H.v calls
    // on different streams never share one), allocated on first use, a
    // larger one beside a smaller one when the schedule needs it, and kept
    // until the basis is freed
    struct PairWs {
        hipStream_t st;
        void *p;
    };
    std::vector<PairWs> pair_ws;
    std::mutex pair_mu;   // guards pair_ws
    // Sector order (bases created with a conserved charge): rows sorted by
    // the charge sum_j q_j v_j, reference order kept inside a sector.
    // d.perm (device) maps reference index -> row, iperm row -> reference
    // index; IF_ref is IF in the reference order; sec_off[s] .. sec_off[s+1]
    // are the rows of the s-th sector (charges ascending); vt: two vectors
    // of nc entries for moving vectors between the orders (allocated on
    // first use).
    int32_t *iperm = nullptr;
    int32_t *IF_ref = nullptr;
    std::vector<int32_t> charge;
    std::vector<int64_t> sec_off;
    cplx *vt[2] = {nullptr, nullptr};
    bool sectored() const { return d.perm != nullptr; }
};

// Folded coefficient tables of a Hamiltonian (device memory), produced on
// the host from Ho and Hint by the rule folding of hamiltonianMatrix.h:
// 1494-1963.  For a row with occupations v (r1[q] = sqrt(v_q+1),
// r2[q] = sqrt((v_q+1)(v_q+2))):
//   diag      = sum_k dk[k] v_k + sum_k d1[k] v_k(v_k-1) + sum_{k<s} d2[k,s] v_k v_s
//   1-jump k->l (v_k>=1, l!=k): sqrt(v_k) r1[l] (A0[k,l] + sum_s v_s A[k,l,s])
//   OT (k; q<=l, q,l != k, v_k>=2): sqrt(v_k(v_k-1)) * ot[k][q+l*M] * (q==l ? r2[q] : r1[q] r1[l])
//   TT (k<s; q<=l not in {k,s}):  sqrt(v_k v_s) * tt[p][q+l*M] * (same)
struct HamCoefDev {
    cplx *dk;    // M
    cplx *d1;    // M
    cplx *d2;    // M*M  (k + s*M, k < s)
    cplx *A0;    // M*M  (k + l*M)
    cplx *A;     // M*M*M  ((k + l*M)*M + s)
    cplx *ot;    // M * M*M
    cplx *tt;    // P * M*M, P = M(M-1)/2 pairs in lexicographic order
    cplx *otc;   // M * packed_entries(M, false): ot in packed-slot order
    cplx *ttc;   // P * packed_entries(M, true): tt in packed-slot order
    uint8_t *nz_1j;  // M*M: 1 if any coefficient of the jump k->l is non-zero
    uint8_t *nz_ot;  // M*M*M: per (k, e)
    uint8_t *nz_tt;  // P*M*M: per (p, e)
    // zero-skipping support for the specialised kernels
    uint32_t *m1j;   // M: bit l set if the jump k->l has a non-zero coefficient
    int32_t *ot_off; // M+1: non-zero list of OT chunk k is nz_e/nz_c[ot_off[k] .. ot_off[k+1])
    int32_t *tt_off; // P+1: same for TT chunk p
    int32_t *nz_e;   // entry index q + l*M of each listed term
    cplx *nz_c;      // its folded coefficient
    // The same operator through the pair space (schedule 6, 3 <= M <= 12;
    // null otherwise): H = sum_{k,l} h1[k][l] a*_k a_l
    //                    + sum_{P,Q} pw[P][Q] a*_{P1} a*_{P2} a_{Q1} a_{Q2}
    // over unordered orbital pairs P, Q (index p = hi(hi+1)/2 + lo), stored
    // pw[Q * PWS + P] and ph1[l * HS + k], HS = pair_h1_stride(M) (column of an annihilated pair or
    // orbital contiguous, PWS = M(M+1)/2 rounded up to 16, zero padded)
    cplx *pw;
    cplx *ph1;
    // factorized pair form (schedule 7; null / 0 unless h1 is diagonal and
    // pw is block diagonal with rank-one blocks): pw[P][Q] = pza[P] pzb[Q]
    // for P, Q in one block; pzblk[P] = block of P (-1: none); pzord lists
    // the pairs of block r at pzoff[r] .. pzoff[r+1]; pzR blocks
    cplx *pza;
    cplx *pzb;
    int32_t *pzblk;
    int32_t *pzord;
    int32_t *pzoff;
    int pzR;
    int real;        // 1 if every coefficient above has a zero imaginary part
    int dense;       // 1 if at least half of the double-jump terms are non-zero:
                     // walking the lists would cost more than it skips
    uint64_t zsig;   // hash of the zero pattern (m1j, ot_off, tt_off, nz_e):
                     // equal for Hamiltonians whose non-zero terms coincide
};

// the same tables d bytes further on (instance x of a batch, or a copy of
// the block elsewhere)
__host__ __device__ inline HamCoefDev shift_coef(HamCoefDev c, ptrdiff_t d)
{
    auto sh = [d](auto *p) { return (decltype(p))((char *)p + d); };
    c.dk = sh(c.dk);
    c.d1 = sh(c.d1);
    c.d2 = sh(c.d2);
    c.A0 = sh(c.A0);
    c.A = sh(c.A);
    c.ot = sh(c.ot);
    c.tt = sh(c.tt);
    c.otc = sh(c.otc);
    c.ttc = sh(c.ttc);
    c.nz_1j = sh(c.nz_1j);
    c.nz_ot = sh(c.nz_ot);
    c.nz_tt = sh(c.nz_tt);
    c.m1j = sh(c.m1j);
    c.ot_off = sh(c.ot_off);
    c.tt_off = sh(c.tt_off);
    c.nz_e = sh(c.nz_e);
    c.nz_c = sh(c.nz_c);
    if (c.pw) c.pw = sh(c.pw);
    if (c.ph1) c.ph1 = sh(c.ph1);
    if (c.pza) {
        c.pza = sh(c.pza);
        c.pzb = sh(c.pzb);
        c.pzblk = sh(c.pzblk);
        c.pzord = sh(c.pzord);
        c.pzoff = sh(c.pzoff);
    }
    return c;
}

// pair-space tables of schedule 6: M(M+1)/2 pairs, padded to 16
__host__ __device__ constexpr int pair_count(int M) { return M * (M + 1) / 2; }
__host__ __device__ constexpr int pair_stride(int M) { return (pair_count(M) + 15) / 16 * 16; }
// row stride of the one-body table ph1 (h1[k][l] at l * stride + k): 16 up to
// 16 orbitals, 32 beyond
__host__ __device__ constexpr int pair_h1_stride(int M) { return M <= 16 ? 16 : 32; }
// the larger orbital c2 of pair P = c2 (c2 + 1) / 2 + c1 (c1 <= c2)
__host__ __device__ constexpr int pair_hi(int P)
{
    int c2 = 0;
    while ((c2 + 1) * (c2 + 2) / 2 <= P) c2++;
    return c2;
}
// rows of the pair-amplitude matrix X of the pair-space H.v: the pairs,
// rounded up to four (the rows past the pairs stay zero)
__host__ __device__ constexpr int pair_x_rows(int M) { return (pair_count(M) + 3) / 4 * 4; }
// row stride of the pair-space matrices over nu states: nu rounded up to 16
__host__ __device__ constexpr int64_t pair_ld(int64_t nu) { return (nu + 15) / 16 * 16; }
__host__ __device__ inline int pair_of2(int a, int b)
{
    return a < b ? b * (b + 1) / 2 + a : a * (a + 1) / 2 + b;
}
#define MCPI_PAIR_MAXM 12
// the factorized pair form (schedule 7) has no product in LDS: larger M
#define MCPI_PAIRZ_MAXM 24
// its several-member launch (batched groups) covers the M of the
// specialised kernels
#define MCPI_PAIRZ_MULTI_MAXM 13

struct mcpi_ham_s {
    mcpi_basis basis;
    HamCoefDev c;
    void *coef_block;   // single device allocation holding the tables
    size_t coef_bytes;  // its size (the same for every Hamiltonian with M orbitals)
    cplx *Ho_dev;       // M*M flat, k + M*l
    cplx *Hint_dev;     // M^4
    int skip_zero;
    int schedule;
    int tuned[2];       // autotuned schedule per skip_zero value (0 = not tuned)
};

namespace mcpi {

__host__ __device__ inline int pair_index(int k, int s, int M)
{
    // lexicographic index of k<s among M(M-1)/2 pairs
    return k * (2 * M - k - 1) / 2 + (s - k - 1);
}

// kernels launched from host code of other translation units
mcpi_status launch_apply(mcpi_ham h, const cplx *C, cplx *out, hipStream_t st);
int ham_schedule(mcpi_ham h);
bool has_fast_kernel(int N, int M);
// pattern tables of basis b for the zero pattern of c; built when absent if
// may_build and device memory allows; nullptr when not available
const PatternTables *basis_pattern(mcpi_basis b, const HamCoefDev &c, bool may_build, hipStream_t st);
mcpi_status launch_apply_batch(const BasisDev &b, int maps, const HamCoefDev &c0, int skip, int sched, int nbatch,
                               size_t cstride, const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st);
mcpi_status launch_apply_multi(const BasisDev &b, const HamCoefDev &c0, int sched, int nbatch, size_t cstride,
                               const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st);
// schedule 5 kernel (acth_comp.hip); MCPI_ERR_STATE when not applicable
mcpi_status launch_comp(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, hipStream_t st);
// schedules: 1 generic thread-per-row kernel, 2 specialised kernel reading
// every stored map, 3 specialised kernel ranking double jumps on the fly,
// 5 every term with the double jumps of a block compacted into full
// wavefronts (Map and packed tables; 2 when the basis has no packed tables)
mcpi_status launch_apply_raw(const BasisDev &b, int maps, const HamCoefDev &c, int skip_zero,
                             int schedule, const cplx *C, cplx *out, hipStream_t st);
mcpi_status build_coefficients(int M, const cplx *Ho, const cplx *Hint, HamCoefDev *c, void **block,
                               size_t *block_bytes = nullptr);
mcpi_status basis_upload_raw(int N, int M, const int32_t *IF, const int32_t *NCmat,
                             const int32_t *Map, const int32_t *MapOT, const int32_t *MapTT,
                             const int *strideOT, const int *strideTT, int64_t ot_len,
                             int64_t tt_len, mcpi_basis *out);
// out[r] = in[idx[r]] for r < n (gather through an index table), on stream st
mcpi_status gather_rows(int64_t n, const int32_t *idx, const cplx *in, cplx *out, hipStream_t st);
// out[idx[r]] = in[r] for r < n (scatter through an index table)
mcpi_status scatter_rows(int64_t n, const int32_t *idx, const cplx *in, cplx *out, hipStream_t st);
// the basis's two order-conversion vectors (allocated on first use)
mcpi_status basis_vectors(mcpi_basis b, cplx **v0, cplx **v1);
// rank table B and NCmat of (N, M) in device memory, for kernels that run on
// caller-owned tables (free both with hipFree)
mcpi_status rank_tables_dev(int N, int M, int32_t **B, int32_t **NCmat);
// w[0 .. n) = a[0 .. n) widened to int64, on stream st
mcpi_status widen_strides(int64_t n, const int32_t *a, int64_t *w, hipStream_t st);
// schedule 6 (acth_pair.hip): the pair-space H.v applies (3 <= M <= 12,
// reference row order, nc < 2^31); workspace entries; launch into the
// workspace ws of pair_workspace_entries(N, M) complex entries
bool pair_applicable(const BasisDev &b, const HamCoefDev &c);
size_t pair_workspace_entries(int N, int M);
mcpi_status launch_pair(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, cplx *ws,
                        hipStream_t st);
// schedule 7 (acth_pair.hip): the factorized pair form applies (as 6, and
// the Hamiltonian's pair matrix factorizes, N >= 2); launch into a
// workspace of (blocks) x pair_ld(NC(N-2, M)) complex entries
bool pairz_applicable(const BasisDev &b, const HamCoefDev &c);
mcpi_status launch_pairz(const BasisDev &b, const HamCoefDev &c, const cplx *C, cplx *out, cplx *ws,
                         hipStream_t st);
// the same for chunks x mv (mv 2 or 4) Hamiltonians of one basis whose
// pair matrices share the blocks and b factors (M <=
// MCPI_PAIRZ_MULTI_MAXM), mv members per thread: member m's vectors at
// C + m mC and out + m mC, its coefficient block mK bytes after member
// m - 1's (member 0's tables in c); a workspace of chunks x mv x (blocks)
// x pair_ld(NC(N-2, M)) complex entries
mcpi_status launch_pairz_multi(const BasisDev &b, const HamCoefDev &c, int mv, int chunks, const cplx *C, cplx *out,
                               int64_t mC, int64_t mK, cplx *ws, hipStream_t st);
// build the requested maps of a basis whose IF table is in place
mcpi_status basis_build_parts(mcpi_basis b, int want_map, int want_ot, int want_tt);

}  // namespace mcpi
