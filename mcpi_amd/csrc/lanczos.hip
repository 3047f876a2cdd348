// Lanczos ground states on the device, one instance or many in lockstep.
//
// Reference: lanczos, groundStateLieb.c:217-344 (three-term recurrence with
// one pass of classical Gram-Schmidt against every previous vector),
// carrDot / carrMod (73-101) and ground (348-471: tridiagonal eigenproblem on
// the host, Ritz vector back-transform 454-458).
//
// K instances run together.  Each has its own Hamiltonian, vectors and
// scalars (LzInst descriptor, LanczosState); instances that share a basis may
// differ in Ho/Hint.  Every kernel below serves all K instances in one launch
// (instance = blockIdx.y, blockIdx.z for the Gram-Schmidt dots), so one
// iteration is the same short chain of launches for any K:
//   1 fin_beta      beta_i = ||HC|| from block partials, breakdown test
//   2 scale         v_{i+1} = HC / beta_i  (into lvec[i+1] and the H.v input)
//   3 H.v           HC = H v_{i+1}: one launch per group of instances that
//                   share a basis (the group's row blocks are interleaved so
//                   its instances read each table row from one XCD's L2)
//   4 sub_dot       HC -= beta_i v_i;  partial <v_{i+1}, HC>
//   5 fin_alpha     alpha_{i+1}
//   6 sub_alpha     HC -= alpha v_{i+1}
//     gemv_dots     partial o_t = <v_t, HC>, t <= i+1
//   7 fin_ortho     o_t
//   8 reorth_norm   HC -= sum_t v_t o_t;  partial ||HC||^2 for the next step
//   9 advance       i += 1
// Every scalar lives in device memory and the pointers never change, so the
// chain is captured once in a hipGraph and replayed lm-1 times with no host
// round trip.  An instance that breaks down stops updating (its kernels
// return at once) while the others go on.  Reductions are two-stage (block
// partials, then one block per instance) and deterministic.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "mcpi_internal.h"

namespace mcpi {

struct LanczosState {
    int it;        // current iteration index i
    int broke;     // 1 once a breakdown was detected
    int steps;     // size of the tridiagonal built so far
    int lm;
    double maxb;
    double tol;
    double beta;   // beta_i of the current iteration
    cplx alpha;    // alpha of the current iteration
};

// one instance: sizes and device buffers
struct LzInst {
    int64_t n;            // basis size nc
    const cplx *C0;       // start vector
    cplx *lvec;           // lm * n, row t = v_t
    cplx *HC, *vin;       // H.v output and input
    cplx *diag, *offd;    // lm each (real parts used)
    cplx *ortho;          // lm Gram-Schmidt coefficients
    double *partial;      // 2 * nb block partials
    double *partial_o;    // 2 * lm * nbg dot partials, t-major
    const double *z;      // lm Ritz coefficients (after the host solve)
    cplx *Cout;           // Ritz vector (may be null)
    cplx *rpart;          // small bases: per-vector-group partial updates of
                          // the Gram-Schmidt pass (ceil(lm / LZ_TG) x n), null
                          // for the single-pass update
    const int32_t *iperm; // sector-ordered basis: reference index of each row
                          // (the start vector is gathered through it and the
                          // Ritz vector scattered); null otherwise
    int nb, nbg;          // vector and dot blocks used by this instance
    int steps;            // Ritz length (after the host solve)
};

#define LZ_T 256

__device__ inline double lz_wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// block sum of (a, b) into partial2[2 * blockIdx.x .. +1]
__device__ inline void block_sum2(double a, double b, double *partial2)
{
    __shared__ double red[2][LZ_T / 64];
    a = lz_wave_sum(a);
    b = lz_wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = a; red[1][w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int x = 0; x < LZ_T / 64; x++) { sa += red[0][x]; sb += red[1][x]; }
        partial2[2 * blockIdx.x] = sa;
        partial2[2 * blockIdx.x + 1] = sb;
    }
}

// The descriptors hold generic pointers; vector accesses go through global
// (address space 1) pointers and the per-step coefficients through constant
// (address space 4) ones, so that the compiler emits global and scalar loads
// instead of flat ones.
#define LZ_GLOBAL __attribute__((address_space(1)))
#define LZ_CONST __attribute__((address_space(4)))
template <class T>
__device__ inline LZ_GLOBAL T *gp(T *p)
{
    return (LZ_GLOBAL T *)p;
}
template <class T>
__device__ inline const LZ_CONST T *kp(const T *p)
{
    return (const LZ_CONST T *)p;
}

typedef double lz_d2 __attribute__((ext_vector_type(2)));   // one 16-byte access

__device__ inline cplx ldv(const LZ_GLOBAL cplx *p)
{
    lz_d2 t = *(const LZ_GLOBAL lz_d2 *)p;
    return cmake(t.x, t.y);
}
// streamed basis vectors (read once per pass): non-temporal
__device__ inline cplx ldv_nt(const LZ_GLOBAL cplx *p)
{
    const lz_d2 t = __builtin_nontemporal_load((const LZ_GLOBAL lz_d2 *)p);
    return cmake(t.x, t.y);
}
__device__ inline void stv(LZ_GLOBAL cplx *p, cplx z)
{
    lz_d2 t;
    t.x = z.re;
    t.y = z.im;
    *(LZ_GLOBAL lz_d2 *)p = t;
}

// grid-stride loop over this instance's rows; blocks beyond its nb return
// before any barrier (the test is uniform over the block)
#define LZ_ROWS(I, j) \
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < (I).n; j += (int64_t)(I).nb * LZ_T)

// ---- set-up kernels (before the first iteration) --

// state, and diag/offdiag zeroed (grid (1, K))
__global__ void lz_init(LanczosState *sts, const LzInst *in, int lm, double tol)
{
    LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    for (int x = threadIdx.x; x < lm; x += blockDim.x) {
        I.diag[x] = cmake(0, 0);
        I.offd[x] = cmake(0, 0);
    }
    if (threadIdx.x == 0) {
        st->it = 0;
        st->broke = 0;
        st->steps = lm;
        st->lm = lm;
        st->maxb = 0;
        st->tol = tol;
        st->beta = 0;
        st->alpha = cmake(0, 0);
    }
}

// v_0 = C0 into lvec[0] and the H.v input (C0 may alias lvec)
__global__ __launch_bounds__(LZ_T) void lz_start(const LzInst *in)
{
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb) return;
    LZ_ROWS(I, j)
    {
        cplx x = ldv(gp(I.C0) + (I.iperm ? (int64_t)I.iperm[j] : j));
        stv(gp(I.lvec) + j, x);
        stv(gp(I.vin) + j, x);
    }
}

// partial <vin, HC>
__global__ __launch_bounds__(LZ_T) void lz_dot(const LzInst *in)
{
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb) return;
    double re = 0, im = 0;
    LZ_ROWS(I, j)
    {
        cplx a = ldv(gp(I.vin) + j), b = ldv(gp(I.HC) + j);
        re += a.re * b.re + a.im * b.im;
        im += a.re * b.im - a.im * b.re;
    }
    block_sum2(re, im, I.partial);
}

// alpha from the block partials, into diag[0] (first) or diag[it + 1]
// (grid (1, K))
__global__ void lz_fin_alpha(LanczosState *sts, const LzInst *in, int first)
{
    __shared__ double red[2][LZ_T];
    LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if (st->broke) return;
    double re = 0, im = 0;
    for (int x = threadIdx.x; x < I.nb; x += LZ_T) { re += I.partial[2 * x]; im += I.partial[2 * x + 1]; }
    red[0][threadIdx.x] = re;
    red[1][threadIdx.x] = im;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int x = 0; x < LZ_T; x++) { a += red[0][x]; b += red[1][x]; }
        st->alpha = cmake(a, b);
        I.diag[first ? 0 : st->it + 1] = cmake(a, b);
    }
}

// HC -= alpha v;  partial ||HC||^2   (set-up only: iteration 0 has no
// reorthogonalization)
__global__ __launch_bounds__(LZ_T) void lz_sub_alpha_norm(const LanczosState *sts, const LzInst *in)
{
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb) return;
    const cplx a = sts[blockIdx.y].alpha;
    double s = 0;
    LZ_ROWS(I, j)
    {
        cplx h = ldv(gp(I.HC) + j), x = ldv(gp(I.vin) + j);
        h.re -= a.re * x.re - a.im * x.im;
        h.im -= a.re * x.im + a.im * x.re;
        stv(gp(I.HC) + j, h);
        s += h.re * h.re + h.im * h.im;
    }
    block_sum2(s, 0, I.partial);
}

// ---- iteration kernels --

// 1: beta_i and the breakdown test (groundStateLieb.c:287-292); grid (1, K)
__global__ void lz_fin_beta(LanczosState *sts, const LzInst *in)
{
    __shared__ double red[LZ_T];
    LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if (st->broke) return;
    double s = 0;
    for (int x = threadIdx.x; x < I.nb; x += LZ_T) s += I.partial[2 * x];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int x = 0; x < LZ_T; x++) t += red[x];
        double b = sqrt(t);
        const int i = st->it;
        I.offd[i] = cmake(b, 0);
        if (b > st->maxb) st->maxb = b;
        st->beta = b;
        if (b == 0.0 || b / st->maxb < st->tol) {
            // invariant Krylov space of dimension i+1: keep alpha_0..alpha_i
            st->broke = 1;
            st->steps = i + 1;
        }
    }
}

// 2: v_{i+1} = HC / beta_i, stored in lvec[i+1] and in the H.v input buffer
__global__ __launch_bounds__(LZ_T) void lz_scale(const LanczosState *sts, const LzInst *in)
{
    const LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb || st->broke) return;
    const double inv = 1.0 / st->beta;
    LZ_GLOBAL cplx *dst = gp(I.lvec) + (int64_t)(st->it + 1) * I.n;
    LZ_ROWS(I, j)
    {
        cplx h = ldv(gp(I.HC) + j);
        h.re *= inv;
        h.im *= inv;
        stv(dst + j, h);
        stv(gp(I.vin) + j, h);
    }
}

// 4: HC -= beta_i v_i;  partial <v_{i+1}, HC>
__global__ __launch_bounds__(LZ_T) void lz_sub_dot(const LanczosState *sts, const LzInst *in)
{
    const LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb || st->broke) return;
    const double b = st->beta;
    const LZ_GLOBAL cplx *vi = gp(I.lvec) + (int64_t)st->it * I.n;
    double re = 0, im = 0;
    LZ_ROWS(I, j)
    {
        cplx h = ldv(gp(I.HC) + j), x = ldv(vi + j), y = ldv(gp(I.vin) + j);
        h.re -= b * x.re;
        h.im -= b * x.im;
        stv(gp(I.HC) + j, h);
        re += y.re * h.re + y.im * h.im;
        im += y.re * h.im - y.im * h.re;
    }
    block_sum2(re, im, I.partial);
}

// 6a: HC -= alpha v_{i+1}
__global__ __launch_bounds__(LZ_T) void lz_sub_alpha(const LanczosState *sts, const LzInst *in)
{
    const LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb || st->broke) return;
    const cplx a = st->alpha;
    LZ_ROWS(I, j)
    {
        cplx x = ldv(gp(I.HC) + j), y = ldv(gp(I.vin) + j);
        x.re -= a.re * y.re - a.im * y.im;
        x.im -= a.re * y.im + a.im * y.re;
        stv(gp(I.HC) + j, x);
    }
}

// 6b: partial o_t = <v_t, HC> for t = 0..i+1.  Block (s, g, instance) owns
// row segment s (LZ_R rows per thread, kept in registers) and basis vectors
// t in [g*TG, (g+1)*TG); waves reduce with shuffles and park their sums in
// LDS, one barrier per block.  TG is 16, or 8 for a chain with few row
// segments (dot_group): fewer vectors per block give more blocks at a given
// step.  Over steps 4..121 on MI355X: 646,646 rows 5.9 / 5.5 / 4.5 TB/s
// with 8 / 16 / 32 vectors per block, 18 instances of 43,758 rows
// 5.6 / 5.1 / 4.7, 13.0 M rows 6.3 / 6.4 / 6.4; 352,716 rows 4.8 / 4.3
// with 8 / 16.
#define LZ_R 8
#define LZ_SEG ((int64_t)LZ_T * LZ_R)
// vector groups of the split Gram-Schmidt update (small bases)
#define LZ_TG 32
template <int TG>
__global__ __launch_bounds__(LZ_T) void lz_gemv_dots(const LanczosState *sts, const LzInst *in)
{
    __shared__ double red[TG][LZ_T / 64][2];
    const LanczosState *st = sts + blockIdx.z;
    const LzInst &I = in[blockIdx.z];
    if ((int)blockIdx.x >= I.nbg || st->broke) return;
    const int nt = st->it + 2;
    const int t0 = blockIdx.y * TG;
    if (t0 >= nt) return;
    const int t1 = min(nt, t0 + TG);
    const int64_t n = I.n;
    const int64_t base = (int64_t)blockIdx.x * LZ_SEG + threadIdx.x;
    cplx h[LZ_R];
#pragma unroll
    for (int r = 0; r < LZ_R; r++) {
        int64_t j = base + (int64_t)r * LZ_T;
        h[r] = (j < n) ? ldv(gp(I.HC) + j) : cmake(0, 0);
    }
    const int w = threadIdx.x >> 6;
    for (int tt = t0; tt < t1; tt++) {
        const LZ_GLOBAL cplx *vt = gp(I.lvec) + (int64_t)tt * n;
        double re = 0, im = 0;
        // small chains (TG = 8): rows past n load row n - 1 and meet h = 0,
        // with no branch, so all LZ_R loads are in flight together; large
        // ones keep the branch (the loads then wait one by one, which with
        // many resident waves measured faster: 13.0 M rows, 200 steps)
#pragma unroll
        for (int r = 0; r < LZ_R; r++) {
            const int64_t j = base + (int64_t)r * LZ_T;
            if (TG == 8 || j < n) {
                const cplx x = ldv_nt(vt + (j < n ? j : n - 1));
                re += x.re * h[r].re + x.im * h[r].im;
                im += x.re * h[r].im - x.im * h[r].re;
            }
        }
        re = lz_wave_sum(re);
        im = lz_wave_sum(im);
        if ((threadIdx.x & 63) == 0) {
            red[tt - t0][w][0] = re;
            red[tt - t0][w][1] = im;
        }
    }
    __syncthreads();
    const int tl = threadIdx.x;
    if (tl < t1 - t0) {
        double sa = 0, sb = 0;
        for (int x = 0; x < LZ_T / 64; x++) { sa += red[tl][x][0]; sb += red[tl][x][1]; }
        int64_t at = 2 * ((int64_t)(t0 + tl) * I.nbg + blockIdx.x);
        I.partial_o[at] = sa;
        I.partial_o[at + 1] = sb;
    }
}

// 7: o_t = sum over the nbg dot blocks.  Block (t, instance) owns one basis
// vector; the partials of t are contiguous (t-major), each thread adds a
// strided subset and the block reduces in a fixed order, so the result does
// not depend on scheduling.  Grid (lm, K); blocks with t >= it + 2 exit.
__global__ __launch_bounds__(LZ_T) void lz_fin_ortho(const LanczosState *sts, const LzInst *in)
{
    __shared__ double red[2][LZ_T / 64];
    const LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    const int tt = blockIdx.x;
    if (st->broke || tt >= st->it + 2) return;
    const LZ_GLOBAL double *p = (const LZ_GLOBAL double *)I.partial_o + 2 * (int64_t)tt * I.nbg;
    double re = 0, im = 0;
    for (int x = threadIdx.x; x < I.nbg; x += LZ_T) {
        re += p[2 * x];
        im += p[2 * x + 1];
    }
    re = lz_wave_sum(re);
    im = lz_wave_sum(im);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = re; red[1][w] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int x = 0; x < LZ_T / 64; x++) { sa += red[0][x]; sb += red[1][x]; }
        I.ortho[tt] = cmake(sa, sb);
    }
}

// 8: HC -= sum_t v_t o_t;  partial ||HC||^2 for the next beta.  A thread
// owns RR rows of each block step (rows j + r * LZ_T) and walks the basis
// vectors two at a time.  With RR = 2 every row's load is issued
// unconditionally, so a lane keeps 2 * RR independent 16-byte loads
// outstanding (a branch per row made the compiler wait for each load
// before the next: 4.6 against 6.3 TB/s at 352,716 rows, where few waves
// are resident).  The o_t are the
// same for every lane: they are read as uniform (scalar-cache) loads.  RR
// is 4, or 2 when the chain's instances have too few rows for 4 to fill
// the device (update_rows).
template <int RR>
__global__ __launch_bounds__(LZ_T) void lz_reorth_norm(const LanczosState *sts, const LzInst *in)
{
    const LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb || st->broke || I.rpart) return;
    const int nt = st->it + 2;
    const int64_t n = I.n;
    const LZ_CONST cplx *o = kp(I.ortho);
    const LZ_GLOBAL cplx *V = gp(I.lvec);
    LZ_GLOBAL cplx *HC = gp(I.HC);
    double s = 0;
    for (int64_t base = (int64_t)blockIdx.x * LZ_T * RR + threadIdx.x; base < n;
         base += (int64_t)I.nb * LZ_T * RR) {
        cplx h[RR];
#pragma unroll
        for (int r = 0; r < RR; r++) {
            const int64_t j = base + (int64_t)r * LZ_T;
            h[r] = (j < n) ? ldv(HC + j) : cmake(0, 0);
        }
        // RR = 2 (small chains): rows past n load row n - 1, never stored,
        // with no branch, so the loads of both unrolled vectors are in flight
        // together; RR = 4 keeps the branch (faster with many resident waves)
#pragma unroll 2
        for (int tt = 0; tt < nt; tt++) {
            const cplx ot = cmake(o[tt].re, o[tt].im);
            const LZ_GLOBAL cplx *vt = V + (int64_t)tt * n;
#pragma unroll
            for (int r = 0; r < RR; r++) {
                const int64_t j = base + (int64_t)r * LZ_T;
                if (RR == 2 || j < n) {
                    const cplx x = ldv_nt(vt + (j < n ? j : n - 1));
                    h[r].re -= x.re * ot.re - x.im * ot.im;
                    h[r].im -= x.re * ot.im + x.im * ot.re;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RR; r++) {
            const int64_t j = base + (int64_t)r * LZ_T;
            if (j < n) {
                stv(HC + j, h[r]);
                s += h[r].re * h[r].re + h[r].im * h[r].im;
            }
        }
    }
    block_sum2(s, 0, I.partial);
}

// 8 on small bases.  With few rows, a pass in which every thread walks all
// i+2 basis vectors for its rows has too few threads to keep the memory
// busy (N=5, M=21: 52 blocks for up to 6,641 vectors).  The update is split
// like the dots: block (s, g, instance) sums v_t o_t over the vector group g
// for the rows of segment s into rpart[g] (8a), then one pass subtracts the
// groups' partial sums from HC and forms the norm partials (8b).
__global__ __launch_bounds__(LZ_T) void lz_reorth_part(const LanczosState *sts, const LzInst *in)
{
    const LanczosState *st = sts + blockIdx.z;
    const LzInst &I = in[blockIdx.z];
    if (!I.rpart || (int)blockIdx.x >= I.nbg || st->broke) return;
    const int nt = st->it + 2;
    const int t0 = blockIdx.y * LZ_TG;
    if (t0 >= nt) return;
    const int t1 = min(nt, t0 + LZ_TG);
    const int64_t n = I.n;
    const int64_t base = (int64_t)blockIdx.x * LZ_SEG + threadIdx.x;
    const LZ_CONST cplx *o = kp(I.ortho);
    cplx p[LZ_R];
#pragma unroll
    for (int r = 0; r < LZ_R; r++) p[r] = cmake(0, 0);
    for (int tt = t0; tt < t1; tt++) {
        const cplx ot = cmake(o[tt].re, o[tt].im);
        const LZ_GLOBAL cplx *vt = gp(I.lvec) + (int64_t)tt * n;
#pragma unroll
        for (int r = 0; r < LZ_R; r++) {   // rows past n: row n - 1, never stored
            const int64_t j = base + (int64_t)r * LZ_T;
            const cplx x = ldv_nt(vt + (j < n ? j : n - 1));
            p[r].re += x.re * ot.re - x.im * ot.im;
            p[r].im += x.re * ot.im + x.im * ot.re;
        }
    }
    LZ_GLOBAL cplx *dst = gp(I.rpart) + (int64_t)blockIdx.y * n;
#pragma unroll
    for (int r = 0; r < LZ_R; r++) {
        const int64_t j = base + (int64_t)r * LZ_T;
        if (j < n) stv(dst + j, p[r]);
    }
}

__global__ __launch_bounds__(LZ_T) void lz_reorth_sum(const LanczosState *sts, const LzInst *in)
{
    const LanczosState *st = sts + blockIdx.y;
    const LzInst &I = in[blockIdx.y];
    if (!I.rpart || (int)blockIdx.x >= I.nb || st->broke) return;
    const int ng = (st->it + 2 + LZ_TG - 1) / LZ_TG;
    const int64_t n = I.n;
    double s = 0;
    LZ_ROWS(I, j)
    {
        cplx h = ldv(gp(I.HC) + j);
        for (int g = 0; g < ng; g++) {
            const cplx x = ldv(gp(I.rpart) + (int64_t)g * n + j);
            h.re -= x.re;
            h.im -= x.im;
        }
        stv(gp(I.HC) + j, h);
        s += h.re * h.re + h.im * h.im;
    }
    block_sum2(s, 0, I.partial);
}

// 9: one thread per instance
__global__ void lz_advance(LanczosState *sts, int K)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < K && !sts[x].broke) sts[x].it += 1;
}

// Ritz vector C = sum_k lvec[k] z[k]  (groundStateLieb.c:454-458); z read
// as uniform loads
__global__ __launch_bounds__(LZ_T) void lz_ritz(const LzInst *in)
{
    const LzInst &I = in[blockIdx.y];
    if ((int)blockIdx.x >= I.nb || !I.Cout) return;
    const int steps = I.steps;
    const LZ_CONST double *z = kp(I.z);
    LZ_ROWS(I, j)
    {
        double re = 0, im = 0;
        for (int k = 0; k < steps; k++) {
            cplx x = ldv(gp(I.lvec) + (int64_t)k * I.n + j);
            re += x.re * z[k];
            im += x.im * z[k];
        }
        stv(gp(I.Cout) + (I.iperm ? (int64_t)I.iperm[j] : j), cmake(re, im));
    }
}

// ------------------------------------------- tridiagonal eigenproblem --
// The reference hands (Re alpha, Re beta) to LAPACKE_dstev and keeps the
// lowest eigenpair (groundStateLieb.c:435-458).  Only that pair is needed
// here: the eigenvalue by bisection on Sturm counts, the eigenvector by
// inverse iteration.

// Eigenvalues of the symmetric tridiagonal T = (d, e) below x: by Sylvester's
// law of inertia, the negative pivots of the LDL^T factorization of T - xI
// (Golub & Van Loan, sec. 8.4.1).  e2 holds the squared off-diagonal.  A
// pivot smaller than pivmin in magnitude is taken as -pivmin (the safeguard
// of LAPACK's dstebz), which keeps the count monotone in x.
static int sturm_count(int n, const double *d, const double *e2, double x, double pivmin)
{
    int below = 0;
    double piv = 1.0;
    for (int i = 0; i < n; i++) {
        piv = (d[i] - x) - (i ? e2[i - 1] / piv : 0.0);
        if (fabs(piv) < pivmin) piv = -pivmin;
        below += piv < 0.0;
    }
    return below;
}

// Lowest eigenvalue of T by bisection: count(lo) = 0 and count(hi) >= 1 are
// kept while the bracket halves, until lo and hi are neighbouring doubles.
// Start: a Gershgorin lower bound and the smallest diagonal entry (every
// d_i is a Rayleigh quotient, so lambda_min <= min d), each widened by a
// few ulps so that the two counts hold at the start.
static double lowest_eigenvalue(int n, const double *d, const double *e)
{
    if (n == 1) return d[0];
    std::vector<double> e2(n > 1 ? n - 1 : 1, 0.0);
    double big = 1.0;
    for (int i = 0; i + 1 < n; i++) {
        e2[i] = e[i] * e[i];
        big = std::max(big, e2[i]);
    }
    const double pivmin = 2.2250738585072014e-308 * big;
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; i++) {
        const double radius = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - radius);
        hi = std::min(hi, d[i]);
    }
    const double pad = 4.0 * 2.220446049250313e-16 * std::max(fabs(lo), fabs(hi)) + pivmin;
    lo -= pad;
    hi += pad;
    for (int it = 0; it < 4096; it++) {
        const double mid = lo + 0.5 * (hi - lo);
        if (!(mid > lo && mid < hi)) break;
        if (sturm_count(n, d, e2.data(), mid, pivmin) >= 1) hi = mid;
        else lo = mid;
    }
    return lo + 0.5 * (hi - lo);
}

static void tridiag_inverse_iteration(int n, const double *d, const double *e, double lam, double *z)
{
    // (T - sigma I) x = b with sigma just below lam; Gaussian elimination with
    // partial pivoting on the tridiagonal (three bands after fill-in)
    double scale = 0;
    for (int i = 0; i < n; i++) scale = std::max(scale, fabs(d[i]) + (i < n - 1 ? fabs(e[i]) : 0));
    const double sigma = lam - 1e-13 * (fabs(lam) + scale + 1e-300);
    std::vector<double> a(n), b(n), c(n), c2(n);
    std::vector<int> swp(n, 0);
    for (int i = 0; i < n; i++) z[i] = 1.0 + 1e-3 * ((i * 7919) % 101);
    for (int sweep = 0; sweep < 3; sweep++) {
        // rows: sub a[i] (row i+1, col i), diag b, super c, super2 c2
        for (int i = 0; i < n; i++) {
            b[i] = d[i] - sigma;
            c[i] = (i < n - 1) ? e[i] : 0;
            a[i] = (i < n - 1) ? e[i] : 0;
            c2[i] = 0;
        }
        for (int i = 0; i < n - 1; i++) {
            if (fabs(b[i]) >= fabs(a[i])) {
                swp[i] = 0;
                double f = (b[i] != 0) ? a[i] / b[i] : 0;
                a[i] = f;
                b[i + 1] -= f * c[i];
            } else {
                swp[i] = 1;
                double f = b[i] / a[i];
                b[i] = a[i];
                a[i] = f;
                double t = c[i];
                c[i] = b[i + 1];
                b[i + 1] = t - f * b[i + 1];
                if (i < n - 2) {
                    c2[i] = c[i + 1];
                    c[i + 1] = -f * c[i + 1];
                }
            }
        }
        for (int i = 0; i < n; i++)
            if (b[i] == 0) b[i] = 1e-300;
        for (int i = 0; i < n - 1; i++) {
            if (!swp[i]) z[i + 1] -= a[i] * z[i];
            else {
                double t = z[i];
                z[i] = z[i + 1];
                z[i + 1] = t - a[i] * z[i + 1];
            }
        }
        z[n - 1] /= b[n - 1];
        if (n > 1) z[n - 2] = (z[n - 2] - c[n - 2] * z[n - 1]) / b[n - 2];
        for (int i = n - 3; i >= 0; i--) z[i] = (z[i] - c[i] * z[i + 1] - c2[i] * z[i + 2]) / b[i];
        double nrm = 0;
        for (int i = 0; i < n; i++) nrm += z[i] * z[i];
        nrm = sqrt(nrm);
        for (int i = 0; i < n; i++) z[i] /= nrm;
    }
}

mcpi_status tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z)
{
    MCPI_ARG(n >= 1 && d && lambda, "bad tridiagonal");
    std::vector<double> ee(n, 0.0);
    for (int i = 0; i < n - 1; i++) ee[i] = e ? e[i] : 0.0;
    for (int i = 0; i < n; i++)
        if (!std::isfinite(d[i]) || !std::isfinite(ee[i])) {
            set_error("tridiagonal has a non-finite entry");
            return MCPI_ERR_STATE;
        }
    const double lam = lowest_eigenvalue(n, d, ee.data());
    *lambda = lam;
    if (z) tridiag_inverse_iteration(n, d, ee.data(), lam, z);
    return MCPI_OK;
}

// ---------------------------------------------------------------- host --

// blocks of LZ_T threads with rr rows each that cover n rows (at least one;
// above the cap the vector kernels stride)
static int64_t row_blocks(int64_t n, int rr)
{
    return std::max<int64_t>(1, std::min<int64_t>((n + (int64_t)LZ_T * rr - 1) / ((int64_t)LZ_T * rr), 8192));
}

// the Gram-Schmidt update of an instance with fewer than 256 blocks of 4
// rows per thread is split over vector groups (lz_reorth_part /
// lz_reorth_sum) when there are more vectors than one group
static bool split_update(int64_t n, int lm) { return row_blocks(n, 4) < 256 && lm > LZ_TG; }

// rows per thread of the single-pass update for a chain of instances: 2
// when 4 leave fewer than 1024 blocks in all (4 per CU), else 4.  Measured
// on MI355X over steps 4..121: one instance of 646,646 rows 6.7 TB/s with 2
// (632 blocks of 4 rows: 4.3), 18 of 43,758 rows 6.2 (4.6); one of
// 1,352,078 rows 6.5 with 4 (2 rows, 2,641 blocks: 5.6), 24 of 646,646
// rows 6.0 with 4 (5.9).
static int update_rows(const std::vector<int64_t> &ns)
{
    int64_t b4 = 0;
    for (int64_t n : ns) b4 += row_blocks(n, 4);
    return b4 < 1024 ? 2 : 4;
}

// basis vectors per block of the dot pass for a chain: 8 below 4096 row
// segments in all, else 16
static int dot_group(const std::vector<int64_t> &ns)
{
    int64_t sg = 0;
    for (int64_t n : ns) sg += (n + LZ_SEG - 1) / LZ_SEG;
    return sg < 4096 ? 8 : 16;
}

// One device allocation carved into aligned pieces.
struct Arena {
    char *base = nullptr;
    size_t used = 0, cap = 0;
    bool overflow = false;
    ~Arena()
    {
        if (base) (void)hipFree(base);
    }
    static size_t up(size_t x) { return (x + 255) & ~(size_t)255; }
    void reserve(size_t bytes) { cap += up(bytes); }
    // pieces must be taken in the sizes they were reserved in
    template <class T>
    T *take(size_t count)
    {
        if (used + up(sizeof(T) * count) > cap) {
            overflow = true;
            return nullptr;
        }
        T *p = (T *)(base + used);
        used += up(sizeof(T) * count);
        return p;
    }
    mcpi_status allocate()
    {
        hipError_t e = hipMalloc((void **)&base, cap ? cap : 256);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("cannot allocate the Lanczos workspace (lm * nc complex per instance)");
            return MCPI_ERR_ALLOC;
        }
        return MCPI_OK;
    }
};

// Instances that share a basis: H.v for all of them is one launch of the
// specialised kernel when they agree on zero-skipping and schedule and a
// specialisation exists; otherwise each instance is launched on its own.
// A batched group keeps its members' coefficient blocks and H.v input and
// output vectors side by side at fixed strides (the batched kernel's layout).
struct HvGroup {
    mcpi_basis basis = nullptr;
    std::vector<int> members;
    bool batched = false;
    bool multi_ok = false;        // members share the zero pattern: multi-vector kernel allowed
    bool pzmv_ok = false;         // the factorized pair form several members at a time allowed
    int mode = 1;                 // 1 batched specialised kernel, 2 multi-vector kernel,
                                  // 3 each member on its own with its tuned schedule,
                                  // 4 the factorized pair form four members at a time
    int sched = 2, skip = 0;
    size_t cbytes = 0;            // coefficient block size = stride
    HamCoefDev c0{};              // member 0's tables inside `coefs`
    char *coefs = nullptr;        // members' coefficient blocks, cbytes apart
    cplx *vin = nullptr, *HC = nullptr;   // members' vectors, nc apart
    cplx *zws = nullptr;          // mode 4: Z tables of four members
};

// mode 4 applies to a batched group when every member has the factorized
// pair form with the same blocks and b factors (Lieb-Liniger couplings of
// one basis: only a and the diagonal one-body part differ) and its tables
// at the same offsets of its coefficient block
static bool pairz_multi_ok(const HvGroup &g, const mcpi_ham *hams)
{
    const BasisDev &b = g.basis->d;
    const mcpi_ham h0 = hams[g.members[0]];
    const HamCoefDev &c0 = h0->c;
    if (g.basis->sectored() || b.M > MCPI_PAIRZ_MULTI_MAXM || !pairz_applicable(b, c0) || h0->coef_bytes % 16)
        return false;
    const int NP = pair_count(b.M), R = c0.pzR;
    auto off = [](const mcpi_ham h, const void *p) { return (const char *)p - (const char *)h->coef_block; };
    auto tables = [&](const HamCoefDev &c, std::vector<char> &t) {
        t.resize(sizeof(cplx) * NP + sizeof(int32_t) * (2 * NP + R + 1));
        char *p = t.data();
        if (hipMemcpy(p, c.pzb, sizeof(cplx) * NP, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(p + sizeof(cplx) * NP, c.pzblk, sizeof(int32_t) * NP, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(p + sizeof(cplx) * NP + sizeof(int32_t) * NP, c.pzoff, sizeof(int32_t) * (R + 1),
                      hipMemcpyDeviceToHost) != hipSuccess)
            return false;
        const int32_t *zoff = (const int32_t *)(p + sizeof(cplx) * NP + sizeof(int32_t) * NP);
        const int nord = zoff[R];
        if (nord < 0 || nord > NP) return false;
        return hipMemcpy(p + sizeof(cplx) * NP + sizeof(int32_t) * (NP + R + 1), c.pzord, sizeof(int32_t) * nord,
                         hipMemcpyDeviceToHost) == hipSuccess;
    };
    std::vector<char> t0, t;
    if (!tables(c0, t0)) {
        (void)hipGetLastError();
        return false;
    }
    for (int x : g.members) {
        const mcpi_ham h = hams[x];
        const HamCoefDev &c = h->c;
        if (!pairz_applicable(b, c) || c.pzR != R || h->coef_bytes != h0->coef_bytes ||
            off(h, c.pza) != off(h0, c0.pza) || off(h, c.pzb) != off(h0, c0.pzb) ||
            off(h, c.pzblk) != off(h0, c0.pzblk) || off(h, c.pzord) != off(h0, c0.pzord) ||
            off(h, c.pzoff) != off(h0, c0.pzoff) || off(h, c.ph1) != off(h0, c0.ph1))
            return false;
        if (x == g.members[0]) continue;
        if (!tables(c, t)) {
            (void)hipGetLastError();
            return false;
        }
        if (t != t0) return false;
    }
    return true;
}

// Z tables of every member in a four-member chunk (at least four)
static size_t pairz_multi_bytes(const HvGroup &g, const mcpi_ham *hams)
{
    const BasisDev &b = g.basis->d;
    const size_t zm = std::max<size_t>(4, g.members.size() / 4 * 4);
    return sizeof(cplx) * zm * (size_t)hams[g.members[0]]->c.pzR * (size_t)pair_ld(nc_of(b.N - 2, b.M));
}

static mcpi_status launch_group(const HvGroup &g, const mcpi_ham *hams, const std::vector<LzInst> &hin,
                                hipStream_t s)
{
    if (g.batched && g.mode == 3) {
        // each member's own H.v (its autotuned schedule, the factorized
        // pair form where it applies); the members share the basis's pair
        // workspace on this stream one after the other
        for (int x : g.members) MCPI_CHECK(launch_apply(hams[x], hin[x].vin, hin[x].HC, s));
        return MCPI_OK;
    }
    if (g.batched && g.mode == 4) {
        // the factorized pair form four members at a time, every chunk of
        // four in one launch pair (then two, then one), from the group's
        // copies of the coefficient blocks
        const BasisDev &bd = g.basis->d;
        const int m = (int)g.members.size();
        int k = 0;
        while (k < m) {
            const int mv = m - k >= 4 ? 4 : (m - k >= 2 ? 2 : 1);
            const int chunks = mv == 4 ? (m - k) / 4 : 1;
            const HamCoefDev ck = shift_coef(g.c0, (ptrdiff_t)(k * g.cbytes));
            const size_t o = (size_t)k * (size_t)bd.nc;
            if (mv == 1)
                MCPI_CHECK(launch_pairz(bd, ck, g.vin + o, g.HC + o, g.zws, s));
            else
                MCPI_CHECK(launch_pairz_multi(bd, ck, mv, chunks, g.vin + o, g.HC + o, bd.nc, (int64_t)g.cbytes,
                                              g.zws, s));
            k += mv * chunks;
        }
        return MCPI_OK;
    }
    if (g.batched) {
        BasisDev bd = g.basis->d;
        int sched = g.sched;
        if (sched == 4) {
            // pattern tables of the group's common zero pattern (built by the
            // tuner before any capture)
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            MCPI_HIP(hipStreamIsCapturing(s, &cs));
            const PatternTables *pt =
                basis_pattern(g.basis, hams[g.members[0]]->c, cs == hipStreamCaptureStatusNone, s);
            if (pt) {
                bd.patOT = pt->tabOT;
                bd.patTT = pt->tabTT;
                bd.ppoffOT = pt->offOT;
                bd.ppoffTT = pt->offTT;
            }
            sched = 2;
        }
        if (g.mode == 2)
            return launch_apply_multi(bd, g.c0, sched, (int)g.members.size(), g.cbytes, g.vin, g.HC, bd.nc, s);
        return launch_apply_batch(bd, g.basis->maps, g.c0, g.skip, sched, (int)g.members.size(), g.cbytes, g.vin,
                                  g.HC, bd.nc, s);
    }
    for (int x : g.members) MCPI_CHECK(launch_apply(hams[x], hin[x].vin, hin[x].HC, s));
    return MCPI_OK;
}

// Time the candidate (kernel, schedule) pairs on the whole group and keep
// the fastest: the batched specialised kernel with the double-jump targets
// read (2) or ranked (3), and, when the members share their zero pattern,
// the multi-vector kernel with the same two schedules; where the factorized
// pair form applies, also the members one by one with their own tuned H.v.
// fixed: a schedule the caller set on the Hamiltonians (0: free).
// kernel choice for batched groups: 0 = timed, 1 = batched specialised
// kernel only, 2 = multi-vector kernel where the group allows it, 3 = the
// factorized pair form several members at a time where the group allows
// it (timed otherwise)
static std::atomic<int> g_batched_kernel{0};

mcpi_status set_batched_kernel(int kind)
{
    MCPI_ARG(kind >= 0 && kind <= 3, "kind must be 0, 1, 2 or 3");
    g_batched_kernel = kind;
    return MCPI_OK;
}

static mcpi_status tune_group(HvGroup &g, const mcpi_ham *hams, const std::vector<LzInst> &hin, int fixed,
                              hipStream_t s)
{
    std::vector<std::pair<int, int>> cand;   // (mode, sched)
    const bool full = g.basis->maps == MCPI_MAPS_FULL;
    if (g.basis->sectored()) {
        // rows in sector order: only the pattern tables hold their targets
        if (!basis_pattern(g.basis, hams[g.members[0]]->c, true, s)) {
            set_error("sector-ordered basis: the pattern tables could not be built (device memory short)");
            return MCPI_ERR_STATE;
        }
        fixed = 4;
    }
    if (g_batched_kernel == 3 && g.pzmv_ok && !fixed) {
        g.mode = 4;
        g.sched = 0;
        return MCPI_OK;
    }
    for (int mode = 1; mode <= 2; mode++) {
        if (mode == 2 && !g.multi_ok) continue;
        if (g_batched_kernel == 1 && mode == 2) continue;
        if (g_batched_kernel == 2 && mode == 1 && g.multi_ok) continue;
        for (int sc = 2; sc <= 4; sc++) {
            if (fixed && sc != fixed) continue;
            if (sc == 3 && !full && mode == 1) continue;   // without stored doubles 2 already ranks
            if (sc == 2 && !full && mode == 2) continue;
            // 4: pattern tables, for a zero pattern all members share
            if (sc == 4 && !(g.multi_ok && basis_pattern(g.basis, hams[g.members[0]]->c, true, s))) continue;
            cand.push_back({mode, sc});
        }
    }
    if (cand.empty()) cand.push_back({1, fixed ? fixed : 2});
    // the members one by one, each with its own tuned H.v: pays one launch
    // pair per member, and wins where the factorized pair form applies
    if (!fixed && (g_batched_kernel == 0 || g_batched_kernel == 3) && !g.basis->sectored() &&
        pairz_applicable(g.basis->d, hams[g.members[0]]->c))
        cand.push_back({3, 0});
    // the same four members at a time: their row and pair arithmetic shared
    if (!fixed && (g_batched_kernel == 0 || g_batched_kernel == 3) && g.pzmv_ok) cand.push_back({4, 0});
    if (cand.size() == 1) {
        g.mode = cand[0].first;
        g.sched = cand[0].second;
        return MCPI_OK;
    }
    hipEvent_t e0, e1;
    MCPI_HIP(hipEventCreate(&e0));
    MCPI_HIP(hipEventCreate(&e1));
    float best = 0;
    std::pair<int, int> pick = cand[0];
    mcpi_status rc = MCPI_OK;
    for (size_t x = 0; x < cand.size(); x++) {
        g.mode = cand[x].first;
        g.sched = cand[x].second;
        rc = launch_group(g, hams, hin, s);   // warm-up
        if (rc != MCPI_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = MCPI_ERR_HIP; break; }
        for (int r = 0; r < 2 && rc == MCPI_OK; r++) rc = launch_group(g, hams, hin, s);
        if (rc != MCPI_OK || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            if (rc == MCPI_OK) rc = MCPI_ERR_HIP;
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (x == 0 || ms < best) { best = ms; pick = cand[x]; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == MCPI_ERR_HIP) set_error("H.v schedule timing failed");
    g.mode = pick.first;
    g.sched = pick.second;
    return rc;
}

// The choice of tune_group, kept across calls for the same group: the
// basis, the member handles with their schedule and zero-skipping settings,
// the kernel kind and the fixed schedule.  A handle freed and reallocated at
// the same address may map to another Hamiltonian, so a cached choice is
// used only while it still applies to the group (every mode computes the
// same H.v; only the time differs).  At most 256 groups are kept.
struct TuneKey {
    const void *basis;
    std::vector<uintptr_t> members;   // handle, schedule, skip_zero per member
    int kind, fixed;
    bool operator<(const TuneKey &o) const
    {
        if (basis != o.basis) return basis < o.basis;
        if (kind != o.kind) return kind < o.kind;
        if (fixed != o.fixed) return fixed < o.fixed;
        return members < o.members;
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, std::pair<int, int>> g_tune_cache;

static bool tuned_choice_applies(const HvGroup &g, const mcpi_ham *hams, int mode, int sched)
{
    const bool full = g.basis->maps == MCPI_MAPS_FULL;
    switch (mode) {
    case 1: return sched >= 2 && sched <= 4 && (sched != 3 || full) && (sched != 4 || g.multi_ok);
    case 2: return g.multi_ok && sched >= 2 && sched <= 4 && (sched != 2 || full);
    case 3: return pairz_applicable(g.basis->d, hams[g.members[0]]->c);
    case 4: return g.pzmv_ok;
    default: return false;
    }
}

static mcpi_status tune_group_cached(HvGroup &g, const mcpi_ham *hams, const std::vector<LzInst> &hin, int fixed,
                                     hipStream_t s)
{
    if (g.basis->sectored()) return tune_group(g, hams, hin, fixed, s);
    TuneKey key{g.basis, {}, g_batched_kernel.load(), fixed};
    for (int x : g.members) {
        key.members.push_back((uintptr_t)hams[x]);
        key.members.push_back((uintptr_t)hams[x]->schedule);
        key.members.push_back((uintptr_t)hams[x]->skip_zero);
    }
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end() && tuned_choice_applies(g, hams, it->second.first, it->second.second)) {
            g.mode = it->second.first;
            g.sched = it->second.second;
            return MCPI_OK;
        }
    }
    MCPI_CHECK(tune_group(g, hams, hin, fixed, s));
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (g_tune_cache.size() >= 256) g_tune_cache.clear();
    g_tune_cache[key] = {g.mode, g.sched};
    return MCPI_OK;
}

// K instances of length lm.  lvec_ext[x] (may be null, or the array null):
// caller's lm*n basis buffer.  diag_h/offd_h: host, K*lm (row x = instance
// x).  Cout (may be null, or entries null): Ritz vectors; E0 (may be null
// only when Cout is): lowest eigenvalues.  steps: K.
static mcpi_status lanczos_core(int K, const mcpi_ham *hams, const cplx *const *C0, cplx *const *lvec_ext, int lm,
                                cplx *diag_h, cplx *offd_h, int *steps, cplx *const *Cout, double *E0, hipStream_t s)
{
    MCPI_ARG(K >= 0 && K <= 65535, "number of instances out of range");
    MCPI_ARG(lm >= 1 && lm <= (1 << 20), "Lanczos length must be in [1, 2^20]");
    if (K == 0) return MCPI_OK;
    const bool want_e = E0 != nullptr;

    // ---- H.v groups by basis
    std::vector<HvGroup> groups;
    std::vector<int> batched_member(K, -1);   // index within a batched group
    for (int x = 0; x < K; x++) {
        MCPI_ARG(hams[x] && C0[x], "null instance");
        auto it = std::find_if(groups.begin(), groups.end(),
                               [&](const HvGroup &g) { return g.basis == hams[x]->basis; });
        if (it == groups.end()) {
            groups.emplace_back();
            groups.back().basis = hams[x]->basis;
            it = groups.end() - 1;
        }
        it->members.push_back(x);
    }
    for (HvGroup &g : groups) {
        const mcpi_ham h0 = hams[g.members[0]];
        bool ok = g.members.size() > 1 && has_fast_kernel(g.basis->d.N, g.basis->d.M);
        bool real = true, dense = true, same_zeros = true;
        for (int x : g.members) {
            const mcpi_ham h = hams[x];
            ok = ok && h->skip_zero == h0->skip_zero && h->schedule == h0->schedule && h->schedule != 1 &&
                 h->coef_bytes == h0->coef_bytes;
            real = real && h->c.real;
            dense = dense && h->c.dense;
            same_zeros = same_zeros && h->c.zsig == h0->c.zsig && !h->c.dense;
        }
        // a sector-ordered basis reads its pattern tables, one zero pattern per launch
        if (g.basis->sectored() && !same_zeros) ok = false;
        g.batched = ok;
        if (ok) {
            g.skip = h0->skip_zero && !dense;
            g.multi_ok = g.skip && same_zeros && g.basis->d.Map != nullptr;
            g.cbytes = h0->coef_bytes;
            g.c0.real = real ? 1 : 0;
            g.pzmv_ok = pairz_multi_ok(g, hams);
            for (size_t k = 0; k < g.members.size(); k++) batched_member[g.members[k]] = (int)k;
        }
    }

    // ---- workspace
    Arena A;
    A.reserve(sizeof(LanczosState) * K);
    A.reserve(sizeof(LzInst) * K);
    std::vector<LzInst> hin(K);
    int nbmax = 1, nbgmax = 1;
    std::vector<int64_t> ns(K);
    for (int x = 0; x < K; x++) ns[x] = hams[x]->basis->d.nc;
    const int rr = update_rows(ns), dtg = dot_group(ns);
    for (int x = 0; x < K; x++) {
        const int64_t n = ns[x];
        LzInst &I = hin[x];
        I.n = n;
        I.nb = (int)row_blocks(n, rr);
        I.nbg = (int)((n + LZ_SEG - 1) / LZ_SEG);
        nbmax = std::max(nbmax, I.nb);
        nbgmax = std::max(nbgmax, I.nbg);
        if (!(lvec_ext && lvec_ext[x])) A.reserve(sizeof(cplx) * n * (size_t)lm);
        if (batched_member[x] < 0)
            for (int k = 0; k < 2; k++) A.reserve(sizeof(cplx) * n);
        for (int k = 0; k < 3; k++) A.reserve(sizeof(cplx) * lm);
        A.reserve(sizeof(double) * 2 * I.nb);
        A.reserve(sizeof(double) * 2 * (size_t)I.nbg * lm);
        A.reserve(sizeof(double) * lm);
        if (split_update(n, lm)) A.reserve(sizeof(cplx) * n * (size_t)((lm + LZ_TG - 1) / LZ_TG));
    }
    for (HvGroup &g : groups) {
        if (!g.batched) continue;
        const size_t m = g.members.size();
        A.reserve(g.cbytes * m);
        for (int k = 0; k < 2; k++) A.reserve(sizeof(cplx) * (size_t)g.basis->d.nc * m);
        if (g.pzmv_ok) A.reserve(pairz_multi_bytes(g, hams));
    }
    MCPI_CHECK(A.allocate());
    bool any_split = false;
    LanczosState *sts = A.take<LanczosState>(K);
    LzInst *din = A.take<LzInst>(K);
    for (HvGroup &g : groups) {
        if (!g.batched) continue;
        const size_t m = g.members.size();
        g.coefs = A.take<char>(g.cbytes * m);
        g.vin = A.take<cplx>((size_t)g.basis->d.nc * m);
        g.HC = A.take<cplx>((size_t)g.basis->d.nc * m);
        if (g.pzmv_ok) g.zws = (cplx *)A.take<char>(pairz_multi_bytes(g, hams));
        if (A.overflow) break;
        for (size_t k = 0; k < m; k++) {
            const mcpi_ham h = hams[g.members[k]];
            MCPI_HIP(hipMemcpyAsync(g.coefs + k * g.cbytes, h->coef_block, g.cbytes, hipMemcpyDeviceToDevice, s));
        }
        const mcpi_ham h0 = hams[g.members[0]];
        const int real = g.c0.real;
        g.c0 = shift_coef(h0->c, (ptrdiff_t)((uintptr_t)g.coefs - (uintptr_t)h0->coef_block));
        g.c0.real = real;
    }
    for (int x = 0; x < K; x++) {
        LzInst &I = hin[x];
        const int64_t n = I.n;
        I.C0 = C0[x];
        I.lvec = (lvec_ext && lvec_ext[x]) ? lvec_ext[x] : A.take<cplx>(n * (size_t)lm);
        if (batched_member[x] < 0) {
            I.HC = A.take<cplx>(n);
            I.vin = A.take<cplx>(n);
        }
        I.diag = A.take<cplx>(lm);
        I.offd = A.take<cplx>(lm);
        I.ortho = A.take<cplx>(lm);
        I.partial = A.take<double>(2 * I.nb);
        I.partial_o = A.take<double>(2 * (size_t)I.nbg * lm);
        double *z = A.take<double>(lm);
        I.z = z;
        I.rpart = split_update(n, lm) ? A.take<cplx>(n * (size_t)((lm + LZ_TG - 1) / LZ_TG)) : nullptr;
        any_split = any_split || I.rpart != nullptr;
        I.Cout = (Cout && Cout[x]) ? Cout[x] : nullptr;
        I.iperm = hams[x]->basis->iperm;
        I.steps = 0;
    }
    for (HvGroup &g : groups) {
        if (!g.batched || !g.vin) continue;
        for (size_t k = 0; k < g.members.size(); k++) {
            hin[g.members[k]].vin = g.vin + k * (size_t)g.basis->d.nc;
            hin[g.members[k]].HC = g.HC + k * (size_t)g.basis->d.nc;
        }
    }
    if (A.overflow) {
        set_error("internal error: Lanczos workspace layout");
        return MCPI_ERR_STATE;
    }
    // a start vector gathered in place would read rows already overwritten:
    // gather from a copy in the (still unused) H.v output instead
    for (int x = 0; x < K; x++) {
        LzInst &I = hin[x];
        if (I.iperm && (const void *)I.C0 == (const void *)I.lvec) {
            MCPI_HIP(hipMemcpyAsync(I.HC, I.C0, sizeof(cplx) * I.n, hipMemcpyDeviceToDevice, s));
            I.C0 = I.HC;
        }
    }
    MCPI_HIP(hipMemcpyAsync(din, hin.data(), sizeof(LzInst) * K, hipMemcpyHostToDevice, s));

    const dim3 blk(LZ_T), gv(nbmax, K), g1(1, K);
    const dim3 gd(nbgmax, (lm + dtg - 1) / dtg, K), gpart(nbgmax, (lm + LZ_TG - 1) / LZ_TG, K);
    const dim3 go(lm, K);
    const int ga = (K + 63) / 64;

    // ---- set-up: v_0, schedules, HC = H v_0, alpha_0, HC -= alpha_0 v_0
    hipLaunchKernelGGL(lz_init, g1, dim3(64), 0, s, sts, (const LzInst *)din, lm, 1e-15);
    hipLaunchKernelGGL(lz_start, gv, blk, 0, s, (const LzInst *)din);
    MCPI_HIP(hipGetLastError());
    for (HvGroup &g : groups) {
        if (!g.batched) continue;
        const mcpi_ham h0 = hams[g.members[0]];
        MCPI_CHECK(tune_group_cached(g, hams, hin, (h0->schedule >= 2 && h0->schedule <= 4) ? h0->schedule : 0, s));
    }
    for (const HvGroup &g : groups) MCPI_CHECK(launch_group(g, hams, hin, s));
    hipLaunchKernelGGL(lz_dot, gv, blk, 0, s, (const LzInst *)din);
    hipLaunchKernelGGL(lz_fin_alpha, g1, blk, 0, s, sts, (const LzInst *)din, 1);
    hipLaunchKernelGGL(lz_sub_alpha_norm, gv, blk, 0, s, (const LanczosState *)sts, (const LzInst *)din);
    MCPI_HIP(hipGetLastError());

    if (lm > 1) {
        const LanczosState *cst = sts;
        const LzInst *cin = din;
        auto enqueue = [&]() -> mcpi_status {
            hipLaunchKernelGGL(lz_fin_beta, g1, blk, 0, s, sts, cin);
            hipLaunchKernelGGL(lz_scale, gv, blk, 0, s, cst, cin);
            for (const HvGroup &g : groups) MCPI_CHECK(launch_group(g, hams, hin, s));
            hipLaunchKernelGGL(lz_sub_dot, gv, blk, 0, s, cst, cin);
            hipLaunchKernelGGL(lz_fin_alpha, g1, blk, 0, s, sts, cin, 0);
            hipLaunchKernelGGL(lz_sub_alpha, gv, blk, 0, s, cst, cin);
            if (dtg == 8)
                hipLaunchKernelGGL(lz_gemv_dots<8>, gd, blk, 0, s, cst, cin);
            else
                hipLaunchKernelGGL(lz_gemv_dots<16>, gd, blk, 0, s, cst, cin);
            hipLaunchKernelGGL(lz_fin_ortho, go, blk, 0, s, cst, cin);
            if (rr == 2)
                hipLaunchKernelGGL(lz_reorth_norm<2>, gv, blk, 0, s, cst, cin);
            else
                hipLaunchKernelGGL(lz_reorth_norm<4>, gv, blk, 0, s, cst, cin);
            if (any_split) {
                hipLaunchKernelGGL(lz_reorth_part, gpart, blk, 0, s, cst, cin);
                hipLaunchKernelGGL(lz_reorth_sum, gv, blk, 0, s, cst, cin);
            }
            hipLaunchKernelGGL(lz_advance, dim3(ga), dim3(64), 0, s, sts, K);
            MCPI_HIP(hipGetLastError());
            return MCPI_OK;
        };
        // capture one iteration and replay it; fall back to plain launches
        // if the runtime cannot capture on this stream
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        bool graphed = false;
        if (s != nullptr && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            mcpi_status st = enqueue();
            hipError_t e = hipStreamEndCapture(s, &graph);
            if (st == MCPI_OK && e == hipSuccess && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess)
                graphed = true;
            else
                (void)hipGetLastError();
        }
        if (graphed) {
            mcpi_status rc = MCPI_OK;
            for (int i = 0; i < lm - 1 && rc == MCPI_OK; i++) {
                hipError_t e = hipGraphLaunch(exec, s);
                if (e != hipSuccess) rc = hip_fail(e, "hipGraphLaunch", __FILE__, __LINE__);
            }
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            MCPI_CHECK(rc);
        } else {
            if (graph) (void)hipGraphDestroy(graph);
            for (int i = 0; i < lm - 1; i++) MCPI_CHECK(enqueue());
        }
    }

    // ---- results to the host
    std::vector<LanczosState> hs(K);
    std::vector<cplx> dg((size_t)K * lm), od((size_t)K * lm);
    MCPI_HIP(hipMemcpyAsync(hs.data(), sts, sizeof(LanczosState) * K, hipMemcpyDeviceToHost, s));
    for (int x = 0; x < K; x++) {
        MCPI_HIP(hipMemcpyAsync(&dg[(size_t)x * lm], hin[x].diag, sizeof(cplx) * lm, hipMemcpyDeviceToHost, s));
        MCPI_HIP(hipMemcpyAsync(&od[(size_t)x * lm], hin[x].offd, sizeof(cplx) * lm, hipMemcpyDeviceToHost, s));
    }
    MCPI_HIP(hipStreamSynchronize(s));
    for (int x = 0; x < K; x++) {
        const int st = hs[x].broke ? hs[x].steps : lm;
        steps[x] = st;
        if (!hs[x].broke) od[(size_t)x * lm + lm - 1] = cmake(0, 0);
    }
    if (diag_h) std::memcpy(diag_h, dg.data(), sizeof(cplx) * K * lm);
    if (offd_h) std::memcpy(offd_h, od.data(), sizeof(cplx) * K * lm);
    // a caller's Krylov basis of a sector-ordered basis goes back to the
    // reference order, row by row through the H.v output
    for (int x = 0; x < K; x++) {
        const LzInst &I = hin[x];
        if (!(I.iperm && lvec_ext && lvec_ext[x])) continue;
        for (int t = 0; t < steps[x]; t++) {
            MCPI_CHECK(scatter_rows(I.n, I.iperm, I.lvec + (size_t)t * I.n, I.HC, s));
            MCPI_HIP(hipMemcpyAsync(I.lvec + (size_t)t * I.n, I.HC, sizeof(cplx) * I.n, hipMemcpyDeviceToDevice, s));
        }
        MCPI_HIP(hipStreamSynchronize(s));
    }
    if (!want_e) return MCPI_OK;

    // ---- tridiagonal eigenproblems on the host, Ritz vectors on the device
    bool any_ritz = false;
    std::vector<double> zall((size_t)K * lm, 0.0);
    for (int x = 0; x < K; x++) {
        const int n = steps[x];
        std::vector<double> d(n), e(n);
        for (int k = 0; k < n; k++) {
            d[k] = dg[(size_t)x * lm + k].re;
            e[k] = od[(size_t)x * lm + k].re;
        }
        MCPI_CHECK(tridiag_lowest(n, d.data(), e.data(), &E0[x], hin[x].Cout ? &zall[(size_t)x * lm] : nullptr));
        hin[x].steps = n;
        any_ritz = any_ritz || hin[x].Cout;
    }
    if (any_ritz) {
        for (int x = 0; x < K; x++)
            if (hin[x].Cout)
                MCPI_HIP(hipMemcpyAsync((void *)hin[x].z, &zall[(size_t)x * lm], sizeof(double) * lm,
                                        hipMemcpyHostToDevice, s));
        MCPI_HIP(hipMemcpyAsync(din, hin.data(), sizeof(LzInst) * K, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(lz_ritz, gv, blk, 0, s, (const LzInst *)din);
        MCPI_HIP(hipGetLastError());
        MCPI_HIP(hipStreamSynchronize(s));
    }
    return MCPI_OK;
}

mcpi_status ground_run(mcpi_ham h, int lm, const cplx *C0, cplx *Cout, double *E0, int *steps_out,
                       hipStream_t s, cplx *lvec_ext)
{
    MCPI_ARG(h && C0 && Cout && E0, "bad argument");
    int st = 0;
    cplx *const lv[1] = {lvec_ext};
    cplx *const co[1] = {Cout};
    MCPI_CHECK(lanczos_core(1, &h, &C0, lv, lm, nullptr, nullptr, &st, co, E0, s));
    if (steps_out) *steps_out = st;
    return MCPI_OK;
}

mcpi_status lanczos_api(mcpi_ham h, int lm, const cplx *C0, cplx *lvec, cplx *diag, cplx *offd, int *steps,
                        hipStream_t s)
{
    MCPI_ARG(h && C0 && lvec && diag && offd && steps, "bad argument");
    cplx *const lv[1] = {lvec};
    return lanczos_core(1, &h, &C0, lv, lm, diag, offd, steps, nullptr, nullptr, s);
}

mcpi_status lanczos_batched_api(int K, const mcpi_ham *hams, const cplx *const *C0, int lm, double *E0,
                                int *steps, cplx *const *Cout, hipStream_t s)
{
    MCPI_ARG(K >= 0 && (K == 0 || (hams && C0 && E0 && steps)), "bad argument");
    return lanczos_core(K, hams, C0, nullptr, lm, nullptr, nullptr, steps, Cout, E0, s);
}

}  // namespace mcpi
