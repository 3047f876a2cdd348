// Lanczos ground state on the device.
//
// Reference: lanczos, groundStateLieb.c:217-344 (three-term recurrence with
// one pass of classical Gram-Schmidt against every previous vector),
// carrDot / carrMod (73-101) and ground (348-471: tridiagonal eigenproblem on
// the host, Ritz vector back-transform).
//
// One iteration i (0 <= i < lm-1) is a fixed chain of kernels that read the
// step index and the scalars from device memory:
//   1 fin_beta      beta_i = ||HC|| from block partials, breakdown test
//   2 scale         v_{i+1} = HC / beta_i  (into lvec[i+1] and the H.v input)
//   3 H.v           HC = H v_{i+1}
//   4 sub_dot       HC -= beta_i v_i;  partial <v_{i+1}, HC>
//   5 fin_scalar    alpha_{i+1}
//   6 sub_alpha     HC -= alpha v_{i+1}
//     gemv_dots     partial o_t = <v_t, HC>, t <= i+1
//   7 fin_ortho     o_t
//   8 reorth_norm   HC -= sum_t v_t o_t;  partial ||HC||^2 for the next step
//   9 advance       i += 1
// The pointers never change between iterations, so the chain is captured
// once in a hipGraph and replayed lm-1 times (no host round trip per step).
// Reductions are two-stage (block partials, then one block), deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

#define LZ_T 256

__device__ inline double lz_wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// block sum of (a, b) into partial[blockIdx] (2 doubles)
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

__device__ inline cplx ldv(const cplx *p)
{
    double2 t = *reinterpret_cast<const double2 *>(p);
    return cmake(t.x, t.y);
}
__device__ inline void stv(cplx *p, cplx z)
{
    double2 t;
    t.x = z.re;
    t.y = z.im;
    *reinterpret_cast<double2 *>(p) = t;
}

// ---- set-up kernels (before the first iteration) --

__global__ void lz_init(LanczosState *st, int lm, double tol)
{
    st->it = 0;
    st->broke = 0;
    st->steps = lm;
    st->lm = lm;
    st->maxb = 0;
    st->tol = tol;
    st->beta = 0;
    st->alpha = cmake(0, 0);
}

// partial <x, y> (x conjugated)
__global__ __launch_bounds__(LZ_T) void lz_dot(int64_t n, const cplx *__restrict__ x,
                                              const cplx *__restrict__ y, double *partial)
{
    double re = 0, im = 0;
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        cplx a = ldv(x + j), b = ldv(y + j);
        re += a.re * b.re + a.im * b.im;
        im += a.re * b.im - a.im * b.re;
    }
    block_sum2(re, im, partial);
}

// sum block partials into a complex scalar; which: 0 -> alpha (diag[it+1] or
// diag[0] when first), 1 -> first alpha into diag[0]
__global__ void lz_fin_alpha(LanczosState *st, int nb, const double *partial, cplx *diag, int first)
{
    __shared__ double red[2][LZ_T];
    if (st->broke) return;
    double re = 0, im = 0;
    for (int x = threadIdx.x; x < nb; x += LZ_T) { re += partial[2 * x]; im += partial[2 * x + 1]; }
    red[0][threadIdx.x] = re;
    red[1][threadIdx.x] = im;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int x = 0; x < LZ_T; x++) { a += red[0][x]; b += red[1][x]; }
        st->alpha = cmake(a, b);
        diag[first ? 0 : st->it + 1] = cmake(a, b);
    }
}

// HC -= alpha v;  partial ||HC||^2   (set-up only: iteration 0 has no reorthogonalization)
__global__ __launch_bounds__(LZ_T) void lz_sub_alpha_norm(int64_t n, const LanczosState *st,
                                                        const cplx *__restrict__ v, cplx *HC,
                                                        double *partial)
{
    const cplx a = st->alpha;
    double s = 0;
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        cplx h = ldv(HC + j), x = ldv(v + j);
        h.re -= a.re * x.re - a.im * x.im;
        h.im -= a.re * x.im + a.im * x.re;
        stv(HC + j, h);
        s += h.re * h.re + h.im * h.im;
    }
    block_sum2(s, 0, partial);
}

// ---- iteration kernels --

// 1: beta_i and the breakdown test (groundStateLieb.c:287-292)
__global__ void lz_fin_beta(LanczosState *st, int nb, const double *partial, cplx *offdiag)
{
    __shared__ double red[LZ_T];
    if (st->broke) return;
    double s = 0;
    for (int x = threadIdx.x; x < nb; x += LZ_T) s += partial[2 * x];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int x = 0; x < LZ_T; x++) t += red[x];
        double b = sqrt(t);
        const int i = st->it;
        offdiag[i] = cmake(b, 0);
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
__global__ __launch_bounds__(LZ_T) void lz_scale(int64_t n, const LanczosState *st, const cplx *__restrict__ HC,
                                               cplx *lvec, cplx *vin)
{
    if (st->broke) return;
    const double inv = 1.0 / st->beta;
    cplx *dst = lvec + (int64_t)(st->it + 1) * n;
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        cplx h = ldv(HC + j);
        h.re *= inv;
        h.im *= inv;
        stv(dst + j, h);
        stv(vin + j, h);
    }
}

// 4: HC -= beta_i v_i;  partial <v_{i+1}, HC>
__global__ __launch_bounds__(LZ_T) void lz_sub_dot(int64_t n, const LanczosState *st, const cplx *lvec,
                                                 const cplx *__restrict__ vin, cplx *HC, double *partial)
{
    if (st->broke) return;
    const double b = st->beta;
    const cplx *vi = lvec + (int64_t)st->it * n;
    double re = 0, im = 0;
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        cplx h = ldv(HC + j), x = ldv(vi + j), y = ldv(vin + j);
        h.re -= b * x.re;
        h.im -= b * x.im;
        stv(HC + j, h);
        re += y.re * h.re + y.im * h.im;
        im += y.re * h.im - y.im * h.re;
    }
    block_sum2(re, im, partial);
}

// 6a: HC -= alpha v_{i+1}
__global__ __launch_bounds__(LZ_T) void lz_sub_alpha(int64_t n, const LanczosState *st,
                                                   const cplx *__restrict__ vin, cplx *HC)
{
    if (st->broke) return;
    const cplx a = st->alpha;
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        cplx x = ldv(HC + j), y = ldv(vin + j);
        x.re -= a.re * y.re - a.im * y.im;
        x.im -= a.re * y.im + a.im * y.re;
        stv(HC + j, x);
    }
}

// 6b: partial o_t = <v_t, HC> for t = 0..i+1.  Block (s, g) owns row
// segment s (LZ_R rows per thread, kept in registers) and basis vectors
// t in [g*LZ_TG, (g+1)*LZ_TG); waves reduce with shuffles and park their sums
// in LDS, one barrier per block.
#define LZ_R 8
#define LZ_TG 32
#define LZ_SEG ((int64_t)LZ_T * LZ_R)
__global__ __launch_bounds__(LZ_T) void lz_gemv_dots(int64_t n, const LanczosState *st, const cplx *lvec,
                                                   const cplx *__restrict__ HC, double *partial_o, int lm)
{
    __shared__ double red[LZ_TG][LZ_T / 64][2];
    if (st->broke) return;
    const int nt = st->it + 2;
    const int t0 = blockIdx.y * LZ_TG;
    if (t0 >= nt) return;
    const int t1 = min(nt, t0 + LZ_TG);
    const int64_t base = (int64_t)blockIdx.x * LZ_SEG + threadIdx.x;
    cplx h[LZ_R];
#pragma unroll
    for (int r = 0; r < LZ_R; r++) {
        int64_t j = base + (int64_t)r * LZ_T;
        h[r] = (j < n) ? ldv(HC + j) : cmake(0, 0);
    }
    const int w = threadIdx.x >> 6;
    for (int tt = t0; tt < t1; tt++) {
        const cplx *vt = lvec + (int64_t)tt * n;
        double re = 0, im = 0;
#pragma unroll
        for (int r = 0; r < LZ_R; r++) {
            int64_t j = base + (int64_t)r * LZ_T;
            if (j < n) {
                cplx x = ldv(vt + j);
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
        int64_t at = 2 * ((int64_t)blockIdx.x * lm + t0 + tl);
        partial_o[at] = sa;
        partial_o[at + 1] = sb;
    }
}

// 7: o_t = sum over blocks (one thread per t)
__global__ void lz_fin_ortho(const LanczosState *st, int nb, const double *partial_o, int lm, cplx *ortho)
{
    if (st->broke) return;
    const int nt = st->it + 2;
    for (int tt = threadIdx.x; tt < nt; tt += blockDim.x) {
        double re = 0, im = 0;
        for (int x = 0; x < nb; x++) {
            re += partial_o[2 * ((int64_t)x * lm + tt)];
            im += partial_o[2 * ((int64_t)x * lm + tt) + 1];
        }
        ortho[tt] = cmake(re, im);
    }
}

// 8: HC -= sum_t v_t o_t;  partial ||HC||^2 for the next beta
__global__ __launch_bounds__(LZ_T) void lz_reorth_norm(int64_t n, const LanczosState *st, const cplx *lvec,
                                                     const cplx *__restrict__ ortho, cplx *HC,
                                                     double *partial)
{
    extern __shared__ double so[];
    if (st->broke) return;
    const int nt = st->it + 2;
    for (int x = threadIdx.x; x < nt; x += LZ_T) {
        cplx o = ortho[x];
        so[2 * x] = o.re;
        so[2 * x + 1] = o.im;
    }
    __syncthreads();
    double s = 0;
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        cplx h = ldv(HC + j);
#pragma unroll 4
        for (int tt = 0; tt < nt; tt++) {
            cplx x = ldv(lvec + (int64_t)tt * n + j);
            double ore = so[2 * tt], oim = so[2 * tt + 1];
            h.re -= x.re * ore - x.im * oim;
            h.im -= x.re * oim + x.im * ore;
        }
        stv(HC + j, h);
        s += h.re * h.re + h.im * h.im;
    }
    block_sum2(s, 0, partial);
}

// 9
__global__ void lz_advance(LanczosState *st)
{
    if (!st->broke) st->it += 1;
}

// Ritz vector C = sum_k lvec[k] z[k]  (groundStateLieb.c:454-458)
__global__ __launch_bounds__(LZ_T) void lz_ritz(int64_t n, int steps, const cplx *lvec, const double *z, cplx *C)
{
    extern __shared__ double sz[];
    for (int x = threadIdx.x; x < steps; x += LZ_T) sz[x] = z[x];
    __syncthreads();
    for (int64_t j = (int64_t)blockIdx.x * LZ_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * LZ_T) {
        double re = 0, im = 0;
        for (int k = 0; k < steps; k++) {
            cplx x = ldv(lvec + (int64_t)k * n + j);
            re += x.re * sz[k];
            im += x.im * sz[k];
        }
        stv(C + j, cmake(re, im));
    }
}

// ---------------------------------------------------------------- host --

struct LanczosBuffers {
    LanczosState *st = nullptr;
    cplx *HC = nullptr, *vin = nullptr, *diag = nullptr, *offd = nullptr, *ortho = nullptr;
    double *partial = nullptr, *partial_o = nullptr;
    int nb = 0, nbg = 0;
    ~LanczosBuffers()
    {
        void *p[] = {st, HC, vin, diag, offd, ortho, partial, partial_o};
        for (void *x : p)
            if (x) (void)hipFree(x);
    }
};

static int vec_blocks(int64_t n)
{
    int64_t b = (n + LZ_T * 4 - 1) / (LZ_T * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1024));
}

static mcpi_status lanczos_run(mcpi_ham h, int lm, const cplx *C0, cplx *lvec, cplx *diag_h, cplx *offd_h,
                               int *steps_out, hipStream_t s)
{
    const int64_t n = h->basis->d.nc;
    LanczosBuffers B;
    B.nb = vec_blocks(n);
    B.nbg = (int)((n + LZ_SEG - 1) / LZ_SEG);
    MCPI_HIP(hipMalloc(&B.st, sizeof(LanczosState)));
    MCPI_HIP(hipMalloc(&B.HC, sizeof(cplx) * n));
    MCPI_HIP(hipMalloc(&B.vin, sizeof(cplx) * n));
    MCPI_HIP(hipMalloc(&B.diag, sizeof(cplx) * lm));
    MCPI_HIP(hipMalloc(&B.offd, sizeof(cplx) * lm));
    MCPI_HIP(hipMalloc(&B.ortho, sizeof(cplx) * lm));
    MCPI_HIP(hipMalloc(&B.partial, sizeof(double) * 2 * B.nb));
    MCPI_HIP(hipMalloc(&B.partial_o, sizeof(double) * 2 * (size_t)B.nbg * lm));
    std::vector<cplx> zeros(lm, cmake(0, 0));
    MCPI_HIP(hipMemcpyAsync(B.diag, zeros.data(), sizeof(cplx) * lm, hipMemcpyHostToDevice, s));
    MCPI_HIP(hipMemcpyAsync(B.offd, zeros.data(), sizeof(cplx) * lm, hipMemcpyHostToDevice, s));

    // set-up: v_0, HC = H v_0, alpha_0, HC -= alpha_0 v_0, ||HC||^2 partials
    hipLaunchKernelGGL(lz_init, dim3(1), dim3(1), 0, s, B.st, lm, 1e-15);
    if (lvec != C0) MCPI_HIP(hipMemcpyAsync(lvec, C0, sizeof(cplx) * n, hipMemcpyDeviceToDevice, s));
    MCPI_HIP(hipMemcpyAsync(B.vin, C0, sizeof(cplx) * n, hipMemcpyDeviceToDevice, s));
    MCPI_CHECK(launch_apply(h, B.vin, B.HC, s));
    hipLaunchKernelGGL(lz_dot, dim3(B.nb), dim3(LZ_T), 0, s, n, (const cplx *)B.vin, (const cplx *)B.HC, B.partial);
    hipLaunchKernelGGL(lz_fin_alpha, dim3(1), dim3(LZ_T), 0, s, B.st, B.nb, (const double *)B.partial, B.diag, 1);
    hipLaunchKernelGGL(lz_sub_alpha_norm, dim3(B.nb), dim3(LZ_T), 0, s, n, (const LanczosState *)B.st,
                       (const cplx *)B.vin, B.HC, B.partial);
    MCPI_HIP(hipGetLastError());

    if (lm > 1) {
        auto enqueue = [&]() -> mcpi_status {
            hipLaunchKernelGGL(lz_fin_beta, dim3(1), dim3(LZ_T), 0, s, B.st, B.nb, (const double *)B.partial, B.offd);
            hipLaunchKernelGGL(lz_scale, dim3(B.nb), dim3(LZ_T), 0, s, n, (const LanczosState *)B.st,
                               (const cplx *)B.HC, lvec, B.vin);
            MCPI_CHECK(launch_apply(h, B.vin, B.HC, s));
            hipLaunchKernelGGL(lz_sub_dot, dim3(B.nb), dim3(LZ_T), 0, s, n, (const LanczosState *)B.st,
                               (const cplx *)lvec, (const cplx *)B.vin, B.HC, B.partial);
            hipLaunchKernelGGL(lz_fin_alpha, dim3(1), dim3(LZ_T), 0, s, B.st, B.nb, (const double *)B.partial, B.diag, 0);
            hipLaunchKernelGGL(lz_sub_alpha, dim3(B.nb), dim3(LZ_T), 0, s, n, (const LanczosState *)B.st,
                               (const cplx *)B.vin, B.HC);
            hipLaunchKernelGGL(lz_gemv_dots, dim3(B.nbg, (lm + LZ_TG - 1) / LZ_TG), dim3(LZ_T), 0, s, n,
                               (const LanczosState *)B.st, (const cplx *)lvec, (const cplx *)B.HC,
                               B.partial_o, lm);
            hipLaunchKernelGGL(lz_fin_ortho, dim3(1), dim3(LZ_T), 0, s, (const LanczosState *)B.st, B.nbg,
                               (const double *)B.partial_o, lm, B.ortho);
            hipLaunchKernelGGL(lz_reorth_norm, dim3(B.nb), dim3(LZ_T), sizeof(double) * 2 * lm, s, n,
                               (const LanczosState *)B.st, (const cplx *)lvec, (const cplx *)B.ortho, B.HC,
                               B.partial);
            hipLaunchKernelGGL(lz_advance, dim3(1), dim3(1), 0, s, B.st);
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
            for (int i = 0; i < lm - 1; i++) MCPI_HIP(hipGraphLaunch(exec, s));
            MCPI_HIP(hipGraphExecDestroy(exec));
            MCPI_HIP(hipGraphDestroy(graph));
        } else {
            if (graph) (void)hipGraphDestroy(graph);
            for (int i = 0; i < lm - 1; i++) MCPI_CHECK(enqueue());
        }
    }
    LanczosState hs;
    MCPI_HIP(hipMemcpyAsync(&hs, B.st, sizeof hs, hipMemcpyDeviceToHost, s));
    MCPI_HIP(hipMemcpyAsync(diag_h, B.diag, sizeof(cplx) * lm, hipMemcpyDeviceToHost, s));
    MCPI_HIP(hipMemcpyAsync(offd_h, B.offd, sizeof(cplx) * lm, hipMemcpyDeviceToHost, s));
    MCPI_HIP(hipStreamSynchronize(s));
    *steps_out = hs.broke ? hs.steps : lm;
    if (!hs.broke && lm >= 1) offd_h[lm - 1] = cmake(0, 0);
    return MCPI_OK;
}

// ------------------------------------------- tridiagonal eigenproblem --
// Eigenvalues of the symmetric tridiagonal (d, e) by the implicit QL method
// with Wilkinson-type shifts (the tql1 scheme of EISPACK, Golub & Van Loan
// sec. 8.3), then the lowest eigenvector by inverse iteration.  Stands in for
// LAPACKE_dstev (groundStateLieb.c:435).
static int tridiag_eigvals(int n, double *d, double *e)
{
    if (n <= 0) return 0;
    e[n - 1] = 0.0;
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) return -1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            bool underflow = false;
            for (i = m - 1; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
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
    std::vector<double> dd(d, d + n), ee(n, 0.0);
    for (int i = 0; i < n - 1; i++) ee[i] = e ? e[i] : 0.0;
    if (tridiag_eigvals(n, dd.data(), ee.data()) != 0) {
        set_error("tridiagonal QL did not converge");
        return MCPI_ERR_STATE;
    }
    double lam = dd[0];
    for (int i = 1; i < n; i++) lam = std::min(lam, dd[i]);
    *lambda = lam;
    if (z) {
        std::vector<double> ev(n - 1 > 0 ? n - 1 : 1, 0.0);
        for (int i = 0; i < n - 1; i++) ev[i] = e ? e[i] : 0.0;
        tridiag_inverse_iteration(n, d, ev.data(), lam, z);
    }
    return MCPI_OK;
}

mcpi_status ground_run(mcpi_ham h, int lm, const cplx *C0, cplx *Cout, double *E0, int *steps_out,
                       hipStream_t s, cplx *lvec_ext)
{
    const int64_t n = h->basis->d.nc;
    MCPI_ARG(lm >= 1, "Lanczos length must be >= 1");
    cplx *lvec = lvec_ext;
    if (!lvec) {
        hipError_t e = hipMalloc(&lvec, sizeof(cplx) * n * (size_t)lm);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("cannot allocate the Lanczos basis (lm * nc complex)");
            return MCPI_ERR_ALLOC;
        }
    }
    std::vector<cplx> diag(lm), offd(lm);
    int steps = 0;
    mcpi_status st = lanczos_run(h, lm, C0, lvec, diag.data(), offd.data(), &steps, s);
    if (st == MCPI_OK) {
        std::vector<double> d(steps), e(steps), z(steps);
        for (int k = 0; k < steps; k++) { d[k] = diag[k].re; e[k] = offd[k].re; }
        double lam = 0;
        st = tridiag_lowest(steps, d.data(), e.data(), &lam, z.data());
        if (st == MCPI_OK) {
            double *zd = nullptr;
            hipError_t err = hipMalloc(&zd, sizeof(double) * steps);
            if (err == hipSuccess) err = hipMemcpyAsync(zd, z.data(), sizeof(double) * steps, hipMemcpyHostToDevice, s);
            if (err == hipSuccess) {
                hipLaunchKernelGGL(lz_ritz, dim3(vec_blocks(n)), dim3(LZ_T), sizeof(double) * steps, s, n, steps,
                                   (const cplx *)lvec, (const double *)zd, Cout);
                err = hipStreamSynchronize(s);
            }
            if (zd) (void)hipFree(zd);
            if (err != hipSuccess) st = hip_fail(err, "Ritz vector", __FILE__, __LINE__);
            *E0 = lam;
            if (steps_out) *steps_out = steps;
        }
    }
    if (!lvec_ext) (void)hipFree(lvec);
    return st;
}

mcpi_status lanczos_api(mcpi_ham h, int lm, const cplx *C0, cplx *lvec, cplx *diag, cplx *offd, int *steps,
                        hipStream_t s)
{
    return lanczos_run(h, lm, C0, lvec, diag, offd, steps, s);
}

}  // namespace mcpi
