// Host side of libmcpi_hip.so: error state, device helpers, Hamiltonian
// handles, and the reference-shaped entry points (include/mcpi_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mcpi_internal.h"

namespace mcpi {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

mcpi_status hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_err = buf;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? MCPI_ERR_ALLOC : MCPI_ERR_HIP;
}

mcpi_status ground_run(mcpi_ham h, int lm, const cplx *C0, cplx *Cout, double *E0, int *steps_out,
                       hipStream_t s, cplx *lvec_ext);
mcpi_status lanczos_api(mcpi_ham h, int lm, const cplx *C0, cplx *lvec, cplx *diag, cplx *offd, int *steps,
                        hipStream_t s);
mcpi_status tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z);
mcpi_status lanczos_batched_api(int K, const mcpi_ham *hams, const cplx *const *C0, int lm, double *E0,
                                int *steps, cplx *const *Cout, hipStream_t s);
mcpi_status set_batched_kernel(int kind);

// One library stream (default priority) for the Lanczos entry points, which
// capture their iteration as a hipGraph (the legacy NULL stream cannot be
// captured).  It does not synchronize with the NULL stream by itself, so
// every call first makes it wait for the work already enqueued there (a
// start vector just written by mcpi_hip_applyH(..., NULL), say).  The calls
// that use it hold g_lz_mu: two threads must not enqueue onto it while one
// of them captures.
static std::mutex g_lz_mu;

static mcpi_status lib_stream(hipStream_t *out)
{
    static hipStream_t s = nullptr;
    static hipEvent_t ev = nullptr;
    static std::once_flag once;
    static hipError_t init = hipSuccess;
    std::call_once(once, [] {
        init = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (init == hipSuccess) init = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    });
    if (init != hipSuccess) return hip_fail(init, "library stream", __FILE__, __LINE__);
    MCPI_HIP(hipEventRecord(ev, nullptr));
    MCPI_HIP(hipStreamWaitEvent(s, ev, 0));
    *out = s;
    return MCPI_OK;
}

}  // namespace mcpi

using namespace mcpi;

// ------------------------------------------------------------ extended --
extern "C" {

const char *mcpi_hip_last_error(void) { return g_err.c_str(); }
const char *mcpi_hip_version(void) { return "mcpi_hip 0.1 (gfx950)"; }

mcpi_status mcpi_hip_malloc(void **dptr, size_t bytes)
{
    MCPI_ARG(dptr, "null pointer");
    MCPI_HIP(hipMalloc(dptr, bytes ? bytes : 16));
    return MCPI_OK;
}
mcpi_status mcpi_hip_free(void *dptr)
{
    if (dptr) MCPI_HIP(hipFree(dptr));
    return MCPI_OK;
}
mcpi_status mcpi_hip_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    MCPI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MCPI_OK;
}
mcpi_status mcpi_hip_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    MCPI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MCPI_OK;
}
mcpi_status mcpi_hip_memset(void *dptr, int value, size_t bytes)
{
    MCPI_HIP(hipMemset(dptr, value, bytes));
    return MCPI_OK;
}
mcpi_status mcpi_hip_synchronize(void)
{
    MCPI_HIP(hipDeviceSynchronize());
    return MCPI_OK;
}
mcpi_status mcpi_hip_device_info(char *name, size_t name_len, int *cu_count, size_t *total_mem)
{
    int dev = 0;
    MCPI_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    MCPI_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (total_mem) *total_mem = p.totalGlobalMem;
    return MCPI_OK;
}

mcpi_status mcpi_hip_ham_create(mcpi_basis b, const mcpi_cplx *Ho, const mcpi_cplx *Hint, mcpi_ham *out)
{
    MCPI_ARG(b && Ho && Hint && out, "null argument");
    const int M = b->d.M;
    mcpi_ham h = new mcpi_ham_s();
    memset(&h->c, 0, sizeof h->c);
    h->basis = b;
    h->skip_zero = 1;
    h->schedule = 0;
    h->tuned[0] = h->tuned[1] = 0;
    h->Ho_dev = h->Hint_dev = nullptr;
    if (b->sectored()) {
        // the Hamiltonian must not couple two sectors
        const int32_t *q = b->charge.data();
        const cplx *ho = (const cplx *)Ho, *hi = (const cplx *)Hint;
        bool ok = true;
        for (int k = 0; k < M && ok; k++)
            for (int l = 0; l < M && ok; l++) {
                const cplx z = ho[k + (size_t)M * l];
                ok = (z.re == 0 && z.im == 0) || q[k] == q[l];
            }
        for (size_t x = 0; x < (size_t)M * M * M * M && ok; x++) {
            const cplx z = hi[x];
            const int a = (int)(x % M), c = (int)(x / M % M), e = (int)(x / M / M % M), f = (int)(x / M / M / M);
            ok = (z.re == 0 && z.im == 0) || q[a] + q[c] == q[e] + q[f];
        }
        if (!ok) {
            delete h;
            MCPI_ARG(false, "the Hamiltonian does not conserve the charge of this sector-ordered basis");
        }
    }
    mcpi_status st = build_coefficients(M, (const cplx *)Ho, (const cplx *)Hint, &h->c, &h->coef_block,
                                        &h->coef_bytes);
    // sector order: the zero-skipping kernel on the pattern tables is the one
    // H.v, so its term lists are walked even when most terms are non-zero
    if (st == MCPI_OK && b->sectored()) h->c.dense = 0;
    if (st == MCPI_OK) {
        const size_t nHo = (size_t)M * M, nHi = nHo * nHo;
        hipError_t e = hipMalloc(&h->Ho_dev, sizeof(cplx) * nHo);
        if (e == hipSuccess) e = hipMalloc(&h->Hint_dev, sizeof(cplx) * nHi);
        if (e == hipSuccess) e = hipMemcpy(h->Ho_dev, Ho, sizeof(cplx) * nHo, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->Hint_dev, Hint, sizeof(cplx) * nHi, hipMemcpyHostToDevice);
        if (e != hipSuccess) st = hip_fail(e, "Hamiltonian upload", __FILE__, __LINE__);
    }
    if (st != MCPI_OK) {
        mcpi_hip_ham_destroy(h);
        return st;
    }
    *out = h;
    return MCPI_OK;
}

mcpi_status mcpi_hip_ham_destroy(mcpi_ham h)
{
    if (!h) return MCPI_OK;
    (void)hipDeviceSynchronize();
    if (h->coef_block) (void)hipFree(h->coef_block);
    if (h->Ho_dev) (void)hipFree(h->Ho_dev);
    if (h->Hint_dev) (void)hipFree(h->Hint_dev);
    delete h;
    return MCPI_OK;
}

mcpi_status mcpi_hip_ham_set_skip_zero(mcpi_ham h, int on)
{
    MCPI_ARG(h, "null handle");
    MCPI_ARG(on || !h->basis->sectored(), "a sector-ordered basis evaluates the non-zero terms only");
    h->skip_zero = on ? 1 : 0;
    return MCPI_OK;
}

mcpi_status mcpi_hip_ham_set_schedule(mcpi_ham h, int schedule)
{
    MCPI_ARG(h, "null handle");
    MCPI_ARG(schedule >= 0 && schedule <= 7, "schedule must be 0 (auto), 1, 2, 3, 4, 5, 6 or 7");
    MCPI_ARG(schedule == 0 || schedule == 4 || !h->basis->sectored(),
             "a sector-ordered basis runs schedule 4 (pattern tables) only");
    h->schedule = schedule;
    return MCPI_OK;
}

mcpi_status mcpi_hip_ham_get_schedule(mcpi_ham h, int *schedule)
{
    MCPI_ARG(h && schedule, "null argument");
    *schedule = mcpi::ham_schedule(h);
    return MCPI_OK;
}

mcpi_status mcpi_hip_applyH(mcpi_ham h, const mcpi_cplx *C_dev, mcpi_cplx *out_dev, void *stream)
{
    MCPI_ARG(h && C_dev && out_dev, "null argument");
    MCPI_ARG((const void *)C_dev != (const void *)out_dev, "input and output must not alias");
    const hipStream_t st = (hipStream_t)stream;
    mcpi_basis b = h->basis;
    if (!b->sectored()) return launch_apply(h, (const cplx *)C_dev, (cplx *)out_dev, st);
    // reference order in and out: gather into row order, H.v, scatter back
    cplx *v0 = nullptr, *v1 = nullptr;
    MCPI_CHECK(basis_vectors(b, &v0, &v1));
    MCPI_CHECK(gather_rows(b->d.nc, b->iperm, (const cplx *)C_dev, v0, st));
    MCPI_CHECK(launch_apply(h, v0, v1, st));
    return scatter_rows(b->d.nc, b->iperm, v1, (cplx *)out_dev, st);
}

mcpi_status mcpi_hip_applyH_rows(mcpi_ham h, const mcpi_cplx *C_dev, mcpi_cplx *out_dev, void *stream)
{
    MCPI_ARG(h && C_dev && out_dev, "null argument");
    MCPI_ARG((const void *)C_dev != (const void *)out_dev, "input and output must not alias");
    return launch_apply(h, (const cplx *)C_dev, (cplx *)out_dev, (hipStream_t)stream);
}

mcpi_status mcpi_hip_conserved_charge(int M, const mcpi_cplx *Ho, const mcpi_cplx *Hint, int32_t *charge)
{
    MCPI_ARG(M >= 1 && M <= MCPI_MAXM && Ho && Hint && charge, "bad argument");
    // candidate: the orbital index (plane-wave momentum up to a shift)
    const cplx *ho = (const cplx *)Ho, *hi = (const cplx *)Hint;
    bool ok = true;
    for (int k = 0; k < M && ok; k++)
        for (int l = 0; l < M && ok; l++) {
            const cplx z = ho[k + (size_t)M * l];
            ok = (z.re == 0 && z.im == 0) || k == l;
        }
    for (size_t x = 0; x < (size_t)M * M * M * M && ok; x++) {
        const cplx z = hi[x];
        const int a = (int)(x % M), c = (int)(x / M % M), e = (int)(x / M / M % M), f = (int)(x / M / M / M);
        ok = (z.re == 0 && z.im == 0) || a + c == e + f;
    }
    if (!ok) {
        set_error("no conserved charge found: the couplings do not conserve the orbital-index sum");
        return MCPI_ERR_STATE;
    }
    for (int j = 0; j < M; j++) charge[j] = j;
    return MCPI_OK;
}

mcpi_status mcpi_hip_applyH_timed(mcpi_ham h, const mcpi_cplx *C_dev, mcpi_cplx *out_dev, int reps,
                                  void *stream, double *avg_ms)
{
    MCPI_ARG(h && C_dev && out_dev && reps >= 1 && avg_ms, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    MCPI_HIP(hipEventCreate(&e0));
    MCPI_HIP(hipEventCreate(&e1));
    MCPI_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) MCPI_CHECK(launch_apply(h, (const cplx *)C_dev, (cplx *)out_dev, s));
    MCPI_HIP(hipEventRecord(e1, s));
    MCPI_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MCPI_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    MCPI_HIP(hipEventDestroy(e0));
    MCPI_HIP(hipEventDestroy(e1));
    return MCPI_OK;
}

mcpi_status mcpi_hip_applyH_timed_each(mcpi_ham h, const mcpi_cplx *C_dev, mcpi_cplx *out_dev, int reps,
                                       void *stream, double *ms_each)
{
    MCPI_ARG(h && C_dev && out_dev && reps >= 1 && ms_each, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::vector<hipEvent_t> ev(reps + 1, nullptr);
    mcpi_status st = MCPI_OK;
    for (int r = 0; r <= reps && st == MCPI_OK; r++)
        if (hipEventCreate(&ev[r]) != hipSuccess) st = hip_fail(hipGetLastError(), "hipEventCreate", __FILE__, __LINE__);
    for (int r = 0; r < reps && st == MCPI_OK; r++) {
        if (hipEventRecord(ev[r], s) != hipSuccess) st = hip_fail(hipGetLastError(), "hipEventRecord", __FILE__, __LINE__);
        if (st == MCPI_OK) st = launch_apply(h, (const cplx *)C_dev, (cplx *)out_dev, s);
    }
    if (st == MCPI_OK && (hipEventRecord(ev[reps], s) != hipSuccess || hipEventSynchronize(ev[reps]) != hipSuccess))
        st = hip_fail(hipGetLastError(), "event", __FILE__, __LINE__);
    for (int r = 0; r < reps && st == MCPI_OK; r++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[r], ev[r + 1]) != hipSuccess)
            st = hip_fail(hipGetLastError(), "hipEventElapsedTime", __FILE__, __LINE__);
        ms_each[r] = ms;
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    return st;
}

}  // extern "C"

namespace {

// State of the CUDA-shaped H.v kept between calls: the rank table and NCmat
// of the last (N, M), the folded coefficients of the last Ho/Hint (reused
// while the caller's device arrays hold the same values entry for entry)
// and int64 copies of int32 strides.  One call at a time (g_dev_mu).
struct DevShapeCache {
    int N = -1, M = -1;
    int32_t *B = nullptr, *NCmat = nullptr;
    std::vector<cplx> ho, hint;
    HamCoefDev c{};
    void *block = nullptr;
    int64_t *w_ot = nullptr, *w_tt = nullptr;
    int64_t nw = 0;
};
DevShapeCache g_devc;
std::mutex g_dev_mu;

mcpi_status dev_shape_apply(int N, int M, const int32_t *Map, const int32_t *MapOT, const int32_t *MapTT,
                            const int64_t *strideOT, const int64_t *strideTT, const int32_t *strideOT32,
                            const int32_t *strideTT32, const int32_t *IF, const cplx *C, const cplx *Ho,
                            const cplx *Hint, cplx *out, hipStream_t s)
{
    MCPI_ARG(M >= 1 && M <= MCPI_MAXM && N >= 0 && N <= MCPI_MAXN, "N or M out of range");
    const int64_t nc = nc_of(N, M);
    MCPI_ARG(nc > 0 && nc <= INT32_MAX, "basis too large");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    DevShapeCache &g = g_devc;
    if (g.N != N || g.M != M) {
        MCPI_HIP(hipDeviceSynchronize());   // earlier launches may still read the old tables
        if (g.B) (void)hipFree(g.B);
        if (g.NCmat) (void)hipFree(g.NCmat);
        g.B = g.NCmat = nullptr;
        g.N = g.M = -1;
        MCPI_CHECK(rank_tables_dev(N, M, &g.B, &g.NCmat));
        g.N = N;
        g.M = M;
    }
    // Ho and Hint are device arrays: read them in stream order and fold them
    // on the host when they differ from the last call's
    const size_t nHo = (size_t)M * M, nHi = nHo * nHo;
    std::vector<cplx> ho(nHo), hi(nHi);
    MCPI_HIP(hipMemcpyAsync(ho.data(), Ho, sizeof(cplx) * nHo, hipMemcpyDeviceToHost, s));
    MCPI_HIP(hipMemcpyAsync(hi.data(), Hint, sizeof(cplx) * nHi, hipMemcpyDeviceToHost, s));
    MCPI_HIP(hipStreamSynchronize(s));
    if (!g.block || g.ho.size() != nHo || memcmp(g.ho.data(), ho.data(), sizeof(cplx) * nHo) != 0 ||
        memcmp(g.hint.data(), hi.data(), sizeof(cplx) * nHi) != 0) {
        MCPI_HIP(hipDeviceSynchronize());
        if (g.block) (void)hipFree(g.block);
        g.block = nullptr;
        g.ho.clear();
        g.hint.clear();
        MCPI_CHECK(build_coefficients(M, ho.data(), hi.data(), &g.c, &g.block));
        g.ho = std::move(ho);
        g.hint = std::move(hi);
    }
    BasisDev b;
    memset(&b, 0, sizeof b);
    b.N = N;
    b.M = M;
    b.nc = nc;
    b.IF = (int32_t *)IF;
    b.Map = (int32_t *)Map;
    b.MapOT = (int32_t *)MapOT;
    b.MapTT = (int32_t *)MapTT;
    b.B = g.B;
    b.NCmat = g.NCmat;
    // every term evaluated, as the CUDA kernel does.  With a kernel
    // specialised on M: single jumps read from Map, double-jump targets
    // ranked from the occupations (the entries MapOT / MapTT hold);
    // otherwise the generic kernel reads every map and the strides.
    if (has_fast_kernel(N, M)) return launch_apply_raw(b, MCPI_MAPS_FULL, g.c, 0, 3, C, out, s);
    if (strideOT32) {
        if (g.nw < nc) {
            MCPI_HIP(hipDeviceSynchronize());
            if (g.w_ot) (void)hipFree(g.w_ot);
            if (g.w_tt) (void)hipFree(g.w_tt);
            g.w_ot = g.w_tt = nullptr;
            g.nw = 0;
            MCPI_HIP(hipMalloc(&g.w_ot, sizeof(int64_t) * nc));
            MCPI_HIP(hipMalloc(&g.w_tt, sizeof(int64_t) * nc));
            g.nw = nc;
        }
        MCPI_CHECK(widen_strides(nc, strideOT32, g.w_ot, s));
        MCPI_CHECK(widen_strides(nc, strideTT32, g.w_tt, s));
        strideOT = g.w_ot;
        strideTT = g.w_tt;
    }
    b.strideOT = (int64_t *)strideOT;
    b.strideTT = (int64_t *)strideTT;
    return launch_apply_raw(b, MCPI_MAPS_FULL, g.c, 0, 1, C, out, s);
}

}  // namespace

extern "C" {

mcpi_status mcpi_hip_applyHconf_dev(int N, int M, const int32_t *Map, const int32_t *MapOT,
                                    const int32_t *MapTT, const int64_t *strideOT, const int64_t *strideTT,
                                    const int32_t *IF, const mcpi_cplx *C, const mcpi_cplx *Ho,
                                    const mcpi_cplx *Hint, mcpi_cplx *out, void *stream)
{
    MCPI_ARG(Map && MapOT && MapTT && strideOT && strideTT && IF && C && Ho && Hint && out, "null argument");
    return dev_shape_apply(N, M, Map, MapOT, MapTT, strideOT, strideTT, nullptr, nullptr, IF, (const cplx *)C,
                           (const cplx *)Ho, (const cplx *)Hint, (cplx *)out, (hipStream_t)stream);
}

mcpi_status mcpi_hip_applyHconf_dev32(int N, int M, const int32_t *Map, const int32_t *MapOT,
                                      const int32_t *MapTT, const int32_t *strideOT, const int32_t *strideTT,
                                      const int32_t *IF, const mcpi_cplx *C, const mcpi_cplx *Ho,
                                      const mcpi_cplx *Hint, mcpi_cplx *out, void *stream)
{
    MCPI_ARG(Map && MapOT && MapTT && strideOT && strideTT && IF && C && Ho && Hint && out, "null argument");
    return dev_shape_apply(N, M, Map, MapOT, MapTT, nullptr, nullptr, strideOT, strideTT, IF, (const cplx *)C,
                           (const cplx *)Ho, (const cplx *)Hint, (cplx *)out, (hipStream_t)stream);
}

mcpi_status mcpi_hip_lanczos(mcpi_ham h, int lm, const mcpi_cplx *C0_dev, mcpi_cplx *lvec_dev,
                             mcpi_cplx *diag, mcpi_cplx *offdiag, int *steps)
{
    MCPI_ARG(h && C0_dev && lvec_dev && diag && offdiag && steps && lm >= 1, "bad argument");
    std::lock_guard<std::mutex> lk(g_lz_mu);
    hipStream_t s = nullptr;
    MCPI_CHECK(lib_stream(&s));
    return lanczos_api(h, lm, (const cplx *)C0_dev, (cplx *)lvec_dev, (cplx *)diag, (cplx *)offdiag, steps, s);
}

mcpi_status mcpi_hip_ground(mcpi_ham h, int lm, const mcpi_cplx *C0_dev, mcpi_cplx *C_out_dev, double *E0,
                            int *steps)
{
    MCPI_ARG(h && C0_dev && C_out_dev && E0 && lm >= 1, "bad argument");
    std::lock_guard<std::mutex> lk(g_lz_mu);
    hipStream_t s = nullptr;
    MCPI_CHECK(lib_stream(&s));
    return ground_run(h, lm, (const cplx *)C0_dev, (cplx *)C_out_dev, E0, steps, s, nullptr);
}

mcpi_status mcpi_hip_tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z)
{
    return tridiag_lowest(n, d, e, lambda, z);
}

mcpi_status mcpi_hip_set_batched_kernel(int kind) { return set_batched_kernel(kind); }

mcpi_status mcpi_hip_lanczos_batched(int n, const mcpi_ham *hams, const mcpi_cplx *const *C0_dev, int lm,
                                     double *E0, int *steps, mcpi_cplx *const *C_out_dev)
{
    std::lock_guard<std::mutex> lk(g_lz_mu);
    hipStream_t s = nullptr;
    MCPI_CHECK(lib_stream(&s));
    return lanczos_batched_api(n, hams, (const cplx *const *)C0_dev, lm, E0, steps, (cplx *const *)C_out_dev, s);
}

}  // extern "C"

// ------------------------------------------------ reference-shaped API --
// These keep the reference's argument meaning and its failure behaviour
// (message + exit on the INT32 guards and on allocation failures).

namespace {

[[noreturn]] void ref_die(const char *msg)
{
    printf("\n\n\n%s\n\n", msg);
    fflush(stdout);
    exit(EXIT_FAILURE);
}

void ref_check(mcpi_status st)
{
    if (st != MCPI_OK) {
        std::string m = "GPU ERROR : " + g_err;
        ref_die(m.c_str());
    }
}

int *ref_ialloc(int64_t n)
{
    int *p = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!p) ref_die("MEMORY ERROR : malloc fail for integer.");
    return p;
}

// NC with the reference's overflow predicate: the interleaved
// multiply/exact-divide sequence of configurationsMap.h:155-201, guarded by
// n > INT_MAX / i before every multiply (same predicate, int64 arithmetic).
int ref_nc(int N, int M)
{
    long n = 1, j = 2;
    if (M > N) {
        for (long i = N + M - 1; i > M - 1; i--) {
            if (n > INT_MAX / i)
                ref_die("INTEGER SIZE ERROR : overflow occurred representing the number of configurations as integers of 32 bits");
            n = n * i;
            if (n % j == 0 && j <= N) { n = n / j; j = j + 1; }
        }
        for (long i = j; i < N + 1; i++) n = n / i;
        return (int)n;
    }
    for (long i = N + M - 1; i > N; i--) {
        if (n > INT_MAX / i)
            ref_die("INTEGER SIZE ERROR : overflow occurred representing the number of configurations as integers of 32 bits");
        n = n * i;
        if (n % j == 0 && j <= M - 1) { n = n / j; j = j + 1; }
    }
    for (long i = j; i < M; i++) n = n / i;
    return (int)n;
}

void count_chunks(int M, int nc, const int *IF, int64_t *n_ot, int64_t *n_tt)
{
    int64_t a = 0, b = 0;
    for (int64_t i = 0; i < nc; i++) {
        int n1 = 0;
        for (int k = 0; k < M; k++) {
            n1 += IF[i * M + k] >= 1;
            a += IF[i * M + k] >= 2;
        }
        b += n1 * (n1 - 1) / 2;
    }
    *n_ot = a;
    *n_tt = b;
}

std::vector<cplx> flat_ho(int M, Cmatrix Ho)
{
    std::vector<cplx> f((size_t)M * M);
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++) memcpy(&f[k + (size_t)M * l], &Ho[k][l], sizeof(cplx));
    return f;
}

// Device state of the reference-shaped entry points, kept between calls so
// that a reference program calling applyHconf_omp inside its own Lanczos
// loop (groundStateLieb.c:297) or a timing loop (performanceTest.c) pays the
// table upload, the rule folding and the H.v schedule tuning once:
//  * the device basis built from the caller's tables, reused while the same
//    host tables are passed again (same N, M, pointers and a sampled
//    fingerprint of their contents; tables are taken to be unchanged while
//    they are passed by the same pointers -- mcpi_hip_release_cached drops
//    everything after a caller edits tables in place);
//  * the Hamiltonian of the last call on that basis with its tuned
//    schedule, reused while Ho and Hint are equal entry for entry to the
//    copies kept here;
//  * two device vectors of nc entries.
// The entry points run one at a time (g_ref_mu, held for the whole call:
// nothing cached may be replaced while a call still uses it).
struct RefCache {
    mcpi_basis b = nullptr;
    int N = -1, M = -1;
    const void *key[8] = {};
    uint64_t fp = 0;
    mcpi_ham h = nullptr;
    std::vector<cplx> ho, hint;   // flat copies the cached Hamiltonian was built from
    cplx *dC = nullptr, *dO = nullptr;
    int64_t nvec = 0;
};
RefCache g_cache;
std::recursive_mutex g_ref_mu;

void drop_cache()
{
    if (g_cache.h) mcpi_hip_ham_destroy(g_cache.h);
    if (g_cache.dC) (void)hipFree(g_cache.dC);
    if (g_cache.dO) (void)hipFree(g_cache.dO);
    if (g_cache.b) mcpi_hip_basis_destroy(g_cache.b);
    g_cache = RefCache();
}

uint64_t fingerprint(const int *a, int64_t n)
{
    uint64_t h = 1469598103934665603ull;
    if (!a || n <= 0) return h;
    int64_t step = std::max<int64_t>(1, n / 64);
    for (int64_t x = 0; x < n; x += step) h = (h ^ (uint64_t)(uint32_t)a[x]) * 1099511628211ull;
    h = (h ^ (uint64_t)(uint32_t)a[n - 1]) * 1099511628211ull;
    return h;
}

// IF == nullptr: a basis generated on the device from (N, M) alone
mcpi_basis ref_basis(int N, int M, const int *IF, const int *NCmat, const int *Map, const int *MapOT,
                     const int *MapTT, const int *strideOT, const int *strideTT)
{
    const int nc = ref_nc(N, M);
    const int64_t MM = (int64_t)M * M;
    int64_t n_ot = 0, n_tt = 0;
    uint64_t fp = 0;
    if (IF) {
        count_chunks(M, nc, IF, &n_ot, &n_tt);
        fp = fingerprint(IF, (int64_t)nc * M) ^ (fingerprint(Map, Map ? nc * MM : 0) * 3) ^
             (fingerprint(MapOT, MapOT ? n_ot * MM : 0) * 5) ^ (fingerprint(MapTT, MapTT ? n_tt * MM : 0) * 7) ^
             (fingerprint(strideOT, strideOT ? nc : 0) * 11) ^ (fingerprint(strideTT, strideTT ? nc : 0) * 13);
    }
    const void *key[8] = {IF, NCmat, Map, MapOT, MapTT, strideOT, strideTT, IF ? nullptr : (const void *)&g_cache};
    // the caller holds g_ref_mu
    if (g_cache.b && g_cache.N == N && g_cache.M == M && g_cache.fp == fp &&
        memcmp(g_cache.key, key, sizeof key) == 0)
        return g_cache.b;
    drop_cache();
    mcpi_basis b = nullptr;
    if (IF)
        ref_check(basis_upload_raw(N, M, IF, NCmat, Map, MapOT, MapTT, strideOT, strideTT, n_ot * MM, n_tt * MM, &b));
    else
        ref_check(mcpi_hip_basis_create(N, M, MCPI_MAPS_NONE, &b));
    g_cache.b = b;
    g_cache.N = N;
    g_cache.M = M;
    g_cache.fp = fp;
    memcpy(g_cache.key, key, sizeof key);
    return b;
}

// the Hamiltonian of (Ho, Hint) on the cached basis b
mcpi_ham ref_ham(mcpi_basis b, Cmatrix Ho, const Carray Hint)
{
    const int M = b->d.M;
    std::vector<cplx> ho = flat_ho(M, Ho);
    const size_t nhi = (size_t)M * M * M * M;
    if (g_cache.h && g_cache.h->basis == b && g_cache.hint.size() == nhi &&
        memcmp(g_cache.ho.data(), ho.data(), sizeof(cplx) * ho.size()) == 0 &&
        memcmp(g_cache.hint.data(), Hint, sizeof(cplx) * nhi) == 0)
        return g_cache.h;
    if (g_cache.h) mcpi_hip_ham_destroy(g_cache.h);
    g_cache.h = nullptr;
    mcpi_ham h = nullptr;
    ref_check(mcpi_hip_ham_create(b, (const mcpi_cplx *)ho.data(), Hint, &h));
    g_cache.h = h;
    g_cache.ho = std::move(ho);
    g_cache.hint.assign((const cplx *)Hint, (const cplx *)Hint + nhi);
    return h;
}

// the two cached device vectors, nc entries each
void ref_vectors(int64_t nc, cplx **dC, cplx **dO)
{
    if (g_cache.nvec != nc) {
        if (g_cache.dC) (void)hipFree(g_cache.dC);
        if (g_cache.dO) (void)hipFree(g_cache.dO);
        g_cache.dC = g_cache.dO = nullptr;
        g_cache.nvec = 0;
        const size_t bytes = sizeof(cplx) * (size_t)std::max<int64_t>(nc, 1);
        if (hipMalloc(&g_cache.dC, bytes) != hipSuccess || hipMalloc(&g_cache.dO, bytes) != hipSuccess)
            ref_die("GPU MEMORY ERROR : hipMalloc fail for complex");
        g_cache.nvec = nc;
    }
    *dC = g_cache.dC;
    if (dO) *dO = g_cache.dO;
}

void ref_upload(cplx *d, const Carray C, int64_t nc)
{
    if (hipMemcpy(d, C, sizeof(cplx) * nc, hipMemcpyHostToDevice) != hipSuccess) ref_die("GPU copy failed");
}

void ref_apply(mcpi_basis b, const Carray C, Cmatrix Ho, const Carray Hint, Carray out)
{
    const int64_t nc = b->d.nc;
    mcpi_ham h = ref_ham(b, Ho, Hint);
    cplx *dC = nullptr, *dO = nullptr;
    ref_vectors(nc, &dC, &dO);
    ref_upload(dC, C, nc);
    ref_check(launch_apply(h, dC, dO, nullptr));
    if (hipMemcpy(out, dO, sizeof(cplx) * nc, hipMemcpyDeviceToHost) != hipSuccess) ref_die("GPU copy failed");
}

void ref_rho1(mcpi_basis b, const Carray C, Cmatrix rho)
{
    const int M = b->d.M;
    const int64_t nc = b->d.nc;
    cplx *dC = nullptr;
    ref_vectors(nc, &dC, nullptr);
    ref_upload(dC, C, nc);
    std::vector<cplx> r((size_t)M * M);
    ref_check(mcpi_hip_rho1(b, (const mcpi_cplx *)dC, (mcpi_cplx *)r.data()));
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++) memcpy(&rho[k][l], &r[k + (size_t)M * l], sizeof(cplx));
}

void ref_rho2(mcpi_basis b, const Carray C, Carray rho)
{
    const int64_t nc = b->d.nc;
    cplx *dC = nullptr;
    ref_vectors(nc, &dC, nullptr);
    ref_upload(dC, C, nc);
    ref_check(mcpi_hip_rho2(b, (const mcpi_cplx *)dC, rho));
}

}  // namespace

extern "C" {

int NC(int N, int M) { return ref_nc(N, M); }

Iarray setupNCmat(int N, int M)
{
    int *t = ref_ialloc((int64_t)(M + 1) * (N + 1));
    for (int i = 0; i < N + 1; i++) {
        t[i + (N + 1) * 0] = 0;
        t[i + (N + 1) * 1] = 1;
        for (int j = 2; j < M + 1; j++) t[i + (N + 1) * j] = ref_nc(i, j);
    }
    return t;
}

void IndexToFock(int k, int N, int M, Iarray v)
{
    int m = M - 1;
    for (int i = 0; i < M; i++) v[i] = 0;
    while (k > 0) {
        while (k - ref_nc(N, m) < 0) m--;
        k -= ref_nc(N, m);
        v[m]++;
        N--;
    }
    if (N > 0) v[0] += N;
}

int FockToIndex(int N, int M, Iarray NCmat, Iarray v)
{
    // Closed form of the reference's loop: the particles of orbital i add
    // NC(n, i) for n from P_i + 1 to P_{i+1} (P_i = particles below orbital
    // i), and by the hockey-stick identity that sum is
    // NC(P_{i+1}, i+1) - NC(P_i, i+1), read from the caller's table.
    const int col = N + 1;
    int k = 0, below = v[0];
    for (int i = 1; i < M; i++) {
        const int upto = below + v[i];
        k += NCmat[upto + col * (i + 1)] - NCmat[below + col * (i + 1)];
        below = upto;
    }
    return k;
}

Iarray setupFocks(int N, int M)
{
    const int nc = ref_nc(N, M);
    if (nc > INT_MAX / M)
        ref_die("MEMORY ERROR : Because of the size of the hashing table it can't be indexed by 32-bit integers");
    mcpi_basis b = nullptr;
    ref_check(mcpi_hip_basis_create(N, M, MCPI_MAPS_NONE, &b));
    int *IF = ref_ialloc((int64_t)nc * M);
    ref_check(mcpi_hip_basis_copy_if(b, IF));
    mcpi_hip_basis_destroy(b);
    return IF;
}

Iarray OneOneMap(int N, int M, Iarray NCmat, Iarray IF)
{
    const int nc = ref_nc(N, M);
    if (nc > INT_MAX / M / M)
        ref_die("MEMORY ERROR : Because of the size of the jump-mappings they can't be indexed by 32-bit integers");
    mcpi_basis b = nullptr;
    ref_check(basis_upload_raw(N, M, IF, NCmat, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, &b));
    ref_check(basis_build_parts(b, 1, 0, 0));
    int *Map = ref_ialloc((int64_t)nc * M * M);
    ref_check(mcpi_hip_basis_copy_map(b, Map));
    mcpi_hip_basis_destroy(b);
    return Map;
}

static Iarray ref_double_map(int N, int M, Iarray NCmat, Iarray IF, Iarray strideC, bool tt)
{
    const int nc = ref_nc(N, M);
    int64_t n_ot = 0, n_tt = 0;
    count_chunks(M, nc, IF, &n_ot, &n_tt);
    int64_t chunks = tt ? n_tt : n_ot;
    if (chunks > INT_MAX / M / M)
        ref_die("MEMORY ERROR : Because of the size of the jump-mappings they can't be indexed by 32-bit integers");
    mcpi_basis b = nullptr;
    ref_check(basis_upload_raw(N, M, IF, NCmat, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, &b));
    ref_check(basis_build_parts(b, 0, tt ? 0 : 1, tt ? 1 : 0));
    int *map = ref_ialloc(chunks * M * M);
    std::vector<int64_t> s64(nc);
    if (tt) ref_check(mcpi_hip_basis_copy_tt(b, s64.data(), map));
    else ref_check(mcpi_hip_basis_copy_ot(b, s64.data(), map));
    for (int i = 0; i < nc; i++) strideC[i] = (int)s64[i];
    mcpi_hip_basis_destroy(b);
    return map;
}

Iarray TwoTwoMap(int N, int M, Iarray NCmat, Iarray IF, Iarray strideC)
{
    return ref_double_map(N, M, NCmat, IF, strideC, true);
}

Iarray OneTwoMap(int N, int M, Iarray NCmat, Iarray IF, Iarray strideC)
{
    return ref_double_map(N, M, NCmat, IF, strideC, false);
}

void applyHconf_omp(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT, Iarray strideOT, Iarray strideTT,
                    Iarray IF, Carray C, Cmatrix Ho, Carray Hint, Carray out)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(N, M, IF, nullptr, Map, MapOT, MapTT, strideOT, strideTT);
    ref_apply(b, C, Ho, Hint, out);
}

void applyHconf_XX(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT, Iarray strideOT, Iarray strideTT,
                   Iarray IF, Carray C, Cmatrix Ho, Carray Hint, Carray out)
{
    applyHconf_omp(N, M, Map, MapOT, MapTT, strideOT, strideTT, IF, C, Ho, Hint, out);
}

void applyHconf_X(int N, int M, Iarray Map, Iarray MapOT, Iarray strideOT, Iarray NCmat, Iarray IF, Carray C,
                  Cmatrix Ho, Carray Hint, Carray out)
{
    // the two-orbital double jumps are ranked on the fly, as the reference
    // does here (hamiltonianMatrix.h:730-880); the same-orbital ones too
    // (same targets as MapOT holds)
    (void)MapOT;
    (void)strideOT;
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(N, M, IF, NCmat, Map, nullptr, nullptr, nullptr, nullptr);
    ref_apply(b, C, Ho, Hint, out);
}

void applyHconf(int N, int M, Iarray NCmat, Iarray IF, Carray C, Cmatrix Ho, Carray Hint, Carray out)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(N, M, IF, NCmat, nullptr, nullptr, nullptr, nullptr, nullptr);
    ref_apply(b, C, Ho, Hint, out);
}

void OBrho(int N, int M, Iarray NCmat, Carray C, Cmatrix rho)
{
    // no table at all: the basis is generated on the device from (N, M)
    (void)NCmat;
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    ref_rho1(ref_basis(N, M, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr), C, rho);
}

void OBrho_X(int N, int M, Iarray NCmat, Iarray IF, Carray C, Cmatrix rho)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    ref_rho1(ref_basis(N, M, IF, NCmat, nullptr, nullptr, nullptr, nullptr, nullptr), C, rho);
}

void OBrho_XM(int N, int M, Iarray Map, Iarray NCmat, Iarray IF, Carray C, Cmatrix rho)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    ref_rho1(ref_basis(N, M, IF, NCmat, Map, nullptr, nullptr, nullptr, nullptr), C, rho);
}

void TBrho(int N, int M, Iarray NCmat, Iarray IF, Carray C, Carray rho)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    ref_rho2(ref_basis(N, M, IF, NCmat, nullptr, nullptr, nullptr, nullptr, nullptr), C, rho);
}

void TBrho_X(int N, int M, Iarray Map, Iarray MapOT, Iarray strideOT, Iarray NCmat, Iarray IF, Carray C,
             Carray rho)
{
    (void)MapOT;
    (void)strideOT;
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    ref_rho2(ref_basis(N, M, IF, NCmat, Map, nullptr, nullptr, nullptr, nullptr), C, rho);
}

void TBrho_XX(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT, Iarray strideOT, Iarray strideTT,
              Iarray NCmat, Iarray IF, Carray C, Carray rho)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    ref_rho2(ref_basis(N, M, IF, NCmat, Map, MapOT, MapTT, strideOT, strideTT), C, rho);
}

int lanczos(int lm, int Npar, int Morb, Iarray IF, Iarray strideOT, Iarray strideTT, Iarray Map, Iarray MapOT,
            Iarray MapTT, Cmatrix Ho, Carray Hint, Carray diag, Carray offdiag, Cmatrix lvec)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(Npar, Morb, IF, nullptr, Map, MapOT, MapTT, strideOT, strideTT);
    const int64_t nc = b->d.nc;
    mcpi_ham h = ref_ham(b, Ho, Hint);
    cplx *dl = nullptr;
    if (hipMalloc(&dl, sizeof(cplx) * nc * (size_t)lm) != hipSuccess) ref_die("GPU MEMORY ERROR : Lanczos basis");
    ref_upload(dl, lvec[0], nc);
    int steps = 0;
    ref_check(mcpi_hip_lanczos(h, lm, (const mcpi_cplx *)dl, (mcpi_cplx *)dl, diag, offdiag, &steps));
    for (int t = 0; t < steps; t++)
        if (hipMemcpy(lvec[t], dl + (size_t)t * nc, sizeof(cplx) * nc, hipMemcpyDeviceToHost) != hipSuccess)
            ref_die("GPU copy failed");
    (void)hipFree(dl);
    return steps;
}

double ground(int Niter, int Npar, int Morb, Iarray IF, Iarray strideOT, Iarray strideTT, Iarray Map,
              Iarray MapOT, Iarray MapTT, Cmatrix Ho, Carray Hint, Carray C)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(Npar, Morb, IF, nullptr, Map, MapOT, MapTT, strideOT, strideTT);
    const int64_t nc = b->d.nc;
    mcpi_ham h = ref_ham(b, Ho, Hint);
    cplx *dC = nullptr;
    ref_vectors(nc, &dC, nullptr);
    ref_upload(dC, C, nc);
    double E0 = 0;
    int steps = 0;
    ref_check(mcpi_hip_ground(h, Niter, (const mcpi_cplx *)dC, (mcpi_cplx *)dC, &E0, &steps));
    if (hipMemcpy(C, dC, sizeof(cplx) * nc, hipMemcpyDeviceToHost) != hipSuccess) ref_die("GPU copy failed");
    return E0;
}

mcpi_status mcpi_hip_release_cached(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    drop_cache();
    return MCPI_OK;
}

mcpi_cplx carrDot(int n, Carray v1, Carray v2)
{
    double re = 0, im = 0;
    const double *a = (const double *)v1, *b = (const double *)v2;
    for (int i = 0; i < n; i++) {
        re += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
        im += a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
    }
    mcpi_cplx z;
    double *zz = (double *)&z;
    zz[0] = re;
    zz[1] = im;
    return z;
}

double carrMod(int n, Carray v)
{
    double s = 0;
    const double *a = (const double *)v;
    for (int i = 0; i < n; i++) s += a[2 * i] * a[2 * i] + a[2 * i + 1] * a[2 * i + 1];
    return sqrt(s);
}

void setupHo(int Morb, double L, Cmatrix Ho)
{
    const double PI = 3.141592653589793;
    for (int n = 0; n < Morb; n++) {
        double k = 2 * PI * (n - Morb / 2) / L;
        double *d = (double *)&Ho[n][n];
        d[0] = 0.5 * k * k;
        d[1] = 0.0;
        for (int j = n + 1; j < Morb; j++) {
            double *a = (double *)&Ho[n][j], *b = (double *)&Ho[j][n];
            a[0] = a[1] = b[0] = b[1] = 0.0;
        }
    }
}

void setupHint(int Morb, double g, Carray Hint)
{
    const int M = Morb;
    double *h = (double *)Hint;
    for (int n = 0; n < M; n++)
        for (int s = 0; s < M; s++)
            for (int q = 0; q < M; q++)
                for (int l = 0; l < M; l++) {
                    int64_t x = n + (int64_t)M * (s + (int64_t)M * (q + (int64_t)M * l));
                    bool conserve = (l - M / 2) + (q - M / 2) - (s - M / 2) - (n - M / 2) == 0;
                    h[2 * x] = conserve ? g : 0.0;
                    h[2 * x + 1] = 0.0;
                }
}

}  // extern "C"
