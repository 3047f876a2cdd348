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

// one library stream (default priority) for calls that do not pass one
static hipStream_t lib_stream()
{
    static hipStream_t s = nullptr;
    static std::once_flag once;
    std::call_once(once, [] { (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking); });
    return s;
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
    mcpi_status st = build_coefficients(M, (const cplx *)Ho, (const cplx *)Hint, &h->c, &h->coef_block,
                                        &h->coef_bytes);
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
    h->skip_zero = on ? 1 : 0;
    return MCPI_OK;
}

mcpi_status mcpi_hip_ham_set_schedule(mcpi_ham h, int schedule)
{
    MCPI_ARG(h, "null handle");
    MCPI_ARG(schedule >= 0 && schedule <= 5, "schedule must be 0 (auto), 1, 2, 3, 4 or 5");
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
    return launch_apply(h, (const cplx *)C_dev, (cplx *)out_dev, (hipStream_t)stream);
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

mcpi_status mcpi_hip_applyHconf_dev(int N, int M, const int32_t *Map, const int32_t *MapOT,
                                    const int32_t *MapTT, const int64_t *strideOT, const int64_t *strideTT,
                                    const int32_t *IF, const mcpi_cplx *C, const mcpi_cplx *Ho,
                                    const mcpi_cplx *Hint, mcpi_cplx *out, void *stream)
{
    MCPI_ARG(Map && MapOT && MapTT && strideOT && strideTT && IF && C && Ho && Hint && out, "null argument");
    MCPI_ARG(M >= 1 && M <= MCPI_MAXM && N >= 0 && N <= MCPI_MAXN, "N or M out of range");
    const int64_t nc = nc_of(N, M);
    MCPI_ARG(nc > 0 && nc <= INT32_MAX, "basis too large");
    // Ho and Hint are device arrays here: fold them on the host
    std::vector<cplx> ho((size_t)M * M), hi((size_t)M * M * M * M);
    MCPI_HIP(hipMemcpy(ho.data(), Ho, sizeof(cplx) * ho.size(), hipMemcpyDeviceToHost));
    MCPI_HIP(hipMemcpy(hi.data(), Hint, sizeof(cplx) * hi.size(), hipMemcpyDeviceToHost));
    HamCoefDev c;
    void *block = nullptr;
    MCPI_CHECK(build_coefficients(M, ho.data(), hi.data(), &c, &block));
    BasisDev b;
    memset(&b, 0, sizeof b);
    b.N = N;
    b.M = M;
    b.nc = nc;
    b.IF = (int32_t *)IF;
    b.Map = (int32_t *)Map;
    b.MapOT = (int32_t *)MapOT;
    b.MapTT = (int32_t *)MapTT;
    b.strideOT = (int64_t *)strideOT;
    b.strideTT = (int64_t *)strideTT;
    mcpi_status st = launch_apply_raw(b, MCPI_MAPS_FULL, c, 0, 1, (const cplx *)C, (cplx *)out,
                                      (hipStream_t)stream);
    if (st == MCPI_OK) {
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) st = hip_fail(e, "H.v", __FILE__, __LINE__);
    }
    (void)hipFree(block);
    return st;
}

mcpi_status mcpi_hip_lanczos(mcpi_ham h, int lm, const mcpi_cplx *C0_dev, mcpi_cplx *lvec_dev,
                             mcpi_cplx *diag, mcpi_cplx *offdiag, int *steps)
{
    MCPI_ARG(h && C0_dev && lvec_dev && diag && offdiag && steps && lm >= 1, "bad argument");
    return lanczos_api(h, lm, (const cplx *)C0_dev, (cplx *)lvec_dev, (cplx *)diag, (cplx *)offdiag, steps,
                       lib_stream());
}

mcpi_status mcpi_hip_ground(mcpi_ham h, int lm, const mcpi_cplx *C0_dev, mcpi_cplx *C_out_dev, double *E0,
                            int *steps)
{
    MCPI_ARG(h && C0_dev && C_out_dev && E0 && lm >= 1, "bad argument");
    return ground_run(h, lm, (const cplx *)C0_dev, (cplx *)C_out_dev, E0, steps, lib_stream(), nullptr);
}

mcpi_status mcpi_hip_tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z)
{
    return tridiag_lowest(n, d, e, lambda, z);
}

mcpi_status mcpi_hip_set_batched_kernel(int kind) { return set_batched_kernel(kind); }

mcpi_status mcpi_hip_lanczos_batched(int n, const mcpi_ham *hams, const mcpi_cplx *const *C0_dev, int lm,
                                     double *E0, int *steps, mcpi_cplx *const *C_out_dev)
{
    return lanczos_batched_api(n, hams, (const cplx *const *)C0_dev, lm, E0, steps, (cplx *const *)C_out_dev,
                               lib_stream());
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

// device basis from caller tables, cached while the same host tables are
// passed again (a Lanczos loop calling applyHconf_omp every step).  The
// reference-shaped entry points that use it run one at a time (g_ref_mu,
// held for the whole call: the cached basis must not be replaced while a
// call still uses it).
struct RefCache {
    mcpi_basis b = nullptr;
    int N = -1, M = -1;
    const void *key[8] = {};
    uint64_t fp = 0;
};
RefCache g_cache;
std::recursive_mutex g_ref_mu;

uint64_t fingerprint(const int *a, int64_t n)
{
    uint64_t h = 1469598103934665603ull;
    if (!a || n <= 0) return h;
    int64_t step = std::max<int64_t>(1, n / 64);
    for (int64_t x = 0; x < n; x += step) h = (h ^ (uint64_t)(uint32_t)a[x]) * 1099511628211ull;
    h = (h ^ (uint64_t)(uint32_t)a[n - 1]) * 1099511628211ull;
    return h;
}

mcpi_basis ref_basis(int N, int M, const int *IF, const int *NCmat, const int *Map, const int *MapOT,
                     const int *MapTT, const int *strideOT, const int *strideTT)
{
    const int nc = ref_nc(N, M);
    int64_t n_ot = 0, n_tt = 0;
    count_chunks(M, nc, IF, &n_ot, &n_tt);
    const int64_t MM = (int64_t)M * M;
    uint64_t fp = fingerprint(IF, (int64_t)nc * M) ^ (fingerprint(Map, Map ? nc * MM : 0) * 3) ^
                  (fingerprint(MapOT, MapOT ? n_ot * MM : 0) * 5) ^ (fingerprint(MapTT, MapTT ? n_tt * MM : 0) * 7) ^
                  (fingerprint(strideOT, strideOT ? nc : 0) * 11) ^ (fingerprint(strideTT, strideTT ? nc : 0) * 13);
    const void *key[8] = {IF, NCmat, Map, MapOT, MapTT, strideOT, strideTT, nullptr};
    // the caller holds g_ref_mu
    if (g_cache.b && g_cache.N == N && g_cache.M == M && g_cache.fp == fp &&
        memcmp(g_cache.key, key, sizeof key) == 0)
        return g_cache.b;
    if (g_cache.b) mcpi_hip_basis_destroy(g_cache.b);
    g_cache.b = nullptr;
    mcpi_basis b = nullptr;
    ref_check(basis_upload_raw(N, M, IF, NCmat, Map, MapOT, MapTT, strideOT, strideTT, n_ot * MM, n_tt * MM, &b));
    g_cache.b = b;
    g_cache.N = N;
    g_cache.M = M;
    g_cache.fp = fp;
    memcpy(g_cache.key, key, sizeof key);
    return b;
}

struct DevVec {
    cplx *p = nullptr;
    explicit DevVec(int64_t n)
    {
        if (hipMalloc(&p, sizeof(cplx) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess)
            ref_die("GPU MEMORY ERROR : hipMalloc fail for complex");
    }
    ~DevVec() { (void)hipFree(p); }
};

void ref_apply(mcpi_basis b, const Carray C, Cmatrix Ho, const Carray Hint, Carray out)
{
    const int M = b->d.M;
    const int64_t nc = b->d.nc;
    std::vector<cplx> ho = flat_ho(M, Ho);
    mcpi_ham h = nullptr;
    ref_check(mcpi_hip_ham_create(b, (const mcpi_cplx *)ho.data(), Hint, &h));
    DevVec dC(nc), dO(nc);
    if (hipMemcpy(dC.p, C, sizeof(cplx) * nc, hipMemcpyHostToDevice) != hipSuccess) ref_die("GPU copy failed");
    ref_check(launch_apply(h, dC.p, dO.p, nullptr));
    if (hipMemcpy(out, dO.p, sizeof(cplx) * nc, hipMemcpyDeviceToHost) != hipSuccess) ref_die("GPU copy failed");
    mcpi_hip_ham_destroy(h);
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
    int k = 0, col = N + 1;
    for (int i = M - 1; i > 0; i--)
        for (int n = v[i]; n > 0; n--) {
            k += NCmat[N + col * i];
            N--;
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

void applyHconf(int N, int M, Iarray NCmat, Iarray IF, Carray C, Cmatrix Ho, Carray Hint, Carray out)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(N, M, IF, NCmat, nullptr, nullptr, nullptr, nullptr, nullptr);
    ref_apply(b, C, Ho, Hint, out);
}

void OBrho_XM(int N, int M, Iarray Map, Iarray NCmat, Iarray IF, Carray C, Cmatrix rho)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(N, M, IF, NCmat, Map, nullptr, nullptr, nullptr, nullptr);
    const int64_t nc = b->d.nc;
    DevVec dC(nc);
    if (hipMemcpy(dC.p, C, sizeof(cplx) * nc, hipMemcpyHostToDevice) != hipSuccess) ref_die("GPU copy failed");
    std::vector<cplx> r((size_t)M * M);
    ref_check(mcpi_hip_rho1(b, (const mcpi_cplx *)dC.p, (mcpi_cplx *)r.data()));
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++) memcpy(&rho[k][l], &r[k + (size_t)M * l], sizeof(cplx));
}

void TBrho_XX(int N, int M, Iarray Map, Iarray MapOT, Iarray MapTT, Iarray strideOT, Iarray strideTT,
              Iarray NCmat, Iarray IF, Carray C, Carray rho)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(N, M, IF, NCmat, Map, MapOT, MapTT, strideOT, strideTT);
    const int64_t nc = b->d.nc;
    DevVec dC(nc);
    if (hipMemcpy(dC.p, C, sizeof(cplx) * nc, hipMemcpyHostToDevice) != hipSuccess) ref_die("GPU copy failed");
    ref_check(mcpi_hip_rho2(b, (const mcpi_cplx *)dC.p, rho));
}

int lanczos(int lm, int Npar, int Morb, Iarray IF, Iarray strideOT, Iarray strideTT, Iarray Map, Iarray MapOT,
            Iarray MapTT, Cmatrix Ho, Carray Hint, Carray diag, Carray offdiag, Cmatrix lvec)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(Npar, Morb, IF, nullptr, Map, MapOT, MapTT, strideOT, strideTT);
    const int64_t nc = b->d.nc;
    std::vector<cplx> ho = flat_ho(Morb, Ho);
    mcpi_ham h = nullptr;
    ref_check(mcpi_hip_ham_create(b, (const mcpi_cplx *)ho.data(), Hint, &h));
    cplx *dl = nullptr;
    if (hipMalloc(&dl, sizeof(cplx) * nc * (size_t)lm) != hipSuccess) ref_die("GPU MEMORY ERROR : Lanczos basis");
    if (hipMemcpy(dl, lvec[0], sizeof(cplx) * nc, hipMemcpyHostToDevice) != hipSuccess) ref_die("GPU copy failed");
    int steps = 0;
    ref_check(mcpi_hip_lanczos(h, lm, (const mcpi_cplx *)dl, (mcpi_cplx *)dl, diag, offdiag, &steps));
    for (int t = 0; t < steps; t++)
        if (hipMemcpy(lvec[t], dl + (size_t)t * nc, sizeof(cplx) * nc, hipMemcpyDeviceToHost) != hipSuccess)
            ref_die("GPU copy failed");
    (void)hipFree(dl);
    mcpi_hip_ham_destroy(h);
    return steps;
}

double ground(int Niter, int Npar, int Morb, Iarray IF, Iarray strideOT, Iarray strideTT, Iarray Map,
              Iarray MapOT, Iarray MapTT, Cmatrix Ho, Carray Hint, Carray C)
{
    std::lock_guard<std::recursive_mutex> lk(g_ref_mu);
    mcpi_basis b = ref_basis(Npar, Morb, IF, nullptr, Map, MapOT, MapTT, strideOT, strideTT);
    const int64_t nc = b->d.nc;
    std::vector<cplx> ho = flat_ho(Morb, Ho);
    mcpi_ham h = nullptr;
    ref_check(mcpi_hip_ham_create(b, (const mcpi_cplx *)ho.data(), Hint, &h));
    DevVec dC(nc);
    if (hipMemcpy(dC.p, C, sizeof(cplx) * nc, hipMemcpyHostToDevice) != hipSuccess) ref_die("GPU copy failed");
    double E0 = 0;
    int steps = 0;
    ref_check(mcpi_hip_ground(h, Niter, (const mcpi_cplx *)dC.p, (mcpi_cplx *)dC.p, &E0, &steps));
    if (hipMemcpy(C, dC.p, sizeof(cplx) * nc, hipMemcpyDeviceToHost) != hipSuccess) ref_die("GPU copy failed");
    mcpi_hip_ham_destroy(h);
    return E0;
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
