// trhepd_propagate_batch_noalloc: the small-N slice loop (N <= 48) on caller-owned device
// buffers, with the library's own work (the signed gather map, the run's counters, the
// evanescent growth rate and the status flags) in a fixed file-scope device array instead
// of memory allocated per call.  For callers that cannot allocate device memory.
//
// Every context on a device shares that device's array: calls that overlap on one device
// (two host threads, or two contexts whose streams run concurrently) race on it.  Each call
// clears what must start at zero before using it.
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace {

constexpr int NOALLOC_MAX_NB = 48;
constexpr size_t OFF_GMAX = 64;    // unsigned long long: bits of max |Im gamma|
constexpr size_t OFF_GMAP = 256;   // int [NOALLOC_MAX_NB^2]: signed gather map, row-major
constexpr size_t NOALLOC_BYTES = OFF_GMAP + sizeof(int) * NOALLOC_MAX_NB * NOALLOC_MAX_NB;
// counter slots ([0] renormalisations and [1] zero pivots are written by the slice kernel)
constexpr int C_NONFINITE = 13, C_BADMAP = 14;

}  // namespace

// the work area (one copy per device, held until the process ends).  Host code never uses
// this name as an address: the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace trhepd {
namespace {

// signed gather map (row-major): upper i<j -> +iv(j,i), lower i>j -> -iv(i,j), diagonal 0;
// an index outside 1..nv is flagged and replaced by 1, so that no later read leaves the table
__global__ void k_na_map(const int *__restrict__ iv, int nb, int nv, int *__restrict__ gm, int *__restrict__ bad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * nb) return;
    const int i = e / nb, j = e % nb;
    int g = 0;
    if (i < j) g = iv[j + i * nb];
    else if (i > j) g = -iv[i + j * nb];
    if (i != j && (g == 0 || g > nv || g < -nv)) {
        atomicOr(bad, 1);
        g = 1;
    }
    gm[e] = g;
}

// max |Im gamma| over beams and angles, as the bit pattern of a non-negative double (whose
// unsigned order is the numeric order)
__global__ void k_na_gmax(const double2 *__restrict__ gamma, int n, unsigned long long *__restrict__ gmax)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double v = e < n ? fabs(gamma[e].y) : 0.0;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMax(gmax, (unsigned long long)__double_as_longlong(v));
}

// flags a NaN / Inf in n doubles
__global__ void k_na_finite(const double *__restrict__ x, size_t n, int *__restrict__ flag)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        if (!isfinite(x[e])) {
            atomicOr(flag, 1);
            return;
        }
}

}  // namespace

void noalloc_attach(trhepd_ctx *ctx)
{
    // the context's device is current (trhepd_ctx_create): this device's copy of the array
    void *p = nullptr;
    ctx->noalloc_err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    ctx->noalloc = ctx->noalloc_err == hipSuccess ? p : nullptr;
}

}  // namespace trhepd

extern "C" int trhepd_propagate_batch_noalloc(trhepd_ctx *ctx, int table_id, int nb, int nang, const int *iv,
                                              const double *gamma, double *f, double *f2, int nb0,
                                              const double *wnsga2)
{
    using namespace trhepd;
    if (!ctx) return TRHEPD_E_ARG;
    if (f2 && !wnsga2) return fail(ctx, TRHEPD_E_ARG, "propagate_batch_noalloc: f2 needs wnsga2");
    if (nang == 0 && nb >= 1) {
        ctx->stats = Stats();
        return TRHEPD_OK;
    }
    Table *t = find_table(ctx, table_id);
    if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "no table %d", table_id);
    if (nb < 1 || nang < 1 || !iv || !gamma || !f || nb0 < 1 || nb0 > nb)
        return fail(ctx, TRHEPD_E_ARG, "propagate_batch_noalloc: bad arguments");
    if (nb > NOALLOC_MAX_NB)
        return fail(ctx, TRHEPD_E_ARG,
                    "propagate_batch_noalloc runs N <= %d (its fixed work area holds a %d x %d map; N = %d): "
                    "use trhepd_propagate_batch",
                    NOALLOC_MAX_NB, NOALLOC_MAX_NB, NOALLOC_MAX_NB, nb);
    if (t->method != TRHEPD_LF && t->method != TRHEPD_SP4 && t->method != TRHEPD_SP6 && t->method != TRHEPD_RK4)
        return fail(ctx, TRHEPD_E_ARG,
                    "propagate_batch_noalloc runs int_method 4-7 (the Magnus methods and rk4gs need "
                    "per-batch working memory; got %d)", t->method);
    if (!ctx->noalloc)
        return hip_check(ctx, ctx->noalloc_err != hipSuccess ? ctx->noalloc_err : hipErrorInvalidSymbol,
                         "hipGetSymbolAddress(work area of propagate_batch_noalloc)");
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    char *base = (char *)ctx->noalloc;
    int *cnt = (int *)base;
    unsigned long long *gmax = (unsigned long long *)(base + OFF_GMAX);
    int *gm = (int *)(base + OFF_GMAP);
    hipStream_t st = ctx->stream;
    const int ng = nb * nang;
    // counters, flags and the growth-rate maximum start at zero (a fill kernel: no memset node)
    launch_fill_i32(st, cnt, (int)(OFF_GMAX / 4) + 2, 0);
    hipLaunchKernelGGL(k_na_map, dim3((nb * nb + 255) / 256), dim3(256), 0, st, iv, nb, t->nv, gm, cnt + C_BADMAP);
    hipLaunchKernelGGL(k_na_gmax, dim3((ng + 255) / 256), dim3(256), 0, st, (const double2 *)gamma, ng, gmax);
    TRY_HIP(ctx, hipGetLastError());
    struct {
        int c[16];
        unsigned long long g;
    } h;
    static_assert(sizeof(h) == OFF_GMAX + 8, "host image of the work area head");
    TRY_HIP(ctx, hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, st));
    TRY_HIP(ctx, hipStreamSynchronize(st));
    if (h.c[C_BADMAP]) return fail(ctx, TRHEPD_E_ARG, "iv holds an index outside 1..nv=%d", t->nv);
    Batch shape;
    double gim;
    std::memcpy(&gim, &h.g, sizeof gim);
    shape.gim_max = gim;
    const int period = batch_period(ctx, shape, *t);
    ctx->stats = Stats();
    TRY_HIP(ctx, hipEventRecord(ctx->ev0, st));
    int rc = run_small(ctx, st, period, *t, nb, nang, gm, (const double2 *)gamma, (double2 *)f, f2, nb0, wnsga2, cnt);
    if (rc) return rc;
    TRY_HIP(ctx, hipEventRecord(ctx->ev1, st));
    const size_t nf = (size_t)2 * nb * nb * nang;
    hipLaunchKernelGGL(k_na_finite, dim3((unsigned)std::min<size_t>((nf + 255) / 256, 1024)), dim3(256), 0, st,
                       (const double *)f, nf, cnt + C_NONFINITE);
    if (f2)
        hipLaunchKernelGGL(k_na_finite, dim3((unsigned)std::min<size_t>((ng + 255) / 256, 1024)), dim3(256), 0, st,
                           (const double *)f2, (size_t)ng, cnt + C_NONFINITE);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipMemcpyAsync(h.c, cnt, sizeof h.c, hipMemcpyDeviceToHost, st));
    TRY_HIP(ctx, hipStreamSynchronize(st));
    float ms = 0;
    TRY_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.n_renorm = h.c[0];
    ctx->stats.n_products = (long long)nang * t->ns * rkn_coef(t->method).s;
    ctx->stats.device_ms = ms;
    if (h.c[1]) return fail(ctx, TRHEPD_E_SINGULAR, "exact zero pivot in a renormalisation LU");
    if (h.c[C_NONFINITE]) return fail(ctx, TRHEPD_E_NONFINITE, "non-finite value in f or f2");
    return TRHEPD_OK;
}
