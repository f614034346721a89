// Kernel benchmark of the int8 digit-sliced stage product (k_stage_i8l) on its own, at the
// shape of one stage of a batched curve: A angles x N beams, ND digits per operand, the
// workgroup tile given on the command line (default: the tile the autotuner picks for the
// N = 521, 40-angle curve).  It exists so that hardware counters can be read for this kernel
// alone, in a process that runs nothing else: every launch is the same mode-0 stage
// (P += alpha (Q M0 + Q diag(d)), Qn = Q + beta P, row maxima of Qn), on seeded operands
// (digits uniform in [-127, 127], FP64 state uniform in [-1, 1]).  The access pattern of the
// kernel does not depend on the values.  With copy_mib = M > 0 it then runs one plain device
// copy of M MiB, whose counter readings calibrate the byte counters (M MiB read, M MiB written).
//
// Usage: stage_i8_bench [angles N digits iters copy_mib wc wr pf]
//        (defaults 40 521 6 20 0 3 4 2)
// Prints per-launch time from HIP events and the algorithmic rates.
#include "../csrc/mixed.hip"

#include <cstdlib>
#include <cstring>

using namespace trhepd;

#define CK(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_));   \
            exit(1);                                                                            \
        }                                                                                       \
    } while (0)

namespace {

__device__ __forceinline__ unsigned mix(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// digits uniform in [-127, 127] over the whole buffer, padding included (the values do not
// change the kernel's work)
__global__ void k_fill_digits(int8_t *p, size_t n, unsigned seed)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        p[e] = (int8_t)((int)(mix((unsigned)e * 2654435761u + seed) % 255u) - 127);
}

// complex entries uniform in [-1, 1]^2 on the N x N part of each angle's NP x NP matrix
__global__ void k_fill_state(double2 *p, int A, int N, int NP, unsigned seed)
{
    const size_t n = (size_t)A * NP * NP;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t r = e % ((size_t)NP * NP);
        const int i = (int)(r % NP), j = (int)(r / NP);
        const bool v = i < N && j < N;
        p[e] = v ? make_double2(mix((unsigned)e + seed) / 2147483648.0 - 1.0, mix((unsigned)e + seed + 1) / 2147483648.0 - 1.0)
                 : make_double2(0, 0);
    }
}

// n complex values uniform in [0, 2)^2 (the per-angle conj(gamma^2) of the diagonal term)
__global__ void k_fill_c(double2 *p, size_t n, unsigned seed)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        p[e] = make_double2(mix((unsigned)e + seed) / 2147483648.0, mix((unsigned)e + seed + 1) / 2147483648.0);
}

__global__ void k_fill_int(int *p, size_t n, unsigned seed)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        p[e] = (int)(mix((unsigned)e + seed) % 7u) - 3;
}

__global__ void k_copy(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) b[e] = a[e];
}

template <class T>
T *dev(size_t n)
{
    void *p = nullptr;
    CK(hipMalloc(&p, n * sizeof(T) + 256));
    CK(hipMemset(p, 0, n * sizeof(T) + 256));
    return (T *)p;
}

}  // namespace

int main(int argc, char **argv)
{
    auto arg = [&](int i, int d) { return argc > i ? atoi(argv[i]) : d; };
    const int A = arg(1, 40), N = arg(2, 521), ND = arg(3, 6), iters = arg(4, 20), copy_mib = arg(5, 0);
    const int wc = arg(6, 3), wr = arg(7, 4), pf = arg(8, 2);
    if (A < 1 || N < 49 || N > 600 || ND < 5 || ND > 7 || iters < 1) {
        fprintf(stderr, "stage_i8_bench: angles >= 1, 49 <= N <= 600, 5 <= digits <= 7, iters >= 1\n");
        return 2;
    }
    const int NP = (N + 15) / 16 * 16;
    const MixedGeo g = mixed_geo(A, N, NP, ND, wc, wr, pf);
    const size_t mat = (size_t)NP * NP, nst = mat * A;
    printf("stage_i8_bench: A=%d N=%d NP=%d ND=%d tile wc=%d wr=%d pf=%d (KC=%d RBP=%d CBP=%d)\n", A, N, NP, ND, g.WC,
           g.WR, g.PF, g.KC, g.RBP, g.CBP);
    fflush(stdout);

    int8_t *qd = dev<int8_t>(mixed_qd_bytes(g)), *md = dev<int8_t>(mixed_md_bytes(g));
    int *eq = dev<int>((size_t)g.RBP * 16), *em = dev<int>((size_t)g.CBP * 16);
    double2 *Q = dev<double2>(nst), *Qn = dev<double2>(nst), *P = dev<double2>(nst), *g2c = dev<double2>((size_t)A * NP);
    VVi *col = dev<VVi>(1);
    double *colsum = dev<double>((size_t)A * (NP / 16) * NP), *diagabs = dev<double>((size_t)A * NP);
    unsigned long long *rmax = dev<unsigned long long>((size_t)g.RBP * 16);
    hipLaunchKernelGGL(k_fill_digits, dim3(4096), dim3(256), 0, 0, qd, mixed_qd_bytes(g), 1u);
    hipLaunchKernelGGL(k_fill_digits, dim3(1024), dim3(256), 0, 0, md, mixed_md_bytes(g), 2u);
    hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, 0, eq, (size_t)g.RBP * 16, 3u);
    hipLaunchKernelGGL(k_fill_int, dim3(8), dim3(256), 0, 0, em, (size_t)g.CBP * 16, 4u);
    hipLaunchKernelGGL(k_fill_state, dim3(4096), dim3(256), 0, 0, Q, A, N, NP, 5u);
    hipLaunchKernelGGL(k_fill_state, dim3(4096), dim3(256), 0, 0, P, A, N, NP, 6u);
    hipLaunchKernelGGL(k_fill_c, dim3(64), dim3(256), 0, 0, g2c, (size_t)A * NP, 7u);
    VVi c0;
    c0.v = make_double2(0.3, -0.2);
    c0.vi = make_double2(0.05, 0.01);
    CK(hipMemcpy(col, &c0, sizeof(c0), hipMemcpyHostToDevice));
    CK(hipDeviceSynchronize());

    MixedStage S;
    memset(&S, 0, sizeof(S));
    S.g = g;
    S.qd = qd;
    S.eq = eq;
    S.md = md;
    S.em = em;
    S.Q = Q;
    S.Qn = Qn;
    S.P = P;
    S.g2c = g2c;
    S.col = col;
    S.alpha = 1e-3;
    S.beta = -1e-3;
    S.mode = 0;
    S.colsum = colsum;
    S.diagabs = diagabs;
    S.rmax = rmax;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    // 30 untimed launches first: the clocks settle under this load within a few of them
    for (int w = 0; w < 30; ++w) launch_stage_i8(0, S);
    CK(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; ++it) launch_stage_i8(0, S);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipGetLastError());
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double us = ms * 1000.0 / iters;
    // algorithmic work and bytes of one launch: 3 ND (ND + 1) N^3 int8 ops per angle; the Q
    // digits (3 ND bytes per element), one stage's M0 digits, and the FP64 state read (Q, P)
    // and written (P, Qn), 64 bytes per element
    const double ops = 3.0 * ND * (ND + 1) * (double)N * N * N * A;
    const double bytes = (double)mixed_qd_bytes(g) + (double)mixed_md_bytes(g) + 64.0 * nst;
    printf("launches %d (after 30 warm-up), %.1f us per launch, %.3f POP/s = %.3f of the 5.0 POP/s int8 peak\n",
           iters + 30, us, ops / us * 1e-9, ops / us * 1e-9 / 5.0);
    printf("algorithmic bytes per launch %.4g (digits %.4g, FP64 state %.4g): %.2f TB/s\n", bytes,
           (double)(mixed_qd_bytes(g) + mixed_md_bytes(g)), 64.0 * nst, bytes / us * 1e-6);
    fflush(stdout);
    if (copy_mib > 0) {
        const size_t n = (size_t)copy_mib * 1024 * 1024 / sizeof(double2);
        double2 *a = dev<double2>(n), *b = dev<double2>(n);
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
        CK(hipDeviceSynchronize());
        CK(hipGetLastError());
        printf("copy kernel: %d MiB read, %d MiB written\n", copy_mib, copy_mib);
    }
    return 0;
}
