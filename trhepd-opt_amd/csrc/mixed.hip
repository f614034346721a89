// Mixed-precision stage products of the batched large-N path (TRHEPD_PREC_MIXED).
//
// The RKN kick P_a += c (Q_a M0 + Q_a diag(d_a)) keeps its FP64 state, FP64 diagonal term
// and FP64 epilogue; only the dense product Q M0 changes arithmetic.  Each row of Q (each
// column of M0) is scaled by a power of two 2^-e so that its largest component, taken over
// the real part, the imaginary part and their sum, is below 127/128, and split into five
// signed base-2^7 digits
//     x = sum_{p=0..4} d_p 2^{-7(p+1)} + r,   d_0 in [-127, 127], d_p in [-64, 64], |r| <= 2^-36.
// The digit products d_p(Q) d_q(M) are exact int8 x int8 -> int32 MFMAs; the pairs of one
// level s = p + q (p + q <= 4, fifteen pairs) share one int32 accumulator, whose magnitude
// stays below 2^25 for k <= 600, so every level sum over k is exact.  Levels are combined
// in FP64 with weights 2^{-7s}.  The complex product uses three real products
// (T1 = Qr Mr, T2 = Qi Mi, T3 = (Qr+Qi)(Mr+Mi); re = T1 - T2, im = T3 - T1 - T2), all under
// the same row and column scales, so the combination is exact on the integer sums.
// The normwise error of one product is about 2^-34 sum_k |Q(i,k)| |M(k,j)|.
//
// Operand roles: the MFMA's A operand is the M0 digit fragment (its row = output column j)
// and its B operand the Q digit fragment (its column = output row i), so that the 16 lanes
// of a fragment row walk down a column of the column-major state and the epilogue's loads
// and stores are contiguous.
//
// Reference arithmetic replaced: the zgemm of hp_qinth / hp_pinth, surf_prkn.f90:222-276.
#include <algorithm>

#include "internal.hpp"

namespace trhepd {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4i mfma_i8(v4i a, v4i b, v4i c)
{
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 w_upper(const VVi &t) { return make_double2(-(t.v.x + t.vi.x), t.v.y + t.vi.y); }
__device__ __forceinline__ double2 w_lower(const VVi &t) { return make_double2(-t.v.x + t.vi.x, -t.v.y + t.vi.y); }

// exponent e with max * 2^-e < 127/128 (0 for an all-zero row)
__device__ __forceinline__ int scale_exp(double mx)
{
    if (!(mx > 0.0)) return 0;
    int e;
    const double f = frexp(mx, &e);
    return f >= 127.0 / 128.0 ? e + 1 : e;
}

__device__ __forceinline__ double cmax3(double2 v)
{
    return fmax(fabs(v.x), fmax(fabs(v.y), fabs(v.x + v.y)));
}

// the ND leading digits of x (|x| < 127/128)
template <int ND>
__device__ __forceinline__ void digits(double x, int *d)
{
    double y = x * 128.0;
#pragma unroll
    for (int p = 0; p < ND; ++p) {
        const double r = rint(y);
        d[p] = (int)r;
        y = (y - r) * 128.0;
    }
}

__host__ __device__ __forceinline__ size_t md_stage_bytes(const MixedGeo &g)
{
    return (size_t)3 * g.ND * g.CBP * g.KC * 1024;
}

// One workgroup splits one 16-vector block (16 rows of Q or 16 columns of M0) over all k.
// Thread (c, g): vector c of the block, k groups g, g+16, ... of 16 consecutive k.  The
// element accessor `get(c, k)` returns the complex entry (zero outside the matrix).
template <int ND, class Get>
__device__ __forceinline__ void split_block(const MixedGeo &g, Get get, int blk, int nblkp, bool vec_valid,
                                            int8_t *__restrict__ out, int *__restrict__ eout)
{
    const int t = threadIdx.x, c = t & 15, gq = t >> 4;
    const int ng = g.KC * 4;
    __shared__ double smax[16][17];
    double mx = 0.0;
    if (vec_valid)
        for (int grp = gq; grp < ng; grp += 16)
#pragma unroll 4
            for (int kk = 0; kk < 16; ++kk) mx = fmax(mx, cmax3(get(c, grp * 16 + kk)));
    smax[c][gq] = mx;
    __syncthreads();
    if (gq == 0) {
        double m = 0.0;
        for (int u = 0; u < 16; ++u) m = fmax(m, smax[c][u]);
        smax[c][16] = m;
        eout[blk * 16 + c] = scale_exp(m);
    }
    __syncthreads();
    const int e = scale_exp(smax[c][16]);
    const size_t plane = (size_t)nblkp * g.KC * 64;   // v4i per digit plane
    v4i *o = (v4i *)out;
    for (int grp = gq; grp < ng; grp += 16) {
        int w[3][ND][4];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int p = 0; p < ND; ++p)
#pragma unroll
                for (int u = 0; u < 4; ++u) w[m][p][u] = 0;
        if (vec_valid) {
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) {
                const double2 v = get(c, grp * 16 + kk);
                const int wd = kk >> 2, b = kk & 3;
                int d[3][ND];
                digits<ND>(ldexp(v.x, -e), d[0]);
                digits<ND>(ldexp(v.y, -e), d[1]);
                digits<ND>(ldexp(v.x + v.y, -e), d[2]);
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int p = 0; p < ND; ++p) w[m][p][wd] |= (d[m][p] & 0xff) << (8 * b);
            }
        }
        const int kc = grp >> 2, lane = c + 16 * (grp & 3);
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int p = 0; p < ND; ++p)
                o[(size_t)(m * ND + p) * plane + ((size_t)blk * g.KC + kc) * 64 + lane] =
                    v4i{w[m][p][0], w[m][p][1], w[m][p][2], w[m][p][3]};
    }
}

template <int ND>
__global__ __launch_bounds__(256) void k_qsplit(MixedGeo g, const double2 *__restrict__ Q, int8_t *__restrict__ qd,
                                                int *__restrict__ eq)
{
    const int rb = blockIdx.x;
    const int c = threadIdx.x & 15;
    const int R = rb * 16 + c;
    const bool rv = rb < g.RB && (R % g.NP) < g.N;
    const double2 *q = Q + (size_t)(R / g.NP) * g.mat + (R % g.NP);
    auto get = [&](int, int k) -> double2 { return k < g.N ? q[(size_t)k * g.NP] : make_double2(0, 0); };
    split_block<ND>(g, get, rb, g.RBP, rv, qd, eq);
}

template <int ND>
__global__ __launch_bounds__(256) void k_msplit(MixedGeo g, const int *__restrict__ gmap, const VVi *__restrict__ col0,
                                                int nv, int8_t *__restrict__ md, int *__restrict__ em)
{
    const int cb = blockIdx.x, st = blockIdx.y;
    const int c = threadIdx.x & 15;
    const int j = cb * 16 + c;
    const VVi *col = col0 + (size_t)st * nv;
    auto get = [&](int, int k) -> double2 {
        if (k >= g.N || k == j) return make_double2(0, 0);
        const int s = gmap[k * g.N + j];
        return s > 0 ? w_upper(col[s - 1]) : w_lower(col[-s - 1]);
    };
    split_block<ND>(g, get, cb, g.CBP, j < g.N, md + (size_t)st * md_stage_bytes(g), em + (size_t)st * g.CBP * 16);
}

struct StageLaunch {
    MixedStage S;
    int ntiles, ncolt, tiles_per_xcd;
};

// One wave computes a 32 x 32 output tile (two 16-row blocks of Q times two 16-column
// blocks of M0); the four waves of a workgroup stack in rows and share the M0 fragments.
// The three real products run one after the other, each over all k with its fifteen
// digit pairs accumulated per level, then folded into the FP64 complex accumulators.
template <int ND>
__global__ __launch_bounds__(256, ND <= 5 ? 2 : 1) void k_stage_i8(StageLaunch Lc)
{
    const MixedStage &S = Lc.S;
    const MixedGeo g = S.g;
    // XCD-aware tile order: the column tiles of one row tile run on the same XCD
    const int L = blockIdx.x;
    const int t = (L & 7) * Lc.tiles_per_xcd + (L >> 3);
    if (t >= Lc.ntiles) return;
    const int rt = t / Lc.ncolt, ct = t % Lc.ncolt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int rb0 = rt * 8 + wave * 2;
    if (rb0 >= g.RBP) return;
    const int cb0 = ct * 2;
    const size_t qplane = (size_t)g.RBP * g.KC * 64, mplane = (size_t)g.CBP * g.KC * 64;
    const v4i *__restrict__ QD = (const v4i *)S.qd + (size_t)rb0 * g.KC * 64 + lane;
    const v4i *__restrict__ MD = (const v4i *)S.md + (size_t)cb0 * g.KC * 64 + lane;
    double re[2][2][4], im[2][2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) re[rb][cb][r] = im[rb][cb][r] = 0.0;
    for (int m = 0; m < 3; ++m) {
        v4i acc[ND][2][2];
#pragma unroll
        for (int s = 0; s < ND; ++s)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) acc[s][rb][cb] = v4i{0, 0, 0, 0};
        const v4i *qa = QD + (size_t)m * ND * qplane;
        const v4i *mb = MD + (size_t)m * ND * mplane;
        for (int kc = 0; kc < g.KC; ++kc) {
            v4i a[ND][2];
#pragma unroll
            for (int p = 0; p < ND; ++p)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) a[p][rb] = qa[p * qplane + ((size_t)rb * g.KC + kc) * 64];
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                v4i b[2];
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) b[cb] = mb[q * mplane + ((size_t)cb * g.KC + kc) * 64];
#pragma unroll
                for (int p = 0; p + q < ND; ++p)
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                        for (int cb = 0; cb < 2; ++cb)
                            acc[p + q][rb][cb] = mfma_i8(b[cb], a[p][rb], acc[p + q][rb][cb]);
            }
        }
        const double sr = m == 2 ? 0.0 : (m == 0 ? 1.0 : -1.0);   // sign into re
        const double si = m == 2 ? 1.0 : -1.0;                     // sign into im
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double T = (double)acc[ND - 1][rb][cb][r];
#pragma unroll
                    for (int s = ND - 2; s >= 0; --s) T = fma(T, 0x1p-7, (double)acc[s][rb][cb][r]);
                    re[rb][cb][r] = fma(sr, T, re[rb][cb][r]);
                    im[rb][cb][r] = fma(si, T, im[rb][cb][r]);
                }
    }
    // FP64 epilogue: lane (l16, kq) holds row i = block row + l16, columns 4 kq + r
    const double2 w0 = w_upper(S.col[0]);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        if (rb0 + rb >= g.RB) continue;
        const int R = (rb0 + rb) * 16 + l16;
        const int a = R / g.NP, i = R % g.NP;
        const double alpha = S.coef ? S.coef[a] : S.alpha;
        const int eqr = S.eq[R] - 14;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            double csum[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                csum[r] = 0.0;
                const int j = (cb0 + cb) * 16 + 4 * kq + r;
                if (i < g.N && j < g.N) {
                    const double sc = ldexp(1.0, eqr + S.em[j]);
                    const double2 c2 = S.g2c[(size_t)a * g.NP + j];
                    const double2 dj = make_double2(w0.x - c2.x, w0.y - c2.y);
                    const size_t idx = (size_t)a * g.mat + i + (size_t)j * g.NP;
                    const double2 q = S.Q[idx];
                    const double2 qd = cmul(q, dj);
                    double2 p = S.P[idx];
                    p.x += alpha * (re[rb][cb][r] * sc + qd.x);
                    p.y += alpha * (im[rb][cb][r] * sc + qd.y);
                    S.P[idx] = p;
                    if (S.mode != 2) {
                        const double2 qn = make_double2(q.x + S.beta * p.x, q.y + S.beta * p.y);
                        S.Qn[idx] = qn;
                        if (S.mode == 1) {
                            if (i == j) S.diagabs[(size_t)a * g.NP + j] = hypot(qn.x, qn.y);
                            else csum[r] = sqrt(qn.x * qn.x + qn.y * qn.y);
                        }
                    }
                }
            }
            if (S.mode == 1) {
                // column sums over the 16 rows of the block: lanes l16 = 0..15 of each kq
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double v = csum[r];
                    v += __shfl_xor(v, 1);
                    v += __shfl_xor(v, 2);
                    v += __shfl_xor(v, 4);
                    v += __shfl_xor(v, 8);
                    const int j = (cb0 + cb) * 16 + 4 * kq + r;
                    if (l16 == 0 && j < g.NP) S.colsum[(size_t)(R >> 4) * g.NP + j] = v;
                }
            }
        }
    }
}

}  // namespace

MixedGeo mixed_geo(int A, int N, int NP, int ND)
{
    MixedGeo g;
    g.ND = ND;
    g.A = A;
    g.N = N;
    g.NP = NP;
    g.mat = (size_t)NP * NP;
    g.KC = (N + 63) / 64;
    g.RB = A * NP / 16;
    g.RBP = (g.RB + 1) / 2 * 2;
    g.CBP = (NP / 16 + 1) / 2 * 2;
    return g;
}

size_t mixed_qd_bytes(const MixedGeo &g) { return (size_t)3 * g.ND * g.RBP * g.KC * 1024; }
size_t mixed_md_bytes(const MixedGeo &g) { return md_stage_bytes(g); }

void launch_msplit(hipStream_t st, const MixedGeo &g, const int *gmap, const VVi *col0, int nv, int nst, int8_t *md,
                   int *em)
{
    switch (g.ND) {
    case 5: hipLaunchKernelGGL(k_msplit<5>, dim3(g.CBP, nst), dim3(256), 0, st, g, gmap, col0, nv, md, em); break;
    case 6: hipLaunchKernelGGL(k_msplit<6>, dim3(g.CBP, nst), dim3(256), 0, st, g, gmap, col0, nv, md, em); break;
    default: hipLaunchKernelGGL(k_msplit<7>, dim3(g.CBP, nst), dim3(256), 0, st, g, gmap, col0, nv, md, em); break;
    }
}

void launch_qsplit(hipStream_t st, const MixedGeo &g, const double2 *Q, int8_t *qd, int *eq)
{
    switch (g.ND) {
    case 5: hipLaunchKernelGGL(k_qsplit<5>, dim3(g.RBP), dim3(256), 0, st, g, Q, qd, eq); break;
    case 6: hipLaunchKernelGGL(k_qsplit<6>, dim3(g.RBP), dim3(256), 0, st, g, Q, qd, eq); break;
    default: hipLaunchKernelGGL(k_qsplit<7>, dim3(g.RBP), dim3(256), 0, st, g, Q, qd, eq); break;
    }
}

void launch_stage_i8(hipStream_t st, const MixedStage &S)
{
    StageLaunch Lc;
    Lc.S = S;
    const int rowt = (S.g.RBP + 7) / 8;
    Lc.ncolt = S.g.CBP / 2;
    Lc.ntiles = rowt * Lc.ncolt;
    Lc.tiles_per_xcd = (Lc.ntiles + 7) / 8;
    const dim3 grid(8 * Lc.tiles_per_xcd), blk(256);
    switch (S.g.ND) {
    case 5: hipLaunchKernelGGL(k_stage_i8<5>, grid, blk, 0, st, Lc); break;
    case 6: hipLaunchKernelGGL(k_stage_i8<6>, grid, blk, 0, st, Lc); break;
    default: hipLaunchKernelGGL(k_stage_i8<7>, grid, blk, 0, st, Lc); break;
    }
}

}  // namespace trhepd
