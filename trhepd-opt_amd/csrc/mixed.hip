// Mixed-precision stage products of the batched large-N path (TRHEPD_PREC_MIXED).
//
// The RKN kick P_a += c (Q_a M0 + Q_a diag(d_a)) keeps its FP64 state, FP64 diagonal term
// and FP64 epilogue; only the dense product Q M0 changes arithmetic.  Each row of Q (each
// column of M0) is scaled by a power of two 2^-e so that its largest component, taken over
// the real part, the imaginary part and their sum, is below 1/2, and split into ND signed
// base-254 digits
//     x = sum_{p<ND} d_p 254^{-(p+1)} + r,   d_p in [-127, 127],  |r| <= 254^-ND / 2
// (each digit is the rounded remainder times 254, which uses the whole int8 range).
// The digit products d_p(Q) d_q(M) are exact int8 x int8 -> int32 MFMAs; the pairs of one
// level s = p + q (p + q < ND) share one int32 accumulator, whose magnitude stays below
// 7 * 127^2 * 576 < 2^26 for k <= 576, so every level sum over k is exact.  Levels are
// combined in FP64 with weights 254^-s.  The complex product uses three real products
// (T1 = Qr Mr, T2 = Qi Mi, T3 = (Qr+Qi)(Mr+Mi); re = T1 - T2, im = T3 - T1 - T2), all under
// the same row and column scales, so the combination is exact on the integer sums.
// The normwise error of one product is about ND 254^-ND sum_k |Q(i,k)| |M(k,j)|.
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

// Workgroup barrier ordering LDS only.  The stage kernel's waves share nothing but LDS
// inside the k loop, so the barrier need not wait for global loads in flight (the Q and M0
// fragments of later chunks), as a full __syncthreads() does on this target.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// exponent e with max * 2^-e < 1/2 (0 for an all-zero row)
__device__ __forceinline__ int scale_exp(double mx)
{
    if (!(mx > 0.0)) return 0;
    int e;
    (void)frexp(mx, &e);   // mx = f 2^e, f in [1/2, 1)
    return e + 1;
}

__device__ __forceinline__ double cmax3(double2 v)
{
    return fmax(fabs(v.x), fmax(fabs(v.y), fabs(v.x + v.y)));
}

// the ND leading base-254 digits of x (|x| < 1/2): every |y| <= 127, so every digit fits
template <int ND>
__device__ __forceinline__ void digits(double x, int *d)
{
    double y = x * 254.0;
#pragma unroll
    for (int p = 0; p < ND; ++p) {
        const double r = rint(y);
        d[p] = (int)r;
        y = (y - r) * 254.0;
    }
}

__host__ __device__ __forceinline__ size_t md_stage_bytes(const MixedGeo &g)
{
    return (size_t)3 * g.ND * g.CBP * g.KC * 1024;
}

// One workgroup splits one 16-vector block (16 rows of Q or 16 columns of M0) over all k.
// Thread (c, g): vector c of the block, k groups g, g+16, ... of 16 consecutive k.  The
// element accessor `get(c, k)` returns the complex entry (zero outside the matrix).  The
// vectors' maxima (of cmax3 over k) come from `vmax` when the producer of the data gathered
// them, else from a first pass over the block.
template <int ND, class Get>
__device__ __forceinline__ void split_block(const MixedGeo &g, Get get, int blk, int nblkp, bool vec_valid,
                                            int8_t *__restrict__ out, int *__restrict__ eout,
                                            const unsigned long long *__restrict__ vmax)
{
    const int t = threadIdx.x, c = t & 15, gq = t >> 4;
    const int ng = g.KC * 4;
    int e;
    if (vmax) {
        e = vec_valid ? scale_exp(__longlong_as_double((long long)vmax[blk * 16 + c])) : 0;
        if (gq == 0) eout[blk * 16 + c] = e;
    } else {
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
        e = scale_exp(smax[c][16]);
    }
    // fragment ((p nblkp + blk) 3 + m) KC + kc of digit p, block blk, real operand m (re, im,
    // re + im), k chunk kc: the 3 KC chunks a stage product walks are consecutive
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
                o[((((size_t)p * nblkp + blk) * 3 + m) * g.KC + kc) * 64 + lane] =
                    v4i{w[m][p][0], w[m][p][1], w[m][p][2], w[m][p][3]};
    }
}

template <int ND>
__global__ __launch_bounds__(256) void k_qsplit(MixedGeo g, const double2 *__restrict__ Q, int8_t *__restrict__ qd,
                                                int *__restrict__ eq, const unsigned long long *__restrict__ vmax,
                                                unsigned long long *__restrict__ clear)
{
    const int rb = blockIdx.x;
    if (clear && threadIdx.x < 16) clear[rb * 16 + threadIdx.x] = 0;
    const int c = threadIdx.x & 15;
    const int R = rb * 16 + c;
    const bool rv = rb < g.RB && (R % g.NP) < g.N;
    const double2 *q = Q + (size_t)(R / g.NP) * g.mat + (R % g.NP);
    auto get = [&](int, int k) -> double2 { return k < g.N ? q[(size_t)k * g.NP] : make_double2(0, 0); };
    split_block<ND>(g, get, rb, g.RBP, rv, qd, eq, vmax);
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
    split_block<ND>(g, get, cb, g.CBP, j < g.N, md + (size_t)st * md_stage_bytes(g), em + (size_t)st * g.CBP * 16,
                    nullptr);
}

struct StageLaunch {
    MixedStage S;
    int ntiles, ncolt, tiles_per_xcd;
};

// LDS-tiled form: a workgroup of WR waves computes 16 WR rows (one 16-row block per wave)
// x 16 WC columns.  Per 64-wide k chunk of one real product a wave needs ND fragments of its
// own Q digits and the WC ND fragments of M0 digits that all WR waves share (1 KB each).
// Only the shared M0 fragments go through LDS (global -> registers -> LDS, double-buffered);
// each wave loads its Q fragments straight into registers.  Both are fetched one chunk
// ahead, so the loads of chunk c+1 overlap the MFMAs of chunk c.  The chunks of the three
// real products form one pipelined sequence of 3 KC steps, whose fragments lie consecutively
// in each digit plane, so that a chunk's address is a wave-uniform base plus 1 KB per chunk.
// A four-wave tile is held to 256 VGPRs (two waves per SIMD), so that two workgroups share a
// CU and one's FP64 epilogue, bound by HBM, overlaps the other's int8 k loop.
template <int ND, int WC, int WR, int PF>
__global__ __launch_bounds__(64 * WR, (WR == 4 && ND * WC <= 24) ? 2 : 1) void k_stage_i8l(StageLaunch Lc)
{
    constexpr int NBF = WC * ND, NR = (NBF + WR - 1) / WR;
    extern __shared__ v4i lds[];   // [2][NBF][64]
    const MixedStage &S = Lc.S;
    const MixedGeo g = S.g;
    const int L = blockIdx.x;
    const int t = (L & 7) * Lc.tiles_per_xcd + (L >> 3);
    if (t >= Lc.ntiles) return;
    const int rt = t / Lc.ncolt, ct = t % Lc.ncolt;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l16 = lane & 15, kq = lane >> 4;
    const int rb0 = rt * WR, cb0 = ct * WC;
    const int nchunk = 3 * g.KC;
    // wave-uniform fragment bases (in v4i): Q digit p of this wave's row block, and the M0
    // fragments r of this wave's share of the workgroup's WC ND column fragments
    const size_t qstep = (size_t)g.RBP * nchunk * 64, mstep = (size_t)g.CBP * nchunk * 64;
    const v4i *__restrict__ QD = (const v4i *)S.qd + (size_t)(rb0 + wave) * nchunk * 64;
    const v4i *__restrict__ MD = (const v4i *)S.md;
    size_t mbase[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int f = min(r * WR + wave, NBF - 1), q = f / WC, cb = f % WC;
        mbase[r] = q * mstep + (size_t)(cb0 + cb) * nchunk * 64;
    }
    v4i stage[NR];
    auto gload = [&](int c) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r * WR + wave < NBF) stage[r] = MD[mbase[r] + (size_t)c * 64 + lane];
    };
    auto qload = [&](int c, v4i *dst) {
#pragma unroll
        for (int p = 0; p < ND; ++p) dst[p] = QD[p * qstep + (size_t)c * 64 + lane];
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int f = r * WR + wave;
            if (f < NBF) lds[((size_t)buf * NBF + f) * 64 + lane] = stage[r];
        }
    };
    v4i acc[ND][WC];
    double re[WC][4], im[WC][4];
#pragma unroll
    for (int cb = 0; cb < WC; ++cb) {
#pragma unroll
        for (int s = 0; s < ND; ++s) acc[s][cb] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) re[cb][r] = im[cb][r] = 0.0;
    }
    // One k chunk: the M0 fragments of chunk c+1 are fetched into registers (then LDS), the
    // Q fragments of chunk c+PF into the register set `qfut`, the products of chunk c run
    // on `qcur`.  The M0 loads are issued first, so that waiting for them at the end of the
    // chunk leaves the Q loads in flight: Q gets PF chunks of latency, M0 one.
    // Past the last chunk the loads repeat the last chunk (in bounds, unused), so that every
    // chunk is straight-line code and the compiler's wait counts stay exact.
    auto chunk = [&](int c, v4i(&qcur)[ND], v4i(&qfut)[ND]) {
        gload(min(c + 1, nchunk - 1));
        qload(min(c + PF, nchunk - 1), qfut);
        const v4i *Lb = lds + (size_t)(c & 1) * NBF * 64 + lane;
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            v4i b[WC];
#pragma unroll
            for (int cb = 0; cb < WC; ++cb) b[cb] = Lb[(q * WC + cb) * 64];
#pragma unroll
            for (int p = 0; p + q < ND; ++p)
#pragma unroll
                for (int cb = 0; cb < WC; ++cb) acc[p + q][cb] = mfma_i8(b[cb], qcur[p], acc[p + q][cb]);
        }
        // the last chunk of real product m: fold its level sums into the FP64 results
        const int m = c + 1 == g.KC ? 0 : c + 1 == 2 * g.KC ? 1 : c + 1 == nchunk ? 2 : -1;
        if (m >= 0) {
            const double sr = m == 2 ? 0.0 : (m == 0 ? 1.0 : -1.0);
            const double si = m == 2 ? 1.0 : -1.0;
#pragma unroll
            for (int cb = 0; cb < WC; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double T = (double)acc[ND - 1][cb][r];
#pragma unroll
                    for (int s = ND - 2; s >= 0; --s) T = fma(T, 1.0 / 254.0, (double)acc[s][cb][r]);
                    re[cb][r] = fma(sr, T, re[cb][r]);
                    im[cb][r] = fma(si, T, im[cb][r]);
                }
#pragma unroll
            for (int cb = 0; cb < WC; ++cb)
#pragma unroll
                for (int s = 0; s < ND; ++s) acc[s][cb] = v4i{0, 0, 0, 0};
        }
        sstore((c + 1) & 1);
        lds_barrier();
    };
    // the register sets rotate with the loop unrolled, so no set is copied (a copy would
    // wait for the loads in flight)
    v4i qa[ND], qb[ND];
    gload(0);
    qload(0, qa);
    if constexpr (PF == 1) {
        sstore(0);
        lds_barrier();
        for (int c = 0; c < nchunk; c += 2) {
            chunk(c, qa, qb);
            if (c + 1 < nchunk) chunk(c + 1, qb, qa);
        }
    } else {
        v4i qc[ND];
        qload(1, qb);   // nchunk >= 3
        sstore(0);
        lds_barrier();
        for (int c = 0; c < nchunk; c += 3) {   // nchunk = 3 KC
            chunk(c, qa, qc);
            chunk(c + 1, qb, qa);
            chunk(c + 2, qc, qb);
        }
    }
    // FP64 epilogue: lane (l16, kq) holds row i = block row + l16, columns 4 kq + r
    if (rb0 + wave >= g.RB) return;
    const double2 w0 = w_upper(S.col[0]);
    const int R = (rb0 + wave) * 16 + l16;
    const int a = R / g.NP, i = R % g.NP;
    const double alpha = S.coef ? S.coef[a] : S.alpha;
    const double beta = (S.xid && S.xid[a]) ? 0.0 : S.beta;
    const int eqr = S.eq[R];
    double qmax = 0.0;   // this lane's maximum of cmax3(Qn) (mode 0/1)
    // The Q, P, conj(gamma^2) and column-scale loads of two column blocks are issued together
    // (clamped to the matrix, so unconditional), then their updates run: one memory round
    // trip per pair of blocks rather than one per element.
#pragma unroll
    for (int cp = 0; cp < WC; cp += 2) {
        double2 qv[2][4], pv[2][4], c2v[2][4];
        int emv[2][4];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (cp + e >= WC) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = (cb0 + cp + e) * 16 + 4 * kq + r, jj = min(j, g.NP - 1);
                const size_t idx = (size_t)a * g.mat + i + (size_t)jj * g.NP;
                qv[e][r] = S.Q[idx];
                pv[e][r] = S.P[idx];
                c2v[e][r] = S.g2c[(size_t)a * g.NP + jj];
                emv[e][r] = S.em[j];
            }
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int cb = cp + e;
            if (cb >= WC) continue;
            double csum[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                csum[r] = 0.0;
                const int j = (cb0 + cb) * 16 + 4 * kq + r;
                if (i < g.N && j < g.N) {
                    const double sc = ldexp(1.0 / (254.0 * 254.0), eqr + emv[e][r]);
                    const double2 dj = make_double2(w0.x - c2v[e][r].x, w0.y - c2v[e][r].y);
                    const size_t idx = (size_t)a * g.mat + i + (size_t)j * g.NP;
                    const double2 q = qv[e][r];
                    const double2 qd = cmul(q, dj);
                    double2 p = pv[e][r];
                    p.x += alpha * (re[cb][r] * sc + qd.x);
                    p.y += alpha * (im[cb][r] * sc + qd.y);
                    S.P[idx] = p;
                    if (S.mode != 2) {
                        const double2 qn = make_double2(q.x + beta * p.x, q.y + beta * p.y);
                        S.Qn[idx] = qn;
                        qmax = fmax(qmax, cmax3(qn));
                        if (S.mode == 1) {
                            if (i == j) S.diagabs[(size_t)a * g.NP + j] = hypot(qn.x, qn.y);
                            else csum[r] = sqrt(qn.x * qn.x + qn.y * qn.y);
                        }
                    }
                }
            }
            if (S.mode == 1) {
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
    // the row maxima of Qn over this tile's columns, for the next stage's digit split
    if (S.mode != 2 && S.rmax) {
        qmax = fmax(qmax, __shfl_xor(qmax, 16));
        qmax = fmax(qmax, __shfl_xor(qmax, 32));
        if (kq == 0 && i < g.N) atomicMax(&S.rmax[R], (unsigned long long)__double_as_longlong(qmax));
    }
}

}  // namespace

MixedGeo mixed_geo(int A, int N, int NP, int ND, int wc, int wr, int pf)
{
    MixedGeo g;
    g.ND = ND;
    g.A = A;
    g.N = N;
    g.NP = NP;
    g.mat = (size_t)NP * NP;
    g.KC = (N + 63) / 64;
    g.RB = A * NP / 16;
    g.RBP = (g.RB + 7) / 8 * 8;
    // column blocks per workgroup: given, or the width in {4, 3, 2} that pads the least
    // (wider on ties)
    const int nb = NP / 16;
    g.WC = 4;
    for (int w = 3; w >= 2; --w)
        if ((nb + w - 1) / w * w < (nb + g.WC - 1) / g.WC * g.WC) g.WC = w;
    if (wc >= 2 && wc <= 4) g.WC = wc;
    g.CBP = (nb + g.WC - 1) / g.WC * g.WC;
    // row blocks per workgroup: eight (two waves per SIMD, at most 256 VGPRs each) unless four
    // are asked for, or the 7-digit 4-block accumulators need the registers of one wave per
    // SIMD.  The LDS holds only the M0 fragments, 2 WC ND KB.
    g.WR = (wr == 4 || (ND == 7 && g.WC == 4)) ? 4 : 8;
    // a Q prefetch two chunks deep needs a third set of Q registers: tiles whose registers
    // have room for it
    g.PF = (pf == 2 && mixed_pf2_fits(ND, g.WC)) ? 2 : 1;
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

void launch_qsplit(hipStream_t st, const MixedGeo &g, const double2 *Q, int8_t *qd, int *eq,
                   const unsigned long long *vmax, unsigned long long *clear)
{
    switch (g.ND) {
    case 5: hipLaunchKernelGGL(k_qsplit<5>, dim3(g.RBP), dim3(256), 0, st, g, Q, qd, eq, vmax, clear); break;
    case 6: hipLaunchKernelGGL(k_qsplit<6>, dim3(g.RBP), dim3(256), 0, st, g, Q, qd, eq, vmax, clear); break;
    default: hipLaunchKernelGGL(k_qsplit<7>, dim3(g.RBP), dim3(256), 0, st, g, Q, qd, eq, vmax, clear); break;
    }
}

template <int ND, int WC, int WR, int PF>
static void stage_i8_on(hipStream_t st, const StageLaunch &Lc)
{
    constexpr int NF = WC * ND;
    const size_t lds = (size_t)2 * NF * 1024;
    static_assert(2 * NF <= 160, "tile exceeds the LDS");
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_stage_i8l<ND, WC, WR, PF>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    hipLaunchKernelGGL((k_stage_i8l<ND, WC, WR, PF>), dim3(8 * Lc.tiles_per_xcd), dim3(64 * WR), lds, st, Lc);
}

template <int ND, int WC>
static void stage_i8_wr(hipStream_t st, const StageLaunch &Lc)
{
    if constexpr (mixed_pf2_fits(ND, WC)) {
        if (Lc.S.g.PF == 2)
            return Lc.S.g.WR == 8 ? stage_i8_on<ND, WC, 8, 2>(st, Lc) : stage_i8_on<ND, WC, 4, 2>(st, Lc);
    }
    if constexpr (!(ND == 7 && WC == 4)) {
        if (Lc.S.g.WR == 8) return stage_i8_on<ND, WC, 8, 1>(st, Lc);
    }
    stage_i8_on<ND, WC, 4, 1>(st, Lc);
}

template <int ND>
static void stage_i8_nd(hipStream_t st, const StageLaunch &Lc)
{
    switch (Lc.S.g.WC) {
    case 2: stage_i8_wr<ND, 2>(st, Lc); break;
    case 3: stage_i8_wr<ND, 3>(st, Lc); break;
    default: stage_i8_wr<ND, 4>(st, Lc); break;
    }
}

void launch_stage_i8(hipStream_t st, const MixedStage &S)
{
    StageLaunch Lc;
    Lc.S = S;
    const int rowt = S.g.RBP / S.g.WR;
    Lc.ncolt = S.g.CBP / S.g.WC;
    Lc.ntiles = rowt * Lc.ncolt;
    Lc.tiles_per_xcd = (Lc.ntiles + 7) / 8;
    switch (S.g.ND) {
    case 5: stage_i8_nd<5>(st, Lc); break;
    case 6: stage_i8_nd<6>(st, Lc); break;
    default: stage_i8_nd<7>(st, Lc); break;
    }
}

}  // namespace trhepd
