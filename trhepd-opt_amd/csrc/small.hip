// Fused slice loop for small beam counts (N <= 48): one workgroup owns one angle for
// the whole integration.  The ratio-form state [X^H | Y^H] (here Q | P) and the slice
// matrix M live in LDS; the stage products Q*M run on v_mfma_f64_16x16x4f64 (four real
// products per complex block), the renormalisation is an in-LDS LU with partial pivoting.
//
// Reference: hp_int_spa surf_prkn.f90:278-330 (RKN), lf_int_spa surf_magnus.f90:46-110
// (Stormer-Verlet), rk4_int surf_rk.f90:27-91 (classical RK4), invstransh/stransh matcomp.f90:140-196, vvi2mate matcomp.f90:253-279,
// cond_est matcomp.f90:498-528, matAbdivB matcomp.f90:53-69, intensity surflf.f90:143-151.
//
// The end-of-slice kick of the RKN schemes is shared with the first stage of the next
// slice (psave, surf_prkn.f90:306-315).  cond_est reads only Q, which that kick does not
// change, so the renormalisation decision is taken first and the two kicks are applied
// as one product with coefficient 2*c_b1 when no renormalisation follows.
#include "device.hpp"
#include "internal.hpp"

namespace trhepd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int NP>
struct SLds {
    static constexpr int LD = NP + 1;
    double qr[NP * LD], qi[NP * LD];   // Q = X^H, column-major
    double pr[NP * LD], pi[NP * LD];   // P = Y^H, column-major
    double mr[NP * LD], mi[NP * LD];   // M, row-major
    int gm[NP * NP];                   // signed gather map, row-major
    double g2r[NP], g2i[NP];           // conj(gamma^2)
    double gcr[NP], gci[NP];           // conj(gamma)
    double red0[NP], red1[NP];
    int piv;
    int flag;
    // NP = 48: each wave's share of the ninth output block of a stage product (d4 re, d4 im
    // per lane)
    double part[NP == 48 ? 4 * 8 * 64 : 1];
};

struct SmallArgs {
    const VVi *table;
    int nv, ns, N, method;
    double dz;
    const double *steps;    // slice thicknesses of an RKN table built from a slice list, or null
    RknCoef co;
    const int *gmap;        // (N*N) row-major signed map
    const double2 *gamma;   // [nang][N]
    double2 *f;             // [nang][N*N] column-major
    double *f2;             // [nang][N] or null
    int nb0;                // 1-based
    const double *wnsga2;   // [nang]
    int *counters;          // [0] renorm count, [1] singular flag
    int period;             // renormalisation period (0: Gershgorin rule)
};

__device__ __forceinline__ double2 cdiv(double ar, double ai, double br, double bi)
{
    // Smith's algorithm
    if (fabs(br) >= fabs(bi)) {
        const double r = bi / br, den = br + bi * r;
        return make_double2((ar + ai * r) / den, (ai - ar * r) / den);
    }
    const double r = br / bi, den = bi + br * r;
    return make_double2((ar * r + ai) / den, (ai * r - ar) / den);
}

// M(k, j) of column c into LDS (vvi2mate): upper -conj(v+vi), lower -v+vi,
// diagonal -conj(v+vi)[1] - conj(gamma_j^2).
template <int NP>
__device__ void gather_m(SLds<NP> &S, const VVi *__restrict__ col, int N)
{
    constexpr int LD = SLds<NP>::LD;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int k = e / N, j = e - k * N;
        const int g = S.gm[e];
        double vr, vi_;
        if (g > 0) {
            const VVi t = col[g - 1];
            vr = -(t.v.x + t.vi.x);
            vi_ = t.v.y + t.vi.y;
        } else if (g < 0) {
            const VVi t = col[-g - 1];
            vr = -t.v.x + t.vi.x;
            vi_ = -t.v.y + t.vi.y;
        } else {
            const VVi t = col[0];
            vr = -(t.v.x + t.vi.x) - S.g2r[j];
            vi_ = (t.v.y + t.vi.y) - S.g2i[j];
        }
        S.mr[k * LD + j] = vr;
        S.mi[k * LD + j] = vi_;
    }
}

// The thread's elements of an N x N matrix, e = tid + 256 q = i + j N (workgroups of 256
// threads), worked out once per integration: ij = i + 256 j (-1 past N^2), and for the
// slice matrix the table entry of M(i, j) and its form, gk = 4 (|g| - 1) + 2 (upper), + 1
// (lower), + 0 (diagonal, entry 0).
template <int NP>
struct Elems {
    static constexpr int E = (NP * NP + 255) / 256;
    int ij[E], gk[E];
};

template <int NP>
__device__ __forceinline__ void elems_init(const SLds<NP> &S, int N, Elems<NP> &el)
{
#pragma unroll
    for (int q = 0; q < Elems<NP>::E; ++q) {
        const int e = threadIdx.x + 256 * q;
        const bool in = e < N * N;
        const int i = e % N, j = e / N;
        const int g = in ? S.gm[i * N + j] : 0;
        el.ij[q] = in ? i + 256 * j : -1;
        el.gk[q] = g > 0 ? 4 * (g - 1) + 2 : (g < 0 ? 4 * (-g - 1) + 1 : 0);
    }
}

// The next slice matrix of the RKN loop, loaded one stage ahead: the table entries of the
// thread's elements of M are held in registers while the current stage runs, and m_commit
// writes M into LDS with the values gather_m would.  Only LDS-only barriers lie between
// the two, so the loads stay in flight across them.
template <int NP>
struct MNext {
    VVi t[Elems<NP>::E];
};

template <int NP>
__device__ __forceinline__ void m_issue(const VVi *__restrict__ col, const Elems<NP> &el, MNext<NP> &m)
{
    // every load is issued (past N^2 entry 0, not used): a load under a branch would be
    // waited for where the branches join
#pragma unroll
    for (int q = 0; q < Elems<NP>::E; ++q) m.t[q] = col[el.gk[q] >> 2];
}

template <int NP>
__device__ __forceinline__ void m_commit(SLds<NP> &S, const Elems<NP> &el, const MNext<NP> &m)
{
    constexpr int LD = SLds<NP>::LD;
#pragma unroll
    for (int q = 0; q < Elems<NP>::E; ++q) {
        if (el.ij[q] < 0) continue;
        const int i = el.ij[q] & 255, j = el.ij[q] >> 8, form = el.gk[q] & 3;
        const VVi t = m.t[q];
        double vr, vi_;
        if (form == 2) {
            vr = -(t.v.x + t.vi.x);
            vi_ = t.v.y + t.vi.y;
        } else if (form == 1) {
            vr = -t.v.x + t.vi.x;
            vi_ = -t.v.y + t.vi.y;
        } else {
            vr = -(t.v.x + t.vi.x) - S.g2r[j];
            vi_ = (t.v.y + t.vi.y) - S.g2i[j];
        }
        S.mr[i * LD + j] = vr;
        S.mi[i * LD + j] = vi_;
    }
}

// P += alpha * Q * M on the matrix cores; each wave owns output blocks blk = wave + 4t.
// At NP = 48 (nine blocks on four waves) wave w owns blocks w and w + 4 and the k steps
// 3w .. 3w + 2 of the ninth, whose four partial products are added in wave order.
template <int NP>
__device__ void gemm_acc(SLds<NP> &S, int N, double alpha)
{
    constexpr int LD = SLds<NP>::LD;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nbk = (N + 15) >> 4, ks = (N + 3) >> 2;
    const int kq = lane >> 4, l16 = lane & 15;
    // the product of block (bi, bj) over k steps [s0, s1)
    auto prod = [&](int bi, int bj, int s0, int s1, d4 &cr, d4 &ci) {
        const int ar = bi * 16 + l16, bc = bj * 16 + l16;
        cr = d4{0, 0, 0, 0};
        ci = d4{0, 0, 0, 0};
        if constexpr (NP <= 32) {
            // all k steps unrolled to the padded size, so that the LDS loads issue ahead of
            // the products (s0 is 0 here; steps from s1 on are skipped).  At NP = 48 the
            // operands of 12 steps would not fit beside the kernel's other registers.
            double a_r[NP / 4], a_i[NP / 4], b_r[NP / 4], b_i[NP / 4];
#pragma unroll
            for (int s = 0; s < NP / 4; ++s) {
                const int kk = s * 4 + kq;
                a_r[s] = S.qr[ar + kk * LD];
                a_i[s] = S.qi[ar + kk * LD];
                b_r[s] = S.mr[kk * LD + bc];
                b_i[s] = S.mi[kk * LD + bc];
            }
#pragma unroll
            for (int s = 0; s < NP / 4; ++s) {
                if (s >= s1) break;
                cr = mfma4(a_r[s], b_r[s], cr);
                cr = mfma4(-a_i[s], b_i[s], cr);
                ci = mfma4(a_r[s], b_i[s], ci);
                ci = mfma4(a_i[s], b_r[s], ci);
            }
        } else {
            // the operands of step s + 1 are loaded before the products of step s
            auto ld = [&](int s, double &a_r, double &a_i, double &b_r, double &b_i) {
                const int kk = s * 4 + kq;
                a_r = S.qr[ar + kk * LD];
                a_i = S.qi[ar + kk * LD];
                b_r = S.mr[kk * LD + bc];
                b_i = S.mi[kk * LD + bc];
            };
            double a_r, a_i, b_r, b_i;
            ld(min(s0, NP / 4 - 1), a_r, a_i, b_r, b_i);
            for (int s = s0; s < s1; ++s) {
                double n_r, n_i, m_r, m_i;
                ld(min(s + 1, NP / 4 - 1), n_r, n_i, m_r, m_i);
                cr = mfma4(a_r, b_r, cr);
                cr = mfma4(-a_i, b_i, cr);
                ci = mfma4(a_r, b_i, ci);
                ci = mfma4(a_i, b_r, ci);
                a_r = n_r;
                a_i = n_i;
                b_r = m_r;
                b_i = m_i;
            }
        }
    };
    auto add = [&](int bi, int bj, const d4 &cr, const d4 &ci) {
        const int j = bj * 16 + l16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = bi * 16 + kq + 4 * r;
            if (i < N && j < N) {
                S.pr[i + j * LD] += alpha * cr[r];
                S.pi[i + j * LD] += alpha * ci[r];
            }
        }
    };
    if (NP == 48 && nbk == 3 && nw == 4) {
        d4 cr, ci;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int blk = wave + 4 * t;
            prod(blk / 3, blk % 3, 0, ks, cr, ci);
            add(blk / 3, blk % 3, cr, ci);
        }
        prod(2, 2, min(3 * wave, ks), min(3 * wave + 3, ks), cr, ci);
        double *pw = S.part + (size_t)wave * 8 * 64;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            pw[r * 64 + lane] = cr[r];
            pw[(4 + r) * 64 + lane] = ci[r];
        }
        lds_barrier();
        // wave w adds element r = w of every lane's share, in wave order
        const int r = wave;
        double sr = S.part[r * 64 + lane], si = S.part[(4 + r) * 64 + lane];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            sr += S.part[(size_t)w * 8 * 64 + r * 64 + lane];
            si += S.part[(size_t)w * 8 * 64 + (4 + r) * 64 + lane];
        }
        const int i = 32 + kq + 4 * r, j = 32 + l16;
        if (i < N && j < N) {
            S.pr[i + j * LD] += alpha * sr;
            S.pi[i + j * LD] += alpha * si;
        }
        return;
    }
    for (int blk = wave; blk < nbk * nbk; blk += nw) {
        const int bi = blk / nbk, bj = blk - bi * nbk;
        d4 cr, ci;
        prod(bi, bj, 0, ks, cr, ci);
        add(bi, bj, cr, ci);
    }
}

// Output blocks a wave owns (blk = wave + 4t): at most MT per wave.  Element (r) of block
// t of a lane is (bi*16 + kq + 4r, bj*16 + l16), the MFMA output layout; the RK4 stages
// keep their per-element values in registers in this ownership.
template <int NP>
struct Own {
    static constexpr int NB = NP / 16, MT = (NB * NB + 3) / 4;
};

// K = Q M (or K = M when ident) for the wave's blocks
template <int NP>
__device__ void qm_blocks(SLds<NP> &S, int N, bool ident, double2 (&K)[Own<NP>::MT][4])
{
    constexpr int LD = SLds<NP>::LD;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nbk = (N + 15) >> 4, ks = (N + 3) >> 2;
    const int kq = lane >> 4, l16 = lane & 15;
#pragma unroll
    for (int t = 0; t < Own<NP>::MT; ++t) {
        const int blk = wave + nw * t;
        if (blk >= nbk * nbk) break;
        const int bi = blk / nbk, bj = blk - bi * nbk;
        if (ident) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = bi * 16 + kq + 4 * r, j = bj * 16 + l16;
                K[t][r] = make_double2(S.mr[i * LD + j], S.mi[i * LD + j]);
            }
            continue;
        }
        const int ar = bi * 16 + l16, bc = bj * 16 + l16;
        d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
        for (int s = 0; s < ks; ++s) {
            const int kk = s * 4 + kq;
            const double a_r = S.qr[ar + kk * LD], a_i = S.qi[ar + kk * LD];
            const double b_r = S.mr[kk * LD + bc], b_i = S.mi[kk * LD + bc];
            cr = mfma4(a_r, b_r, cr);
            cr = mfma4(-a_i, b_i, cr);
            ci = mfma4(a_r, b_i, ci);
            ci = mfma4(a_i, b_r, ci);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) K[t][r] = make_double2(cr[r], ci[r]);
    }
}

// One classical RK4 slice of X' = Y, Y' = X M in the ratio form (rk4_int, surf_rk.f90:27-91):
// K_a = X_a M at X_1 = X, X_2 = X + h/2 Y_1, X_3 = X + h/2 Y_2, X_4 = X + h Y_3 with
// Y_1 = Y, Y_2 = Y + h/2 K_1, Y_3 = Y + h/2 K_2, Y_4 = Y + h K_3; then
// X += h/6 (Y_1 + 2 Y_2 + 2 Y_3 + Y_4), Y += h/6 (K_1 + 2 K_2 + 2 K_3 + K_4).
// Q holds the current X_a (updated in place between barriers), P holds Y; the sums and the
// stage values stay in registers.  M0 is column c0 of the table, M1 c0+1, M2 c0+2.
template <int NP>
__device__ void rk4_slice(SLds<NP> &S, const VVi *__restrict__ table, int nv, int c0, int N, double h, bool xid)
{
    constexpr int LD = SLds<NP>::LD, MT = Own<NP>::MT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nbk = (N + 15) >> 4;
    const int kq = lane >> 4, l16 = lane & 15;
    double2 K[MT][4], ax[MT][4], ay[MT][4], y2[MT][4], y3[MT][4];
    auto each = [&](auto &&fn) {
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const int blk = wave + nw * t;
            if (blk >= nbk * nbk) break;
            const int bi = blk / nbk, bj = blk - bi * nbk;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = bi * 16 + kq + 4 * r, j = bj * 16 + l16;
                if (i < N && j < N) fn(t, r, i + j * LD);
            }
        }
    };
    auto cadd = [](double2 a, double c, double2 b) { return make_double2(a.x + c * b.x, a.y + c * b.y); };
    // stage 1
    gather_m<NP>(S, table + (size_t)c0 * nv, N);
    __syncthreads();
    qm_blocks<NP>(S, N, xid, K);
    __syncthreads();
    each([&](int t, int r, int e) {
        const double2 y = make_double2(S.pr[e], S.pi[e]);
        ax[t][r] = y;
        ay[t][r] = K[t][r];
        y2[t][r] = cadd(y, h / 2, K[t][r]);
        S.qr[e] += (h / 2) * y.x;          // X_2
        S.qi[e] += (h / 2) * y.y;
    });
    gather_m<NP>(S, table + (size_t)(c0 + 1) * nv, N);
    __syncthreads();
    // stage 2
    qm_blocks<NP>(S, N, false, K);
    __syncthreads();
    each([&](int t, int r, int e) {
        const double2 y = make_double2(S.pr[e], S.pi[e]);
        ax[t][r] = cadd(ax[t][r], 2.0, y2[t][r]);
        ay[t][r] = cadd(ay[t][r], 2.0, K[t][r]);
        y3[t][r] = cadd(y, h / 2, K[t][r]);
        S.qr[e] += (h / 2) * (y2[t][r].x - y.x);   // X_3 = X + h/2 Y_2
        S.qi[e] += (h / 2) * (y2[t][r].y - y.y);
    });
    __syncthreads();
    // stage 3
    qm_blocks<NP>(S, N, false, K);
    __syncthreads();
    each([&](int t, int r, int e) {
        const double2 y = make_double2(S.pr[e], S.pi[e]);
        ax[t][r] = cadd(ax[t][r], 2.0, y3[t][r]);
        ay[t][r] = cadd(ay[t][r], 2.0, K[t][r]);
        const double2 y4 = cadd(y, h, K[t][r]);
        S.qr[e] += h * y3[t][r].x - (h / 2) * y2[t][r].x;   // X_4 = X + h Y_3
        S.qi[e] += h * y3[t][r].y - (h / 2) * y2[t][r].y;
        ax[t][r] = cadd(ax[t][r], 1.0, y4);
    });
    gather_m<NP>(S, table + (size_t)(c0 + 2) * nv, N);
    __syncthreads();
    // stage 4 and the update
    qm_blocks<NP>(S, N, false, K);
    __syncthreads();
    each([&](int t, int r, int e) {
        ay[t][r] = cadd(ay[t][r], 1.0, K[t][r]);
        const double2 x = make_double2(S.qr[e] - h * y3[t][r].x, S.qi[e] - h * y3[t][r].y);   // X
        S.qr[e] = x.x + (h / 6) * ax[t][r].x;
        S.qi[e] = x.y + (h / 6) * ax[t][r].y;
        S.pr[e] += (h / 6) * ay[t][r].x;
        S.pi[e] += (h / 6) * ay[t][r].y;
    });
    __syncthreads();
}

// [Q | P] <- [LU | Q^{-1} P]: zgetrf(N, 2N) with |re|+|im| partial pivoting, then
// back substitution with the upper factor (matAbdivB, matcomp.f90:53-69).  The back
// substitution keeps P in registers (N <= 64: lane i holds row i, wave w columns w, w + 4,
// ...) and needs no barrier.
template <int NP>
__device__ void lu_solve(SLds<NP> &S, int N)
{
    static_assert(NP <= 64, "a row per lane in the back substitution");
    constexpr int LD = SLds<NP>::LD;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nt = blockDim.x, nw = nt >> 6;
    const bool row = lane < N;
    for (int k = 0; k < N; ++k) {
        if (wave == 0) {
            double best = -1.0;
            int bi = N;
            for (int i = k + lane; i < N; i += 64) {
                const double v = fabs(S.qr[i + k * LD]) + fabs(S.qi[i + k * LD]);
                if (v > best) { best = v; bi = i; }
            }
            pivot_wave(best, bi);   // the first maximum, whatever the order of combination
            if (lane == 0) {
                S.piv = bi;
                if (best == 0.0) S.flag = 1;
            }
        }
        __syncthreads();
        const int p = S.piv;
        if (p != k) {
            for (int c = tid; c < 2 * N; c += nt) {
                double *re = c < N ? S.qr : S.pr, *im = c < N ? S.qi : S.pi;
                const int cc = c < N ? c : c - N;
                const double tr = re[k + cc * LD], ti = im[k + cc * LD];
                re[k + cc * LD] = re[p + cc * LD];
                im[k + cc * LD] = im[p + cc * LD];
                re[p + cc * LD] = tr;
                im[p + cc * LD] = ti;
            }
        }
        __syncthreads();
        const double pkr = S.qr[k + k * LD], pki = S.qi[k + k * LD];
        if (pkr != 0.0 || pki != 0.0) {
            const double2 rp = cdiv(1.0, 0.0, pkr, pki);
            for (int i = k + 1 + tid; i < N; i += nt) {
                const double lr = S.qr[i + k * LD], li = S.qi[i + k * LD];
                S.qr[i + k * LD] = lr * rp.x - li * rp.y;
                S.qi[i + k * LD] = lr * rp.y + li * rp.x;
            }
        }
        __syncthreads();
        const int nr = N - 1 - k, ncq = N - 1 - k, nc = ncq + N;
        for (int e = tid; e < nr * nc; e += nt) {
            const int i = k + 1 + e % nr, c = e / nr;
            const double lr = S.qr[i + k * LD], li = S.qi[i + k * LD];
            double *re, *im;
            int cc;
            if (c < ncq) { re = S.qr; im = S.qi; cc = k + 1 + c; }
            else { re = S.pr; im = S.pi; cc = c - ncq; }
            const double ur = re[k + cc * LD], ui = im[k + cc * LD];
            re[i + cc * LD] -= lr * ur - li * ui;
            im[i + cc * LD] -= lr * ui + li * ur;
        }
        __syncthreads();
    }
    // back substitution: X = U^{-1} P by columns, P(k, j) / U(k, k) as a product with the
    // reciprocal of the pivot
    constexpr int JQ = (NP + 3) / 4;
    double xr[JQ], xi[JQ];
#pragma unroll
    for (int q = 0; q < JQ; ++q) {
        const int j = wave + nw * q;
        xr[q] = (row && j < N) ? S.pr[lane + j * LD] : 0.0;
        xi[q] = (row && j < N) ? S.pi[lane + j * LD] : 0.0;
    }
    const double2 rd = row ? cdiv(1.0, 0.0, S.qr[lane * (LD + 1)], S.qi[lane * (LD + 1)]) : make_double2(0, 0);
    for (int k = N - 1; k >= 0; --k) {
        const double rr = readlane_f64(rd.x, k), ri = readlane_f64(rd.y, k);
        const double ur = lane < k ? S.qr[lane + k * LD] : 0.0, ui = lane < k ? S.qi[lane + k * LD] : 0.0;
#pragma unroll
        for (int q = 0; q < JQ; ++q) {
            const double pr = readlane_f64(xr[q], k), pi = readlane_f64(xi[q], k);
            const double x_r = pr * rr - pi * ri, x_i = pr * ri + pi * rr;
            if (lane == k) {
                xr[q] = x_r;
                xi[q] = x_i;
            } else if (lane < k) {
                xr[q] -= x_r * ur - x_i * ui;
                xi[q] -= x_r * ui + x_i * ur;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < JQ; ++q) {
        const int j = wave + nw * q;
        if (row && j < N) {
            S.pr[lane + j * LD] = xr[q];
            S.pi[lane + j * LD] = xi[q];
        }
    }
    __syncthreads();
}

// The whole integration of angle `pt` of one batch by one workgroup.
// co: the integrator's coefficients, read where the kernel arguments hold them (a local copy
// of their dynamically indexed array would live in scratch memory)
template <int NP>
__device__ __forceinline__ void small_body(const SmallArgs &A, const RknCoef &co, const int pt, SLds<NP> &S)
{
    constexpr int LD = SLds<NP>::LD;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int N = A.N;
    const double dzu = A.dz;
    for (int e = tid; e < NP * LD; e += nt) {
        S.qr[e] = 0; S.qi[e] = 0; S.pr[e] = 0; S.pi[e] = 0; S.mr[e] = 0; S.mi[e] = 0;
    }
    for (int e = tid; e < N * N; e += nt) S.gm[e] = A.gmap[e];
    for (int j = tid; j < N; j += nt) {
        const double2 g = A.gamma[(size_t)pt * N + j];
        S.g2r[j] = g.x * g.x - g.y * g.y;    // conj(g*g)
        S.g2i[j] = -(g.x * g.y + g.y * g.x);
        S.gcr[j] = g.x;
        S.gci[j] = -g.y;
    }
    if (tid == 0) S.flag = 0;
    __syncthreads();
    // invstransh: X = I + F^H, Y = I - F^H
    const double2 *F = A.f + (size_t)pt * N * N;
    for (int e = tid; e < N * N; e += nt) {
        const int i = e % N, j = e / N;
        const double2 fji = F[j + i * N];
        const double d = (i == j) ? 1.0 : 0.0;
        S.qr[i + j * LD] = d + fji.x;
        S.qi[i + j * LD] = -fji.y;
        S.pr[i + j * LD] = d - fji.x;
        S.pi[i + j * LD] = fji.y;
    }
    __syncthreads();
    lu_solve<NP>(S, N);
    for (int e = tid; e < N * N; e += nt) {
        const int i = e % N, j = e / N;
        // (0,1) * Y(i,j) * conj(g_i)
        const double yr = -S.pi[i + j * LD], yi = S.pr[i + j * LD];
        S.pr[i + j * LD] = yr * S.gcr[i] - yi * S.gci[i];
        S.pi[i + j * LD] = yr * S.gci[i] + yi * S.gcr[i];
    }
    __syncthreads();

    int xid = 1, nren = 0;
    if (A.method == TRHEPD_LF) {   // X^H = I + (dz/2) Y^H  (surf_magnus.f90:63-67)
        for (int e = tid; e < N * N; e += nt) {
            const int i = e % N, j = e / N;
            S.qr[i + j * LD] = (dzu / 2) * S.pr[i + j * LD] + (i == j ? 1.0 : 0.0);
            S.qi[i + j * LD] = (dzu / 2) * S.pi[i + j * LD];
        }
        __syncthreads();
        xid = 0;
    }
    const int s = co.s;
    const double cb1 = co.bab[0];
    const bool rkn = A.method == TRHEPD_SP4 || A.method == TRHEPD_SP6;
    if (A.method == TRHEPD_RK4) {   // X^H = I (surf_rk.f90:46-47)
        for (int e = tid; e < N * N; e += nt) {
            const int i = e % N, j = e / N;
            S.qr[i + j * LD] = (i == j ? 1.0 : 0.0);
            S.qi[i + j * LD] = 0.0;
        }
        __syncthreads();
    }
    Elems<NP> el;
    elems_init<NP>(S, N, el);
    MNext<NP> mpre;
    int mcol = 0;   // the table column whose entries mpre holds (RKN)
    const int mlast = s * A.ns;   // the last column of an RKN table
    auto next_m = [&]() {   // past the last column the last one again (not used)
        m_issue<NP>(A.table + (size_t)min(mcol + 1, mlast) * A.nv, el, mpre);
        ++mcol;
    };
    if (rkn) m_issue<NP>(A.table, el, mpre);
    // f(i, j) for each of the thread's elements, as `for (e = tid; e < N * N; e += 256)`
    auto each = [&](auto f) {
#pragma unroll
        for (int q = 0; q < Elems<NP>::E; ++q)
            if (el.ij[q] >= 0) f(el.ij[q] & 255, el.ij[q] >> 8);
    };
    for (int l = 1; l <= A.ns; ++l) {
        // this slice's thickness, and the next one's (the last slice's past the end)
        const double dz = A.steps ? A.steps[l - 1] : dzu;
        const double dzn = A.steps ? A.steps[l < A.ns ? l : A.ns - 1] : dzu;
        if (A.method == TRHEPD_RK4) {
            rk4_slice<NP>(S, A.table, A.nv, 2 * (l - 1), N, dz, xid);
        } else if (A.method == TRHEPD_LF) {
            // Stormer-Verlet: kick with M(l), drift dz (dz/2 on the last slice)
            const VVi *col = A.table + (size_t)(l - 1) * A.nv;
            gather_m<NP>(S, col, N);
            __syncthreads();
            if (xid) {
                for (int e = tid; e < N * N; e += nt) {
                    const int i = e % N, j = e / N;
                    S.pr[i + j * LD] += dz * S.mr[i * LD + j];
                    S.pi[i + j * LD] += dz * S.mi[i * LD + j];
                }
            } else {
                gemm_acc<NP>(S, N, dz);
            }
            __syncthreads();
            const double dz2 = (l == A.ns) ? dz / 2 : dz;
            for (int e = tid; e < N * N; e += nt) {
                const int i = e % N, j = e / N;
                if (xid) {
                    S.qr[i + j * LD] = dz2 * S.pr[i + j * LD] + (i == j ? 1.0 : 0.0);
                    S.qi[i + j * LD] = dz2 * S.pi[i + j * LD];
                } else {
                    S.qr[i + j * LD] += dz2 * S.pr[i + j * LD];
                    S.qi[i + j * LD] += dz2 * S.pi[i + j * LD];
                }
            }
            __syncthreads();
        } else {
            // M of table column s (l - 1) + k - 1 for stage k, s l for the end kick: the
            // columns are used in order, each loaded one stage ahead (mcol: the column whose
            // entries mpre holds).  After a renormalisation the first kick's column is the
            // previous end kick's, still in LDS.
            for (int k = 1; k <= s; ++k) {
                const double cb = co.bab[2 * k - 2], ca = co.bab[2 * k - 1];
                if (k == 1) {
                    if (xid) {
                        if (mcol == 0) {   // the first slice: M not yet in LDS
                            m_commit<NP>(S, el, mpre);
                            next_m();
                            lds_barrier();
                        }
                        each([&](int i, int j) {
                            const double pr = S.pr[i + j * LD] + (cb * dz) * S.mr[i * LD + j];
                            const double pi = S.pi[i + j * LD] + (cb * dz) * S.mi[i * LD + j];
                            S.pr[i + j * LD] = pr;
                            S.pi[i + j * LD] = pi;
                            S.qr[i + j * LD] = (ca * dz) * pr + (i == j ? 1.0 : 0.0);
                            S.qi[i + j * LD] = (ca * dz) * pi;
                        });
                    } else {
                        each([&](int i, int j) {
                            S.qr[i + j * LD] += (ca * dz) * S.pr[i + j * LD];
                            S.qi[i + j * LD] += (ca * dz) * S.pi[i + j * LD];
                        });
                    }
                    lds_barrier();
                } else {
                    m_commit<NP>(S, el, mpre);   // column s (l - 1) + k - 1
                    next_m();
                    lds_barrier();
                    gemm_acc<NP>(S, N, cb * dz);
                    lds_barrier();
                    each([&](int i, int j) {
                        S.qr[i + j * LD] += (ca * dz) * S.pr[i + j * LD];
                        S.qi[i + j * LD] += (ca * dz) * S.pi[i + j * LD];
                    });
                    lds_barrier();
                }
            }
        }
        // cond_est of Q (Gershgorin, matcomp.f90:498-528), needed only by the reference's rule;
        // each column's sum in row order, its loads issued eight rows at a time
        for (int j = tid; j < N && A.period <= 0; j += nt) {
            double t = 0.0;
            for (int i0 = 0; i0 < N; i0 += 8) {
                double re[8], im[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int i = min(i0 + u, N - 1);
                    re[u] = S.qr[i + j * LD];
                    im[u] = S.qi[i + j * LD];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int i = i0 + u;
                    if (i < N && i != j) t += sqrt(re[u] * re[u] + im[u] * im[u]);
                }
            }
            const double dj = hypot(S.qr[j + j * LD], S.qi[j + j * LD]);
            S.red0[j] = dj - t;
            S.red1[j] = dj + t;
        }
        if (rkn) {
            m_commit<NP>(S, el, mpre);   // column s l, the end kick's
            next_m();
            lds_barrier();
        } else {
            __syncthreads();
        }
        double d = 0.0;
        if (A.period <= 0) {
            double mn = hypot(S.qr[0], S.qi[0]), mx = mn;
            for (int j = 0; j < N; ++j) {
                mn = fmin(mn, S.red0[j]);
                mx = fmax(mx, S.red1[j]);
            }
            d = (mn < 0.0) ? 1e100 : mx / mn;
        }
        const bool renorm = (A.period > 0 ? (l % A.period == 0) : (d > 1e3)) || (l == A.ns);
        if (rkn) {
            // the end kick, merged with the next slice's first kick unless the angle renormalises
            gemm_acc<NP>(S, N, renorm ? cb1 * dz : cb1 * dz + cb1 * dzn);
            lds_barrier();
        }
        if (renorm) {
            lu_solve<NP>(S, N);
            xid = 1;
            ++nren;
            if (A.method == TRHEPD_RK4) {   // matsetid: X^H = I (surf_rk.f90:82-84)
                for (int e = tid; e < N * N; e += nt) {
                    const int i = e % N, j = e / N;
                    S.qr[i + j * LD] = (i == j ? 1.0 : 0.0);
                    S.qi[i + j * LD] = 0.0;
                }
                __syncthreads();
            }
        } else {
            xid = 0;
        }
    }
    // stransh: X = conj(G) - iY, Y' = conj(G) + iY, Y'' = X^{-1} Y', F = Y''^H
    for (int e = tid; e < N * N; e += nt) {
        const int i = e % N, j = e / N;
        const double yr = S.pr[i + j * LD], yi = S.pi[i + j * LD];
        const double dr = (i == j) ? S.gcr[i] : 0.0, di = (i == j) ? S.gci[i] : 0.0;
        S.qr[i + j * LD] = dr + yi;     // conj(g) + (0,-1) y
        S.qi[i + j * LD] = di - yr;
        S.pr[i + j * LD] = dr - yi;     // conj(g) + (0, 1) y
        S.pi[i + j * LD] = di + yr;
    }
    __syncthreads();
    lu_solve<NP>(S, N);
    double2 *Fo = A.f + (size_t)pt * N * N;
    for (int e = tid; e < N * N; e += nt) {
        const int i = e % N, j = e / N;
        Fo[i + j * N] = make_double2(S.pr[j + i * LD], -S.pi[j + i * LD]);
    }
    if (A.f2) {
        const int c0 = A.nb0 - 1;
        const double w = A.wnsga2[pt];
        for (int i = tid; i < N; i += nt) {
            const double sr = S.gcr[i];   // Re(gamma_i)
            const double fr = S.pr[c0 + i * LD], fi = -S.pi[c0 + i * LD];
            A.f2[(size_t)pt * N + i] = (sr > 1e-10) ? (fr * fr + fi * fi) * w / sr : 0.0;
        }
    }
    if (tid == 0) {
        atomicAdd(&A.counters[0], nren);
        if (S.flag) atomicOr(&A.counters[1], 1);
    }
}

// one workgroup per angle of one batch
template <int NP>
__global__ __launch_bounds__(256) void k_small(SmallArgs A)
{
    __shared__ SLds<NP> S;
    small_body<NP>(A, A.co, blockIdx.x, S);
}

// Batches that differ only in their table, arrays and slice count (several structures with
// the same beams, integrator and step) in one launch: one workgroup per (batch, angle), the
// batch found from the first global workgroup of each.  The per-batch part travels in the
// kernel arguments, so a launch needs no device memory of its own.
constexpr int SMALL_MULTI = 24;
struct SmallPart {
    const VVi *table;
    const int *gmap;
    const double2 *gamma;
    double2 *f;
    double *f2;
    const double *wnsga2;
    int *counters;
    int nv, ns, period, first;
};
struct SmallMultiArgs {
    SmallArgs base;   // the fields every batch shares (N, method, dz, coefficients, nb0)
    int nbatch;
    SmallPart part[SMALL_MULTI];
};
static_assert(sizeof(SmallMultiArgs) <= 4096, "kernel arguments of the fused launch exceed 4 KiB");

template <int NP>
__global__ __launch_bounds__(256) void k_small_multi(SmallMultiArgs M)
{
    __shared__ SLds<NP> S;
    const int blk = blockIdx.x;
    int k = 0;
    for (int u = 1; u < M.nbatch; ++u)
        if (blk >= M.part[u].first) k = u;
    const SmallPart &P = M.part[k];
    SmallArgs A = M.base;
    A.table = P.table;
    A.gmap = P.gmap;
    A.gamma = P.gamma;
    A.f = P.f;
    A.f2 = P.f2;
    A.wnsga2 = P.wnsga2;
    A.counters = P.counters;
    A.nv = P.nv;
    A.ns = P.ns;
    A.period = P.period;
    small_body<NP>(A, M.base.co, blk - P.first, S);
}

// hp_int_spa1d (surf_prkn.f90:485-524): the scalar N = 1 form used by int_method 6.
__global__ void k_spa1d(SmallArgs A, int npts)
{
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= npts) return;
    const int s = A.co.s;
    const double dz = A.dz;
    const double2 g = A.gamma[pt];
    const double2 gc = make_double2(g.x, -g.y);
    const double2 g2c = make_double2(g.x * g.x - g.y * g.y, -(2.0 * g.x * g.y));
    double2 f = A.f[pt];
    // p = i (1 - conj f) / (1 + conj f) * conj g
    const double2 num = make_double2(1.0 - f.x, f.y), den = make_double2(1.0 + f.x, -f.y);
    double2 r = cdiv(num.x, num.y, den.x, den.y);
    r = make_double2(-r.y, r.x);
    double2 p = make_double2(r.x * gc.x - r.y * gc.y, r.x * gc.y + r.y * gc.x);
    double2 q = make_double2(1.0, 0.0);
    for (int l = 1; l <= A.ns; ++l) {
        for (int k = 1; k <= s; ++k) {
            const double cb = (l != 1 && k == 1) ? 2 * A.co.bab[0] : A.co.bab[2 * k - 2];
            const double ca = A.co.bab[2 * k - 1];
            const VVi t = A.table[(size_t)(s * (l - 1) + k - 1) * A.nv];
            const double2 m = make_double2(-(t.v.x + t.vi.x) - g2c.x, (t.v.y + t.vi.y) - g2c.y);
            const double2 mq = make_double2(m.x * q.x - m.y * q.y, m.x * q.y + m.y * q.x);
            p = make_double2(p.x + cb * dz * mq.x, p.y + cb * dz * mq.y);
            q = make_double2(q.x + ca * dz * p.x, q.y + ca * dz * p.y);
        }
        p = cdiv(p.x, p.y, q.x, q.y);
        q = make_double2(1.0, 0.0);
    }
    const VVi t = A.table[(size_t)(s * A.ns) * A.nv];
    const double2 m = make_double2(-(t.v.x + t.vi.x) - g2c.x, (t.v.y + t.vi.y) - g2c.y);
    p = make_double2(p.x + A.co.bab[0] * dz * (m.x * q.x - m.y * q.y), p.y + A.co.bab[0] * dz * (m.x * q.y + m.y * q.x));
    // f = (conj g + i p) / (conj g - i p)
    const double2 a = make_double2(gc.x - p.y, gc.y + p.x), b = make_double2(gc.x + p.y, gc.y - p.x);
    f = cdiv(a.x, a.y, b.x, b.y);
    A.f[pt] = f;
    if (A.f2) {
        const double sr = g.x;
        A.f2[pt] = (sr > 1e-10) ? (f.x * f.x + f.y * f.y) * A.wnsga2[pt] / sr : 0.0;
    }
}

}  // namespace

int run_small(trhepd_ctx *ctx, hipStream_t stream, int period, const Table &t, int nb, int nang, const int *d_gmap,
              const double2 *d_gamma, double2 *d_f, double *d_f2, int nb0, const double *d_wnsga2,
              int *d_counters)
{
    SmallArgs A;
    A.table = t.d;
    A.nv = t.nv;
    A.ns = t.ns;
    A.N = nb;
    A.method = t.method;
    A.dz = t.dz;
    A.steps = t.d_steps;
    A.co = rkn_coef(t.method);
    A.gmap = d_gmap;
    A.gamma = d_gamma;
    A.f = d_f;
    A.f2 = d_f2;
    A.nb0 = nb0;
    A.wnsga2 = d_wnsga2;
    A.counters = d_counters;
    A.period = period;
    if (nb == 1 && t.method == TRHEPD_SP6) {
        if (t.d_steps) return fail(ctx, TRHEPD_E_ARG, "the one-beam sp6 form runs uniform slices only");
        hipLaunchKernelGGL(k_spa1d, dim3((nang + 63) / 64), dim3(64), 0, stream, A, nang);
    } else if (nb <= 16) {
        hipLaunchKernelGGL(k_small<16>, dim3(nang), dim3(256), 0, stream, A);
    } else if (nb <= 32) {
        hipLaunchKernelGGL(k_small<32>, dim3(nang), dim3(256), 0, stream, A);
    } else if (nb <= 48) {
        hipLaunchKernelGGL(k_small<48>, dim3(nang), dim3(256), 0, stream, A);
    } else {
        return fail(ctx, TRHEPD_E_ARG, "small path supports N <= 48 (got %d)", nb);
    }
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

}  // namespace trhepd

namespace trhepd {

bool small_fusable(int nb, int method) { return nb >= 1 && nb <= 48 && !(nb == 1 && method == TRHEPD_SP6); }

int run_small_multi(trhepd_ctx *ctx, hipStream_t stream, const std::vector<SmallBatch> &bs)
{
    if (bs.empty()) return TRHEPD_OK;
    const int nb = bs[0].nb;
    for (size_t i0 = 0; i0 < bs.size(); i0 += SMALL_MULTI) {
        const size_t i1 = std::min(bs.size(), i0 + SMALL_MULTI);
        SmallMultiArgs M{};
        const Table &t0 = *bs[i0].t;
        M.base.N = nb;
        M.base.method = t0.method;
        M.base.dz = t0.dz;
        M.base.steps = t0.d_steps;
        M.base.co = rkn_coef(t0.method);
        M.base.nb0 = bs[i0].nb0;
        M.nbatch = (int)(i1 - i0);
        int blocks = 0;
        for (size_t i = i0; i < i1; ++i) {
            const SmallBatch &b = bs[i];
            if (b.nb != nb || b.t->method != t0.method || b.t->dz != t0.dz || b.t->steps != t0.steps ||
                b.nb0 != bs[i0].nb0 || !small_fusable(b.nb, b.t->method))
                return fail(ctx, TRHEPD_E_ARG, "fused small-N launch: batches differ in N, method, dz or nb0");
            SmallPart &P = M.part[i - i0];
            P.table = b.t->d;
            P.gmap = b.gmap;
            P.gamma = b.gamma;
            P.f = b.f;
            P.f2 = b.f2;
            P.wnsga2 = b.w;
            P.counters = b.cnt;
            P.nv = b.t->nv;
            P.ns = b.t->ns;
            P.period = b.period;
            P.first = blocks;
            blocks += b.nang;
        }
        if (nb <= 16) hipLaunchKernelGGL(k_small_multi<16>, dim3(blocks), dim3(256), 0, stream, M);
        else if (nb <= 32) hipLaunchKernelGGL(k_small_multi<32>, dim3(blocks), dim3(256), 0, stream, M);
        else hipLaunchKernelGGL(k_small_multi<48>, dim3(blocks), dim3(256), 0, stream, M);
        TRY_HIP(ctx, hipGetLastError());
    }
    return TRHEPD_OK;
}

}  // namespace trhepd
