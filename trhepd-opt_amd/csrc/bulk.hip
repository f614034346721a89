// Bulk reflection of the layered crystal (blkref, bulk.f90:7-124): for every glancing angle
// and every group of coupled beams, the transfer matrix of one unit layer (the product of the
// slice transfer matrices of trmat, trmatg.f90:6-59, and the column phases of the layer
// shift) and the layer recursion R_l = (T11 R + T12)(T21 R + T22)^-1 (gcmi3, gcmi.f90:5-15)
// until |R_ii|^2 settles.
//
// trmat forms the slice matrix from the eigen-decomposition H U = U K^2 of the slice
// Hamiltonian H = v + vi + diag(gamma^2):  T = B D B^-1 with B = [[tau, rho], [rho, tau]],
// tau = (Gamma U + U K), rho = (Gamma U - U K), D = diag(e^{i K dz}, e^{-i K dz}).  Since
// B = Cm V with Cm = [[Gamma, I], [Gamma, -I]] and V = [[U, U], [U K, -U K]], and
// V D V^-1 = exp(dz A) for A = i [[0, I], [H, 0]], the slice matrix is
//     T = Cm E Cm^-1,   E = exp(dz A) = [[c(H), i s(H)], [i H s(H), c(H)]],
//     c(H) = sum_k (-dz^2 H)^k / (2k)!,   s(H) = dz sum_k (-dz^2 H)^k / (2k+1)!,
// which needs no eigen-solver.  All slices of a group share Gamma, so the unit-layer matrix
// is Cm (E_ns ... E_1) Cm^-1: the conjugations collapse to two diagonal scalings.  One
// workgroup handles one (angle, group) item with its matrices in global memory (L2
// resident); the products run on v_mfma_f64_16x16x4f64.
#include <algorithm>

#include "internal.hpp"

namespace trhepd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    const double r = b.x / b.y, den = b.y + b.x * r;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

struct BulkItem {
    int ang, grp, n, boff;   // angle, group, beams, first beam of the group (sorted order)
    size_t ws;               // workspace offset (double2)
    size_t out;              // output offset (double2): the n x n block, column-major
};

struct BulkArgs {
    const BulkItem *items;
    const VVi *table;        // [ns][nv]
    int nv, ns;
    double dz;
    const int *ivg;          // blkghk's iv(nbgm, nb): column = beam k, row = l - first beam of group
    int nbgm, nb;
    const double2 *gamma;    // [nang][nb] (sorted beam order)
    const double2 *phase;    // [nb] column phases of the layer shift, or null
    int ml;
    double epsb;
    double2 *ws;
    double2 *out;
    int *layers;             // [nitems] layers used
};

// C = A B (+ C when acc) for NP x NP complex matrices (column-major, leading dimension ld,
// NP a multiple of 16), whole workgroup; no barrier inside
__device__ void wg_gemm(int NPr, int NPc, int NPk, const double2 *A, int lda, const double2 *B, int ldb, double2 *C,
                        int ldc, bool acc)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int nbr = NPr >> 4, nbc = NPc >> 4, ks = NPk >> 2;
    for (int blk = wave; blk < nbr * nbc; blk += nw) {
        const int bi = blk % nbr, bj = blk / nbr;
        d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
        const double2 *ap = A + bi * 16 + l16;
        const double2 *bp = B + (size_t)(bj * 16 + l16) * ldb;
        for (int s = 0; s < ks; ++s) {
            const int k = s * 4 + kq;
            const double2 a = ap[(size_t)k * lda], b = bp[k];
            cr = mfma4(a.x, b.x, cr);
            ci = mfma4(a.x, b.y, ci);
            cr = mfma4(-a.y, b.y, cr);
            ci = mfma4(a.y, b.x, ci);
        }
        double2 *cp = C + (size_t)(bj * 16 + l16) * ldc + bi * 16 + kq;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double2 v = make_double2(cr[r], ci[r]);
            if (acc) {
                const double2 o = cp[4 * r];
                v.x += o.x;
                v.y += o.y;
            }
            cp[4 * r] = v;
        }
    }
}

// Solves M Y = B in place (B <- Y) for n x n M with |re|+|im| partial pivoting; M is
// destroyed.  Whole workgroup; returns with a barrier done.
__device__ void wg_solve(int n, double2 *M, double2 *B, int ld, double *sbest, int *sidx)
{
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
    for (int c = 0; c < n; ++c) {
        double best = -1.0;
        int bi = n;
        for (int r = c + tid; r < n; r += nt) {
            const double2 v = M[(size_t)c * ld + r];
            const double m = fabs(v.x) + fabs(v.y);
            if (m > best) {
                best = m;
                bi = r;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0) {
            sbest[wave] = best;
            sidx[wave] = bi;
        }
        __syncthreads();
        best = sbest[0];
        bi = sidx[0];
        for (int w = 1; w < (nt >> 6); ++w)
            if (sbest[w] > best || (sbest[w] == best && sidx[w] < bi)) {
                best = sbest[w];
                bi = sidx[w];
            }
        __syncthreads();
        if (bi != c) {   // swap rows c and bi of M (columns c..n-1) and B
            for (int j = c + tid; j < n; j += nt) {
                const double2 t = M[(size_t)j * ld + c];
                M[(size_t)j * ld + c] = M[(size_t)j * ld + bi];
                M[(size_t)j * ld + bi] = t;
            }
            for (int j = tid; j < n; j += nt) {
                const double2 t = B[(size_t)j * ld + c];
                B[(size_t)j * ld + c] = B[(size_t)j * ld + bi];
                B[(size_t)j * ld + bi] = t;
            }
        }
        __syncthreads();
        const double2 p = M[(size_t)c * ld + c];
        const double2 rp = (p.x != 0.0 || p.y != 0.0) ? cdiv(make_double2(1, 0), p) : make_double2(0, 0);
        // eliminate below the pivot: rows r > c, columns c+1..n-1 of M and all of B
        const int rows = n - c - 1, cols = n - c - 1;
        for (int e = tid; e < rows * (cols + n); e += nt) {
            const int r = c + 1 + e % rows, q = e / rows;
            const double2 l = cmul(M[(size_t)c * ld + r], rp);
            if (q < cols) {
                const int j = c + 1 + q;
                const double2 u = M[(size_t)j * ld + c];
                double2 v = M[(size_t)j * ld + r];
                const double2 lu = cmul(l, u);
                v.x -= lu.x;
                v.y -= lu.y;
                M[(size_t)j * ld + r] = v;
            } else {
                const int j = q - cols;
                const double2 u = B[(size_t)j * ld + c];
                double2 v = B[(size_t)j * ld + r];
                const double2 lu = cmul(l, u);
                v.x -= lu.x;
                v.y -= lu.y;
                B[(size_t)j * ld + r] = v;
            }
        }
        __syncthreads();
    }
    // back substitution, one right-hand-side column per thread
    for (int j = tid; j < n; j += nt) {
        double2 *y = B + (size_t)j * ld;
        for (int r = n - 1; r >= 0; --r) {
            double2 v = y[r];
            for (int k = r + 1; k < n; ++k) {
                const double2 t = cmul(M[(size_t)k * ld + r], y[k]);
                v.x -= t.x;
                v.y -= t.y;
            }
            y[r] = cdiv(v, M[(size_t)r * ld + r]);
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_bulk(BulkArgs A)
{
    __shared__ double sbest[4];
    __shared__ int sidx[4];
    __shared__ double snorm;
    const BulkItem it = A.items[blockIdx.x];
    const int n = it.n, tid = threadIdx.x, nt = blockDim.x;
    const int NP = (n + 15) & ~15, N2 = 2 * NP;
    double2 *w = A.ws + it.ws;
    double2 *H = w, *X = H + NP * NP, *Cc = X + NP * NP, *Ss = Cc + NP * NP, *Tm = Ss + NP * NP,
            *HS = Tm + NP * NP;
    double2 *E = HS + NP * NP, *P = E + N2 * N2, *Pn = P + N2 * N2;
    const double2 *g = A.gamma + (size_t)it.ang * A.nb + it.boff;
    const double dz = A.dz;
    // P = I
    for (int e = tid; e < N2 * N2; e += nt) P[e] = make_double2((e % N2) == (e / N2) ? 1.0 : 0.0, 0.0);
    for (int l = 0; l < A.ns; ++l) {
        const VVi *col = A.table + (size_t)l * A.nv;
        // H of slice l (trmatg.f90:20-29) and its row-sum norm
        for (int e = tid; e < NP * NP; e += nt) {
            const int i = e % NP, j = e / NP;
            double2 h = make_double2(0, 0);
            if (i < n && j < n) {
                if (i == j) {
                    const double2 gg = cmul(g[i], g[i]);
                    h = make_double2(col[0].v.x + col[0].vi.x + gg.x, col[0].v.y + col[0].vi.y + gg.y);
                } else if (i > j) {   // H(i,j), i > j: v + vi at iv(i, j)
                    const VVi t = col[A.ivg[(size_t)(it.boff + j) * A.nbgm + i] - 1];
                    h = make_double2(t.v.x + t.vi.x, t.v.y + t.vi.y);
                } else {              // H(i,j), i < j: conj(v - vi) at iv(j, i)
                    const VVi t = col[A.ivg[(size_t)(it.boff + i) * A.nbgm + j] - 1];
                    h = make_double2(t.v.x - t.vi.x, -(t.v.y - t.vi.y));
                }
            }
            H[e] = h;
        }
        if (tid == 0) snorm = 0.0;
        __syncthreads();
        if (tid < NP) {
            double s = 0.0;
            for (int j = 0; j < NP; ++j) s += fabs(H[(size_t)j * NP + tid].x) + fabs(H[(size_t)j * NP + tid].y);
            atomicMax((unsigned long long *)&snorm, __double_as_longlong(s));
        }
        __syncthreads();
        // scaling: h = dz / 2^sq with |h^2 H| <= 1, then 2^sq squarings of E
        const double xn = snorm * dz * dz;
        int sq = 0;
        while (xn / (double)(1ull << (2 * sq)) > 1.0 && sq < 30) ++sq;
        const double h = dz / (double)(1ull << sq), h2 = h * h;
        // series degree: x^K / (2K)! below 1e-18 with x = h^2 |H| <= 1
        const double x = xn / (double)(1ull << (2 * sq));
        int K = 1;
        {
            double term = x / 2.0;
            while (term > 1e-18 && K < 20) {
                ++K;
                term *= x / ((2.0 * K - 1) * (2.0 * K));
            }
        }
        for (int e = tid; e < NP * NP; e += nt) {
            const double2 v = H[e];
            X[e] = make_double2(h2 * v.x, h2 * v.y);
            const bool d = (e % NP) == (e / NP);
            Cc[e] = make_double2(d ? 1.0 : 0.0, 0.0);
            Ss[e] = make_double2(d ? 1.0 : 0.0, 0.0);
        }
        __syncthreads();
        // Horner: c = I - X/((2k-1)2k) (I - ...), s = I - X/(2k(2k+1)) (I - ...)
        for (int k = K; k >= 1; --k) {
            wg_gemm(NP, NP, NP, X, NP, Cc, NP, Tm, NP, false);
            __syncthreads();
            const double fc = 1.0 / ((2.0 * k - 1) * (2.0 * k));
            for (int e = tid; e < NP * NP; e += nt) {
                const double2 t = Tm[e];
                const bool d = (e % NP) == (e / NP);
                Cc[e] = make_double2((d ? 1.0 : 0.0) - fc * t.x, -fc * t.y);
            }
            __syncthreads();
            wg_gemm(NP, NP, NP, X, NP, Ss, NP, Tm, NP, false);
            __syncthreads();
            const double fs = 1.0 / ((2.0 * k) * (2.0 * k + 1));
            for (int e = tid; e < NP * NP; e += nt) {
                const double2 t = Tm[e];
                const bool d = (e % NP) == (e / NP);
                Ss[e] = make_double2((d ? 1.0 : 0.0) - fs * t.x, -fs * t.y);
            }
            __syncthreads();
        }
        for (int e = tid; e < NP * NP; e += nt) Ss[e] = make_double2(h * Ss[e].x, h * Ss[e].y);
        __syncthreads();
        wg_gemm(NP, NP, NP, H, NP, Ss, NP, HS, NP, false);
        __syncthreads();
        // E = [[c, i s], [i H s, c]]
        for (int e = tid; e < N2 * N2; e += nt) {
            const int i = e % N2, j = e / N2;
            const int bi = i >= NP, bj = j >= NP, ii = i - bi * NP, jj = j - bj * NP;
            const size_t q = (size_t)jj * NP + ii;
            double2 v;
            if (bi == bj) v = Cc[q];
            else if (!bi) v = make_double2(-Ss[q].y, Ss[q].x);
            else v = make_double2(-HS[q].y, HS[q].x);
            E[e] = v;
        }
        __syncthreads();
        for (int r = 0; r < sq; ++r) {   // E <- E^2
            wg_gemm(N2, N2, N2, E, N2, E, N2, Pn, N2, false);
            __syncthreads();
            for (int e = tid; e < N2 * N2; e += nt) E[e] = Pn[e];
            __syncthreads();
        }
        // P <- E P
        wg_gemm(N2, N2, N2, E, N2, P, N2, Pn, N2, false);
        __syncthreads();
        double2 *t = P;
        P = Pn;
        Pn = t;
    }
    // unit-layer matrix T = Cm P Cm^-1 (elementwise in the diagonal Gamma), then the column
    // phases of the layer shift (bulk.f90:83-93); T goes to Pn
    for (int e = tid; e < N2 * N2; e += nt) {
        const int i = e % N2, j = e / N2;
        const int bi = i >= NP, bj = j >= NP, ii = i - bi * NP, jj = j - bj * NP;
        double2 v = make_double2(0, 0);
        if (ii < n && jj < n) {
            // (Cm P)(ii, .) rows: top = Gamma P1. + P2., bottom = Gamma P1. - P2.
            auto cp = [&](int col) {
                const double2 p1 = P[(size_t)col * N2 + ii], p2 = P[(size_t)col * N2 + NP + ii];
                const double2 gp = cmul(g[ii], p1);
                return bi ? make_double2(gp.x - p2.x, gp.y - p2.y) : make_double2(gp.x + p2.x, gp.y + p2.y);
            };
            // Cm^-1 = 1/2 [[Gamma^-1, Gamma^-1], [I, -I]]
            const double2 a = cdiv(cp(jj), g[jj]), b = cp(NP + jj);
            v = bj ? make_double2(0.5 * (a.x - b.x), 0.5 * (a.y - b.y)) : make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
            if (A.phase) v = cmul(v, A.phase[it.boff + jj]);
        }
        Pn[e] = v;
    }
    __syncthreads();
    // layer recursion (bulk.f90:95-119): R = (T11 R + T12)(T21 R + T22)^-1
    double2 *T = Pn;
    double2 *R = H, *t1 = X, *t2 = Cc, *Mt = Ss;
    double2 *r20 = E;   // |R_ii|^2 of the previous layer (real part)
    int used = A.ml;
    for (int l = 1; l <= A.ml; ++l) {
        for (int e = tid; e < NP * NP; e += nt) {
            const int i = e % NP, j = e / NP;
            t1[e] = T[(size_t)(NP + j) * N2 + i];
            t2[e] = T[(size_t)(NP + j) * N2 + NP + i];
        }
        __syncthreads();
        if (l > 1) {
            wg_gemm(NP, NP, NP, T, N2, R, NP, t1, NP, true);
            wg_gemm(NP, NP, NP, T + NP, N2, R, NP, t2, NP, true);
            __syncthreads();
        }
        // R = t1 t2^-1: t2^T R^T = t1^T
        for (int e = tid; e < NP * NP; e += nt) {
            const int i = e % NP, j = e / NP;
            Mt[e] = t2[(size_t)i * NP + j];
            Tm[e] = t1[(size_t)i * NP + j];
        }
        __syncthreads();
        wg_solve(n, Mt, Tm, NP, sbest, sidx);
        for (int e = tid; e < NP * NP; e += nt) {
            const int i = e % NP, j = e / NP;
            R[e] = (i < n && j < n) ? Tm[(size_t)i * NP + j] : make_double2(0, 0);
        }
        __syncthreads();
        // judgement on |R_ii|^2 (bulk.f90:104-117)
        if (tid == 0) snorm = 0.0;
        __syncthreads();
        if (l >= 20 && tid < n) {
            const double2 d = R[(size_t)tid * NP + tid];
            const double r2 = d.x * d.x + d.y * d.y;
            if (l > 20) atomicMax((unsigned long long *)&snorm, __double_as_longlong(fabs(r2 - r20[tid].x)));
            r20[tid].x = r2;
        }
        __syncthreads();
        if (l > 20 && snorm < A.epsb) {
            used = l;
            break;
        }
    }
    double2 *o = A.out + it.out;
    for (int e = tid; e < n * n; e += nt) {
        const int i = e % n, j = e / n;
        o[e] = R[(size_t)j * NP + i];
    }
    if (tid == 0) A.layers[blockIdx.x] = used;
}

}  // namespace
}  // namespace trhepd

using namespace trhepd;

extern "C" int trhepd_set_bulk_batched_above(trhepd_ctx *ctx, int n)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (n < 0 || n > 64) return fail(ctx, TRHEPD_E_ARG, "bulk group limit of the one-workgroup kernel is 0..64 (got %d)", n);
    ctx->bulk_one_wg_max = n;
    return TRHEPD_OK;
}

extern "C" int trhepd_bulk_reflect(trhepd_ctx *ctx, int table_id, int ngr, const int *nbg, int nbgm, const int *iv,
                                   int nang, const double *gamma, const double *phase, int ml, double epsb,
                                   double *blocks, int *layers)
{
    if (!ctx) return TRHEPD_E_ARG;
    const Table *t = find_table(ctx, table_id);
    if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "no table %d", table_id);
    if (t->method != TRHEPD_M2 && t->method != TRHEPD_LF)
        return fail(ctx, TRHEPD_E_ARG, "bulk_reflect needs a table of slice midpoints (int_method 1 or 4)");
    if (ngr < 1 || !nbg || !iv || nang < 1 || !gamma || !blocks || ml < 1 || nbgm < 1)
        return fail(ctx, TRHEPD_E_ARG, "bulk_reflect: bad arguments");
    int nb = 0;
    for (int g = 0; g < ngr; ++g) {
        if (nbg[g] < 1 || nbg[g] > nbgm || nbg[g] > 600)
            return fail(ctx, TRHEPD_E_ARG, "group %d has %d beams (1..%d, at most 600)", g + 1, nbg[g], nbgm);
        nb += nbg[g];
    }
    // iv entries used: strict lower triangle of each group, 1..nv
    {
        int b0 = 0;
        for (int g = 0; g < ngr; ++g) {
            for (int k = 0; k < nbg[g]; ++k)
                for (int l = k + 1; l < nbg[g]; ++l) {
                    const int m = iv[(size_t)(b0 + k) * nbgm + l];
                    if (m < 1 || m > t->nv) return fail(ctx, TRHEPD_E_ARG, "iv(%d,%d) = %d outside 1..%d", l + 1, b0 + k + 1, m, t->nv);
                }
            b0 += nbg[g];
        }
    }
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    // groups of up to 64 beams: one workgroup per (angle, group) item; larger groups: the
    // batched path over angles (bulk_large_group), one group after another
    const int ONE_WG_MAX = ctx->bulk_one_wg_max;
    std::vector<BulkItem> items;
    std::vector<int> big;           // groups of the batched path
    std::vector<size_t> goff(ngr);  // offset of a group's block within one angle's output
    size_t per_angle = 0;
    for (int g = 0; g < ngr; ++g) {
        goff[g] = per_angle;
        per_angle += (size_t)nbg[g] * nbg[g];
        if (nbg[g] > ONE_WG_MAX) big.push_back(g);
    }
    size_t ws = 0, out = per_angle * nang;
    for (int a = 0; a < nang; ++a) {
        int b0 = 0;
        for (int g = 0; g < ngr; ++g) {
            const int n = nbg[g], NP = (n + 15) & ~15;
            if (n <= ONE_WG_MAX) {
                items.push_back(BulkItem{a, g, n, b0, ws, (size_t)a * per_angle + goff[g]});
                ws += (size_t)6 * NP * NP + (size_t)3 * 4 * NP * NP;
            }
            b0 += n;
        }
    }
    // largest groups first
    std::vector<BulkItem> order = items;
    std::stable_sort(order.begin(), order.end(), [](const BulkItem &x, const BulkItem &y) { return x.n > y.n; });
    const size_t n_it = order.size();
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_it = take(n_it * sizeof(BulkItem)), o_iv = take((size_t)nbgm * nb * 4),
                 o_g = take((size_t)nang * nb * 16), o_ph = take((size_t)nb * 16), o_ws = take(ws * 16),
                 o_out = take(out * 16), o_ly = take(n_it * 4);
    char *d = nullptr;
    int rc = dev_alloc(ctx, (void **)&d, off);
    if (rc) return rc;
    BulkArgs A;
    A.items = (const BulkItem *)(d + o_it);
    A.table = t->d;
    A.nv = t->nv;
    A.ns = t->ns;
    A.dz = t->dz;
    A.ivg = (const int *)(d + o_iv);
    A.nbgm = nbgm;
    A.nb = nb;
    A.gamma = (const double2 *)(d + o_g);
    A.phase = phase ? (const double2 *)(d + o_ph) : nullptr;
    A.ml = ml;
    A.epsb = epsb;
    A.ws = (double2 *)(d + o_ws);
    A.out = (double2 *)(d + o_out);
    A.layers = (int *)(d + o_ly);
    std::vector<int> ly(n_it);
    hipError_t e = hipMemcpyAsync(d + o_it, order.data(), n_it * sizeof(BulkItem), hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o_iv, iv, (size_t)nbgm * nb * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o_g, gamma, (size_t)nang * nb * 16, hipMemcpyHostToDevice, ctx->stream);
    if (!e && phase) e = hipMemcpyAsync(d + o_ph, phase, (size_t)nb * 16, hipMemcpyHostToDevice, ctx->stream);
    if (!e && n_it) {
        hipLaunchKernelGGL(k_bulk, dim3((unsigned)n_it), dim3(256), 0, ctx->stream, A);
        e = hipGetLastError();
    }
    std::vector<int> lyb((size_t)nang * big.size());
    for (size_t k = 0; k < big.size() && !e && !rc; ++k) {
        const int g = big[k];
        int b0 = 0;
        double gmax2 = 0.0;
        for (int q = 0; q < g; ++q) b0 += nbg[q];
        for (int a = 0; a < nang; ++a)
            for (int i = 0; i < nbg[g]; ++i) {
                const double gr = gamma[2 * ((size_t)a * nb + b0 + i)], gi = gamma[2 * ((size_t)a * nb + b0 + i) + 1];
                gmax2 = std::max(gmax2, gr * gr + gi * gi);
            }
        rc = bulk_large_group(ctx, ctx->stream, *t, nbg[g], b0, nang, A.ivg, nbgm, A.gamma, nb, A.phase, gmax2, ml,
                              epsb, A.out + goff[g], per_angle, lyb.data() + k * nang);
    }
    if (!e && !rc) e = hipMemcpyAsync(blocks, d + o_out, out * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (!e && !rc && n_it) e = hipMemcpyAsync(ly.data(), d + o_ly, n_it * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (!e) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, d, off);
    if (rc) return rc;
    if (e) return hip_check(ctx, e, "bulk_reflect");
    if (layers) {
        for (size_t i = 0; i < n_it; ++i) {
            // items were reordered: report in (angle, group) order
            const BulkItem &it = order[i];
            layers[(size_t)it.ang * ngr + it.grp] = ly[i];
        }
        for (size_t k = 0; k < big.size(); ++k)
            for (int a = 0; a < nang; ++a) layers[(size_t)a * ngr + big[k]] = lyb[k * nang + a];
    }
    for (size_t q = 0; q < 2 * out; ++q)
        if (!std::isfinite(blocks[q])) return fail(ctx, TRHEPD_E_NONFINITE, "non-finite bulk reflection");
    return TRHEPD_OK;
}
