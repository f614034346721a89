// Batched slice loop for large beam counts (48 < N <= 600): all angles of a batch advance
// in lock-step, one kernel per stage.
//
// The slice matrix of stage (l, k) is M_a = M0 + diag(d_a): the off-diagonal part M0 is the
// same for every angle (it depends on the depth only), the diagonal d_a,j = -conj(v+vi)[1]
// - conj(gamma_a,j^2) carries the angle.  So the RKN kick P_a += c Q_a M_a is one tall
// complex GEMM [Q_1; ...; Q_A] * M0 on v_mfma_f64_16x16x4f64 (four real products per
// complex product) plus a column scaling, fused with the drift Q_a += c' P_a in the
// epilogue.  Q is double-buffered (a GEMM reads whole rows of Q while its own tile is
// updated); P is updated in place.
//
// Reference: hp_int_spa surf_prkn.f90:278-330, vvi2mate matcomp.f90:253-279, cond_est
// matcomp.f90:498-528, matAbdivB matcomp.f90:53-69 (zgetrf(N,2N) + ztrsm), invstransh /
// stransh matcomp.f90:140-196, intensity surflf.f90:143-151.
//
// Renormalisation decisions are taken on the device (k_cond); the host reads the flags
// once per slice, while the end-of-slice kick runs, and launches the batched LU only for
// the angles that need it.
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
// -conj(v + vi) of table entry t
__device__ __forceinline__ double2 w_upper(const VVi &t) { return make_double2(-(t.v.x + t.vi.x), t.v.y + t.vi.y); }
// -v + vi
__device__ __forceinline__ double2 w_lower(const VVi &t) { return make_double2(-t.v.x + t.vi.x, -t.v.y + t.vi.y); }

struct Geo {
    int A, N, NP;
    size_t mat;   // NP*NP
};

// ---------------------------------------------------------------------------
// per-angle constants: conj(gamma^2) and conj(gamma), zero padded to NP
__global__ void k_prep(Geo g, const double2 *__restrict__ gamma, double2 *__restrict__ g2c,
                       double2 *__restrict__ gc)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
    if (j >= g.NP) return;
    double2 x = make_double2(0, 0), y = make_double2(0, 0);
    if (j < g.N) {
        const double2 v = gamma[(size_t)a * g.N + j];
        x = make_double2(v.x * v.x - v.y * v.y, -(v.x * v.y + v.y * v.x));
        y = make_double2(v.x, -v.y);
    }
    g2c[(size_t)a * g.NP + j] = x;
    gc[(size_t)a * g.NP + j] = y;
}

// invstransh setup: X = I + F^H, Y = I - F^H (zero padding)
__global__ void k_inv_setup(Geo g, const double2 *__restrict__ f, double2 *__restrict__ Q, double2 *__restrict__ P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    double2 q = make_double2(0, 0), p = make_double2(0, 0);
    if (i < g.N && j < g.N) {
        const double2 fji = f[(size_t)a * g.N * g.N + j + (size_t)i * g.N];
        const double d = i == j ? 1.0 : 0.0;
        q = make_double2(d + fji.x, -fji.y);
        p = make_double2(d - fji.x, fji.y);
    }
    Q[e] = q;
    P[e] = p;
}

// P(i,j) <- (0,1) P(i,j) conj(g_i)
__global__ void k_inv_scale(Geo g, const double2 *__restrict__ gc, double2 *__restrict__ P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    if (i >= g.N || j >= g.N) return;
    const double2 y = P[e];
    P[e] = cmul(make_double2(-y.y, y.x), gc[(size_t)a * g.NP + i]);
}

// stransh setup: X = conj(G) - iY, Y' = conj(G) + iY
__global__ void k_st_setup(Geo g, const double2 *__restrict__ gc, double2 *__restrict__ Q, double2 *__restrict__ P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    if (i >= g.N || j >= g.N) {
        Q[e] = make_double2(0, 0);
        return;
    }
    const double2 y = P[e];
    const double2 d = (i == j) ? gc[(size_t)a * g.NP + i] : make_double2(0, 0);
    Q[e] = make_double2(d.x + y.y, d.y - y.x);
    P[e] = make_double2(d.x - y.y, d.y + y.x);
}

// F(i,j) = conj(Y''(j,i)); intensities of column nb0
__global__ void k_out(Geo g, const double2 *__restrict__ P, const double2 *__restrict__ gc, double2 *__restrict__ f,
                      double *__restrict__ f2, int nb0, const double *__restrict__ wnsga2)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nn = (size_t)g.N * g.N;
    if (e >= nn * g.A) return;
    const int a = (int)(e / nn);
    const int r = (int)(e % nn), i = r % g.N, j = r / g.N;
    const double2 y = P[(size_t)a * g.mat + j + (size_t)i * g.NP];
    f[e] = make_double2(y.x, -y.y);
    if (f2 && j == nb0 - 1) {
        const double sr = gc[(size_t)a * g.NP + i].x;
        f2[(size_t)a * g.N + i] = (sr > 1e-10) ? (y.x * y.x + y.y * y.y) * wnsga2[a] / sr : 0.0;
    }
}

// dense off-diagonal slice matrix M0 (row-major NP x NP), diagonal and padding zero
__global__ void k_gather(Geo g, const int *__restrict__ gmap, const VVi *__restrict__ col, double2 *__restrict__ M0)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int)g.mat) return;
    const int k = e / g.NP, j = e % g.NP;
    double2 m = make_double2(0, 0);
    if (k < g.N && j < g.N && k != j) {
        const int s = gmap[k * g.N + j];
        m = s > 0 ? w_upper(col[s - 1]) : w_lower(col[-s - 1]);
    }
    M0[e] = m;
}

// stage 1 of a slice: after a renormalisation P += a M, Q = I + b P; otherwise Q += b P
// (the kick was merged into the previous end-of-slice kick)
__global__ void k_stage1(Geo g, const int *__restrict__ gmap, const VVi *__restrict__ col,
                         const double2 *__restrict__ g2c, const int *__restrict__ xid, double alpha,
                         double beta, double2 *__restrict__ Q, double2 *__restrict__ P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    if (i >= g.N || j >= g.N) return;
    if (xid[a]) {
        double2 m;
        if (i == j) {
            const double2 w = w_upper(col[0]), c2 = g2c[(size_t)a * g.NP + j];
            m = make_double2(w.x - c2.x, w.y - c2.y);
        } else {
            const int s = gmap[i * g.N + j];
            m = s > 0 ? w_upper(col[s - 1]) : w_lower(col[-s - 1]);
        }
        double2 p = P[e];
        p = make_double2(p.x + alpha * m.x, p.y + alpha * m.y);
        P[e] = p;
        Q[e] = make_double2(beta * p.x + (i == j ? 1.0 : 0.0), beta * p.y);
    } else {
        const double2 p = P[e], q = Q[e];
        Q[e] = make_double2(q.x + beta * p.x, q.y + beta * p.y);
    }
}

// ---------------------------------------------------------------------------
// stage GEMM: P += alpha_a (Q M0 + Q diag(d_a)); mode 0/1 also Qn = Q + beta P
// (mode 1 adds the Gershgorin column sums of Qn), mode 2 leaves Q alone.
struct StageArgs {
    Geo g;
    const double2 *Q;
    double2 *Qn;
    double2 *P;
    const double2 *M0;
    const double2 *g2c;
    const VVi *col;
    double alpha, beta;
    const double *coef;   // per-angle alpha (mode 2) or null
    int mode;
    double *colsum;       // [A*NP/16][NP]
    double *diagabs;      // [A*NP]
    int ntiles, ncolt, tiles_per_xcd;
};

template <int BN>
__global__ __launch_bounds__(256) void k_stage_gemm(StageArgs S)
{
    constexpr int NB = BN / 16;
    const Geo g = S.g;
    // XCD-aware tile order: the 8 XCDs take contiguous tile ranges, so the column tiles
    // of one row tile (which share the A rows) run on the same XCD and its L2
    const int L = blockIdx.x;
    const int t = (L & 7) * S.tiles_per_xcd + (L >> 3);
    if (t >= S.ntiles) return;
    const int rt = t / S.ncolt, ct = t % S.ncolt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int rows = g.A * g.NP;
    const int r0 = rt * 128 + wave * 32;
    if (r0 >= rows) return;
    const int c0 = ct * BN;
    // A-fragment bases of the two 16-row blocks (each inside one angle)
    size_t abase[2];
    bool rvalid[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const int R = r0 + rb * 16 + l16;
        rvalid[rb] = (r0 + rb * 16) < rows;
        const int Rc = rvalid[rb] ? R : 0;
        abase[rb] = (size_t)(Rc / g.NP) * g.mat + (Rc % g.NP);
    }
    d4 cr[2][NB], ci[2][NB];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            cr[rb][cb] = d4{0, 0, 0, 0};
            ci[rb][cb] = d4{0, 0, 0, 0};
        }
    const int ksteps = (g.N + 3) >> 2;
    const double2 *__restrict__ Qp = S.Q;
    const double2 *__restrict__ Mp = S.M0;
    for (int s = 0; s < ksteps; ++s) {
        const int k = s * 4 + kq;
        double2 av[2], bv[NB];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = Qp[abase[rb] + (size_t)k * g.NP];
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const int j = c0 + cb * 16 + l16;
            bv[cb] = (j < g.NP) ? Mp[(size_t)k * g.NP + j] : make_double2(0, 0);
        }
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) {
                cr[rb][cb] = mfma4(av[rb].x, bv[cb].x, cr[rb][cb]);
                cr[rb][cb] = mfma4(-av[rb].y, bv[cb].y, cr[rb][cb]);
                ci[rb][cb] = mfma4(av[rb].x, bv[cb].y, ci[rb][cb]);
                ci[rb][cb] = mfma4(av[rb].y, bv[cb].x, ci[rb][cb]);
            }
    }
    const double2 w0 = w_upper(S.col[0]);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        if (!rvalid[rb]) continue;
        const int Rb = r0 + rb * 16;            // first row of the 16-row block
        const int a = Rb / g.NP, ib = Rb % g.NP;
        const double alpha = S.coef ? S.coef[a] : S.alpha;
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const int j = c0 + cb * 16 + l16;
            double csum = 0.0;
            if (j < g.N) {
                const double2 c2 = S.g2c[(size_t)a * g.NP + j];
                const double2 dj = make_double2(w0.x - c2.x, w0.y - c2.y);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = ib + kq + 4 * r;
                    if (i < g.N) {
                        const size_t idx = (size_t)a * g.mat + i + (size_t)j * g.NP;
                        const double2 q = S.Q[idx];
                        const double2 qd = cmul(q, dj);
                        double2 p = S.P[idx];
                        p.x += alpha * (cr[rb][cb][r] + qd.x);
                        p.y += alpha * (ci[rb][cb][r] + qd.y);
                        S.P[idx] = p;
                        if (S.mode != 2) {
                            const double2 qn = make_double2(q.x + S.beta * p.x, q.y + S.beta * p.y);
                            S.Qn[idx] = qn;
                            if (S.mode == 1) {
                                if (i == j) S.diagabs[(size_t)a * g.NP + j] = hypot(qn.x, qn.y);
                                else csum += sqrt(qn.x * qn.x + qn.y * qn.y);
                            }
                        }
                    }
                }
            }
            if (S.mode == 1) {
                csum += __shfl_xor(csum, 16);
                csum += __shfl_xor(csum, 32);
                if (kq == 0 && j < g.NP) S.colsum[(size_t)(Rb >> 4) * g.NP + j] = csum;
            }
        }
    }
}

// Gershgorin estimate per angle; renormalise when > 1e3 or on the last slice.
__global__ void k_cond(Geo g, const double *__restrict__ colsum, const double *__restrict__ diagabs, int last,
                       double cb1dz, double *__restrict__ coef, int *__restrict__ xid, int *__restrict__ ren,
                       int *__restrict__ counters)
{
    __shared__ double smn[256], smx[256];
    const int a = blockIdx.x, tid = threadIdx.x;
    const int nblk = g.NP / 16;
    double mn = 1e308, mx = -1e308;
    for (int j = tid; j < g.N; j += blockDim.x) {
        double t = 0.0;
        for (int b = 0; b < nblk; ++b) t += colsum[((size_t)a * nblk + b) * g.NP + j];
        const double dj = diagabs[(size_t)a * g.NP + j];
        mn = fmin(mn, dj - t);
        mx = fmax(mx, dj + t);
    }
    smn[tid] = mn;
    smx[tid] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) {
            smn[tid] = fmin(smn[tid], smn[tid + o]);
            smx[tid] = fmax(smx[tid], smx[tid + o]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double d00 = diagabs[(size_t)a * g.NP];
        const double m0 = fmin(d00, smn[0]), m1 = fmax(d00, smx[0]);
        const double d = (m0 < 0.0) ? 1e100 : m1 / m0;
        const int r = (d > 1e3) || last;
        ren[a] = r;
        xid[a] = r;
        coef[a] = r ? cb1dz : 2.0 * cb1dz;
        if (r) atomicAdd(&counters[0], 1);
    }
}

// ---------------------------------------------------------------------------
// batched LU of [X | Y] (N x 2N, X = Q block, Y = P block) with |re|+|im| partial
// pivoting, then Y <- U^{-1} Y.  Panel width 16; `list` names the angles to solve.
struct LuArgs {
    Geo g;
    double2 *Q, *P;
    int *piv;          // [A][NP]
    const int *list;
    int *flag;         // singular pivot flag
};

__device__ __forceinline__ double2 *colp(const LuArgs &L, int a, int c)
{
    return (c < L.g.N) ? L.Q + (size_t)a * L.g.mat + (size_t)c * L.g.NP
                       : L.P + (size_t)a * L.g.mat + (size_t)(c - L.g.N) * L.g.NP;
}

// factor the panel of columns [k0, k0+kw) in LDS
__global__ __launch_bounds__(256) void k_lu_panel(LuArgs L, int k0)
{
    extern __shared__ double2 pan[];   // [rows][16]
    __shared__ double rbest[4];
    __shared__ int ribest[4];
    const int a = L.list[blockIdx.x];
    const int N = L.g.N, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kw = min(16, N - k0), rows = N - k0;
    for (int e = tid; e < rows * 16; e += 256) {
        const int r = e >> 4, c = e & 15;
        pan[e] = (c < kw) ? colp(L, a, k0 + c)[k0 + r] : make_double2(0, 0);
    }
    __syncthreads();
    for (int c = 0; c < kw; ++c) {
        double best = -1.0;
        int bi = rows;
        for (int r = c + tid; r < rows; r += 256) {
            const double2 v = pan[r * 16 + c];
            const double m = fabs(v.x) + fabs(v.y);
            if (m > best) { best = m; bi = r; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) { rbest[wave] = best; ribest[wave] = bi; }
        __syncthreads();
        best = rbest[0];
        bi = ribest[0];
        for (int w = 1; w < 4; ++w)
            if (rbest[w] > best || (rbest[w] == best && ribest[w] < bi)) { best = rbest[w]; bi = ribest[w]; }
        if (tid == 0) {
            L.piv[(size_t)a * L.g.NP + k0 + c] = k0 + bi;
            if (best == 0.0) *L.flag = 1;
        }
        if (bi != c && tid < 16) {
            const double2 t0 = pan[c * 16 + tid];
            pan[c * 16 + tid] = pan[bi * 16 + tid];
            pan[bi * 16 + tid] = t0;
        }
        __syncthreads();
        const double2 pk = pan[c * 16 + c];
        if (pk.x != 0.0 || pk.y != 0.0) {
            const double2 rp = cdiv(make_double2(1, 0), pk);
            for (int r = c + 1 + tid; r < rows; r += 256) pan[r * 16 + c] = cmul(pan[r * 16 + c], rp);
        }
        __syncthreads();
        const int nc = kw - 1 - c;
        if (nc > 0)
            for (int e = tid; e < (rows - c - 1) * nc; e += 256) {
                const int r = c + 1 + e / nc, cc = c + 1 + e % nc;
                const double2 l = pan[r * 16 + c], u = pan[c * 16 + cc];
                double2 v = pan[r * 16 + cc];
                v.x -= l.x * u.x - l.y * u.y;
                v.y -= l.x * u.y + l.y * u.x;
                pan[r * 16 + cc] = v;
            }
        __syncthreads();
    }
    for (int e = tid; e < rows * 16; e += 256) {
        const int r = e >> 4, c = e & 15;
        if (c < kw) colp(L, a, k0 + c)[k0 + r] = pan[e];
    }
}

// for one 16-column tile right of the panel: row swaps, U12 = L11^{-1} A12, A22 -= L21 U12
__global__ __launch_bounds__(256) void k_lu_update(LuArgs L, int k0, int nqt)
{
    __shared__ double2 U[16][16];   // [k][j]
    const int a = L.list[blockIdx.y];
    const int N = L.g.N, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kw = min(16, N - k0);
    // tile columns: Q columns [k0+kw, N) in tiles of 16, then P columns [0, N)
    const int tix = blockIdx.x;
    int cstart, cw;
    if (tix < nqt) {
        cstart = k0 + kw + tix * 16;
        cw = min(16, N - cstart);
    } else {
        cstart = N + (tix - nqt) * 16;
        cw = min(16, 2 * N - cstart);
    }
    if (tid < 16) {
        for (int r = 0; r < 16; ++r) U[r][tid] = make_double2(0, 0);
        if (tid < cw) {
            double2 *col = colp(L, a, cstart + tid);
            for (int c = 0; c < kw; ++c) {
                const int p = L.piv[(size_t)a * L.g.NP + k0 + c];
                if (p != k0 + c) {
                    const double2 t0 = col[k0 + c];
                    col[k0 + c] = col[p];
                    col[p] = t0;
                }
            }
            double2 u[16];
            for (int r = 0; r < kw; ++r) {
                double2 x = col[k0 + r];
                for (int rr = 0; rr < r; ++rr) {
                    const double2 lv = colp(L, a, k0 + rr)[k0 + r];
                    x.x -= lv.x * u[rr].x - lv.y * u[rr].y;
                    x.y -= lv.x * u[rr].y + lv.y * u[rr].x;
                }
                u[r] = x;
                col[k0 + r] = x;
                U[r][tid] = x;
            }
        }
    }
    __syncthreads();
    const int l16 = lane & 15, kq = lane >> 4;
    const int first = k0 + kw;
    const int nrb = (N - first + 15) / 16;
    const int ks = (kw + 3) >> 2;
    for (int rb = wave; rb < nrb; rb += 4) {
        const int i0 = first + rb * 16;
        const int ia = i0 + l16;
        d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
        for (int s = 0; s < ks; ++s) {
            const int k = s * 4 + kq;
            double2 av = make_double2(0, 0);
            if (k < kw && ia < N) av = colp(L, a, k0 + k)[ia];
            const double2 bv = U[k][l16];
            cr = mfma4(av.x, bv.x, cr);
            cr = mfma4(-av.y, bv.y, cr);
            ci = mfma4(av.x, bv.y, ci);
            ci = mfma4(av.y, bv.x, ci);
        }
        if (l16 < cw) {
            double2 *col = colp(L, a, cstart + l16);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + kq + 4 * r;
                if (i < N) {
                    double2 v = col[i];
                    v.x -= cr[r];
                    v.y -= ci[r];
                    col[i] = v;
                }
            }
        }
    }
}

// Y <- U^{-1} Y for one 16-column tile of Y, all block rows (independent per tile)
__global__ __launch_bounds__(256) void k_lu_back(LuArgs L)
{
    __shared__ double2 X[16][16];
    const int a = L.list[blockIdx.y];
    const int N = L.g.N, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j0 = blockIdx.x * 16;
    const int cw = min(16, N - j0);
    const int nblk = (N + 15) / 16;
    const int l16 = lane & 15, kq = lane >> 4;
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int r0 = kb * 16, rw = min(16, N - r0);
        if (tid < 16) {
            for (int r = 0; r < 16; ++r) X[r][tid] = make_double2(0, 0);
            if (tid < cw) {
                double2 *y = colp(L, a, N + j0 + tid);
                double2 x[16];
                for (int r = rw - 1; r >= 0; --r) {
                    double2 v = y[r0 + r];
                    for (int rr = r + 1; rr < rw; ++rr) {
                        const double2 u = colp(L, a, r0 + rr)[r0 + r];
                        v.x -= u.x * x[rr].x - u.y * x[rr].y;
                        v.y -= u.x * x[rr].y + u.y * x[rr].x;
                    }
                    x[r] = cdiv(v, colp(L, a, r0 + r)[r0 + r]);
                    y[r0 + r] = x[r];
                    X[r][tid] = x[r];
                }
            }
        }
        __syncthreads();
        const int ks = (rw + 3) >> 2;
        for (int rb = wave; rb < kb; rb += 4) {
            const int i0 = rb * 16, ia = i0 + l16;
            d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
            for (int s = 0; s < ks; ++s) {
                const int k = s * 4 + kq;
                double2 av = make_double2(0, 0);
                if (k < rw) av = colp(L, a, r0 + k)[ia];
                const double2 bv = X[k][l16];
                cr = mfma4(av.x, bv.x, cr);
                cr = mfma4(-av.y, bv.y, cr);
                ci = mfma4(av.x, bv.y, ci);
                ci = mfma4(av.y, bv.x, ci);
            }
            if (l16 < cw) {
                double2 *y = colp(L, a, N + j0 + l16);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + kq + 4 * r;
                    double2 v = y[i];
                    v.x -= cr[r];
                    v.y -= ci[r];
                    y[i] = v;
                }
            }
        }
        __syncthreads();
    }
}

int lu_solve(trhepd_ctx *ctx, LuArgs &L, int npts)
{
    if (npts == 0) return TRHEPD_OK;
    const int N = L.g.N;
    for (int k0 = 0; k0 < N; k0 += 16) {
        const int kw = std::min(16, N - k0);
        const size_t lds = (size_t)(N - k0) * 16 * sizeof(double2);
        hipLaunchKernelGGL(k_lu_panel, dim3(npts), dim3(256), lds, ctx->stream, L, k0);
        const int nqt = (N - k0 - kw + 15) / 16, npt = (N + 15) / 16;
        hipLaunchKernelGGL(k_lu_update, dim3(nqt + npt, npts), dim3(256), 0, ctx->stream, L, k0, nqt);
    }
    hipLaunchKernelGGL(k_lu_back, dim3((N + 15) / 16, npts), dim3(256), 0, ctx->stream, L);
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

template <int BN>
void launch_stage(trhepd_ctx *ctx, StageArgs &S)
{
    const int rowt = (S.g.A * S.g.NP + 127) / 128;
    S.ncolt = (S.g.NP + BN - 1) / BN;
    S.ntiles = rowt * S.ncolt;
    S.tiles_per_xcd = (S.ntiles + 7) / 8;
    hipLaunchKernelGGL(k_stage_gemm<BN>, dim3(8 * S.tiles_per_xcd), dim3(256), 0, ctx->stream, S);
}

}  // namespace

int run_batched(trhepd_ctx *ctx, const Table &t, int nb, int nang, const int *d_gmap, const double2 *d_gamma,
                double2 *d_f, double *d_f2, int nb0, const double *d_wnsga2, int *d_counters)
{
    if (t.method != TRHEPD_SP4 && t.method != TRHEPD_SP6)
        return fail(ctx, TRHEPD_E_ARG, "large-N device path supports int_method 5 and 6 (got %d)", t.method);
    if (nb > 600) return fail(ctx, TRHEPD_E_ARG, "N = %d exceeds the device path limit of 600 beams", nb);
    Geo g;
    g.A = nang;
    g.N = nb;
    g.NP = (nb + 15) / 16 * 16;
    g.mat = (size_t)g.NP * g.NP;
    const size_t mb = g.mat * g.A * sizeof(double2);
    void *pQ0, *pQ1, *pP, *pM, *pg2, *pgc, *pcs, *pda, *pco, *pfl, *ppv, *plist, *pall;
    int rc;
    if ((rc = scratch(ctx, 10, mb, &pQ0)) || (rc = scratch(ctx, 11, mb, &pQ1)) || (rc = scratch(ctx, 12, mb, &pP)) ||
        (rc = scratch(ctx, 13, g.mat * sizeof(double2), &pM)) ||
        (rc = scratch(ctx, 14, (size_t)g.A * g.NP * sizeof(double2), &pg2)) ||
        (rc = scratch(ctx, 15, (size_t)g.A * g.NP * sizeof(double2), &pgc)) ||
        (rc = scratch(ctx, 16, (size_t)g.A * (g.NP / 16) * g.NP * 8, &pcs)) ||
        (rc = scratch(ctx, 17, (size_t)g.A * g.NP * 8, &pda)) || (rc = scratch(ctx, 18, (size_t)g.A * 8, &pco)) ||
        (rc = scratch(ctx, 19, (size_t)g.A * 8, &pfl)) || (rc = scratch(ctx, 20, (size_t)g.A * g.NP * 4, &ppv)) ||
        (rc = scratch(ctx, 21, (size_t)g.A * 4, &plist)) || (rc = scratch(ctx, 22, (size_t)g.A * 4, &pall)))
        return rc;
    int *d_xid = (int *)pfl, *d_ren = (int *)pfl + g.A;
    std::vector<int> all(g.A);
    for (int i = 0; i < g.A; ++i) all[i] = i;
    TRY_HIP(ctx, hipMemcpyAsync(pall, all.data(), g.A * 4, hipMemcpyHostToDevice, ctx->stream));
    int *h_ren = nullptr;
    TRY_HIP(ctx, hipHostMalloc((void **)&h_ren, (size_t)g.A * 8 + 64, hipHostMallocDefault));
    struct Pin {
        int *p;
        ~Pin() { (void)hipHostFree(p); }
    } pin{h_ren};
    hipEvent_t evf;
    TRY_HIP(ctx, hipEventCreateWithFlags(&evf, hipEventDisableTiming));
    struct Ev {
        hipEvent_t e;
        ~Ev() { (void)hipEventDestroy(e); }
    } evguard{evf};

    const size_t ntot = g.mat * g.A;
    const dim3 eb(256), eg((unsigned)((ntot + 255) / 256));
    hipLaunchKernelGGL(k_prep, dim3((g.NP + 255) / 256, g.A), dim3(256), 0, ctx->stream, g, d_gamma,
                       (double2 *)pg2, (double2 *)pgc);
    double2 *Q = (double2 *)pQ0, *Qn = (double2 *)pQ1, *P = (double2 *)pP;
    hipLaunchKernelGGL(k_inv_setup, eg, eb, 0, ctx->stream, g, (const double2 *)d_f, Q, P);
    TRY_HIP(ctx, hipMemsetAsync(pM, 0, g.mat * sizeof(double2), ctx->stream));
    LuArgs L{g, Q, P, (int *)ppv, (const int *)pall, d_counters + 1};
    // k_lu_panel needs up to N*16*16 bytes of dynamic LDS
    TRY_HIP(ctx, hipFuncSetAttribute((const void *)k_lu_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(g.N * 16 * sizeof(double2))));
    if ((rc = lu_solve(ctx, L, g.A))) return rc;
    hipLaunchKernelGGL(k_inv_scale, eg, eb, 0, ctx->stream, g, (const double2 *)pgc, P);
    std::vector<int> ones(g.A, 1);
    TRY_HIP(ctx, hipMemcpyAsync(d_xid, ones.data(), g.A * 4, hipMemcpyHostToDevice, ctx->stream));

    const RknCoef co = rkn_coef(t.method);
    const int s = co.s;
    StageArgs S;
    S.g = g;
    S.M0 = (const double2 *)pM;
    S.g2c = (const double2 *)pg2;
    S.colsum = (double *)pcs;
    S.diagabs = (double *)pda;
    const int gthreads = 256, ggrid = (int)((g.mat + gthreads - 1) / gthreads);
    auto stage = [&](int mode, const VVi *col, double alpha, double beta, const double *coef) {
        hipLaunchKernelGGL(k_gather, dim3(ggrid), dim3(gthreads), 0, ctx->stream, g, d_gmap, col, (double2 *)pM);
        S.Q = Q;
        S.Qn = Qn;
        S.P = P;
        S.col = col;
        S.alpha = alpha;
        S.beta = beta;
        S.coef = coef;
        S.mode = mode;
        prof_mark(ctx);
        launch_stage<48>(ctx, S);
        prof_mark(ctx);
        if (mode != 2) std::swap(Q, Qn);
    };
    for (int l = 1; l <= t.ns; ++l) {
        const VVi *col1 = t.d + (size_t)(s * (l - 1)) * t.nv;
        hipLaunchKernelGGL(k_stage1, eg, eb, 0, ctx->stream, g, d_gmap, col1, (const double2 *)pg2,
                           (const int *)d_xid, co.bab[0] * t.dz, co.bab[1] * t.dz, Q, P);
        for (int k = 2; k <= s; ++k) {
            const VVi *col = t.d + (size_t)(s * (l - 1) + k - 1) * t.nv;
            stage(k == s ? 1 : 0, col, co.bab[2 * k - 2] * t.dz, co.bab[2 * k - 1] * t.dz, nullptr);
        }
        hipLaunchKernelGGL(k_cond, dim3(g.A), dim3(256), 0, ctx->stream, g, (const double *)pcs, (const double *)pda,
                           l == t.ns ? 1 : 0, co.bab[0] * t.dz, (double *)pco, d_xid, d_ren, d_counters);
        TRY_HIP(ctx, hipMemcpyAsync(h_ren, d_ren, g.A * 4, hipMemcpyDeviceToHost, ctx->stream));
        TRY_HIP(ctx, hipEventRecord(evf, ctx->stream));
        stage(2, t.d + (size_t)(s * l) * t.nv, 0.0, 0.0, (const double *)pco);
        TRY_HIP(ctx, hipGetLastError());
        TRY_HIP(ctx, hipEventSynchronize(evf));
        // compact list of the angles to renormalise, staged in pinned memory: the copy is
        // queued behind this slice's kernels and the next slice rewrites the staging area
        // only after waiting on its own event, which follows this copy in stream order
        int *lst = h_ren + g.A, nl = 0;
        for (int i = 0; i < g.A; ++i)
            if (h_ren[i]) lst[nl++] = i;
        if (nl) {
            TRY_HIP(ctx, hipMemcpyAsync(plist, lst, (size_t)nl * 4, hipMemcpyHostToDevice, ctx->stream));
            LuArgs Lr{g, Q, P, (int *)ppv, (const int *)plist, d_counters + 1};
            if ((rc = lu_solve(ctx, Lr, nl))) return rc;
        }
    }
    hipLaunchKernelGGL(k_st_setup, eg, eb, 0, ctx->stream, g, (const double2 *)pgc, Q, P);
    LuArgs Lf{g, Q, P, (int *)ppv, (const int *)pall, d_counters + 1};
    if ((rc = lu_solve(ctx, Lf, g.A))) return rc;
    const size_t nout = (size_t)g.N * g.N * g.A;
    hipLaunchKernelGGL(k_out, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, ctx->stream, g, (const double2 *)P,
                       (const double2 *)pgc, d_f, d_f2, nb0, d_wnsga2);
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

}  // namespace trhepd
