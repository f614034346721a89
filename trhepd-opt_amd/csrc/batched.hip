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
#include <cmath>
#include <initializer_list>
#include <type_traits>
#include <utility>
#include <vector>

#include <cstdio>
#include <cstdlib>

#include "device.hpp"
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
// fills (kernels rather than memsets, so that a run can be captured into a HIP graph)
__global__ void k_fill_i32(int *__restrict__ p, int n, int v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = v;
}
__global__ void k_zero2(double2 *__restrict__ p, size_t n)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = make_double2(0, 0);
}

// per-angle constants: conj(gamma^2) and conj(gamma), zero padded to NP
// Q = I + beta P on the listed angles (the identity after a renormalisation; the leapfrog
// start Q = I + dz/2 P of surf_magnus.f90:63-67); padding stays zero
__global__ void k_qreset(Geo g, const int *__restrict__ list, double beta, double2 *__restrict__ Q,
                         const double2 *__restrict__ P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat) return;
    const size_t idx = (size_t)list[blockIdx.y] * g.mat + e;
    const int i = (int)(e % g.NP), j = (int)(e / g.NP);
    double2 q = make_double2(0, 0);
    if (i < g.N && j < g.N) {
        const double2 p = P[idx];
        q = make_double2(beta * p.x + (i == j ? 1.0 : 0.0), beta * p.y);
    }
    Q[idx] = q;
}

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
// (rk4_intgs: f = Y^T, no conjugate, surf_rk.f90:154-155 with matAdivB2)
__global__ void k_out(Geo g, const double2 *__restrict__ P, const double2 *__restrict__ gc, double2 *__restrict__ f,
                      double *__restrict__ f2, int nb0, const double *__restrict__ wnsga2, int conjugate)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nn = (size_t)g.N * g.N;
    if (e >= nn * g.A) return;
    const int a = (int)(e / nn);
    const int r = (int)(e % nn), i = r % g.N, j = r / g.N;
    const double2 y = P[(size_t)a * g.mat + j + (size_t)i * g.NP];
    f[e] = conjugate ? make_double2(y.x, -y.y) : y;
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
                         double beta, int only_xid, double2 *__restrict__ Q, double2 *__restrict__ P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    if (only_xid && !xid[a]) return;   // its drift ran in the previous slice's end kick
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
    const int *xid;       // mode 0: angles whose drift is skipped (beta 0: Qn = Q), or null
    int mode;
    double *colsum;       // [A*NP/16][NP]
    double *diagabs;      // [A*NP]
    int ntiles, ncolt, tiles_per_xcd;
};

template <int BN>
__global__ __launch_bounds__(256, BN <= 48 ? 2 : 1) void k_stage_gemm(StageArgs S)
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
    int jb[NB];
    bool jv[NB];
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
        jb[cb] = c0 + cb * 16 + l16;
        jv[cb] = jb[cb] < g.NP;
    }
    // operands of k-step s+3 are loaded while the matrix cores work on step s (four
    // register sets rotate with the loop unrolled by four, so no set is copied while its
    // loads are in flight; past the last step the loads repeat it, unused); the four real
    // products of each complex block are issued term by term so that consecutive MFMAs never
    // depend on each other
    auto load = [&](int st, double2 (&av)[2], double2 (&bv)[NB]) {
        const int k = min(st, ksteps - 1) * 4 + kq;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = Qp[abase[rb] + (size_t)k * g.NP];
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) bv[cb] = jv[cb] ? Mp[(size_t)k * g.NP + jb[cb]] : make_double2(0, 0);
    };
    auto step = [&](const double2 (&av)[2], const double2 (&bv)[NB]) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) cr[rb][cb] = mfma4(av[rb].x, bv[cb].x, cr[rb][cb]);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) ci[rb][cb] = mfma4(av[rb].x, bv[cb].y, ci[rb][cb]);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) cr[rb][cb] = mfma4(-av[rb].y, bv[cb].y, cr[rb][cb]);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) ci[rb][cb] = mfma4(av[rb].y, bv[cb].x, ci[rb][cb]);
    };
    double2 a0[2], b0[NB], a1[2], b1[NB], a2[2], b2[NB], a3[2], b3[NB];
    load(0, a0, b0);
    load(1, a1, b1);
    load(2, a2, b2);
    for (int s = 0; s < ksteps; s += 4) {
        load(s + 3, a3, b3);
        step(a0, b0);
        load(s + 4, a0, b0);
        if (s + 1 < ksteps) step(a1, b1);
        load(s + 5, a1, b1);
        if (s + 2 < ksteps) step(a2, b2);
        load(s + 6, a2, b2);
        if (s + 3 < ksteps) step(a3, b3);
    }
    const double2 w0 = w_upper(S.col[0]);
    // Epilogue.  The Q, P and conj(gamma^2) loads of EB column blocks are issued together
    // (clamped to the matrix, so unconditional), then their updates run: one memory round
    // trip per EB blocks rather than one per element.  EB = 2 where the registers allow two
    // waves per SIMD, else 1.
    constexpr int EB = NB <= 2 ? 2 : 1;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        if (!rvalid[rb]) continue;
        const int Rb = r0 + rb * 16;            // first row of the 16-row block
        const int a = Rb / g.NP, ib = Rb % g.NP;
        const double alpha = S.coef ? S.coef[a] : S.alpha;
        const double beta = (S.xid && S.xid[a]) ? 0.0 : S.beta;
#pragma unroll
        for (int cp = 0; cp < NB; cp += EB) {
            double2 qv[EB][4], pv[EB][4], c2v[EB];
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                if (cp + e >= NB) continue;
                const int jj = min(c0 + (cp + e) * 16 + l16, g.NP - 1);
                c2v[e] = S.g2c[(size_t)a * g.NP + jj];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const size_t idx = (size_t)a * g.mat + (ib + kq + 4 * r) + (size_t)jj * g.NP;
                    qv[e][r] = S.Q[idx];
                    pv[e][r] = S.P[idx];
                }
            }
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                const int cb = cp + e;
                if (cb >= NB) continue;
                const int j = c0 + cb * 16 + l16;
                double csum = 0.0;
                if (j < g.N) {
                    const double2 dj = make_double2(w0.x - c2v[e].x, w0.y - c2v[e].y);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = ib + kq + 4 * r;
                        if (i < g.N) {
                            const size_t idx = (size_t)a * g.mat + i + (size_t)j * g.NP;
                            const double2 q = qv[e][r];
                            const double2 qd = cmul(q, dj);
                            double2 p = pv[e][r];
                            p.x += alpha * (cr[rb][cb][r] + qd.x);
                            p.y += alpha * (ci[rb][cb][r] + qd.y);
                            S.P[idx] = p;
                            if (S.mode != 2) {
                                const double2 qn = make_double2(q.x + beta * p.x, q.y + beta * p.y);
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
}

// Gershgorin estimate per angle; renormalise when > 1e3 or on the last slice.
// coef: the end kick's weight per angle, cend (c_b1 times this slice's thickness) where the
// angle renormalises, else cmerged (that plus c_b1 times the next slice's thickness: the
// next slice's first kick merged in)
__global__ void k_cond(Geo g, const double *__restrict__ colsum, const double *__restrict__ diagabs, int last,
                       int forced, double cend, double cmerged, double *__restrict__ coef, int *__restrict__ xid,
                       int *__restrict__ ren, int *__restrict__ counters)
{
    __shared__ double smn[256], smx[256];
    const int a = blockIdx.x, tid = threadIdx.x;
    const int nblk = g.NP / 16;
    double mn = 1e308, mx = -1e308;
    for (int j = tid; j < g.N; j += blockDim.x) {
        // the column's partial sums are loaded together, then added in block order
        constexpr int MAXB = 40;   // NP <= 640
        double part[MAXB];
#pragma unroll
        for (int b = 0; b < MAXB; ++b) part[b] = b < nblk ? colsum[((size_t)a * nblk + b) * g.NP + j] : 0.0;
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < MAXB; ++b)
            if (b < nblk) t += part[b];
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
        const int r = (forced >= 0 ? forced : (d > 1e3)) || last;
        ren[a] = r;
        xid[a] = r;
        coef[a] = r ? cend : cmerged;
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
    double2 *dinv;     // [npts][nblk][16*16] inverses of U's diagonal blocks
    int *perm;         // [2][A][LU_PERM] a panel's row interchanges as one permutation, by panel parity
};

// Layout of one angle's perm record: [0] the number n <= 32 of rows the panel's interchanges
// touch, [8 + i] the row that slot i writes, [40 + i] the row whose old content it receives.
// Slots 0..kw-1 are the panel's own rows k0..k0+kw-1, in order.  Panel k0 writes the record
// set (k0 / 16) & 1, so that the next panel can be factored while the trailing update of
// this one still reads it.
constexpr int LU_PERM = 72;

__device__ __forceinline__ int *perm_rec(const LuArgs &L, int k0, int a)
{
    return L.perm + ((size_t)((k0 >> 4) & 1) * L.g.A + a) * LU_PERM;
}

__device__ __forceinline__ double2 *colp(const LuArgs &L, int a, int c)
{
    return (c < L.g.N) ? L.Q + (size_t)a * L.g.mat + (size_t)c * L.g.NP
                       : L.P + (size_t)a * L.g.mat + (size_t)(c - L.g.N) * L.g.NP;
}

// Factors the panel [k0, k0+16) (rows k0..N-1) of every listed angle: one workgroup per
// angle, the panel held in registers (thread t owns panel rows t, t + 256 and t + 512, all
// 16 columns).  Per column: the pivot search (|re|+|im|, the first maximum), the row
// interchange (the pivot row is published in LDS to every thread, the displaced row goes to
// the pivot's owner), then each thread scales and updates its own rows.  LDS holds only the
// published rows and the per-wave search results, double-buffered by column parity, so
// that two barriers per column suffice.  Writes the factored panel and the pivots back.
constexpr int LU_RPT = 3;   // rows per thread: panels of up to 768 rows


struct PanelLds {
    double2 prow[2][16], crow[2][16];
    double wbest[2][4];
    int wbi[2][4];
    int pidx[2], cidx[2];   // original panel rows of the published rows
    int ntouch;
};

// A thread's panel rows are three separate arrays, visited through `rows3` with the row
// number as a constant: no row is ever addressed by a run-time index, so all three stay
// in registers.
struct PanelRows {
    double2 (&r0)[16], (&r1)[16], (&r2)[16];
};
template <class F> __device__ __forceinline__ void rows3(const PanelRows &v, F f)
{
    f(std::integral_constant<int, 0>{}, v.r0);
    f(std::integral_constant<int, 1>{}, v.r1);
    f(std::integral_constant<int, 2>{}, v.r2);
}

// column C of the panel (a template, so that every column index is a constant)
template <int C>
__device__ __forceinline__ void panel_column(int rows, const PanelRows &v, int (&oi)[LU_RPT], int (&pv)[16],
                                             bool &sing, PanelLds &sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int s = C & 1;
    double best = -1.0;
    int bi = rows;
    rows3(v, [&](auto J, double2 (&x)[16]) {
        const int r = tid + 256 * J;
        if (r >= C && r < rows) {
            const double m = fabs(x[C].x) + fabs(x[C].y);
            if (m > best) {
                best = m;
                bi = r;
            }
        }
    });
    pivot_wave(best, bi);
    if (lane == 0) {
        sh.wbest[s][wave] = best;
        sh.wbi[s][wave] = bi;
    }
    lds_barrier();
    {
        // the four waves' results, read together, combined without branches
        double wb[4];
        int wi[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            wb[w] = sh.wbest[s][w];
            wi[w] = sh.wbi[s][w];
        }
        best = wb[0];
        bi = wi[0];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            const bool better = wb[w] > best || (wb[w] == best && wi[w] < bi);
            best = better ? wb[w] : best;
            bi = better ? wi[w] : bi;
        }
    }
    if (bi >= rows) bi = C;   // no candidate compares (NaN column): no interchange
    pv[C] = bi;   // written to memory after the last column, so no barrier waits for a store
    if (!(best > 0.0)) sing = true;
    // publish the pivot row; row C (thread C's first row) goes to the pivot row's owner
    rows3(v, [&](auto J, double2 (&x)[16]) {
        if (tid + 256 * J == bi) {
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) sh.prow[s][cc] = x[cc];
            sh.pidx[s] = oi[J];
        }
    });
    if (tid == C && bi != C) {
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) sh.crow[s][cc] = v.r0[cc];
        sh.cidx[s] = oi[0];
    }
    lds_barrier();
    if (bi != C) {
        if (tid == C) {
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) v.r0[cc] = sh.prow[s][cc];
            oi[0] = sh.pidx[s];
        }
        rows3(v, [&](auto J, double2 (&x)[16]) {
            if (tid + 256 * J == bi) {
#pragma unroll
                for (int cc = 0; cc < 16; ++cc) x[cc] = sh.crow[s][cc];
                oi[J] = sh.cidx[s];
            }
        });
    }
    const double2 pk = sh.prow[s][C];
    const bool nz = (pk.x != 0.0 || pk.y != 0.0);
    const double2 rp = nz ? cdiv(make_double2(1, 0), pk) : make_double2(0, 0);
    const double2 *u = sh.prow[s];
    // Each row below C is scaled and updated; for the other rows the multiplier is zero, so
    // that the update is the same straight-line code for every row and leaves them exactly
    // as they are.  Columns beyond kw are zero in every row and stay zero.
    double2 l[LU_RPT];
    rows3(v, [&](auto J, double2 (&x)[16]) {
        const int r = tid + 256 * J;
        const bool below = r > C && r < rows;
        l[J] = below ? x[C] : make_double2(0, 0);
        if (nz) l[J] = cmul(l[J], rp);
        if (below) x[C] = l[J];
    });
    const bool third = rows > 512;   // uniform: whether any thread has a third row
#pragma unroll
    for (int cc = C + 1; cc < 16; ++cc) {
        const double2 uc = u[cc];
        rows3(v, [&](auto J, double2 (&x)[16]) {
            if (J < 2 || third) {
                x[cc].x -= l[J].x * uc.x - l[J].y * uc.y;
                x[cc].y -= l[J].x * uc.y + l[J].y * uc.x;
            }
        });
    }
}

template <int C>
__device__ __forceinline__ void panel_columns(int kw, int rows, const PanelRows &v, int (&oi)[LU_RPT], int (&pv)[16],
                                              bool &sing, PanelLds &sh)
{
    // kw is uniform, so every thread takes the same branches to the barriers
    if (C < kw) panel_column<C>(rows, v, oi, pv, sing, sh);
    if constexpr (C + 1 < 16) panel_columns<C + 1>(kw, rows, v, oi, pv, sing, sh);
}

// Factors panel k0 of angle a with all 256 threads of the workgroup (see k_lu_panel).
__device__ __forceinline__ void factor_panel(const LuArgs &L, int a, int k0, PanelLds &sh)
{
    const int N = L.g.N, tid = threadIdx.x;
    const int kw = min(16, N - k0), rows = N - k0;
    double2 v0[16], v1[16], v2[16];
    const PanelRows v{v0, v1, v2};
    rows3(v, [&](auto J, double2 (&x)[16]) {
        const int r = tid + 256 * J;
#pragma unroll
        for (int c = 0; c < 16; ++c) x[c] = (c < kw && r < rows) ? colp(L, a, k0 + c)[k0 + r] : make_double2(0, 0);
    });
    int oi[LU_RPT];   // the original panel row of each row the thread holds
#pragma unroll
    for (int j = 0; j < LU_RPT; ++j) oi[j] = tid + 256 * j;
    if (tid == 0) sh.ntouch = kw;
    int pv[16];
    bool sing = false;
    panel_columns<0>(kw, rows, v, oi, pv, sing, sh);
    if (tid == 0) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c < kw) L.piv[(size_t)a * L.g.NP + k0 + c] = k0 + pv[c];
        if (sing) *L.flag = 1;
    }
    // the interchanges as one permutation for the trailing updates: the panel's own rows
    // take slots 0..kw-1, every other row whose content changed one of the following slots
    int *pm = perm_rec(L, k0, a);
    rows3(v, [&](auto J, double2 (&x)[16]) {
        const int r = tid + 256 * J;
        if (r < rows) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c < kw) colp(L, a, k0 + c)[k0 + r] = x[c];
            int slot = -1;
            if (r < kw) slot = r;
            else if (oi[J] != r) slot = atomicAdd(&sh.ntouch, 1);
            if (slot >= 0) {
                pm[8 + slot] = k0 + r;
                pm[40 + slot] = k0 + oi[J];
            }
        }
    });
    lds_barrier();   // the slot counter is complete
    if (tid == 0) pm[0] = sh.ntouch;
}

// the first panel of every listed angle; the later ones are factored inside k_lu_step
__global__ __launch_bounds__(256) void k_lu_panel(LuArgs L, int k0)
{
    __shared__ PanelLds sh;
    factor_panel(L, L.list[blockIdx.x], k0, sh);
}

// One elimination step for a 64-column tile right of panel [k0, k0+kw): the panel's row
// interchanges (gathered as one permutation of the <= 32 rows they touch), U12 =
// L11^{-1} A12, A22 -= L21 U12 on the matrix cores.  The next panel is the first 16 columns
// of tile 0: once tile 0's update is done, its workgroup factors that panel, while the other
// tiles of the step still update theirs.
constexpr int LU_TILE = 64;

__global__ __launch_bounds__(256, 2) void k_lu_step(LuArgs L, int k0)
{
    __shared__ double2 swp[32 * LU_TILE];
    __shared__ double2 U[16][LU_TILE];
    __shared__ double2 L11[16][16];
    __shared__ int trow[32], tsrc[32];
    const int a = L.list[blockIdx.y];
    const int N = L.g.N, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kw = min(16, N - k0), first = k0 + kw;
    // tile 0 is the next panel's 16 columns, the other tiles LU_TILE columns each
    const bool t0 = blockIdx.x == 0;
    const int c0 = t0 ? first : first + 16 + (blockIdx.x - 1) * LU_TILE;
    const int cw = min(t0 ? 16 : LU_TILE, 2 * N - c0);
    // the panel's interchanges, gathered by the panel kernel as one permutation
    const int *pm = perm_rec(L, k0, a);
    const int nt = pm[0];
    if (tid < nt) {
        trow[tid] = pm[8 + tid];
        tsrc[tid] = pm[40 + tid];
    }
    {
        const int r = tid >> 4, c = tid & 15;
        L11[r][c] = (r < kw && c < kw) ? colp(L, a, k0 + c)[k0 + r] : make_double2(0, 0);
    }
    __syncthreads();
    {
        // up to 32 x 64 elements: 8 per thread, loads issued together
        double2 buf[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * 256, i = e / LU_TILE, c = e % LU_TILE;
            buf[q] = (i < nt && c < cw) ? colp(L, a, c0 + c)[tsrc[i]] : make_double2(0, 0);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * 256, i = e / LU_TILE;
            if (i < nt) swp[e] = buf[q];
        }
    }
    __syncthreads();
    for (int e = tid; e < nt * LU_TILE; e += 256) {
        const int i = e / LU_TILE, c = e % LU_TILE;
        if (c < cw && trow[i] >= first) colp(L, a, c0 + c)[trow[i]] = swp[i * LU_TILE + c];
    }
    // U12: forward substitution with unit-lower L11, one column per thread (rows k0.. are
    // the first kw touched rows, in order)
    if (tid < LU_TILE) {
        for (int r = 0; r < 16; ++r) U[r][tid] = make_double2(0, 0);
        if (tid < cw) {
            double2 *col = colp(L, a, c0 + tid);
            for (int r = 0; r < kw; ++r) {
                double2 v = swp[r * LU_TILE + tid];
                for (int rr = 0; rr < r; ++rr) {
                    const double2 l = L11[r][rr], x = U[rr][tid];
                    v.x -= l.x * x.x - l.y * x.y;
                    v.y -= l.x * x.y + l.y * x.x;
                }
                U[r][tid] = v;
                col[k0 + r] = v;
            }
        }
    }
    __syncthreads();
    const int l16 = lane & 15, kq = lane >> 4;
    const int nrb = (N - first + 15) / 16;
    const int ks = (kw + 3) >> 2;
    const int ncb = (cw + 15) / 16;
    // A22 -= L21 U12: wave w owns the 16-column block w of the tile and walks down all row
    // blocks four at a time; the L21 fragments and the C values of the four blocks are loaded
    // together, before their products, and the wave's U12 operands stay in registers.  The
    // product is formed transposed (U12^T as the A operand), so that a lane's results are
    // rows l16 of columns kq + 4 r: 16 lanes cover 16 consecutive rows of a column.
    // (tile 0: all four waves on its one column block, row blocks interleaved, so that it
    // reaches the panel factorisation early)
    const int cb = t0 ? 0 : wave, hs = t0 ? 4 : 1;
    if (cb < ncb) {
        double2 uv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = s * 4 + kq, c = cb * 16 + l16;
            uv[s] = (s < ks && k < kw && c < cw) ? U[k][c] : make_double2(0, 0);
        }
        double2 *col[4];
        bool cok[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = cb * 16 + kq + 4 * r;
            cok[r] = c < cw;
            col[r] = colp(L, a, c0 + (cok[r] ? c : 0));
        }
        // iteration it covers row blocks rs + 2 hs it and rs + 2 hs it + hs; the operands of
        // iteration it + 1 are loaded while iteration it runs (two register sets rotate with
        // the loop unrolled by two; past the last iteration every load is predicated off)
        const int rs = t0 ? wave : 0;
        const int niter = nrb > rs ? (nrb - rs + 2 * hs - 1) / (2 * hs) : 0;
        auto load = [&](int it, double2 (&av)[2][4], double2 (&cv)[2][4]) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ia = first + (rs + it * 2 * hs + h * hs) * 16 + l16;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int k = s * 4 + kq;
                    av[h][s] = (s < ks && k < kw && ia < N) ? colp(L, a, k0 + k)[ia] : make_double2(0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) cv[h][r] = (cok[r] && ia < N) ? col[r][ia] : make_double2(0, 0);
            }
        };
        auto update = [&](int it, const double2 (&av)[2][4], const double2 (&cv)[2][4]) {
            d4 cr[2], ci[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                cr[h] = d4{0, 0, 0, 0};
                ci[h] = d4{0, 0, 0, 0};
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s >= ks) break;
#pragma unroll
                for (int h = 0; h < 2; ++h) cr[h] = mfma4(uv[s].x, av[h][s].x, cr[h]);
#pragma unroll
                for (int h = 0; h < 2; ++h) ci[h] = mfma4(uv[s].x, av[h][s].y, ci[h]);
#pragma unroll
                for (int h = 0; h < 2; ++h) cr[h] = mfma4(-uv[s].y, av[h][s].y, cr[h]);
#pragma unroll
                for (int h = 0; h < 2; ++h) ci[h] = mfma4(uv[s].y, av[h][s].x, ci[h]);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ia = first + (rs + it * 2 * hs + h * hs) * 16 + l16;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (cok[r] && ia < N) {
                        double2 v = cv[h][r];
                        v.x -= cr[h][r];
                        v.y -= ci[h][r];
                        col[r][ia] = v;
                    }
            }
        };
        double2 a0[2][4], c0v[2][4], a1[2][4], c1v[2][4];
        load(0, a0, c0v);
        for (int it = 0; it < niter; it += 2) {
            load(it + 1, a1, c1v);
            update(it, a0, c0v);
            load(it + 2, a0, c0v);
            if (it + 1 < niter) update(it + 1, a1, c1v);
        }
    }
    if (t0 && first < N) {
        __shared__ PanelLds sh;
        __syncthreads();   // tile 0's update of the next panel's columns is visible
        factor_panel(L, a, first, sh);
    }
}

// Inverses of the 16x16 diagonal blocks of U, all blocks of all listed angles in
// parallel: one column of the inverse per thread (back substitution in LDS).
__global__ __launch_bounds__(64) void k_lu_dinv(LuArgs L, double2 *__restrict__ dinv)
{
    __shared__ double2 Ukk[16][17];
    const int a = L.list[blockIdx.y], kb = blockIdx.x;
    const int N = L.g.N, tid = threadIdx.x;
    const int r0 = kb * 16, rw = min(16, N - r0);
    for (int e = tid; e < 256; e += 64) {
        const int r = e >> 4, c = e & 15;
        Ukk[r][c] = (r < rw && c < rw) ? colp(L, a, r0 + c)[r0 + r] : make_double2(0, 0);
    }
    __syncthreads();
    double2 *out = dinv + ((size_t)blockIdx.y * gridDim.x + kb) * 256;   // [col][row]
    if (tid < 16) {
        double2 x[16];
#pragma unroll
        for (int r = 15; r >= 0; --r) {
            double2 v = make_double2(r == tid ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int rr = r + 1; rr < 16; ++rr) {
                const double2 u = Ukk[r][rr];
                v.x -= u.x * x[rr].x - u.y * x[rr].y;
                v.y -= u.x * x[rr].y + u.y * x[rr].x;
            }
            x[r] = (r < rw && tid < rw) ? cdiv(v, Ukk[r][r]) : make_double2(0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) out[tid * 16 + r] = x[r];
    }
}
// Y <- U^{-1} Y by block back substitution, one workgroup of BK_WAVES waves per 16-column
// block of Y, that column of Y held in registers for the whole solve: wave w holds row
// blocks w, w + BK_WAVES, ... (at most BK_SLOTS), each as MFMA accumulators (rows kq + 4 r of
// the lane's column), so that the updates Y_i -= U_ik X accumulate into them in place.
// Step k (k = nblk-1 .. 0): wave k % BK_WAVES forms X = inv(U_kk) Y_k from its registers,
// stores it (the final rows of Y_k) and publishes it in LDS; after an LDS-only barrier every
// wave subtracts U_ik X from its blocks i < k.  The U_ik operands are loaded at the start of
// the step, before X is known, and the barrier does not wait for them; two waves per SIMD
// cover each other's waits.  The blocks move down one register slot every BK_WAVES steps,
// so that the block a step finishes is always its wave's slot 0.
constexpr int BK_WAVES = 8, BK_SLOTS = 5;   // row blocks per column: N <= 640

__global__ __launch_bounds__(64 * BK_WAVES) void k_lu_back(LuArgs L, const double2 *__restrict__ dinv)
{
    __shared__ double xs[2][2][4][64];   // [step parity][re, im][r][lane]
    const int a = L.list[blockIdx.y];
    const int N = L.g.N, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nblk = (N + 15) / 16;
    const int j0 = blockIdx.x * 16, cw = min(16, N - j0);
    const int l16 = lane & 15, kq = lane >> 4;
    double2 *y = colp(L, a, N + j0 + min(l16, cw - 1));
    const int qtop = (nblk - 1) / BK_WAVES;
    // slot i of group q (steps BK_WAVES q + BK_WAVES-1 .. BK_WAVES q) holds row block
    // BK_WAVES (q - i) + wave
    d4 br[BK_SLOTS], bi[BK_SLOTS];
#pragma unroll
    for (int i = 0; i < BK_SLOTS; ++i) {
        const int rb = BK_WAVES * (qtop - i) + wave;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rb * 16 + kq + 4 * r;
            const double2 v = (rb >= 0 && row < N && l16 < cw) ? y[row] : make_double2(0, 0);
            br[i][r] = v.x;
            bi[i][r] = v.y;
        }
    }
    for (int q = qtop; q >= 0; --q) {
#pragma unroll
        for (int wp = BK_WAVES - 1; wp >= 0; --wp) {
            const int kb = BK_WAVES * q + wp;
            if (kb >= nblk) continue;   // uniform
            const int r0 = kb * 16, rw = min(16, N - r0), ks = (rw + 3) >> 2;
            double2 au[BK_SLOTS][4];   // U(block of slot i, block kb)
#pragma unroll
            for (int i = 0; i < BK_SLOTS; ++i) {
                const int rb = BK_WAVES * (q - i) + wave;
                const bool hv = (i > 0 || wave < wp) && rb >= 0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int k = s * 4 + kq;
                    au[i][s] = (hv && s < ks && k < rw) ? colp(L, a, r0 + k)[rb * 16 + l16] : make_double2(0, 0);
                }
            }
            const int par = kb & 1;
            if (wave == wp) {
                const double2 *inv = dinv + ((size_t)blockIdx.y * nblk + kb) * 256;
                d4 xr = {0, 0, 0, 0}, xi = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const double2 iv = inv[(s * 4 + kq) * 16 + l16];   // inv(l16, 4 s + kq)
                    xr = mfma4(iv.x, br[0][s], xr);
                    xi = mfma4(iv.x, bi[0][s], xi);
                    xr = mfma4(-iv.y, bi[0][s], xr);
                    xi = mfma4(iv.y, br[0][s], xi);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    xs[par][0][r][lane] = xr[r];
                    xs[par][1][r][lane] = xi[r];
                    if (l16 < cw && kq + 4 * r < rw) y[r0 + kq + 4 * r] = make_double2(xr[r], xi[r]);
                }
            }
            lds_barrier();
            d4 xr, xi;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                xr[r] = xs[par][0][r][lane];
                xi[r] = xs[par][1][r][lane];
            }
#pragma unroll
            for (int i = 0; i < BK_SLOTS; ++i) {
                const int rb = BK_WAVES * (q - i) + wave;
                if (!((i > 0 || wave < wp) && rb >= 0)) continue;   // uniform per wave
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (s >= ks) break;
                    br[i] = mfma4(-au[i][s].x, xr[s], br[i]);
                    bi[i] = mfma4(-au[i][s].x, xi[s], bi[i]);
                    br[i] = mfma4(au[i][s].y, xi[s], br[i]);
                    bi[i] = mfma4(-au[i][s].y, xr[s], bi[i]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i + 1 < BK_SLOTS; ++i) {
            br[i] = br[i + 1];
            bi[i] = bi[i + 1];
        }
        br[BK_SLOTS - 1] = d4{0, 0, 0, 0};
        bi[BK_SLOTS - 1] = d4{0, 0, 0, 0};
    }
}

// ---------------------------------------------------------------------------
// Magnus integrators (m2/m4/m6_int_spa, surf_magnus.f90:114-339).  Their products involve
// per-angle matrices on both sides (M^2, Q rp, P q), so they run as batched dense products
// of NP x NP column-major matrices (zero padding) plus elementwise combinations.

// dense per-angle slice matrix M_a = M0 + diag(d_a) (vvi2mate, matcomp.f90:253-279)
__global__ void k_build_m(Geo g, const int *__restrict__ gmap, const VVi *__restrict__ col,
                          const double2 *__restrict__ g2c, double2 *__restrict__ M)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    double2 m = make_double2(0, 0);
    if (i < g.N && j < g.N) {
        if (i == j) {
            const double2 w = w_upper(col[0]), c2 = g2c[(size_t)a * g.NP + j];
            m = make_double2(w.x - c2.x, w.y - c2.y);
        } else {
            const int s = gmap[i * g.N + j];
            m = s > 0 ? w_upper(col[s - 1]) : w_lower(col[-s - 1]);
        }
    }
    M[e] = m;
}

// C_a = alpha A_a B_a + beta C_a + ident I for every angle; one wave per 16 x 16 output
// block, four waves per workgroup.  C must not alias A or B.
struct BmmArgs {
    Geo g;
    const double2 *A, *B;
    double2 *C;
    double alpha, beta, ident;
    const int *list;   // angles (blockIdx.y indexes the list), or null for all
    int conjA;         // use A^H instead of A
};

__global__ __launch_bounds__(256) void k_bmm(BmmArgs S)
{
    const Geo g = S.g;
    const int nbl = g.NP / 16;
    const int a = S.list ? S.list[blockIdx.y] : (int)blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int blk = blockIdx.x * 4 + wave;
    if (blk >= nbl * nbl) return;
    const int bi = blk % nbl, bj = blk / nbl;
    // A(row, k) at Ab[k * ks]; with conjA the element is conj(A(k, row))
    const double2 *__restrict__ Ab = S.A + (size_t)a * g.mat +
                                     (S.conjA ? (size_t)(bi * 16 + l16) * g.NP : (size_t)(bi * 16 + l16));
    const size_t ks = S.conjA ? 1 : g.NP;
    const double sgn = S.conjA ? -1.0 : 1.0;
    const double2 *__restrict__ Bb = S.B + (size_t)a * g.mat + (size_t)(bj * 16 + l16) * g.NP;
    d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    // 16 k per iteration: the operands of four MFMA steps are loaded together
    for (int k0 = 0; k0 < g.NP; k0 += 16) {
        double2 av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + 4 * u + kq;
            av[u] = Ab[(size_t)k * ks];
            bv[u] = Bb[k];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            av[u].y *= sgn;
            cr = mfma4(av[u].x, bv[u].x, cr);
            ci = mfma4(av[u].x, bv[u].y, ci);
            cr = mfma4(-av[u].y, bv[u].y, cr);
            ci = mfma4(av[u].y, bv[u].x, ci);
        }
    }
    const int j = bj * 16 + l16;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = bi * 16 + kq + 4 * r;
        if (i >= g.N || j >= g.N) continue;
        const size_t idx = (size_t)a * g.mat + i + (size_t)j * g.NP;
        double2 c = make_double2(S.alpha * cr[r], S.alpha * ci[r]);
        if (S.beta != 0.0) {
            const double2 o = S.C[idx];
            c = make_double2(S.beta * o.x + c.x, S.beta * o.y + c.y);
        }
        if (i == j) c.x += S.ident;
        S.C[idx] = c;
    }
}

// out = c0 x0 + c1 x1 + c2 x2 + c3 x3 + ident I (real coefficients, elementwise, n terms,
// summed left to right); out may be one of the inputs
struct LinArgs {
    Geo g;
    int n;
    double c[4];
    const double2 *x[4];
    double2 *out;
    double ident;
};

__global__ void k_lin(LinArgs L)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.g.mat * L.g.A) return;
    const int r = (int)(e % L.g.mat), i = r % L.g.NP, j = r / L.g.NP;
    double2 v = make_double2(0, 0);
    if (i < L.g.N && j < L.g.N) {
        for (int t = 0; t < L.n; ++t) {
            const double2 x = L.x[t][e];
            v.x += L.c[t] * x.x;
            v.y += L.c[t] * x.y;
        }
        if (i == j) v.x += L.ident;
    }
    L.out[e] = v;
}

// Up to four k_lin outputs in one pass over the elements, formed in order: a term naming an
// earlier output of the set takes that output's new value (what a later k_lin would load),
// any other term the memory as it was before the set; the outputs are stored after all are
// formed, so an output may also be an input.  The arithmetic of every output is k_lin's.
struct LinSet {
    int nout;
    LinArgs o[4];
};

__global__ void k_lin_set(Geo g, LinSet S)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    const bool in = i < g.N && j < g.N;
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v[u] = make_double2(0, 0);
        if (u >= S.nout || !in) continue;
        const LinArgs &L = S.o[u];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t >= L.n) break;
            double2 x;
            if (u > 0 && L.x[t] == S.o[0].out) x = v[0];
            else if (u > 1 && L.x[t] == S.o[1].out) x = v[1];
            else if (u > 2 && L.x[t] == S.o[2].out) x = v[2];
            else x = L.x[t][e];
            v[u].x += L.c[t] * x.x;
            v[u].y += L.c[t] * x.y;
        }
        if (i == j) v[u].x += L.ident;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (u < S.nout) S.o[u].out[e] = v[u];
}

// Gershgorin terms of Q (cond_est, matcomp.f90:498-528) in k_cond's layout: per angle, per
// 16-row block b and column j, the sum of |Q(i,j)| over the block's rows i != j; |Q(j,j)|
__global__ void k_gersh(Geo g, const double2 *__restrict__ Q, double *__restrict__ colsum,
                        double *__restrict__ diagabs)
{
    const int nblk = g.NP / 16;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)g.A * nblk * g.NP) return;
    const int j = (int)(e % g.NP);
    const size_t ab = e / g.NP;   // a * nblk + b
    const int a = (int)(ab / nblk), b = (int)(ab % nblk);
    double s = 0.0;
    if (j < g.N) {
        const double2 *qc = Q + (size_t)a * g.mat + (size_t)j * g.NP;
        for (int i = b * 16; i < b * 16 + 16 && i < g.N; ++i) {
            const double2 q = qc[i];
            if (i == j) diagabs[(size_t)a * g.NP + j] = hypot(q.x, q.y);
            else s += sqrt(q.x * q.x + q.y * q.y);
        }
    }
    colsum[ab * g.NP + j] = s;
}

// ---------------------------------------------------------------------------
// rk4_intgs (surf_rk.f90:93-156): RK4 on [q; p] with left products M q, the Cholesky-QR
// renormalisation of bcgs (matgs.f90:4-37) when the power estimate of ||[q; p]||^2 exceeds
// 1e2 (power_est, matcomp.f90:530-594), and strans + matAdivB2 at the end.

// vvi2matf (matcomp.f90:281-306): upper conj(vi - v), lower -(v + vi), diagonal
// -(v + vi)[1] - gamma_j^2
__global__ void k_build_mf(Geo g, const int *__restrict__ gmap, const VVi *__restrict__ col,
                           const double2 *__restrict__ g2c, double2 *__restrict__ M)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    double2 m = make_double2(0, 0);
    if (i < g.N && j < g.N) {
        if (i == j) {
            const VVi t = col[0];
            const double2 c2 = g2c[(size_t)a * g.NP + j];   // conj(gamma^2)
            m = make_double2(-(t.v.x + t.vi.x) - c2.x, -(t.v.y + t.vi.y) + c2.y);
        } else {
            const int s = gmap[i * g.N + j];
            if (s > 0) {
                const VVi t = col[s - 1];
                m = make_double2(t.vi.x - t.v.x, -(t.vi.y - t.v.y));
            } else {
                const VVi t = col[-s - 1];
                m = make_double2(-(t.v.x + t.vi.x), -(t.v.y + t.vi.y));
            }
        }
    }
    M[e] = m;
}

// invstrans (matcomp.f90:89-115): q = G^-1 (I + F), p = i (F - I); x = (1 + i) / sqrt(2N)
__global__ void k_gs_setup(Geo g, const double2 *__restrict__ gamma, const double2 *__restrict__ f,
                           double2 *__restrict__ Q, double2 *__restrict__ P, double2 *__restrict__ X)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    double2 q = make_double2(0, 0), p = make_double2(0, 0);
    if (i < g.N && j < g.N) {
        const double2 gi = gamma[(size_t)a * g.N + i];
        const double2 ig = cdiv(make_double2(1, 0), gi);
        const double2 fij = f[(size_t)a * g.N * g.N + i + (size_t)j * g.N];
        const double2 t = cmul(ig, fij);
        if (i == j) {
            q = make_double2(ig.x + t.x, ig.y + t.y);
            p = make_double2(-fij.y, -1.0 + fij.x);
        } else {
            q = t;
            p = make_double2(-fij.y, fij.x);
        }
    }
    Q[e] = q;
    P[e] = p;
    if (r < g.NP) {
        const double v = 1.0 / sqrt(2.0 * g.N);
        X[(size_t)a * g.NP + r] = r < g.N ? make_double2(v, v) : make_double2(0, 0);
    }
}

// Blocked Cholesky C = R^H R (upper R written over the upper triangle of C; below the
// diagonal C is left as scratch), right-looking over 16-column blocks.  Step kb:
// k_chol_panel factors the diagonal block (every workgroup redundantly, in LDS) and solves
// the row panel R(kb, c) = R_kk^-H C(kb, c) for its 64 columns c, one column per thread;
// k_chol_update then subtracts R(kb, i)^H R(kb, j) from the trailing blocks i <= j on the
// matrix cores.  The factored diagonal block goes to `diag` ([list][block][col][row]), not
// into C, since other workgroups of the same launch still read C's block; k_chol_diag puts
// all of them into C at the end.
__global__ __launch_bounds__(256) void k_chol_panel(Geo g, const int *__restrict__ list, double2 *__restrict__ C, int kb,
                                                    double2 *__restrict__ diag, int *__restrict__ counters, int count)
{
    __shared__ double2 Rk[16][17];
    const int a = list[blockIdx.y], tid = threadIdx.x;
    const int N = g.N, NP = g.NP, k0 = kb * 16, kw = min(16, N - k0);
    double2 *c = C + (size_t)a * g.mat;
    {
        const int r = tid & 15, cc = tid >> 4;
        Rk[r][cc] = (r <= cc && cc < kw) ? c[(size_t)(k0 + cc) * NP + k0 + r] : make_double2(0, 0);
    }
    __syncthreads();
    // the 16 x 16 diagonal block, column by column: thread cc < 16 owns column cc
    for (int j = 0; j < kw; ++j) {
        if (tid == 0) {
            double d = Rk[j][j].x;
            for (int r = 0; r < j; ++r) d -= Rk[r][j].x * Rk[r][j].x + Rk[r][j].y * Rk[r][j].y;
            Rk[j][j] = make_double2(sqrt(d), 0.0);
        }
        __syncthreads();
        if (tid > j && tid < kw) {
            double2 acc = Rk[j][tid];
            for (int r = 0; r < j; ++r) {
                const double2 u = Rk[r][j], v = Rk[r][tid];
                acc.x -= u.x * v.x + u.y * v.y;   // conj(u) v
                acc.y -= u.x * v.y - u.y * v.x;
            }
            const double d = Rk[j][j].x;
            Rk[j][tid] = make_double2(acc.x / d, acc.y / d);
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        const int r = tid & 15, cc = tid >> 4;
        diag[((size_t)blockIdx.y * (NP / 16) + kb) * 256 + cc * 16 + r] = Rk[r][cc];
        if (kb == 0 && count && tid == 0) atomicAdd(&counters[0], 1);
    }
    // the row panel: columns k0 + 16 + 64 blockIdx.x + tid (tid < 64)
    const int col = k0 + 16 + blockIdx.x * 64 + tid;
    if (tid < 64 && col < N) {
        double2 *cp = c + (size_t)col * NP + k0;
        double2 x[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (j < kw) {
                double2 v = cp[j];
#pragma unroll
                for (int r = 0; r < j; ++r) {
                    const double2 u = Rk[r][j];   // R_kk^H (j, r) = conj(R_kk(r, j))
                    v.x -= u.x * x[r].x + u.y * x[r].y;
                    v.y -= u.x * x[r].y - u.y * x[r].x;
                }
                const double d = Rk[j][j].x;
                x[j] = make_double2(v.x / d, v.y / d);
                cp[j] = x[j];
            }
        }
    }
}

// the factored diagonal blocks into the upper triangle of C (grid: blocks, angles)
__global__ __launch_bounds__(256) void k_chol_diag(Geo g, const int *__restrict__ list, double2 *__restrict__ C,
                                                   const double2 *__restrict__ diag)
{
    const int a = list[blockIdx.y], kb = blockIdx.x, tid = threadIdx.x;
    const int r = tid & 15, cc = tid >> 4, k0 = kb * 16;
    if (r <= cc && k0 + cc < g.N)
        C[(size_t)a * g.mat + (size_t)(k0 + cc) * g.NP + k0 + r] = diag[((size_t)blockIdx.y * gridDim.x + kb) * 256 + cc * 16 + r];
}

// C(i, j) -= sum_r conj(R(k0 + r, i)) R(k0 + r, j) for the trailing 16 x 16 blocks with
// i-block <= j-block: one wave per block (grid: j-blocks, groups of four i-blocks, angles)
__global__ __launch_bounds__(256) void k_chol_update(Geo g, const int *__restrict__ list, double2 *__restrict__ C, int kb)
{
    const int a = list[blockIdx.z];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int N = g.N, NP = g.NP, k0 = kb * 16, kw = min(16, N - k0);
    const int b0 = kb + 1;
    const int bj = b0 + blockIdx.x, bi = b0 + blockIdx.y * 4 + wave;
    if (bi > bj) return;
    double2 *c = C + (size_t)a * g.mat;
    const int ks = (kw + 3) >> 2;
    d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    double2 av[4], bv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = 4 * s + kq;
        av[s] = (s < ks && k < kw) ? c[(size_t)(bi * 16 + l16) * NP + k0 + k] : make_double2(0, 0);
        bv[s] = (s < ks && k < kw) ? c[(size_t)(bj * 16 + l16) * NP + k0 + k] : make_double2(0, 0);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        // conj(a) b = (ar br + ai bi) + i (ar bi - ai br)
        cr = mfma4(av[s].x, bv[s].x, cr);
        cr = mfma4(av[s].y, bv[s].y, cr);
        ci = mfma4(av[s].x, bv[s].y, ci);
        ci = mfma4(-av[s].y, bv[s].x, ci);
    }
    const int j = bj * 16 + l16;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = bi * 16 + kq + 4 * r;
        if (i < N && j < N) {
            double2 *e = c + (size_t)j * NP + i;
            double2 v = *e;
            v.x -= cr[r];
            v.y -= ci[r];
            *e = v;
        }
    }
}

// ztrsm 'R','U','N','N' by blocks: X <- X U^-1 with one wave per 16-row block of X, walking
// the 16-column blocks jb: T = X_jb - sum_{ib<jb} X_ib U_(ib,jb), then X_jb = T inv(U_(jb,jb))
// (the inverses of the diagonal blocks from k_lu_dinv).  The products are formed transposed
// (U^T X^T), so that a lane's results are row l16 of columns kq + 4 r (16 lanes cover 16
// consecutive rows of a column) and are, register for register, the operand of the
// multiplication by the inverse.  Each wave reads back only blocks it has itself finished.
__global__ __launch_bounds__(256) void k_trsm_rb(Geo g, const int *__restrict__ list, const double2 *__restrict__ C,
                                                 const double2 *__restrict__ dinv, double2 *__restrict__ X)
{
    const int a = list[blockIdx.y];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int nblk = g.NP / 16;
    const int rb = blockIdx.x * 4 + wave;
    if (rb >= nblk) return;
    const size_t NP = g.NP;
    const double2 *__restrict__ u = C + (size_t)a * g.mat;
    double2 *x = X + (size_t)a * g.mat + rb * 16 + l16;   // X(rb 16 + l16, c) at x[c NP]
    const double2 *inv0 = dinv + (size_t)blockIdx.y * nblk * 256;
    for (int jb = 0; jb < nblk; ++jb) {
        d4 tr, ti;   // T(l16, jb 16 + kq + 4 r)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double2 v = x[(size_t)(jb * 16 + kq + 4 * r) * NP];
            tr[r] = v.x;
            ti[r] = v.y;
        }
        for (int ib0 = 0; ib0 < jb; ib0 += 4) {
            double2 av[4][4], bv[4][4];   // U(ib 16 + 4 s + kq, jb 16 + l16), X(l16, ib 16 + 4 s + kq)
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int ib = ib0 + h;
                const bool hv = ib < jb;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int k = ib * 16 + 4 * s + kq;
                    av[h][s] = hv ? u[(size_t)(jb * 16 + l16) * NP + k] : make_double2(0, 0);
                    bv[h][s] = hv ? x[(size_t)k * NP] : make_double2(0, 0);
                }
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                if (ib0 + h >= jb) break;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    tr = mfma4(-av[h][s].x, bv[h][s].x, tr);
                    tr = mfma4(av[h][s].y, bv[h][s].y, tr);
                    ti = mfma4(-av[h][s].x, bv[h][s].y, ti);
                    ti = mfma4(-av[h][s].y, bv[h][s].x, ti);
                }
            }
        }
        const double2 *inv = inv0 + (size_t)jb * 256;
        d4 xr = {0, 0, 0, 0}, xi = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const double2 iv = inv[l16 * 16 + 4 * s + kq];   // inv(4 s + kq, l16)
            xr = mfma4(iv.x, tr[s], xr);
            xr = mfma4(-iv.y, ti[s], xr);
            xi = mfma4(iv.y, tr[s], xi);
            xi = mfma4(iv.x, ti[s], xi);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) x[(size_t)(jb * 16 + kq + 4 * r) * NP] = make_double2(xr[r], xi[r]);
    }
}

// power_est / power_step (matcomp.f90:530-594) for every angle, x carried between slices:
// y = [q x; p x], z = q^H y1 + p^H y2, d = |x^H z|, x = z / |z|, until |z - x| < 1e-2 or
// d grows by less than 1e-2 d, at most N steps.  The first steps of a slice run spread over
// 16-row blocks of [q; p] (k_pw_rows: y on the block's rows and the block's share of z,
// many workgroups per angle; k_pw_fin: z summed over the blocks, the normalisation and the
// stopping test, one workgroup per angle); k_power goes on from where they stopped, one
// workgroup per angle, when the estimate has not converged in those steps.
struct PwState {
    double d0, d;   // the previous and the latest estimate
    int it, done;   // steps taken; 1 once the stopping test has passed
};
constexpr int PW_FAST = 3;   // block-split steps per slice
constexpr int PW_MAXN = 640;

// one step's first half for row block blockIdx.x of angle blockIdx.y: y1 = q x, y2 = p x on
// the block's 16 rows, then part[a][w][j] = sum over those rows of conj(q(i,j)) y1(i) +
// conj(p(i,j)) y2(i).  Thread (r, cg): row r of the block, columns cg, cg + 16, ...
__global__ __launch_bounds__(256) void k_pw_rows(Geo g, const double2 *__restrict__ Q, const double2 *__restrict__ P,
                                                 const double2 *__restrict__ X, double2 *__restrict__ part,
                                                 const PwState *__restrict__ pst, int k)
{
    const int a = blockIdx.y, w = blockIdx.x, tid = threadIdx.x;
    if (k > 0 && pst[a].done) return;
    __shared__ double2 sx[PW_MAXN];
    __shared__ double2 red[2][16][17];
    __shared__ double2 yb[2][16];
    const int N = g.N, NP = g.NP;
    const int r = tid & 15, cg = tid >> 4, i = w * 16 + r;
    const double2 *q = Q + (size_t)a * g.mat, *p = P + (size_t)a * g.mat;
    for (int j = tid; j < N; j += 256) sx[j] = X[(size_t)a * NP + j];
    __syncthreads();
    double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
    if (i < N)
        for (int j = cg; j < N; j += 16) {
            const double2 xj = sx[j], qv = q[i + (size_t)j * NP], pv = p[i + (size_t)j * NP];
            s1.x += qv.x * xj.x - qv.y * xj.y;
            s1.y += qv.x * xj.y + qv.y * xj.x;
            s2.x += pv.x * xj.x - pv.y * xj.y;
            s2.y += pv.x * xj.y + pv.y * xj.x;
        }
    red[0][cg][r] = s1;
    red[1][cg][r] = s2;
    __syncthreads();
    if (tid < 32) {
        const int m = tid >> 4, rr = tid & 15;
        double2 s = make_double2(0, 0);
        for (int c = 0; c < 16; ++c) {
            s.x += red[m][c][rr].x;
            s.y += red[m][c][rr].y;
        }
        yb[m][rr] = s;
    }
    __syncthreads();
    const double2 y1 = i < N ? yb[0][r] : make_double2(0, 0), y2 = i < N ? yb[1][r] : make_double2(0, 0);
    double2 *out = part + ((size_t)a * gridDim.x + w) * NP;
    const int nt = (N + 15) / 16;   // the same trip count in every lane: the sums below shuffle
    for (int t = 0; t < nt; ++t) {
        const int j = cg + 16 * t;
        double sx_ = 0.0, sy_ = 0.0;
        if (i < N && j < N) {
            const double2 qv = q[i + (size_t)j * NP], pv = p[i + (size_t)j * NP];
            sx_ = qv.x * y1.x + qv.y * y1.y + pv.x * y2.x + pv.y * y2.y;
            sy_ = qv.x * y1.y - qv.y * y1.x + pv.x * y2.y - pv.y * y2.x;
        }
        // sum over the 16 rows: lanes r of one column group (xor offsets below 16)
        for (int o = 8; o > 0; o >>= 1) {
            sx_ += __shfl_xor(sx_, o);
            sy_ += __shfl_xor(sy_, o);
        }
        if (r == 0 && j < N) out[j] = make_double2(sx_, sy_);
    }
}

// one step's second half, one workgroup per angle: z = the sum of the blocks' parts (in block
// order), d = |x^H z|, x = z / |z|, the stopping test; the state records where it stands
__global__ __launch_bounds__(256) void k_pw_fin(Geo g, const double2 *__restrict__ part, int nw,
                                                double2 *__restrict__ X, PwState *__restrict__ pst, int k)
{
    const int a = blockIdx.x, tid = threadIdx.x, N = g.N, NP = g.NP;
    if (k > 0 && pst[a].done) return;
    __shared__ double red[3][256];
    __shared__ double2 z[PW_MAXN];
    double2 *x = X + (size_t)a * NP;
    double dn = 0.0, crx = 0.0, cry = 0.0;
    for (int j = tid; j < N; j += 256) {
        double2 s = make_double2(0, 0);
        for (int w = 0; w < nw; ++w) {
            const double2 v = part[((size_t)a * nw + w) * NP + j];
            s.x += v.x;
            s.y += v.y;
        }
        z[j] = s;
        dn += s.x * s.x + s.y * s.y;
        const double2 xj = x[j];   // conj(x) z
        crx += xj.x * s.x + xj.y * s.y;
        cry += xj.x * s.y - xj.y * s.x;
    }
    red[0][tid] = dn;
    red[1][tid] = crx;
    red[2][tid] = cry;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o)
            for (int c = 0; c < 3; ++c) red[c][tid] += red[c][tid + o];
        __syncthreads();
    }
    const double d = hypot(red[1][0], red[2][0]);
    const double inv = 1.0 / sqrt(red[0][0]);
    __syncthreads();
    double dp = 0.0;
    for (int j = tid; j < N; j += 256) {
        const double2 zz = make_double2(z[j].x * inv, z[j].y * inv), xj = x[j];
        const double cx = zz.x - xj.x, cy = zz.y - xj.y;
        dp += cx * cx + cy * cy;
        x[j] = zz;
    }
    red[0][tid] = dp;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[0][tid] += red[0][tid + o];
        __syncthreads();
    }
    if (tid == 0) {
        dp = sqrt(red[0][0]);
        const double d0 = k == 0 ? -1.0 : pst[a].d0;
        const int stop = dp < 1e-2 || d - d0 < 1e-2 * d;
        PwState st;
        st.d0 = stop ? d0 : d;
        st.d = d;
        st.it = k + 1;
        st.done = stop;
        pst[a] = st;
    }
}

// the remaining steps (from pst's state after `fast` block-split steps; none when fast is 0)
// and the decision ren[a] = d > 1e2
__global__ __launch_bounds__(256) void k_power(Geo g, const double2 *__restrict__ Q, const double2 *__restrict__ P,
                                               double2 *__restrict__ X, int *__restrict__ ren,
                                               int *__restrict__ xid, const PwState *__restrict__ pst, int fast)
{
    extern __shared__ double2 sh[];   // x[NP], y[2 NP], z[NP]
    __shared__ double red[3][256];
    const int a = blockIdx.x, tid = threadIdx.x, nt = blockDim.x, N = g.N, NP = g.NP;
    double2 *x = sh, *y = sh + NP, *z = sh + 3 * NP;
    const double2 *q = Q + (size_t)a * g.mat, *p = P + (size_t)a * g.mat;
    double d = -1.0, d0 = -1.0;
    int it0 = 0;
    if (fast > 0) {
        const PwState st = pst[a];
        d = st.d;
        d0 = st.d0;
        it0 = st.done ? N : st.it;
    }
    if (it0 < N) {
        for (int i = tid; i < N; i += nt) x[i] = X[(size_t)a * NP + i];
        __syncthreads();
    }
    for (int it = it0; it < N; ++it) {
        for (int i = tid; i < N; i += nt) {
            double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
            for (int j = 0; j < N; ++j) {
                const double2 xj = x[j], qv = q[i + (size_t)j * NP], pv = p[i + (size_t)j * NP];
                s1.x += qv.x * xj.x - qv.y * xj.y;
                s1.y += qv.x * xj.y + qv.y * xj.x;
                s2.x += pv.x * xj.x - pv.y * xj.y;
                s2.y += pv.x * xj.y + pv.y * xj.x;
            }
            y[i] = s1;
            y[NP + i] = s2;
        }
        __syncthreads();
        double dn = 0.0, crx = 0.0, cry = 0.0;
        for (int j = tid; j < N; j += nt) {
            double2 s = make_double2(0, 0);
            for (int i = 0; i < N; ++i) {   // conj(q(i,j)) y1(i)
                const double2 qv = q[i + (size_t)j * NP], yv = y[i];
                s.x += qv.x * yv.x + qv.y * yv.y;
                s.y += qv.x * yv.y - qv.y * yv.x;
            }
            for (int i = 0; i < N; ++i) {   // + conj(p(i,j)) y2(i)
                const double2 pv = p[i + (size_t)j * NP], yv = y[NP + i];
                s.x += pv.x * yv.x + pv.y * yv.y;
                s.y += pv.x * yv.y - pv.y * yv.x;
            }
            z[j] = s;
            dn += s.x * s.x + s.y * s.y;
            const double2 xj = x[j];   // conj(x) z
            crx += xj.x * s.x + xj.y * s.y;
            cry += xj.x * s.y - xj.y * s.x;
        }
        red[0][tid] = dn;
        red[1][tid] = crx;
        red[2][tid] = cry;
        __syncthreads();
        for (int o = nt / 2; o > 0; o >>= 1) {
            if (tid < o)
                for (int k = 0; k < 3; ++k) red[k][tid] += red[k][tid + o];
            __syncthreads();
        }
        d = hypot(red[1][0], red[2][0]);
        const double inv = 1.0 / sqrt(red[0][0]);
        __syncthreads();
        double dp = 0.0;
        for (int j = tid; j < N; j += nt) {
            const double2 zz = make_double2(z[j].x * inv, z[j].y * inv);
            const double cx = zz.x - x[j].x, cy = zz.y - x[j].y;
            dp += cx * cx + cy * cy;
            x[j] = zz;
        }
        red[0][tid] = dp;
        __syncthreads();
        for (int o = nt / 2; o > 0; o >>= 1) {
            if (tid < o) red[0][tid] += red[0][tid + o];
            __syncthreads();
        }
        dp = sqrt(red[0][0]);
        __syncthreads();
        if (dp < 1e-2 || d - d0 < 1e-2 * d) break;
        d0 = d;
    }
    if (it0 < N)
        for (int i = tid; i < N; i += nt) X[(size_t)a * NP + i] = x[i];
    if (tid == 0) {
        ren[a] = d > 1e2;
        xid[a] = d > 1e2;
    }
}

// strans (matcomp.f90:117-138): q' = G q + i p, p' = G q - i p, written transposed for
// matAdivB2's solve p' q'^-1 = (q'^T \ p'^T)^T
__global__ void k_gs_strans(Geo g, const double2 *__restrict__ gamma, const double2 *__restrict__ Q,
                            const double2 *__restrict__ P, double2 *__restrict__ Qt, double2 *__restrict__ Pt)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    const size_t et = (size_t)a * g.mat + j + (size_t)i * g.NP;
    if (i >= g.N || j >= g.N) {
        Qt[et] = make_double2(0, 0);
        Pt[et] = make_double2(0, 0);
        return;
    }
    const double2 gi = gamma[(size_t)a * g.N + i], x = Q[e], y = P[e];
    const double2 gx = cmul(gi, x);
    Qt[et] = make_double2(gx.x - y.y, gx.y + y.x);
    Pt[et] = make_double2(gx.x + y.y, gx.y - y.x);
}

// ---------------------------------------------------------------------------
// Bulk reflection of a group of more than 64 beams (blkref, bulk.f90:51-124), batched over
// angles on the kernels above: the slice matrices E = [[c(H), i s(H)], [i H s(H), c(H)]] as
// in the one-workgroup bulk kernel (Horner series after scaling, squarings), their product,
// the unit-layer matrix T = Cm P Cm^-1 with the column phases, and the layer recursion with a
// right division solved by the batched LU on the transposed system.  2n x 2n matrices use a
// second geometry of 2n rows (compact blocks: rows 0..n-1 on top, n..2n-1 below).

// per slice: the largest angle-independent row sum of |H| (off-diagonal |v + vi| and
// |conj(v - vi)|, plus |v + vi| of the (0,0) component on the diagonal); the host adds
// max |gamma^2|.  One workgroup per slice.
__global__ void k_bulkl_norm(int n, int boff, const int *__restrict__ ivg, int nbgm, const VVi *__restrict__ table,
                             int nv, double *__restrict__ out)
{
    __shared__ double red[256];
    const VVi *col = table + (size_t)blockIdx.x * nv;
    const double d0 = hypot(col[0].v.x + col[0].vi.x, col[0].v.y + col[0].vi.y);
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s = d0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const int m = i > j ? ivg[(size_t)(boff + j) * nbgm + i] : ivg[(size_t)(boff + i) * nbgm + j];
            const VVi t = col[m - 1];
            s += i > j ? hypot(t.v.x + t.vi.x, t.v.y + t.vi.y) : hypot(t.v.x - t.vi.x, t.v.y - t.vi.y);
        }
        mx = fmax(mx, s);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// H of one slice for every angle (trmatg.f90:20-29): v + vi + gamma^2 on the diagonal,
// v + vi at iv(i, j) below it, conj(v - vi) at iv(j, i) above it; padding zero
__global__ void k_bulkl_h(Geo g, int boff, const int *__restrict__ ivg, int nbgm, const VVi *__restrict__ col,
                          const double2 *__restrict__ gamma, int nb, double2 *__restrict__ H)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    double2 h = make_double2(0, 0);
    if (i < g.N && j < g.N) {
        if (i == j) {
            const double2 gi = gamma[(size_t)a * nb + boff + i];
            const double2 gg = cmul(gi, gi);
            h = make_double2(col[0].v.x + col[0].vi.x + gg.x, col[0].v.y + col[0].vi.y + gg.y);
        } else if (i > j) {
            const VVi t = col[ivg[(size_t)(boff + j) * nbgm + i] - 1];
            h = make_double2(t.v.x + t.vi.x, t.v.y + t.vi.y);
        } else {
            const VVi t = col[ivg[(size_t)(boff + i) * nbgm + j] - 1];
            h = make_double2(t.v.x - t.vi.x, -(t.v.y - t.vi.y));
        }
    }
    H[e] = h;
}

// E = [[c, i s], [i HS, c]] (2n-row geometry g2) from c, s, HS (n-row geometry g)
__global__ void k_bulkl_e(Geo g, Geo g2, const double2 *__restrict__ Cc, const double2 *__restrict__ Ss,
                          const double2 *__restrict__ HS, double2 *__restrict__ E)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g2.mat * g2.A) return;
    const int a = (int)(e / g2.mat);
    const int r = (int)(e % g2.mat), i = r % g2.NP, j = r / g2.NP;
    const int n = g.N;
    double2 v = make_double2(0, 0);
    if (i < 2 * n && j < 2 * n) {
        const int bi = i >= n, bj = j >= n, ii = i - bi * n, jj = j - bj * n;
        const size_t q = (size_t)a * g.mat + (size_t)jj * g.NP + ii;
        if (bi == bj) v = Cc[q];
        else if (!bi) v = make_double2(-Ss[q].y, Ss[q].x);
        else v = make_double2(-HS[q].y, HS[q].x);
    }
    E[e] = v;
}

// unit-layer matrix T = Cm P Cm^-1, Cm = [[Gamma, I], [Gamma, -I]], then the column phases of
// the layer shift (bulk.f90:83-93); written as its four n x n blocks (geometry g)
__global__ void k_bulkl_t(Geo g, Geo g2, const double2 *__restrict__ P, const double2 *__restrict__ gamma, int nb,
                          int boff, const double2 *__restrict__ phase, double2 *__restrict__ T11,
                          double2 *__restrict__ T12, double2 *__restrict__ T21, double2 *__restrict__ T22)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), ii = r % g.NP, jj = r / g.NP;
    const int n = g.N;
    double2 v[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    if (ii < n && jj < n) {
        const double2 *Pa = P + (size_t)a * g2.mat;
        const double2 *ga = gamma + (size_t)a * nb + boff;
        const double2 ph = phase ? phase[boff + jj] : make_double2(1, 0);
        for (int bi = 0; bi < 2; ++bi)
            for (int bj = 0; bj < 2; ++bj) {
                // rows of Cm P: top = Gamma P1. + P2., bottom = Gamma P1. - P2.
                auto cp = [&](int col) {
                    const double2 p1 = Pa[(size_t)col * g2.NP + ii], p2 = Pa[(size_t)col * g2.NP + n + ii];
                    const double2 gp = cmul(ga[ii], p1);
                    return bi ? make_double2(gp.x - p2.x, gp.y - p2.y) : make_double2(gp.x + p2.x, gp.y + p2.y);
                };
                // Cm^-1 = 1/2 [[Gamma^-1, Gamma^-1], [I, -I]]
                const double2 x = cdiv(cp(jj), ga[jj]), y = cp(n + jj);
                double2 t = bj ? make_double2(0.5 * (x.x - y.x), 0.5 * (x.y - y.y))
                               : make_double2(0.5 * (x.x + y.x), 0.5 * (x.y + y.y));
                if (phase) t = cmul(t, ph);
                v[bi * 2 + bj] = t;
            }
    }
    T11[e] = v[0];
    T12[e] = v[1];
    T21[e] = v[2];
    T22[e] = v[3];
}

// dst(i, j) = src(j, i) per angle (padding zero)
__global__ void k_transpose(Geo g, const double2 *__restrict__ src, double2 *__restrict__ dst)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.mat * g.A) return;
    const int a = (int)(e / g.mat);
    const int r = (int)(e % g.mat), i = r % g.NP, j = r / g.NP;
    dst[e] = (i < g.N && j < g.N) ? src[(size_t)a * g.mat + (size_t)i * g.NP + j] : make_double2(0, 0);
}

// blkref's judgement (bulk.f90:104-117) after layer l, one workgroup per angle: |R_ii|^2
// is kept from layer 20 on; from layer 21 an angle whose largest change is below epsb is
// finished: its R goes to the output (n x n, column-major) and its layer count is set.
// At the last layer (final != 0) every unfinished angle is written out with layer l.
__global__ void k_bulkl_judge(Geo g, const double2 *__restrict__ R, double *__restrict__ r20, int l, double epsb,
                              int final, int *__restrict__ done, int *__restrict__ layers, double2 *__restrict__ out,
                              size_t out_stride)
{
    __shared__ double red[256];
    __shared__ int fin;
    const int a = blockIdx.x, n = g.N;
    if (done[a]) return;
    const double2 *Ra = R + (size_t)a * g.mat;
    double mx = 0.0;
    if (l >= 20)
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double2 d = Ra[(size_t)i * g.NP + i];
            const double r2 = d.x * d.x + d.y * d.y;
            if (l > 20) mx = fmax(mx, fabs(r2 - r20[(size_t)a * g.NP + i]));
            r20[(size_t)a * g.NP + i] = r2;
        }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) fin = (l > 20 && red[0] < epsb) || final;
    __syncthreads();
    if (!fin) return;
    double2 *o = out + (size_t)a * out_stride;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) o[e] = Ra[(size_t)(e / n) * g.NP + (e % n)];
    if (threadIdx.x == 0) {
        done[a] = 1;
        layers[a] = l;
    }
}

}  // namespace

struct Runner {
    Geo g;
    char *d = nullptr;
    size_t bytes = 0;
    double2 *Q0, *Q1, *P, *M, *g2c, *gc, *dinv;
    double *colsum, *diagabs, *coef;
    int *xid, *ren, *piv, *list, *all, *perm;
    int *h = nullptr;           // pinned: [A] flags, [A] list
    hipEvent_t evf = nullptr;
    double2 *Q, *Qn;            // current / next Q buffer
    int forced = -1;            // this slice's renormalisation: -1 device decides, 0 no, 1 all
    // RKN: the previous slice's end kick also ran the next slice's first drift (Q += c_a1 dz P)
    // for the angles that did not renormalise, so the next first stage only serves the others
    bool fused_prev = false;
    bool fuse_drift = true;     // TRHEPD_FUSE_DRIFT=0 keeps the drift in the slice's first stage
    // sliced-integer stage products (mixed precision): digits of Q and of one slice's M0s
    bool mixed = false;
    MixedGeo mg;
    int8_t *qd = nullptr, *md = nullptr;
    int *eq = nullptr, *em = nullptr;
    // per-row maxima of Q gathered by the stage kernel that wrote it (double-buffered);
    // rmax_ok: rmax[rmax_cur] holds the current Q's, so its digit split needs one pass
    unsigned long long *rmax[2] = {};
    int rmax_cur = 0;
    bool rmax_ok = false;
    int bn = 48;                // FP64 stage GEMM tile width (autotuned)
    // Magnus (int_method 1-3): per-angle work matrices (slice matrices, exponential blocks,
    // the carried rp)
    double2 *W[10] = {};
    double2 *X = nullptr;       // rk4_intgs: power-iteration vector per angle
    double2 *xpart = nullptr;   // rk4_intgs: the row blocks' shares of z, [A][NP/16][NP]
    PwState *pst = nullptr;     // rk4_intgs: power-iteration state per angle
    LuArgs lu(const int *lst) const { return LuArgs{g, Q, P, piv, lst, nullptr, dinv, perm}; }
};

void launch_fill_i32(hipStream_t st, int *p, int n, int v)
{
    hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

// per-angle work matrices of the Magnus methods (0: not a Magnus method)
static int magnus_mats(int method)
{
    return method == TRHEPD_M2 ? 4 : method == TRHEPD_M4 ? 6 : method == TRHEPD_M6 ? 8 : 0;
}
// rk4_intgs: slice matrices, stage q, product, stage p, the two derivative sums, Cholesky factor
constexpr int GS_MATS = 9;
// RK4 on the stage GEMM: the product X_r, the stage operands Q_r / P_r, the derivative sums
constexpr int RK4_MATS = 5;
// methods whose products are all per-angle (FP64 in either precision, no autotuned tiles)
static bool dense_method(int method) { return magnus_mats(method) || method == TRHEPD_RK4GS; }

int runner_create(trhepd_ctx *ctx, Batch &b, const Table &t)
{
    if (t.method < TRHEPD_M2 || t.method > TRHEPD_RK4GS)
        return fail(ctx, TRHEPD_E_ARG, "unknown int_method %d", t.method);
    if (b.nb > 600) return fail(ctx, TRHEPD_E_ARG, "N = %d exceeds the device path limit of 600 beams", b.nb);
    Runner *R = new Runner();
    Geo &g = R->g;
    g.A = b.nang;
    g.N = b.nb;
    g.NP = (b.nb + 15) / 16 * 16;
    g.mat = (size_t)g.NP * g.NP;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t mb = g.mat * g.A * sizeof(double2);
    const size_t oQ0 = take(mb), oQ1 = take(mb), oP = take(mb), oM = take(g.mat * sizeof(double2)),
                 og2 = take((size_t)g.A * g.NP * 16), ogc = take((size_t)g.A * g.NP * 16),
                 odi = take((size_t)g.A * (g.NP / 16) * 256 * sizeof(double2)),
                 ocs = take((size_t)g.A * (g.NP / 16) * g.NP * 8), oda = take((size_t)g.A * g.NP * 8),
                 oco = take((size_t)g.A * 8), oxi = take((size_t)g.A * 4), ore = take((size_t)g.A * 4),
                 opv = take((size_t)g.A * g.NP * 4), oli = take((size_t)g.A * 4), oal = take((size_t)g.A * 4),
                 opm = take((size_t)2 * g.A * LU_PERM * 4);
    const int nw = t.method == TRHEPD_RK4GS ? GS_MATS : t.method == TRHEPD_RK4 ? RK4_MATS : magnus_mats(t.method);
    const size_t omg = nw ? take(nw * mb) : 0;
    const bool gs = t.method == TRHEPD_RK4GS;
    const size_t ox = gs ? take((size_t)g.A * g.NP * sizeof(double2)) : 0,
                 oxp = gs ? take((size_t)g.A * (g.NP / 16) * g.NP * sizeof(double2)) : 0,
                 ops = gs ? take((size_t)g.A * sizeof(PwState)) : 0;
    // the Magnus and rk4gs products are all per-angle, and RK4's stage operands are formed
    // outside the stage kernel: they run in FP64 in either precision
    R->mixed = ctx->precision == TRHEPD_PREC_MIXED && !dense_method(t.method) && t.method != TRHEPD_RK4;
    // TRHEPD_FUSE_DRIFT=0 runs every slice's first drift in its own pass (a test control: both
    // forms give the same bits)
    if (const char *fd = std::getenv("TRHEPD_FUSE_DRIFT")) R->fuse_drift = fd[0] != '0';
    size_t oqd = 0, omd = 0, oeq = 0, oem = 0, orm = 0;
    if (R->mixed) {
        // digit buffers sized for the widest padding any candidate tile needs
        const int s = rkn_coef(t.method).s;
        R->mg = mixed_geo(g.A, g.N, g.NP, ctx->mixed_digits);
        size_t md = 0;
        int cbp = 0;
        for (int wc = 2; wc <= 4; ++wc) {
            const MixedGeo c = mixed_geo(g.A, g.N, g.NP, ctx->mixed_digits, wc);
            md = std::max(md, mixed_md_bytes(c));
            cbp = std::max(cbp, c.CBP);
        }
        oqd = take(mixed_qd_bytes(R->mg));
        omd = take(md * s);
        oeq = take((size_t)R->mg.RBP * 16 * 4);
        oem = take((size_t)cbp * 16 * 4 * s);
        orm = take((size_t)2 * R->mg.RBP * 16 * 8);
    }
    int rc = dev_alloc(ctx, (void **)&R->d, off);
    if (rc) {
        delete R;
        return rc;
    }
    R->bytes = off;
    char *d = R->d;
    R->Q0 = (double2 *)(d + oQ0);
    R->Q1 = (double2 *)(d + oQ1);
    R->P = (double2 *)(d + oP);
    R->M = (double2 *)(d + oM);
    R->g2c = (double2 *)(d + og2);
    R->gc = (double2 *)(d + ogc);
    R->dinv = (double2 *)(d + odi);
    R->colsum = (double *)(d + ocs);
    R->diagabs = (double *)(d + oda);
    R->coef = (double *)(d + oco);
    R->xid = (int *)(d + oxi);
    R->ren = (int *)(d + ore);
    R->piv = (int *)(d + opv);
    R->perm = (int *)(d + opm);
    R->list = (int *)(d + oli);
    R->all = (int *)(d + oal);
    for (int w = 0; w < nw; ++w) R->W[w] = (double2 *)(d + omg + w * mb);
    if (gs) {
        R->X = (double2 *)(d + ox);
        R->xpart = (double2 *)(d + oxp);
        R->pst = (PwState *)(d + ops);
    }
    if (R->mixed) {
        R->qd = (int8_t *)(d + oqd);
        R->md = (int8_t *)(d + omd);
        R->eq = (int *)(d + oeq);
        R->em = (int *)(d + oem);
        R->rmax[0] = (unsigned long long *)(d + orm);
        R->rmax[1] = R->rmax[0] + (size_t)R->mg.RBP * 16;
    }
    hipError_t e = hipHostMalloc((void **)&R->h, (size_t)g.A * 8 + 64, hipHostMallocDefault);
    if (!e) e = hipEventCreateWithFlags(&R->evf, hipEventDisableTiming);
    if (e) {
        dev_free(ctx, R->d, R->bytes);
        if (R->h) (void)hipHostFree(R->h);
        delete R;
        return hip_check(ctx, e, "runner host buffers");
    }
    R->Q = R->Q0;
    R->Qn = R->Q1;
    std::vector<int> all(g.A);
    for (int i = 0; i < g.A; ++i) all[i] = i;
    TRY_HIP(ctx, hipMemsetAsync(R->d, 0, R->bytes, b.stream));
    TRY_HIP(ctx, hipMemcpyAsync(R->all, all.data(), g.A * 4, hipMemcpyHostToDevice, b.stream));
    TRY_HIP(ctx, hipStreamSynchronize(b.stream));
    b.run = R;
    return dense_method(t.method) ? TRHEPD_OK : autotune(ctx, b, t);
}

// ---------------------------------------------------------------------------
// state access for the adaptive slice control (adaptive.hip): P of every angle (after a
// renormalisation Q = I, so P is the whole state), and the step-doubling difference

// per angle: max over the N x N part of max(|re|, |im|) of P - ref, over that of P
__global__ void k_pdiff(Geo g, const double2 *__restrict__ P, const double2 *__restrict__ ref, double *__restrict__ out)
{
    __shared__ double sd[256], sp[256];
    const int a = blockIdx.x, tid = threadIdx.x;
    double dm = 0.0, pm = 0.0;
    for (int e = tid; e < g.N * g.N; e += blockDim.x) {
        const size_t idx = (size_t)a * g.mat + (e % g.N) + (size_t)(e / g.N) * g.NP;
        const double2 p = P[idx], r = ref[idx];
        dm = fmax(dm, fmax(fabs(p.x - r.x), fabs(p.y - r.y)));
        pm = fmax(pm, fmax(fabs(p.x), fabs(p.y)));
    }
    sd[tid] = dm;
    sp[tid] = pm;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) {
            sd[tid] = fmax(sd[tid], sd[tid + o]);
            sp[tid] = fmax(sp[tid], sp[tid + o]);
        }
        __syncthreads();
    }
    if (tid == 0) out[a] = sp[0] > 0.0 ? sd[0] / sp[0] : sd[0];
}

size_t runner_state_bytes(const Batch &b) { return b.run ? b.run->g.mat * b.run->g.A * sizeof(double2) : 0; }

int runner_save_p(trhepd_ctx *ctx, Batch &b, double2 *dst)
{
    TRY_HIP(ctx, hipMemcpyAsync(dst, b.run->P, runner_state_bytes(b), hipMemcpyDeviceToDevice, b.stream));
    return TRHEPD_OK;
}

int runner_load_p(trhepd_ctx *ctx, Batch &b, const double2 *src)
{
    TRY_HIP(ctx, hipMemcpyAsync(b.run->P, src, runner_state_bytes(b), hipMemcpyDeviceToDevice, b.stream));
    return TRHEPD_OK;
}

int runner_p_difference(trhepd_ctx *ctx, Batch &b, const double2 *ref, double *work, double *out)
{
    const Runner &R = *b.run;
    hipLaunchKernelGGL(k_pdiff, dim3(R.g.A), dim3(256), 0, b.stream, R.g, (const double2 *)R.P, ref, work);
    std::vector<double> h(R.g.A);
    TRY_HIP(ctx, hipMemcpyAsync(h.data(), work, (size_t)R.g.A * 8, hipMemcpyDeviceToHost, b.stream));
    TRY_HIP(ctx, hipStreamSynchronize(b.stream));
    double m = 0.0;
    for (double v : h) m = (v > m || v != v) ? v : m;   // a NaN propagates
    *out = m;
    return TRHEPD_OK;
}

void runner_destroy(trhepd_ctx *ctx, Batch &b)
{
    Runner *R = b.run;
    if (!R) return;
    dev_free(ctx, R->d, R->bytes);
    (void)hipHostFree(R->h);
    (void)hipEventDestroy(R->evf);
    delete R;
    b.run = nullptr;
}

static int lu_on(trhepd_ctx *ctx, hipStream_t st, LuArgs L, int npts)
{
    if (npts == 0) return TRHEPD_OK;
    const int N = L.g.N;
    static_assert(LU_RPT == 3 && LU_RPT * 256 >= 600 && BK_WAVES * BK_SLOTS * 16 >= 600, "LU panel rows per workgroup or back-substitution slots below the device path's beam limit");
    hipLaunchKernelGGL(k_lu_panel, dim3(npts), dim3(256), 0, st, L, 0);
    for (int k0 = 0; k0 < N; k0 += 16) {
        const int kw = std::min(16, N - k0);
        const int rest = 2 * N - (k0 + kw) - 16;   // columns right of tile 0
        const int ntile = 1 + (rest > 0 ? (rest + LU_TILE - 1) / LU_TILE : 0);
        hipLaunchKernelGGL(k_lu_step, dim3(ntile, npts), dim3(256), 0, st, L, k0);   // also panel k0 + 16
    }
    const int nblk = (N + 15) / 16;
    hipLaunchKernelGGL(k_lu_dinv, dim3(nblk, npts), dim3(64), 0, st, L, L.dinv);
    hipLaunchKernelGGL(k_lu_back, dim3(nblk, npts), dim3(64 * BK_WAVES), 0, st, L, (const double2 *)L.dinv);
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

static void qreset(Batch &b, const int *list, int npts, double beta)
{
    const Runner &R = *b.run;
    hipLaunchKernelGGL(k_qreset, dim3((unsigned)((R.g.mat + 255) / 256), npts), dim3(256), 0, b.stream, R.g, list,
                       beta, R.Q, (const double2 *)R.P);
}

// index of rp among the Magnus work matrices
static int magnus_rp(int method) { return method == TRHEPD_M2 ? 3 : method == TRHEPD_M4 ? 5 : 4; }

static void bmm(Batch &b, const double2 *A, const double2 *B, double2 *C, double alpha, double beta,
                const int *list = nullptr, int npts = 0, int conjA = 0)
{
    const Geo &g = b.run->g;
    const int nbl = g.NP / 16;
    const BmmArgs S{g, A, B, C, alpha, beta, 0.0, list, conjA};
    hipLaunchKernelGGL(k_bmm, dim3((unsigned)((nbl * nbl + 3) / 4), list ? npts : g.A), dim3(256), 0, b.stream, S);
}

// out = sum of c x over the terms, + ident on the diagonal
static void lin(Batch &b, double2 *out, std::initializer_list<std::pair<double, const double2 *>> terms,
                double ident = 0.0)
{
    const Geo &g = b.run->g;
    LinArgs L{};
    L.g = g;
    for (const auto &t : terms) {
        L.c[L.n] = t.first;
        L.x[L.n] = t.second;
        ++L.n;
    }
    L.out = out;
    L.ident = ident;
    hipLaunchKernelGGL(k_lin, dim3((unsigned)((g.mat * g.A + 255) / 256)), dim3(256), 0, b.stream, L);
}

// several lin() calls in order, in one pass (k_lin_set)
using LinTerms = std::initializer_list<std::pair<double, const double2 *>>;
static void lins(Batch &b, std::initializer_list<std::pair<double2 *, LinTerms>> outs)
{
    const Geo &g = b.run->g;
    LinSet S{};
    for (const auto &o : outs) {
        LinArgs &L = S.o[S.nout++];
        L.g = g;
        for (const auto &t : o.second) {
            L.c[L.n] = t.first;
            L.x[L.n] = t.second;
            ++L.n;
        }
        L.out = o.first;
    }
    hipLaunchKernelGGL(k_lin_set, dim3((unsigned)((g.mat * g.A + 255) / 256)), dim3(256), 0, b.stream, g, S);
}

static void build_m(Batch &b, const VVi *col, double2 *M)
{
    const Runner &R = *b.run;
    hipLaunchKernelGGL(k_build_m, dim3((unsigned)((R.g.mat * R.g.A + 255) / 256)), dim3(256), 0, b.stream, R.g,
                       (const int *)b.gmap, col, (const double2 *)R.g2c, M);
}

// hyperbolic_exp4r (matcomp.f90:388-405): r = a^2, q = h I + h^3/6 a + h^5/120 a^2,
// r = h/2 a - h^3/24 a^2
static void hyp_exp4(Batch &b, double h, const double2 *a, double2 *q, double2 *r)
{
    const double h3 = h * h * h, h5 = h3 * h * h;
    bmm(b, a, a, r, 1.0, 0.0);
    lin(b, q, {{h3 / 6, a}, {h5 / 120, r}}, h);
    lin(b, r, {{h / 2, a}, {-(h3 / 24), r}});
}

// hyperbolic_exp6r (matcomp.f90:407-428), t a scratch matrix
static void hyp_exp6(Batch &b, double h, const double2 *a, double2 *q, double2 *r, double2 *t)
{
    const double cq0 = 1.0 / 6, cq1 = 1.0 / 120, cq2 = 1.0 / 5040, cr0 = 0.5, cr1 = -1.0 / 24, cr2 = 1.0 / 240;
    const double h3 = h * h * h, h5 = h3 * h * h, h7 = h5 * h * h;
    bmm(b, a, a, r, 1.0, 0.0);
    bmm(b, r, a, t, 1.0, 0.0);
    lin(b, q, {{h3 * cq0, a}, {h5 * cq1, r}, {h7 * cq2, t}}, h);
    lin(b, r, {{h * cr0, a}, {h3 * cr1, r}, {h5 * cr2, t}});
}

// One slice of m2/m4/m6_int_spa.  After a renormalisation Q = I exactly, so the
// reference's two branches (P += r + rp / Q = P q + I when Q = I, else the products) are
// the same products here: I X is exact on the matrix cores.
static void magnus_slice(Batch &b, const Table &t, int l)
{
    Runner &R = *b.run;
    double2 **W = R.W;
    const double dz = t.dz;
    double2 *Q = R.Q, *P = R.P;
    if (t.method == TRHEPD_M2) {   // surf_magnus.f90:114-178
        double2 *m = W[0], *r = W[1], *q = W[2], *rp = W[3];
        build_m(b, t.d + (size_t)(l - 1) * t.nv, m);
        hyp_exp4(b, dz, m, q, r);
        lin(b, rp, {{1.0, rp}, {1.0, r}});
        bmm(b, Q, rp, P, 1.0, 1.0);
        lin(b, rp, {{1.0, r}});
        bmm(b, P, q, Q, 1.0, 1.0);
        return;
    }
    const VVi *c0 = t.d + (size_t)(3 * (l - 1)) * t.nv;
    if (t.method == TRHEPD_M4) {   // surf_magnus.f90:180-251
        const double a0 = std::sqrt(15.0) / 36, a1 = 5.0 / 36;
        double2 *m1 = W[0], *m2 = W[1], *m3 = W[2], *r = W[3], *q = W[4], *rp = W[5];
        build_m(b, c0, m1);
        build_m(b, c0 + t.nv, m2);
        build_m(b, c0 + 2 * t.nv, m3);
        lin(b, q, {{1.0, m3}, {-1.0, m1}});
        lin(b, m3, {{1.0, m1}, {-2.0, m2}, {1.0, m3}});
        lin(b, m1, {{1.0, q}});
        hyp_exp4(b, dz, m2, q, r);
        lin(b, rp, {{1.0, rp}, {1.0, r}, {-(a0 * dz), m1}, {a1 * dz, m3}});
        bmm(b, Q, rp, P, 1.0, 1.0);
        bmm(b, P, q, Q, 1.0, 1.0);
        lin(b, rp, {{1.0, r}, {a0 * dz, m1}, {a1 * dz, m3}});
        return;
    }
    // m6, surf_magnus.f90:253-339
    const double a0 = std::sqrt(15.0) / 180, a1 = 1.0 / 18, a2 = 1.0 / 12960;
    const double d0 = 4.0 / (3.0 * std::sqrt(15.0)), d1 = 1.0 / 6;
    double2 *m1 = W[0], *m2 = W[1], *m3 = W[2], *tt = W[3], *rp = W[4], *r = W[5], *q = W[6], *tw = W[7];
    build_m(b, c0, m1);
    build_m(b, c0 + t.nv, m2);
    build_m(b, c0 + 2 * t.nv, m3);
    lin(b, r, {{1.0, m3}, {-1.0, m1}});
    lin(b, q, {{1.0, m1}, {-2.0, m2}, {1.0, m3}});
    lin(b, m1, {{1.0, r}});
    lin(b, m2, {{1.0, m2}, {d1, q}});
    bmm(b, m1, m1, m3, a2 * dz * dz * dz, 0.0);
    lin(b, m3, {{1.0, m3}, {a1 * dz, q}});
    lin(b, tt, {{1.0, m2}, {-d0, m1}});
    hyp_exp6(b, dz / 2, tt, q, r, tw);
    lin(b, rp, {{1.0, rp}, {1.0, r}, {-(a0 * dz), m1}, {1.0, m3}});
    bmm(b, Q, rp, P, 1.0, 1.0);
    bmm(b, P, q, Q, 1.0, 1.0);
    lin(b, rp, {{1.0, r}});
    lin(b, tt, {{1.0, m2}, {d0, m1}});
    hyp_exp6(b, dz / 2, tt, q, r, tw);
    lin(b, rp, {{1.0, rp}, {1.0, r}});
    bmm(b, Q, rp, P, 1.0, 1.0);
    bmm(b, P, q, Q, 1.0, 1.0);
    lin(b, rp, {{1.0, r}, {dz * a0, m1}, {1.0, m3}});
}

// bcgs (matgs.f90:4-37) of the listed angles: C = q^H q + p^H p, C = U^H U, [q; p] <- [q; p] U^-1
static void cqr(Batch &b, const int *list, int npts, bool count)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    double2 *C = R.W[GS_MATS - 1];
    bmm(b, R.Q, R.Q, C, 1.0, 0.0, list, npts, 1);
    bmm(b, R.P, R.P, C, 1.0, 1.0, list, npts, 1);
    const int nblk = g.NP / 16;
    for (int kb = 0; kb < nblk; ++kb) {
        const int rest = g.N - 16 * (kb + 1);   // columns right of the diagonal block
        const int ntile = rest > 0 ? (rest + 63) / 64 : 1;
        hipLaunchKernelGGL(k_chol_panel, dim3(ntile, npts), dim3(256), 0, b.stream, g, list, C, kb, R.dinv, b.cnt,
                           count ? 1 : 0);
        const int m = nblk - kb - 1;
        if (m > 0)
            hipLaunchKernelGGL(k_chol_update, dim3(m, (m + 3) / 4, npts), dim3(256), 0, b.stream, g, list, C, kb);
    }
    hipLaunchKernelGGL(k_chol_diag, dim3(nblk, npts), dim3(256), 0, b.stream, g, list, C, (const double2 *)R.dinv);
    // [q; p] <- [q; p] R^-1: inverses of R's diagonal blocks, then the blocked solves
    LuArgs L{g, C, nullptr, nullptr, list, nullptr, R.dinv, nullptr};
    hipLaunchKernelGGL(k_lu_dinv, dim3(nblk, npts), dim3(64), 0, b.stream, L, R.dinv);
    const dim3 tg((unsigned)((nblk + 3) / 4), npts);
    hipLaunchKernelGGL(k_trsm_rb, tg, dim3(256), 0, b.stream, g, list, (const double2 *)C, (const double2 *)R.dinv, R.Q);
    hipLaunchKernelGGL(k_trsm_rb, tg, dim3(256), 0, b.stream, g, list, (const double2 *)C, (const double2 *)R.dinv, R.P);
}

static void build_mf(Batch &b, const VVi *col, double2 *M)
{
    const Runner &R = *b.run;
    hipLaunchKernelGGL(k_build_mf, dim3((unsigned)((R.g.mat * R.g.A + 255) / 256)), dim3(256), 0, b.stream, R.g,
                       (const int *)b.gmap, col, (const double2 *)R.g2c, M);
}

// one slice of rk4_intgs (surf_rk.f90:114-143)
static void gs_slice(Batch &b, const Table &t, int l)
{
    Runner &R = *b.run;
    double2 **W = R.W;
    double2 *m0 = W[0], *m1 = W[1], *m2 = W[2], *q2 = W[3], *sm = W[4], *ps = W[5], *SQ = W[6], *SP = W[7];
    double2 *Q = R.Q, *P = R.P;
    const double dz = t.dz, h = 0.5 * dz;
    const VVi *c0 = t.d + (size_t)(2 * (l - 1)) * t.nv;
    build_mf(b, c0, m0);
    build_mf(b, c0 + t.nv, m1);
    build_mf(b, c0 + 2 * t.nv, m2);
    bmm(b, m0, Q, sm, 1.0, 0.0);   // s1 = m0 q, r1 = p
    lins(b, {{SQ, {{1.0, P}}}, {SP, {{1.0, sm}}}, {q2, {{1.0, Q}, {h, P}}}, {ps, {{1.0, P}, {h, sm}}}});   // r2
    bmm(b, m1, q2, sm, 1.0, 0.0);      // s2
    lins(b, {{SQ, {{1.0, SQ}, {2.0, ps}}}, {SP, {{1.0, SP}, {2.0, sm}}}, {q2, {{1.0, Q}, {h, ps}}},
             {ps, {{1.0, P}, {h, sm}}}});   // r3
    bmm(b, m1, q2, sm, 1.0, 0.0);      // s3
    lins(b, {{SQ, {{1.0, SQ}, {2.0, ps}}}, {SP, {{1.0, SP}, {2.0, sm}}}, {q2, {{1.0, Q}, {dz, ps}}},
             {ps, {{1.0, P}, {dz, sm}}}});   // r4
    bmm(b, m2, q2, sm, 1.0, 0.0);      // s4
    lins(b, {{SQ, {{1.0, SQ}, {1.0, ps}}}, {SP, {{1.0, SP}, {1.0, sm}}}, {Q, {{1.0, Q}, {dz / 6, SQ}}},
             {P, {{1.0, P}, {dz / 6, SP}}}});
}

int runner_begin(trhepd_ctx *ctx, Batch &b, const Table &t)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    const size_t ntot = g.mat * g.A;
    const dim3 eb(256), eg((unsigned)((ntot + 255) / 256));
    hipLaunchKernelGGL(k_prep, dim3((g.NP + 255) / 256, g.A), dim3(256), 0, b.stream, g, (const double2 *)b.gamma,
                       R.g2c, R.gc);
    R.Q = R.Q0;
    R.Qn = R.Q1;
    R.rmax_ok = false;
    R.fused_prev = false;
    if (t.method == TRHEPD_RK4GS) {   // invstrans, bcgs, x of power_est(-100) (surf_rk.f90:109-111)
        hipLaunchKernelGGL(k_gs_setup, eg, eb, 0, b.stream, g, (const double2 *)b.gamma, (const double2 *)b.fout, R.Q,
                           R.P, R.X);
        cqr(b, R.all, g.A, false);
        TRY_HIP(ctx, hipGetLastError());
        return TRHEPD_OK;
    }
    hipLaunchKernelGGL(k_inv_setup, eg, eb, 0, b.stream, g, (const double2 *)b.fout, R.Q, R.P);
    hipLaunchKernelGGL(k_zero2, dim3((unsigned)((g.mat + 255) / 256)), dim3(256), 0, b.stream, R.M, g.mat);
    LuArgs L = R.lu(R.all);
    L.flag = b.cnt + 1;
    int rc = lu_on(ctx, b.stream, L, g.A);
    if (rc) return rc;
    hipLaunchKernelGGL(k_inv_scale, eg, eb, 0, b.stream, g, (const double2 *)R.gc, R.P);
    hipLaunchKernelGGL(k_fill_i32, dim3((g.A + 255) / 256), dim3(256), 0, b.stream, R.xid, g.A, 1);
    // leapfrog starts from Q = I + dz/2 P (surf_magnus.f90:63-67), the Magnus methods and RK4
    // from Q = I (surf_rk.f90:46-47)
    if (t.method == TRHEPD_LF || t.method == TRHEPD_RK4 || magnus_mats(t.method))
        qreset(b, R.all, g.A, t.method == TRHEPD_LF ? 0.5 * t.dz : 0.0);
    // the Magnus methods carry rp from slice to slice, starting at zero
    if (magnus_mats(t.method))
        hipLaunchKernelGGL(k_zero2, dim3((unsigned)((g.mat * g.A + 255) / 256)), dim3(256), 0, b.stream,
                           R.W[magnus_rp(t.method)], g.mat * g.A);
    return TRHEPD_OK;
}

// the FP64 stage GEMM at the batch's tuned tile width
static void gemm_launch(trhepd_ctx *ctx, Batch &b, StageArgs &S)
{
    const Runner &R = *b.run;
    const Geo &g = R.g;
    const int rowt = (g.A * g.NP + 127) / 128;
    S.ncolt = (g.NP + R.bn - 1) / R.bn;
    S.ntiles = rowt * S.ncolt;
    S.tiles_per_xcd = (S.ntiles + 7) / 8;
    const dim3 grid(8 * S.tiles_per_xcd), blk(256);
    prof_mark(ctx, b.stream);
    if (R.bn == 32) hipLaunchKernelGGL(k_stage_gemm<32>, grid, blk, 0, b.stream, S);
    else if (R.bn == 64) hipLaunchKernelGGL(k_stage_gemm<64>, grid, blk, 0, b.stream, S);
    else hipLaunchKernelGGL(k_stage_gemm<48>, grid, blk, 0, b.stream, S);
    prof_mark(ctx, b.stream);
}

// stage product number `st` of the slice (its digits, on the mixed path, are md[st])
static void stage_launch(trhepd_ctx *ctx, Batch &b, const Table &t, int mode, const VVi *col, double alpha,
                         double beta, const double *coef, int st, const int *xid = nullptr)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    if (R.mixed) {
        // one pass over Q when the previous stage gathered its row maxima; either way the
        // split clears the other buffer, which this stage's kernel fills for the next split
        unsigned long long *nxt = R.rmax[R.rmax_cur ^ 1];
        launch_qsplit(b.stream, R.mg, R.Q, R.qd, R.eq, R.rmax_ok ? R.rmax[R.rmax_cur] : nullptr, nxt);
        MixedStage M;
        M.g = R.mg;
        M.qd = R.qd;
        M.eq = R.eq;
        M.md = R.md + (size_t)st * mixed_md_bytes(R.mg);
        M.em = R.em + (size_t)st * R.mg.CBP * 16;
        M.Q = R.Q;
        M.Qn = R.Qn;
        M.P = R.P;
        M.g2c = R.g2c;
        M.col = col;
        M.alpha = alpha;
        M.beta = beta;
        M.coef = coef;
        M.xid = xid;
        M.mode = mode;
        M.colsum = R.colsum;
        M.diagabs = R.diagabs;
        M.rmax = mode != 2 ? nxt : nullptr;
        prof_mark(ctx, b.stream);
        launch_stage_i8(b.stream, M);
        prof_mark(ctx, b.stream);
        if (mode != 2) {
            std::swap(R.Q, R.Qn);
            R.rmax_cur ^= 1;
            R.rmax_ok = true;
        }
        return;
    }
    const int ggrid = (int)((g.mat + 255) / 256);
    hipLaunchKernelGGL(k_gather, dim3(ggrid), dim3(256), 0, b.stream, g, b.gmap, col, R.M);
    StageArgs S;
    S.g = g;
    S.Q = R.Q;
    S.Qn = R.Qn;
    S.P = R.P;
    S.M0 = R.M;
    S.g2c = R.g2c;
    S.col = col;
    S.alpha = alpha;
    S.beta = beta;
    S.coef = coef;
    S.xid = xid;
    S.mode = mode;
    S.colsum = R.colsum;
    S.diagabs = R.diagabs;
    gemm_launch(ctx, b, S);
    if (mode != 2) std::swap(R.Q, R.Qn);
}

// rk4_int (surf_rk.f90:48-78) on the batched path.  Each stage product X_r = Q_r M(z_r) is the
// stage GEMM's accumulate mode (dst += 1 (Q M0 + Q diag(d_a)), the launch the RKN end kick
// uses) into a zeroed buffer; after a renormalisation Q = I and X_1 = M exactly, as the
// reference's x_is_id branch sets it.  The combinations follow the reference's order:
// t2 = t1 + c k_r with c = dz/2, dz/2, dz, then t1 += (dz/6) (((k1 + 2 k2) + 2 k3) + k4).
static void rk4_product(trhepd_ctx *ctx, Batch &b, const VVi *col, const double2 *Qop, double2 *X)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    hipLaunchKernelGGL(k_zero2, dim3((unsigned)((g.mat * g.A + 255) / 256)), dim3(256), 0, b.stream, X,
                       g.mat * g.A);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((g.mat + 255) / 256)), dim3(256), 0, b.stream, g, b.gmap, col, R.M);
    StageArgs S;
    S.g = g;
    S.Q = Qop;
    S.Qn = X;   // not written in mode 2
    S.P = X;
    S.M0 = R.M;
    S.g2c = R.g2c;
    S.col = col;
    S.alpha = 1.0;
    S.beta = 0.0;
    S.coef = nullptr;
    S.xid = nullptr;
    S.mode = 2;
    S.colsum = R.colsum;
    S.diagabs = R.diagabs;
    gemm_launch(ctx, b, S);
}

static void rk4_slice(trhepd_ctx *ctx, Batch &b, const Table &t, int l)
{
    Runner &R = *b.run;
    double2 *X = R.W[0], *Qs = R.W[1], *Ps = R.W[2], *SQ = R.W[3], *SP = R.W[4];
    const double h = 0.5 * t.dz, dz = t.dz;
    // matrices at z = dz (l - 1), z + dz/2 and z + dz (rk4_int_init, surf_rk.f90:17-24)
    const VVi *c0 = t.d + (size_t)(2 * (l - 1)) * t.nv, *c1 = c0 + t.nv, *c2 = c1 + t.nv;
    rk4_product(ctx, b, c0, R.Q, X);   // k1 = [P, Q m0]
    lins(b, {{SQ, {{1.0, R.P}}}, {SP, {{1.0, X}}}, {Qs, {{1.0, R.Q}, {h, R.P}}}, {Ps, {{1.0, R.P}, {h, X}}}});
    rk4_product(ctx, b, c1, Qs, X);    // k2 = [P2, Q2 m1]
    lins(b, {{SQ, {{1.0, SQ}, {2.0, Ps}}}, {SP, {{1.0, SP}, {2.0, X}}}, {Qs, {{1.0, R.Q}, {h, Ps}}},
             {Ps, {{1.0, R.P}, {h, X}}}});
    rk4_product(ctx, b, c1, Qs, X);    // k3 = [P3, Q3 m1]
    lins(b, {{SQ, {{1.0, SQ}, {2.0, Ps}}}, {SP, {{1.0, SP}, {2.0, X}}}, {Qs, {{1.0, R.Q}, {dz, Ps}}},
             {Ps, {{1.0, R.P}, {dz, X}}}});
    rk4_product(ctx, b, c2, Qs, X);    // k4 = [P4, Q4 m2]
    lins(b, {{SQ, {{1.0, SQ}, {1.0, Ps}}}, {SP, {{1.0, SP}, {1.0, X}}}, {R.Q, {{1.0, R.Q}, {dz / 6, SQ}}},
             {R.P, {{1.0, R.P}, {dz / 6, SP}}}});
}

// Stage-product tile variants: FP64 tile widths 32 / 48 / 64 columns; mixed (column
// blocks, row blocks, Q prefetch depth) per workgroup (2..4, 8 or 4, 1 or 2).
static int n_variants(const Runner &R) { return R.mixed ? 12 : 3; }

static bool apply_variant(Runner &R, int v)
{
    if (!R.mixed) {
        R.bn = v == 0 ? 32 : v == 1 ? 48 : 64;
        return true;
    }
    const int pf = 1 + v / 6, u = v % 6, wc = 2 + u / 2, wr = (u & 1) ? 4 : 8;
    const MixedGeo c = mixed_geo(R.g.A, R.g.N, R.g.NP, R.mg.ND, wc, wr, pf);
    if (c.WR != wr || c.PF != pf) return false;   // that tile does not fit the registers
    R.mg = c;
    return true;
}

// Autotuner: times every variant of the stage product on this batch's own buffers (before
// the first run, so the values are irrelevant and are overwritten by the boundary
// transform) and keeps the fastest; the choice is cached per (precision, N, angles, digits).
int autotune(trhepd_ctx *ctx, Batch &b, const Table &t)
{
    Runner &R = *b.run;
    const std::array<int, 4> key{R.mixed ? 1 : 0, R.g.N, R.g.A, R.mixed ? R.mg.ND : 0};
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) {
        apply_variant(R, it->second);
        return TRHEPD_OK;
    }
    // TRHEPD_STAGE_VARIANT pins one int8 tile variant of the mixed path instead of timing
    // them (a test control: every variant must give the same bits)
    if (const char *pin = R.mixed ? std::getenv("TRHEPD_STAGE_VARIANT") : nullptr) {
        const int v = std::atoi(pin);
        if (v < 0 || v >= n_variants(R) || !apply_variant(R, v))
            return fail(ctx, TRHEPD_E_ARG, "TRHEPD_STAGE_VARIANT=%s does not name a tile of this batch", pin);
        ctx->tuned[key] = v;
        return TRHEPD_OK;
    }
    const bool prof = ctx->prof.on;
    ctx->prof.on = false;
    // TRHEPD_AUTOTUNE_LOG=1 prints every variant's time per launch to stderr
    const char *log = std::getenv("TRHEPD_AUTOTUNE_LOG");
    const bool verbose = log && log[0] == '1';
    int best = -1;
    float best_ms = 0;
    for (int v = 0; v < n_variants(R); ++v) {
        if (!apply_variant(R, v)) continue;
        stage_launch(ctx, b, t, 2, t.d, 0.0, 0.0, R.coef, 0);   // warm-up (code load)
        TRY_HIP(ctx, hipEventRecord(b.ev0, b.stream));
        for (int k = 0; k < 3; ++k) stage_launch(ctx, b, t, 2, t.d, 0.0, 0.0, R.coef, 0);
        TRY_HIP(ctx, hipEventRecord(b.ev1, b.stream));
        TRY_HIP(ctx, hipEventSynchronize(b.ev1));
        float ms = 0;
        TRY_HIP(ctx, hipEventElapsedTime(&ms, b.ev0, b.ev1));
        if (verbose) {
            if (R.mixed)
                fprintf(stderr, "autotune N=%d angles=%d digits=%d variant %d (wc %d wr %d pf %d): %.1f us\n", R.g.N,
                        R.g.A, R.mg.ND, v, R.mg.WC, R.mg.WR, R.mg.PF, ms * 1000.0f / 3);
            else
                fprintf(stderr, "autotune N=%d angles=%d fp64 variant %d (bn %d): %.1f us\n", R.g.N, R.g.A, v, R.bn,
                        ms * 1000.0f / 3);
        }
        if (best < 0 || ms < best_ms) {
            best = v;
            best_ms = ms;
        }
    }
    ctx->prof.on = prof;
    TRY_HIP(ctx, hipGetLastError());
    apply_variant(R, best);
    ctx->tuned[key] = best;
    return TRHEPD_OK;
}

int runner_slice_enqueue(trhepd_ctx *ctx, Batch &b, const Table &t, int l)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    const RknCoef co = rkn_coef(t.method);
    // after an end kick that carried the drift, Q changes outside the stage kernels only for
    // the renormalised angles; with none of them, the maxima the end kick gathered still hold
    const bool rkn = t.method == TRHEPD_SP4 || t.method == TRHEPD_SP6;
    const bool skip_first = rkn && R.fused_prev && R.forced == 0;
    if (!skip_first) R.rmax_ok = false;
    const int s = co.s;
    const size_t ntot = g.mat * g.A;
    const dim3 eb(256), eg((unsigned)((ntot + 255) / 256));
    if (t.method == TRHEPD_LF) {
        // Stormer-Verlet (lf_int_spa, surf_magnus.f90:68-99): kick with the slice-midpoint
        // matrix, then a full drift (a half drift on the last slice); after a
        // renormalisation Q = I and the product reduces to the reference's P += dz M
        const VVi *col = t.d + (size_t)(l - 1) * t.nv;
        if (R.mixed) launch_msplit(b.stream, R.mg, b.gmap, col, t.nv, 1, R.md, R.em);
        stage_launch(ctx, b, t, 1, col, t.dz, l == t.ns ? 0.5 * t.dz : t.dz, nullptr, 0);
    } else if (t.method == TRHEPD_RK4GS) {
        gs_slice(b, t, l);
        // renormalisation: the power estimate (device) or every K-th slice
        const int period = batch_period(ctx, b, t);
        R.forced = period > 0 ? (l % period == 0 ? 1 : 0) : -1;
        if (R.forced < 0) {
            const int nw = g.NP / 16, fast = std::min(PW_FAST, g.N);
            for (int k = 0; k < fast; ++k) {
                hipLaunchKernelGGL(k_pw_rows, dim3(nw, g.A), dim3(256), 0, b.stream, g, (const double2 *)R.Q,
                                   (const double2 *)R.P, (const double2 *)R.X, R.xpart, (const PwState *)R.pst, k);
                hipLaunchKernelGGL(k_pw_fin, dim3(g.A), dim3(256), 0, b.stream, g, (const double2 *)R.xpart, nw, R.X,
                                   R.pst, k);
            }
            hipLaunchKernelGGL(k_power, dim3(g.A), dim3(256), (size_t)4 * g.NP * sizeof(double2), b.stream, g,
                               (const double2 *)R.Q, (const double2 *)R.P, R.X, R.ren, R.xid,
                               (const PwState *)R.pst, fast);
            TRY_HIP(ctx, hipMemcpyAsync(R.h, R.ren, g.A * 4, hipMemcpyDeviceToHost, b.stream));
            TRY_HIP(ctx, hipEventRecord(R.evf, b.stream));
        }
        TRY_HIP(ctx, hipGetLastError());
        return TRHEPD_OK;
    } else if (t.method == TRHEPD_RK4) {
        rk4_slice(ctx, b, t, l);
        const int nblk = g.NP / 16;
        hipLaunchKernelGGL(k_gersh, dim3((unsigned)(((size_t)g.A * nblk * g.NP + 255) / 256)), dim3(256), 0,
                           b.stream, g, (const double2 *)R.Q, R.colsum, R.diagabs);
    } else if (magnus_mats(t.method)) {
        magnus_slice(b, t, l);
        const int nblk = g.NP / 16;
        hipLaunchKernelGGL(k_gersh, dim3((unsigned)(((size_t)g.A * nblk * g.NP + 255) / 256)), dim3(256), 0,
                           b.stream, g, (const double2 *)R.Q, R.colsum, R.diagabs);
    } else {
        const VVi *col1 = t.d + (size_t)(s * (l - 1)) * t.nv;
        const double h = slice_dz(t, l);
        // the s stage products of slice l use table columns s(l-1)+1 .. s l
        if (R.mixed) launch_msplit(b.stream, R.mg, b.gmap, col1 + t.nv, t.nv, s, R.md, R.em);
        if (!skip_first)
            hipLaunchKernelGGL(k_stage1, eg, eb, 0, b.stream, g, b.gmap, col1, (const double2 *)R.g2c,
                               (const int *)R.xid, co.bab[0] * h, co.bab[1] * h, R.fused_prev ? 1 : 0, R.Q, R.P);
        for (int k = 2; k <= s; ++k)
            stage_launch(ctx, b, t, k == s ? 1 : 0, t.d + (size_t)(s * (l - 1) + k - 1) * t.nv, co.bab[2 * k - 2] * h,
                         co.bab[2 * k - 1] * h, nullptr, k - 2);
    }
    // periodic schedule: the decision is known on the host, no flag round trip
    const int period = batch_period(ctx, b, t);
    R.forced = period > 0 ? ((l % period == 0) || l == t.ns ? 1 : 0) : -1;
    const double cend = co.bab[0] * slice_dz(t, l);
    hipLaunchKernelGGL(k_cond, dim3(g.A), dim3(256), 0, b.stream, g, (const double *)R.colsum,
                       (const double *)R.diagabs, l == t.ns ? 1 : 0, R.forced, cend,
                       cend + co.bab[0] * slice_dz(t, l + 1), R.coef, R.xid, R.ren, b.cnt);
    if (R.forced < 0) {
        TRY_HIP(ctx, hipMemcpyAsync(R.h, R.ren, g.A * 4, hipMemcpyDeviceToHost, b.stream));
        TRY_HIP(ctx, hipEventRecord(R.evf, b.stream));
    }
    // RKN: the end-of-slice kick (merged with the next slice's first kick unless the angle
    // renormalises).  Unless every angle renormalises here, it also runs the next slice's
    // first drift Q += c_a1 dz P for the angles that do not (per-angle beta 0 for the others,
    // whose Q the renormalisation then reads unchanged).
    if (rkn) {
        const VVi *col = t.d + (size_t)(s * l) * t.nv;
        const bool fuse = R.fuse_drift && R.forced != 1 && l != t.ns;
        if (fuse) stage_launch(ctx, b, t, 0, col, 0.0, co.bab[1] * slice_dz(t, l + 1), R.coef, s - 1, R.xid);
        else stage_launch(ctx, b, t, 2, col, 0.0, 0.0, R.coef, s - 1);
        R.fused_prev = fuse;
    }
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

int runner_slice_finish(trhepd_ctx *ctx, Batch &b, const Table &t)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    // leapfrog and the Magnus methods restart from Q = I after a renormalisation (the RKN
    // path rebuilds Q in its first stage)
    const bool reset = t.method == TRHEPD_LF || t.method == TRHEPD_RK4 || magnus_mats(t.method);
    const bool gs = t.method == TRHEPD_RK4GS;
    if (gs && R.forced >= 0) {
        if (R.forced) cqr(b, R.all, g.A, true);
        TRY_HIP(ctx, hipGetLastError());
        return TRHEPD_OK;
    }
    if (R.forced >= 0) {
        if (!R.forced) return TRHEPD_OK;
        LuArgs L = R.lu(R.all);
        L.flag = b.cnt + 1;
        const int rc = lu_on(ctx, b.stream, L, g.A);
        if (!rc && reset) qreset(b, R.all, g.A, 0.0);
        return rc;
    }
    TRY_HIP(ctx, hipEventSynchronize(R.evf));
    // compact list of the angles to renormalise, staged in pinned memory: its copy is
    // queued behind this slice's kernels, and the next slice rewrites the staging area only
    // after waiting on its own event, which follows this copy in stream order
    int *lst = R.h + g.A, nl = 0;
    for (int i = 0; i < g.A; ++i)
        if (R.h[i]) lst[nl++] = i;
    if (!nl) return TRHEPD_OK;
    TRY_HIP(ctx, hipMemcpyAsync(R.list, lst, (size_t)nl * 4, hipMemcpyHostToDevice, b.stream));
    if (gs) {
        cqr(b, R.list, nl, true);
        TRY_HIP(ctx, hipGetLastError());
        return TRHEPD_OK;
    }
    LuArgs L = R.lu(R.list);
    L.flag = b.cnt + 1;
    const int rc = lu_on(ctx, b.stream, L, nl);
    if (!rc && reset) qreset(b, R.list, nl, 0.0);
    return rc;
}

int runner_end(trhepd_ctx *ctx, Batch &b, const Table &t)
{
    Runner &R = *b.run;
    const Geo &g = R.g;
    const size_t ntot = g.mat * g.A;
    const dim3 eb(256), eg((unsigned)((ntot + 255) / 256));
    if (t.method == TRHEPD_RK4GS) {   // strans, matAdivB2, f = p (surf_rk.f90:153-155)
        double2 *Qt = R.W[0], *Pt = R.W[1];
        hipLaunchKernelGGL(k_gs_strans, eg, eb, 0, b.stream, g, (const double2 *)b.gamma, (const double2 *)R.Q,
                           (const double2 *)R.P, Qt, Pt);
        LuArgs L{g, Qt, Pt, R.piv, R.all, b.cnt + 1, R.dinv, R.perm};
        int rc = lu_on(ctx, b.stream, L, g.A);
        if (rc) return rc;
        const size_t nout = (size_t)g.N * g.N * g.A;
        hipLaunchKernelGGL(k_out, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, b.stream, g,
                           (const double2 *)Pt, (const double2 *)R.gc, b.fout, b.has_f2 ? b.f2 : nullptr, b.nb0,
                           (const double *)b.w, 0);
        TRY_HIP(ctx, hipGetLastError());
        return TRHEPD_OK;
    }
    // the Magnus methods' last half step (surf_magnus.f90:172, 245, 333)
    if (magnus_mats(t.method)) lin(b, R.P, {{1.0, R.P}, {1.0, R.W[magnus_rp(t.method)]}});
    hipLaunchKernelGGL(k_st_setup, eg, eb, 0, b.stream, g, (const double2 *)R.gc, R.Q, R.P);
    LuArgs L = R.lu(R.all);
    L.flag = b.cnt + 1;
    int rc = lu_on(ctx, b.stream, L, g.A);
    if (rc) return rc;
    const size_t nout = (size_t)g.N * g.N * g.A;
    hipLaunchKernelGGL(k_out, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, b.stream, g, (const double2 *)R.P,
                       (const double2 *)R.gc, b.fout, b.has_f2 ? b.f2 : nullptr, b.nb0, (const double *)b.w, 1);
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

// Bulk reflection of one group of n > 64 beams for all nang angles (blkref, bulk.f90:51-124),
// batched over the angles.  Device inputs as trhepd_bulk_reflect holds them: ivg (nbgm, nb),
// gamma [nang][nb], phase [nb] or null; gmax2 bounds |gamma^2| over the group's beams and the
// angles.  Angle a's n x n block goes to out + a * out_stride; layers[a] the layers it used.
int bulk_large_group(trhepd_ctx *ctx, hipStream_t st, const Table &t, int n, int boff, int nang, const int *ivg,
                     int nbgm, const double2 *gamma, int nb, const double2 *phase, double gmax2, int ml, double epsb,
                     double2 *out, size_t out_stride, int *layers)
{
    if (n > 600) return fail(ctx, TRHEPD_E_ARG, "bulk group of %d beams exceeds the device limit of 600", n);
    Geo g{nang, n, (n + 15) / 16 * 16, 0};
    g.mat = (size_t)g.NP * g.NP;
    Geo g2{nang, 2 * n, (2 * n + 15) / 16 * 16, 0};
    g2.mat = (size_t)g2.NP * g2.NP;
    const size_t mb = g.mat * g.A * sizeof(double2), mb2 = g2.mat * g2.A * sizeof(double2);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    size_t om[11];
    for (size_t &o : om) o = take(mb);
    const size_t oe = take(mb2), op = take(mb2), opn = take(mb2), opiv = take((size_t)g.A * g.NP * 4),
                 odi = take((size_t)g.A * (g.NP / 16) * 256 * sizeof(double2)),
                 opm = take((size_t)2 * g.A * LU_PERM * 4), oflag = take(64), oall = take((size_t)g.A * 4),
                 or20 = take((size_t)g.A * g.NP * 8), odone = take((size_t)g.A * 4), oly = take((size_t)g.A * 4),
                 onorm = take((size_t)t.ns * 8);
    char *d = nullptr;
    int rc = dev_alloc(ctx, (void **)&d, off);
    if (rc) return rc;
    double2 *M[11];
    for (int k = 0; k < 11; ++k) M[k] = (double2 *)(d + om[k]);
    double2 *H = M[0], *X = M[1], *Cc = M[2], *Ss = M[3], *Tm = M[4], *HS = M[5];
    double2 *R = M[6], *t1 = M[7], *t2 = M[8], *Qt = M[9], *Pt = M[10];
    double2 *E = (double2 *)(d + oe), *P = (double2 *)(d + op), *Pn = (double2 *)(d + opn);
    int *flag = (int *)(d + oflag), *all = (int *)(d + oall), *done = (int *)(d + odone), *ly = (int *)(d + oly);
    double *r20 = (double *)(d + or20), *norm = (double *)(d + onorm);
    auto grid = [](size_t n_) { return dim3((unsigned)((n_ + 255) / 256)); };
    auto lin = [&](double2 *o, std::initializer_list<std::pair<double, const double2 *>> terms, double ident) {
        LinArgs L{};
        L.g = g;
        for (const auto &q : terms) {
            L.c[L.n] = q.first;
            L.x[L.n] = q.second;
            ++L.n;
        }
        L.out = o;
        L.ident = ident;
        hipLaunchKernelGGL(k_lin, grid(g.mat * g.A), dim3(256), 0, st, L);
    };
    auto prod = [&](const Geo &gg, const double2 *A, const double2 *B, double2 *C, double beta) {
        const int nbl = gg.NP / 16;
        const BmmArgs S{gg, A, B, C, 1.0, beta, 0.0, nullptr, 0};
        hipLaunchKernelGGL(k_bmm, dim3((unsigned)((nbl * nbl + 3) / 4), gg.A), dim3(256), 0, st, S);
    };
    std::vector<int> seq(g.A);
    for (int a = 0; a < g.A; ++a) seq[a] = a;
    std::vector<double> hn(t.ns);
    hipError_t e = hipMemsetAsync(d, 0, off, st);
    if (!e) e = hipMemcpyAsync(all, seq.data(), (size_t)g.A * 4, hipMemcpyHostToDevice, st);
    if (!e) {
        hipLaunchKernelGGL(k_bulkl_norm, dim3(t.ns), dim3(256), 0, st, n, boff, ivg, nbgm, (const VVi *)t.d, t.nv, norm);
        e = hipMemcpyAsync(hn.data(), norm, (size_t)t.ns * 8, hipMemcpyDeviceToHost, st);
    }
    if (!e) e = hipStreamSynchronize(st);
    if (e) {
        dev_free(ctx, d, off);
        return hip_check(ctx, e, "bulk (large group) setup");
    }
    // P = I (2n x 2n)
    {
        LinArgs L{};
        L.g = g2;
        L.out = P;
        L.ident = 1.0;
        hipLaunchKernelGGL(k_lin, grid(g2.mat * g2.A), dim3(256), 0, st, L);
    }
    const double dz = t.dz;
    for (int l = 0; l < t.ns; ++l) {
        // scaling h = dz / 2^sq with |h^2 H| <= 1 and the series degree, as the one-workgroup kernel
        const double xn = (hn[l] + gmax2) * dz * dz;
        int sq = 0;
        while (xn / (double)(1ull << (2 * sq)) > 1.0 && sq < 30) ++sq;
        const double h = dz / (double)(1ull << sq), h2 = h * h;
        const double x = xn / (double)(1ull << (2 * sq));
        int K = 1;
        for (double term = x / 2.0; term > 1e-18 && K < 20;) {
            ++K;
            term *= x / ((2.0 * K - 1) * (2.0 * K));
        }
        hipLaunchKernelGGL(k_bulkl_h, grid(g.mat * g.A), dim3(256), 0, st, g, boff, ivg, nbgm,
                           (const VVi *)(t.d + (size_t)l * t.nv), gamma, nb, H);
        lin(X, {{h2, H}}, 0.0);
        lin(Cc, {}, 1.0);
        lin(Ss, {}, 1.0);
        for (int k = K; k >= 1; --k) {   // Horner: c = I - X/((2k-1)2k) (...), s = I - X/(2k(2k+1)) (...)
            prod(g, X, Cc, Tm, 0.0);
            lin(Cc, {{-1.0 / ((2.0 * k - 1) * (2.0 * k)), Tm}}, 1.0);
            prod(g, X, Ss, Tm, 0.0);
            lin(Ss, {{-1.0 / ((2.0 * k) * (2.0 * k + 1)), Tm}}, 1.0);
        }
        lin(Ss, {{h, Ss}}, 0.0);
        prod(g, H, Ss, HS, 0.0);
        hipLaunchKernelGGL(k_bulkl_e, grid(g2.mat * g2.A), dim3(256), 0, st, g, g2, (const double2 *)Cc,
                           (const double2 *)Ss, (const double2 *)HS, E);
        for (int r = 0; r < sq; ++r) {   // E <- E^2
            prod(g2, E, E, Pn, 0.0);
            std::swap(E, Pn);
        }
        prod(g2, E, P, Pn, 0.0);   // the running product T_l ... T_1
        std::swap(P, Pn);
    }
    double2 *T11 = H, *T12 = X, *T21 = Cc, *T22 = Ss;
    hipLaunchKernelGGL(k_bulkl_t, grid(g.mat * g.A), dim3(256), 0, st, g, g2, (const double2 *)P, gamma, nb, boff,
                       phase, T11, T12, T21, T22);
    LuArgs L{g, Qt, Pt, (int *)(d + opiv), all, flag, (double2 *)(d + odi), (int *)(d + opm)};
    for (int l = 1; l <= ml && !rc; ++l) {
        // R = (T11 R + T12)(T21 R + T22)^-1, the right division as (t2^T)^-1 t1^T = R^T
        lin(t1, {{1.0, T12}}, 0.0);
        lin(t2, {{1.0, T22}}, 0.0);
        if (l > 1) {
            prod(g, T11, R, t1, 1.0);
            prod(g, T21, R, t2, 1.0);
        }
        hipLaunchKernelGGL(k_transpose, grid(g.mat * g.A), dim3(256), 0, st, g, (const double2 *)t2, Qt);
        hipLaunchKernelGGL(k_transpose, grid(g.mat * g.A), dim3(256), 0, st, g, (const double2 *)t1, Pt);
        rc = lu_on(ctx, st, L, g.A);
        if (rc) break;
        hipLaunchKernelGGL(k_transpose, grid(g.mat * g.A), dim3(256), 0, st, g, (const double2 *)Pt, R);
        hipLaunchKernelGGL(k_bulkl_judge, dim3(g.A), dim3(256), 0, st, g, (const double2 *)R, r20, l, epsb,
                           l == ml ? 1 : 0, done, ly, out, out_stride);
        if (l > 20 && l < ml) {   // stop once every angle has settled
            std::vector<int> hd(g.A);
            e = hipMemcpyAsync(hd.data(), done, (size_t)g.A * 4, hipMemcpyDeviceToHost, st);
            if (!e) e = hipStreamSynchronize(st);
            if (e) {
                rc = hip_check(ctx, e, "bulk (large group) judgement");
                break;
            }
            bool any = false;
            for (int v : hd) any |= v == 0;
            if (!any) break;
        }
    }
    int fl[16] = {};
    if (!rc) e = hipGetLastError();
    if (!rc && !e) e = hipMemcpyAsync(layers, ly, (size_t)g.A * 4, hipMemcpyDeviceToHost, st);
    if (!rc && !e) e = hipMemcpyAsync(fl, flag, 64, hipMemcpyDeviceToHost, st);
    const hipError_t e2 = hipStreamSynchronize(st);
    dev_free(ctx, d, off);
    if (rc) return rc;
    if (e || e2) return hip_check(ctx, e ? e : e2, "bulk (large group)");
    if (fl[0]) return fail(ctx, TRHEPD_E_SINGULAR, "exact zero pivot in a bulk layer solve");
    return TRHEPD_OK;
}

}  // namespace trhepd
