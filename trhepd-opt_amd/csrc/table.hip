// Device kernels of the per-structure setup and the potential table:
//   asfcef    scpot.f90:62-95    electron scattering factors per Fourier component
//   strfac    strfac.f90:68-74   structure-factor phase sums over plane-group orbits
//   table     surfio.f90:29-41 + scpot.f90:115-193 at the depths of the *_int_init
//
// The table is HBM-write bound: one column (depth) per workgroup row, one Fourier
// component per thread, the per-(depth, atom) Gaussian weights staged in LDS.
#include <cmath>

#include "internal.hpp"

namespace trhepd {
namespace {

constexpr int MAB = 6;

__global__ void k_asfcef(int nv, int nelm, const int *__restrict__ igh, const int *__restrict__ igk,
                         double ghx, double ghy, double gky, int nabr, int nabi,
                         const double *__restrict__ a, const double *__restrict__ c,
                         const double *__restrict__ dth, const double *__restrict__ dtk,
                         double *__restrict__ asf, double *__restrict__ ghk)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv) return;
    const double h = (double)igh[k], kk = (double)igk[k];
    const double gy = ghy * h + gky * kk;
    const double s = (ghx * h) * (ghx * h) + gy * gy;
    for (int e = 0; e < nelm; ++e) {
        const double u = dth[e] * ghx * h, w = dth[e] * ghy * h + dtk[e] * gky * kk;
        const double sj = -0.5 * (u * u + w * w);
        double *out = asf + (size_t)nv * MAB * e;
        for (int i = 0; i < MAB; ++i) out[k + (size_t)nv * i] = 0.0;
        for (int i = 0; i < nabr; ++i) out[k + (size_t)nv * i] = a[i + MAB * e] * exp(c[i + MAB * e] * s + sj);
        if (nabi >= 2)
            for (int i = nabr; i < nabr + nabi; ++i)
                out[k + (size_t)nv * i] = a[i + MAB * e] * exp(c[i + MAB * e] * s + sj * 0.5);
        if (nabi == 1 && k == 0) out[(size_t)nv * nabr] = a[nabr + MAB * e];
    }
    ghk[k] = sqrt(s);
}

// st(m, j) = sum over the orbit of atom j of exp(i g_m . (d + r_i))
__global__ void k_strfac(int natm, int nv, const double *__restrict__ gh, const double *__restrict__ gk,
                         const int *__restrict__ ndeg, const double *__restrict__ xb,
                         const double *__restrict__ yb, double dx, double dy, double2 *__restrict__ st)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (m >= nv) return;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < ndeg[j]; ++i) {
        const double gr = gh[m] * (dx + xb[12 * j + i]) + gk[m] * (dy + yb[12 * j + i]);
        double sn, cs;
        sincos(gr, &sn, &cs);
        re += cs;
        im += sn;
    }
    st[m + (size_t)nv * j] = make_double2(re, im);
}

// One atom group of comp_scpot: atoms [a0, a0+na) evaluated at depth zi + zshift,
// components [0, ncomp), structure factors st (ncomp leading dim), mesh factor rmesh.
struct Part {
    int a0, na, ncomp;
    double zshift, rmesh;
    const double2 *st;
};

constexpr int TB_THREADS = 256;
constexpr int MAX_ATOMS = 512;     // per part, weights staged in LDS

__global__ __launch_bounds__(TB_THREADS) void k_table(
    int nv, int nabr, int nabi, const double *__restrict__ depth, const int *__restrict__ ielm,
    const double *__restrict__ z, const double *__restrict__ dchg, const double *__restrict__ ghk,
    const double *__restrict__ ocr, const double *__restrict__ b, const double *__restrict__ asf,
    Part p0, Part p1, Part p2, VVi *__restrict__ out)
{
    // weights w[atom][i] = exp(b_i z2) * ocrmesh (0 when below exp(-70)), i < nabr (+nabi)
    __shared__ double w[MAX_ATOMS * MAB];
    __shared__ double zrel[MAX_ATOMS];
    const int col = blockIdx.y;
    const int k = blockIdx.x * TB_THREADS + threadIdx.x;
    const double zi0 = depth[col];
    double vr = 0.0, vim = 0.0, vir = 0.0, vii = 0.0;
    const Part parts[3] = {p0, p1, p2};
    for (int pi = 0; pi < 3; ++pi) {
        const Part &P = parts[pi];
        const double zi = zi0 + P.zshift;
        __syncthreads();
        for (int t = threadIdx.x; t < P.na * MAB; t += TB_THREADS) {
            const int j = t / MAB, i = t % MAB;
            const int aj = P.a0 + j;
            const int ie = ielm[aj] - 1;
            const double z1 = zi - z[aj];
            const double ex = b[i + MAB * ie] * (z1 * z1);
            double wv = 0.0;
            if (i < nabr + (nabi >= 2 ? nabi : 0) && ex > -70.0) wv = exp(ex) * (P.rmesh * ocr[aj]);
            w[t] = wv;
            if (i == 0) zrel[j] = z1;
        }
        __syncthreads();
        if (k >= P.ncomp) continue;
        for (int j = 0; j < P.na; ++j) {
            const int aj = P.a0 + j;
            const int ie = ielm[aj] - 1;
            const double ocrmesh = P.rmesh * ocr[aj];
            const double *as = asf + (size_t)nv * MAB * ie;
            double tv = 0.0, tvi = 0.0;
            if (fabs(dchg[ie]) > 1e-4) {           // ion term (scpot.f90:139-153)
                const double z1 = -fabs(zrel[j]);
                if (k == 0) {
                    tv = ocrmesh * z1 * dchg[ie];
                } else {
                    const double ex = ghk[k] * z1;
                    tv = (ex > -70.0) ? exp(ex) * ocrmesh / ghk[k] * dchg[ie] : 0.0;
                }
            }
            const double z2 = zrel[j] * zrel[j];
            if (nabi >= 2) {                      // tvi is overwritten per term (scpot.f90:157-165)
                for (int i = nabr; i < nabr + nabi; ++i)
                    if (b[i + MAB * ie] * z2 > -70.0) tvi = as[k + (size_t)nv * i] * w[j * MAB + i];
            }
            const double a0 = as[(size_t)nv * nabr];   // asf(1, nabr+1, ie) = sap
            for (int i = 0; i < nabr; ++i) {
                const double ex = b[i + MAB * ie] * z2;
                if (ex > -70.0) {
                    const double aex = as[k + (size_t)nv * i] * w[j * MAB + i];
                    tv += aex;
                    if (nabi == 1) tvi += fabs(a0 * aex);
                }
            }
            const double2 ct = P.st[k + (size_t)P.ncomp * j];
            vr += tv * ct.x;
            vim += tv * ct.y;
            // (0, tvi) * ct
            vir += -tvi * ct.y;
            vii += tvi * ct.x;
        }
    }
    if (k < nv) {
        VVi o;
        o.v = make_double2(vr, vim);
        o.vi = make_double2(vir, vii);
        out[(size_t)col * nv + k] = o;
    }
}

}  // namespace

int launch_asfcef(trhepd_ctx *ctx, int nv, int nelm, const int *igh, const int *igk, double ghx,
                  double ghy, double gky, int nabr, int nabi, const double *a, const double *c,
                  const double *dth, const double *dtk, double *asf, double *ghk)
{
    const dim3 grid((nv + 255) / 256);
    hipLaunchKernelGGL(k_asfcef, grid, dim3(256), 0, ctx->stream, nv, nelm, igh, igk, ghx, ghy, gky,
                       nabr, nabi, a, c, dth, dtk, asf, ghk);
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

int launch_strfac(trhepd_ctx *ctx, int natm, int nv, const double *gh, const double *gk,
                  const int *ndeg, const double *xb, const double *yb, double dx, double dy, double *st)
{
    if (natm == 0 || nv == 0) return TRHEPD_OK;
    const dim3 grid((nv + 255) / 256, natm);
    hipLaunchKernelGGL(k_strfac, grid, dim3(256), 0, ctx->stream, natm, nv, gh, gk, ndeg, xb, yb, dx,
                       dy, reinterpret_cast<double2 *>(st));
    TRY_HIP(ctx, hipGetLastError());
    return TRHEPD_OK;
}

// Depth of table column c for int_method m: the *_int_init loops
// (surf_prkn.f90:187-196, surf_magnus.f90:18-21 and 37-41, surf_rk.f90:18-24).
static void table_depths(int method, double dz, int ns, std::vector<double> &d)
{
    if (method == TRHEPD_SP4 || method == TRHEPD_SP6) {
        const RknCoef co = rkn_coef(method);
        for (int l = 0; l < ns; ++l) {
            double iz = dz * l;
            for (int k = 0; k < co.s; ++k) {
                d.push_back(iz);
                iz = iz + dz * co.bab[2 * k + 1];
            }
        }
        d.push_back(dz * ns);
    } else if (method == TRHEPD_M2 || method == TRHEPD_LF) {
        for (int l = 0; l < ns; ++l) d.push_back(dz * (l + 0.5));
    } else if (method == TRHEPD_M4 || method == TRHEPD_M6) {
        const double g0 = (5.0 - std::sqrt(15.0)) / 10, g2 = (5.0 + std::sqrt(15.0)) / 10;
        for (int l = 0; l < ns; ++l) {
            d.push_back(dz * (l + g0));
            d.push_back(dz * (l + 0.5));
            d.push_back(dz * (l + g2));
        }
    } else {   // RK4 / RK4GS
        for (int l = 0; l < ns; ++l) {
            const double iz = dz * l;
            d.push_back(iz);
            d.push_back(iz + 0.5 * dz);
        }
        d.push_back(dz * ns);
    }
}

// depths of an RKN table built from a slice list: slice l starts at z0 plus the thicknesses
// of the slices before it (summed in order), its stage depths follow as in hp_int_init
// (surf_prkn.f90:187-195) with that slice's thickness, and the last column is the end of the
// last slice
static void table_depths_steps(int method, const std::vector<double> &steps, double z0, std::vector<double> &d)
{
    const RknCoef co = rkn_coef(method);
    double zs = z0;
    for (const double h : steps) {
        double iz = zs;
        for (int k = 0; k < co.s; ++k) {
            d.push_back(iz);
            iz = iz + h * co.bab[2 * k + 1];
        }
        zs = zs + h;
    }
    d.push_back(zs);
}

int launch_table(trhepd_ctx *ctx, const trhepd_potential *pot, Table &t)
{
    std::vector<double> depth;
    if (t.steps.empty()) table_depths(t.method, t.dz, t.ns, depth);
    else table_depths_steps(t.method, t.steps, t.z0, depth);
    t.ncols = (int)depth.size();
    const int natm = pot->natmb + pot->natms;
    if (pot->natms > MAX_ATOMS || pot->natmb > MAX_ATOMS)
        return fail(ctx, TRHEPD_E_ARG, "table: more than %d atoms in one part", MAX_ATOMS);
    // stage the descriptor on the device (one allocation, released at the end)
    const size_t nasf = (size_t)pot->nv * MAB * pot->nelm;
    const size_t n_st0 = (size_t)pot->nv * pot->natms, n_st1 = (size_t)pot->nvb * pot->natmb;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_depth = take(depth.size() * 8), o_ielm = take(natm * 4), o_z = take(natm * 8),
                 o_dchg = take(pot->nelm * 8), o_ghk = take(pot->nv * 8), o_ocr = take(natm * 8),
                 o_b = take(MAB * pot->nelm * 8), o_asf = take(nasf * 8), o_st0 = take(n_st0 * 16),
                 o_st1 = take(n_st1 * 16), o_st2 = take(n_st1 * 16);
    char *dd = nullptr;
    int rc = dev_alloc(ctx, (void **)&dd, off);
    if (rc) return rc;
    auto up = [&](size_t o, const void *src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dd + o, src, bytes, hipMemcpyHostToDevice, ctx->stream) : hipSuccess;
    };
    hipError_t e = hipSuccess;
    if (!e) e = up(o_depth, depth.data(), depth.size() * 8);
    if (!e) e = up(o_ielm, pot->ielm, natm * 4);
    if (!e) e = up(o_z, pot->z, natm * 8);
    if (!e) e = up(o_dchg, pot->dchg, pot->nelm * 8);
    if (!e) e = up(o_ghk, pot->ghk, pot->nv * 8);
    if (!e) e = up(o_ocr, pot->ocr, natm * 8);
    if (!e) e = up(o_b, pot->b, MAB * pot->nelm * 8);
    if (!e) e = up(o_asf, pot->asf, nasf * 8);
    if (!e) e = up(o_st0, pot->st0, n_st0 * 16);
    if (!e) e = up(o_st1, pot->st1, n_st1 * 16);
    if (!e) e = up(o_st2, pot->st2, n_st1 * 16);
    if (e) {
        dev_free(ctx, dd, off);
        return hip_check(ctx, e, "table upload");
    }
    // the slice list (if any) follows the columns in the same allocation
    const size_t cbytes = (size_t)t.ncols * pot->nv * sizeof(VVi);
    const size_t tbytes = cbytes + t.steps.size() * sizeof(double);
    rc = dev_alloc(ctx, (void **)&t.d, tbytes);
    if (rc) {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx, dd, off);
        return rc;
    }
    t.bytes = tbytes;
    if (!t.steps.empty()) {
        t.d_steps = (const double *)((char *)t.d + cbytes);
        e = hipMemcpyAsync((char *)t.d + cbytes, t.steps.data(), t.steps.size() * sizeof(double),
                           hipMemcpyHostToDevice, ctx->stream);
        if (e) {
            (void)hipStreamSynchronize(ctx->stream);
            dev_free(ctx, dd, off);
            dev_free(ctx, t.d, tbytes);
            t.d = nullptr;
            return hip_check(ctx, e, "table slice list upload");
        }
    }
    t.nv = pot->nv;
    // comp_scpot: surface atoms at zi - cc, top bulk at zi, second bulk at zi + cc
    Part ps{pot->natmb, pot->natms, pot->nv, -pot->cc, 1.0, (const double2 *)(dd + o_st0)};
    Part pb1{0, pot->natmb, pot->nvb, 0.0, pot->smesh, (const double2 *)(dd + o_st1)};
    Part pb2{0, pot->natmb, pot->nvb, pot->cc, pot->smesh, (const double2 *)(dd + o_st2)};
    const dim3 grid((pot->nv + TB_THREADS - 1) / TB_THREADS, t.ncols);
    hipLaunchKernelGGL(k_table, grid, dim3(TB_THREADS), 0, ctx->stream, pot->nv, pot->nabr, pot->nabi,
                       (const double *)(dd + o_depth), (const int *)(dd + o_ielm),
                       (const double *)(dd + o_z), (const double *)(dd + o_dchg),
                       (const double *)(dd + o_ghk), (const double *)(dd + o_ocr),
                       (const double *)(dd + o_b), (const double *)(dd + o_asf), ps, pb1, pb2, t.d);
    e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, dd, off);
    if (e) return hip_check(ctx, e, "k_table launch");
    if (e2) return hip_check(ctx, e2, "k_table");
    return TRHEPD_OK;
}

}  // namespace trhepd
