/*
 * trhepd_oracle.c -- CPU restatement of the sim-trhepd-rheed surface slice loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * Every routine names the reference routine it restates (paths are relative to the
 * reference repository root, all under sim-trhepd-rheed/src/ unless stated).
 * Conventions follow the reference: complex(8) arrays are column-major interleaved
 * (re, im) doubles, index maps (iv, ielm) are 1-based, shapes are Fortran shapes.
 *
 * Pinning (what fixes each function, beyond retyping the formula):
 *   - end-to-end FP64 rocking curves of test/test{23,47,521}/test-results/perf-{4,5,6,7}-*.txt
 *     (lf / sp4 / sp6 / rk4) pin table build, gather, propagators, cond_est, LU, intensity,
 *     structure factors of plane group 14, asfcrr/asfcef, srfghk;
 *   - plane-group orbits of the other 16 groups: closure under the group generators and
 *     the International Tables multiplicities (tests/test_oracle_math.py);
 *   - LU solve: residual against random systems and numpy.linalg.solve;
 *   - cond_est: Varah's bound (an upper bound on cond_1 for column-diagonally-dominant
 *     matrices), the diagonal closed form and the 1e100 branch (tests/test_oracle_math.py);
 *   - hp_int_spa1d: against hp_int_spa run at N=1 on the one-beam Ge(001) sample deck (the
 *     gap falls as dz^4) and its own order-6 convergence;
 *   - Magnus m2/m4/m6 and rk4_intgs: parity unpinned by a reference vector; pinned by the
 *     observed convergence order in dz against sp6 at fine dz (2, 4, 6, 4) and, for
 *     rk4_intgs, by agreement with rk4 (which the perf-7 goldens pin).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;
#define MAB 6
#define AT(a, i, j, ld) ((a)[(size_t)(i) + (size_t)(j) * (size_t)(ld)])

/* ------------------------------------------------------------------------- */
/* descriptor for the potential-table inputs (module mod_scpot3, surfio.f90:18-27) */
typedef struct {
    int nv, nvb, natms, natmb, nelm, nabr, nabi;
    double cc, smesh;
    const int *ielm;      /* natmb+natms, 1-based element index (bulk atoms first) */
    const double *z;      /* natmb+natms, zoffset already added */
    const double *dchg;   /* nelm */
    const double *ghk;    /* nv */
    const double *ocr;    /* natmb+natms */
    const double *b;      /* (MAB, nelm) */
    const double *asf;    /* (nv, MAB, nelm) */
    const zc *st0;        /* (nv, natms)  surface */
    const zc *st1;        /* (nvb, natmb) top bulk */
    const zc *st2;        /* (nvb, natmb) second bulk */
} orc_pot;

/* ------------------------------------------------------------------------- */
/* asfcrr: scpot.f90:11-58 (energy / Debye-Waller correction, in place) */
void orc_asfcrr(double be, double *wn, int nelm, int nabr, int nabi, double *a, double *b,
                double *c, double *dchg, double *dth, double *dtk, const double *dtz,
                const double *elm, double tem, double smesh)
{
    const double pi = atan(1.0) * 4.0, c2m = 511.001, ek = .262466;
    *wn = sqrt(1e3 * be * ek * (1.0 + 0.5 * be / c2m));
    double s = 4.0 * pi / smesh;
    double rel = (1.0 + be / c2m) * s;
    double rat = sqrt(100.0 / be * (1.0 + 50.0 / c2m) / (1.0 + be * 0.5 / c2m)) * s;
    double vib = 11407.0 * tem;
    for (int i = 0; i < nelm; i++) {
        double dwz;
        if (tem > 1e-5) {
            dwz = vib / (dtz[i] * dtz[i] * elm[i]);
            dth[i] = sqrt((vib + vib) / elm[i]) / (4.0 * pi * dth[i]);
            dtk[i] = sqrt((vib + vib) / elm[i]) / (4.0 * pi * dtk[i]);
        } else if (tem < -1e-5) {
            dwz = (8.0 * pi * pi) * dtz[i] * dtz[i];
        } else {
            dwz = dtz[i];
            dth[i] = sqrt(dth[i] * 0.5) / (pi + pi);
            dtk[i] = sqrt(dtk[i] * 0.5) / (pi + pi);
        }
        for (int j = 0; j < nabr; j++) {
            double *aj = &AT(a, j, i, MAB), *bj = &AT(b, j, i, MAB);
            AT(c, j, i, MAB) = *bj / (-16.0 * pi * pi);
            *aj = *aj * sqrt(4.0 * pi / (*bj + dwz)) * rel;
            *bj = -4.0 * pi * pi / (*bj + dwz);
        }
        if (fabs(dchg[i]) > 1e-4) dchg[i] = dchg[i] * 1.88972619 * rel;
        if (nabi >= 2) {
            for (int j = nabr; j < nabr + nabi; j++) {
                double *aj = &AT(a, j, i, MAB), *bj = &AT(b, j, i, MAB);
                AT(c, j, i, MAB) = *bj / (-16.0 * pi * pi);
                *aj = *aj * sqrt(4.0 * pi / *bj) * rat;
                *bj = -4.0 * pi * pi / *bj;
            }
        }
    }
}

/* asfcef: scpot.f90:62-95 */
void orc_asfcef(int nv, int nelm, const int *igh, const int *igk, double ghx, double ghy,
                double gky, int nabr, int nabi, const double *a, const double *c,
                double *asf, const double *dth, const double *dtk, double *ghk)
{
    for (int k = 0; k < nv; k++) {
        double h = (double)igh[k], kk = (double)igk[k];
        double s = (ghx * h) * (ghx * h) + (ghy * h + gky * kk) * (ghy * h + gky * kk);
        for (int j = 0; j < nelm; j++) {
            double u = dth[j] * ghx * h, w = dth[j] * ghy * h + dtk[j] * gky * kk;
            double sj = -0.5 * (u * u + w * w);
            for (int i = 0; i < nabr; i++)
                asf[k + (size_t)nv * (i + MAB * j)] = AT(a, i, j, MAB) * exp(AT(c, i, j, MAB) * s + sj);
            if (nabi >= 2)
                for (int i = nabr; i < nabr + nabi; i++)
                    asf[k + (size_t)nv * (i + MAB * j)] = AT(a, i, j, MAB) * exp(AT(c, i, j, MAB) * s + sj * 0.5);
        }
        ghk[k] = sqrt(s);
    }
    if (nabi == 1)
        for (int j = 0; j < nelm; j++) asf[0 + (size_t)nv * (nabr + MAB * j)] = AT(a, nabr, j, MAB);
}

/* ------------------------------------------------------------------------- */
/* plane-group orbits: strfac.f90:5-76 with the 17 group routines strfac.f90:88-527.
 * Each group is a list of cases; the first case whose predicate holds gives the
 * equivalent positions as affine images (x,y) -> (m00 x + m01 y + t0, m10 x + m11 y + t1). */
static int nr(double v, double t) { return fabs(v - t) < 2e-4; }
static int nr0h(double v) { return nr(v, 0.0) || nr(v, 0.5); }

typedef struct { int m00, m01, m10, m11; double t0, t1; } op2;
#define OP(a, b, c, d, e, f) { a, b, c, d, e, f }
static const op2 OP_ID = OP(1, 0, 0, 1, 0, 0);

static int orbit_apply(const op2 *ops, int n, double x, double y, double *xs, double *ys)
{
    for (int i = 0; i < n; i++) {
        xs[i] = ops[i].m00 * x + ops[i].m01 * y + ops[i].t0;
        ys[i] = ops[i].m10 * x + ops[i].m11 * y + ops[i].t1;
    }
    return n;
}

int orc_orbit(int nsg, double x, double y, double *xs, double *ys)
{
#define RET(...) do { const op2 o_[] = { OP_ID, __VA_ARGS__ }; \
        return orbit_apply(o_, (int)(sizeof o_ / sizeof o_[0]), x, y, xs, ys); } while (0)
#define RET1() do { xs[0] = x; ys[0] = y; return 1; } while (0)
    int ax2 = nr0h(x) && nr0h(y);                 /* on a twofold axis at (0|1/2, 0|1/2) */
    int c3 = (nr(x, 0) && nr(y, 0)) || (nr(x, 1.0 / 3) && nr(y, 2.0 / 3)) || (nr(x, 2.0 / 3) && nr(y, 1.0 / 3));
    int k3 = (nr(x, 1.0 / 3) && nr(y, 2.0 / 3)) || (nr(x, 2.0 / 3) && nr(y, 1.0 / 3));
    int o4 = (nr(x, 0) && nr(y, 0)) || (nr(x, 0.5) && nr(y, 0.5));
    int t4 = (nr(x, 0.5) && nr(y, 0)) || (nr(x, 0) && nr(y, 0.5));
    switch (nsg) {
    case 1: RET1();
    case 2: if (ax2) RET1(); RET(OP(-1, 0, 0, -1, 0, 0));                                   /* p2 */
    case 3: if (nr0h(y)) RET1(); RET(OP(1, 0, 0, -1, 0, 0));                               /* pm */
    case 4: RET(OP(1, 0, 0, -1, 0.5, 0));                                                  /* pg */
    case 5: if (nr0h(y)) RET(OP(1, 0, 0, -1, 0.5, 0.5));                                   /* cm */
        RET(OP(1, 0, 0, -1, 0.5, 0.5), OP(1, 0, 0, -1, 0, 0), OP(1, 0, 0, 1, 0.5, 0.5));
    case 6:                                                                                /* p2mm */
        if (ax2) RET1();
        if (nr0h(x)) RET(OP(1, 0, 0, -1, 0, 0));
        if (nr0h(y)) RET(OP(-1, 0, 0, 1, 0, 0));
        RET(OP(1, 0, 0, -1, 0, 0), OP(-1, 0, 0, 1, 0, 0), OP(-1, 0, 0, -1, 0, 0));
    case 7:                                                                                /* p2mg */
        if (ax2) RET(OP(1, 0, 0, -1, 0, 0.5));
        if (nr(y, 0.25) || nr(y, 0.75)) RET(OP(-1, 0, 0, -1, 0, 0));
        RET(OP(1, 0, 0, -1, 0, 0.5), OP(-1, 0, 0, -1, 0, 0), OP(-1, 0, 0, 1, 0, 0.5));
    case 8:                                                                                /* p2gg */
        if (ax2) RET(OP(-1, 0, 0, 1, 0.5, 0.5));
        RET(OP(-1, 0, 0, 1, 0.5, 0.5), OP(1, 0, 0, -1, 0.5, 0.5), OP(-1, 0, 0, -1, 0, 0));
    case 9:                                                                                /* c2mm */
        if (ax2) RET(OP(-1, 0, 0, 1, 0.5, 0.5));
        if ((nr(x, 0.25) || nr(x, 0.75)) && (nr(y, 0.25) || nr(y, 0.75)))
            RET(OP(1, 0, 0, -1, 0, 0), OP(-1, 0, 0, 1, 0, 0), OP(-1, 0, 0, -1, 0, 0));
        if (nr0h(x)) RET(OP(1, 0, 0, -1, 0, 0), OP(-1, 0, 0, 1, 0.5, 0.5), OP(-1, 0, 0, -1, 0.5, 0.5));
        if (nr0h(y)) RET(OP(-1, 0, 0, 1, 0, 0), OP(1, 0, 0, -1, 0.5, 0.5), OP(-1, 0, 0, -1, 0.5, 0.5));
        RET(OP(1, 0, 0, -1, 0, 0), OP(-1, 0, 0, 1, 0, 0), OP(-1, 0, 0, -1, 0, 0),
            OP(1, 0, 0, 1, 0.5, 0.5), OP(1, 0, 0, -1, 0.5, 0.5), OP(-1, 0, 0, 1, 0.5, 0.5),
            OP(-1, 0, 0, -1, 0.5, 0.5));
    case 10:                                                                               /* p4 */
        if (o4) RET1();
        if (t4) RET(OP(0, 1, -1, 0, 0, 0));
        RET(OP(0, -1, 1, 0, 0, 0), OP(-1, 0, 0, -1, 0, 0), OP(0, 1, -1, 0, 0, 0));
    case 11:                                                                               /* p4mm */
        if (o4) RET1();
        if (t4) RET(OP(0, 1, 1, 0, 0, 0));
        if (nr(y - x, 0) || nr(x + y, 1) || nr0h(x) || nr0h(y))
            RET(OP(0, -1, 1, 0, 0, 0), OP(-1, 0, 0, -1, 0, 0), OP(0, 1, -1, 0, 0, 0));
        RET(OP(0, 1, 1, 0, 0, 0), OP(0, -1, 1, 0, 0, 0), OP(-1, 0, 0, 1, 0, 0), OP(-1, 0, 0, -1, 0, 0),
            OP(0, -1, -1, 0, 0, 0), OP(0, 1, -1, 0, 0, 0), OP(1, 0, 0, -1, 0, 0));
    case 12:                                                                               /* p4gm */
        if (o4) RET(OP(0, -1, -1, 0, 0.5, 0.5));
        if (t4) RET(OP(0, 1, -1, 0, 0, 0));
        if (nr(y - x, -0.5) || nr(y - x, 0.5) || nr(x + y, 0.5) || nr(x + y, 1.5))
            RET(OP(0, -1, 1, 0, 0, 0), OP(-1, 0, 0, -1, 0, 0), OP(0, 1, -1, 0, 0, 0));
        RET(OP(0, -1, 1, 0, 0, 0), OP(-1, 0, 0, -1, 0, 0), OP(0, 1, -1, 0, 0, 0),
            OP(0, -1, -1, 0, 0.5, 0.5), OP(1, 0, 0, -1, 0.5, 0.5), OP(0, 1, 1, 0, 0.5, 0.5),
            OP(-1, 0, 0, 1, 0.5, 0.5));
    case 13:                                                                               /* p3 */
        if (c3) RET1();
        RET(OP(0, -1, 1, -1, 0, 0), OP(-1, 1, -1, 0, 0, 0));
    case 14:                                                                               /* p3m1 */
        if (c3) RET1();
        if (nr(x + x - y, 0) || nr(y + y - x, 0) || nr(x + x - y, 1) || nr(y + y - x, 1) || nr(x + y, 1))
            RET(OP(0, -1, 1, -1, 0, 0), OP(-1, 1, -1, 0, 0, 0));
        RET(OP(1, 0, 1, -1, 0, 0), OP(0, -1, 1, -1, 0, 0), OP(-1, 1, 0, 1, 0, 0),
            OP(-1, 1, -1, 0, 0, 0), OP(0, -1, -1, 0, 0, 0));
    case 15:                                                                               /* p31m */
        if (nr(x, 0) && nr(y, 0)) RET1();
        if (k3) RET(OP(0, 1, 1, 0, 0, 0));
        if (nr(x, 0) || nr(y, 0) || nr(x - y, 0)) RET(OP(0, -1, 1, -1, 0, 0), OP(-1, 1, -1, 0, 0, 0));
        RET(OP(0, 1, 1, 0, 0, 0), OP(0, -1, 1, -1, 0, 0), OP(-1, 0, -1, 1, 0, 0),
            OP(-1, 1, -1, 0, 0, 0), OP(1, -1, 0, -1, 0, 0));
    case 16:                                                                               /* p6 */
        if (nr(x, 0) && nr(y, 0)) RET1();
        if (k3) RET(OP(0, 1, 1, 0, 0, 0));
        if ((nr(x, 0.5) && nr(y, 0)) || (nr(x, 0) && nr(y, 0.5)) || (nr(x, 0.5) && nr(y, 0.5)))
            RET(OP(0, -1, 1, -1, 0, 0), OP(-1, 1, -1, 0, 0, 0));
        RET(OP(1, -1, 1, 0, 0, 0), OP(0, -1, 1, -1, 0, 0), OP(-1, 0, 0, -1, 0, 0),
            OP(-1, 1, -1, 0, 0, 0), OP(0, 1, -1, 1, 0, 0));
    default:                                                                               /* p6mm */
        if (nr(x, 0) && nr(y, 0)) RET1();
        if (k3) RET(OP(0, 1, 1, 0, 0, 0));
        if ((nr(x, 0.5) && nr(y, 0)) || (nr(x, 0) && nr(y, 0.5)) || (nr(x, 0.5) && nr(y, 0.5)))
            RET(OP(0, -1, 1, -1, 0, 0), OP(-1, 1, -1, 0, 0, 0));
        if (nr(x, 0) || nr(y, 0) || nr(x - y, 0))
            RET(OP(0, -1, 1, -1, 0, 0), OP(-1, 1, -1, 0, 0, 0), OP(-1, 0, 0, -1, 0, 0),
                OP(0, 1, -1, 1, 0, 0), OP(1, -1, 1, 0, 0, 0));
        RET(OP(1, 0, 1, -1, 0, 0), OP(1, -1, 1, 0, 0, 0), OP(0, 1, 1, 0, 0, 0), OP(0, -1, 1, -1, 0, 0),
            OP(-1, 1, 0, 1, 0, 0), OP(-1, 0, 0, -1, 0, 0), OP(-1, 0, -1, 1, 0, 0),
            OP(-1, 1, -1, 0, 0, 0), OP(0, -1, -1, 0, 0, 0), OP(0, 1, -1, 1, 0, 0), OP(1, -1, 0, -1, 0, 0));
    }
#undef RET
#undef RET1
}

static double reduce01(double x) { return x - floor(x + 1e-4); }   /* strfac.f90:80-84 */

/* strfac: strfac.f90:5-76.  Returns the number of equivalent positions (0 on bad nsg). */
int orc_strfac(int nsg, int nv, int ma, int mb, int na, int nb, const double *gh,
               const double *gk, double x, double y, double dx, double dy, zc *st)
{
    if (nsg < 1 || nsg > 17) return 0;
    double s = (double)(ma * nb - mb * na);
    double xs[12], ys[12], xb[12], yb[12];
    double x0 = reduce01((nb * x - na * y) / s), y0 = reduce01((ma * y - mb * x) / s);
    int ndeg = orc_orbit(nsg, x0, y0, xs, ys);
    for (int i = 0; i < ndeg; i++) {
        double u = reduce01(xs[i]), w = reduce01(ys[i]);
        xb[i] = ma * u + na * w;
        yb[i] = mb * u + nb * w;
    }
    for (int m = 0; m < nv; m++) {
        zc acc = 0;
        for (int i = 0; i < ndeg; i++) {
            double gr = gh[m] * (dx + xb[i]) + gk[m] * (dy + yb[i]);
            acc += cos(gr) + I * sin(gr);
        }
        st[m] = acc;
    }
    return ndeg;
}

/* srfghk: surfsub.f90:9-41.  igh/igk have room for nvm entries and hold nvb on entry.
 * iv is (nb, nb); only the strict lower triangle iv(l,k), l>k, is set.  Returns nv or -1. */
int orc_srfghk(int nvm, int nvb, int *igh, int *igk, int nb, const int *ih, const int *ik, int *iv)
{
    int nv = nvb;
    for (int k = 0; k < nb - 1; k++)
        for (int l = k + 1; l < nb; l++) {
            int h0 = ih[l] - ih[k], k0 = ik[l] - ik[k], found = 0;
            for (int m = 1; m < nv; m++)
                if (h0 == igh[m] && k0 == igk[m]) { AT(iv, l, k, nb) = m + 1; found = 1; break; }
            if (!found) {
                if (nv + 1 > nvm) return -1;
                igh[nv] = h0; igk[nv] = k0; nv++;
                AT(iv, l, k, nb) = nv;
            }
        }
    return nv;
}

/* ------------------------------------------------------------------------- */
/* scpot: scpot.f90:115-193, accumulating into v, vi (length nv). asf leading dim mv. */
static void scpot(int mv, int nv, zc *v, zc *vi, int natm, const int *ielm, const double *z,
                  double zi, int nabr, int nabi, const double *asf, const double *b,
                  const double *dchg, double rmesh, const double *ghk, const double *ocr, const zc *st)
{
    const double uf = -70.0;
    double *tv = malloc(sizeof(double) * (size_t)nv), *tvi = malloc(sizeof(double) * (size_t)nv);
    for (int j = 0; j < natm; j++) {
        double ocrmesh = rmesh * ocr[j];
        int ie = ielm[j] - 1;
        double z1 = zi - z[j], z2 = z1 * z1;
        if (fabs(dchg[ie]) > 1e-4) {
            z1 = -fabs(z1);
            tv[0] = ocrmesh * z1 * dchg[ie];
            for (int k = 1; k < nv; k++) {
                double ex = ghk[k] * z1;
                tv[k] = (ex > uf) ? exp(ex) * ocrmesh / ghk[k] * dchg[ie] : 0.0;
            }
        } else {
            for (int k = 0; k < nv; k++) tv[k] = 0.0;
        }
        for (int k = 0; k < nv; k++) tvi[k] = 0.0;
        if (nabi >= 2)
            for (int i = nabr; i < nabr + nabi; i++) {
                double ex = AT(b, i, ie, MAB) * z2;
                if (ex > uf) {
                    ex = exp(ex) * ocrmesh;
                    for (int k = 0; k < nv; k++) tvi[k] = asf[k + (size_t)mv * (i + MAB * ie)] * ex;
                }
            }
        for (int i = 0; i < nabr; i++) {
            double ex = AT(b, i, ie, MAB) * z2;
            if (ex > uf) {
                ex = exp(ex) * ocrmesh;
                const double *ai = &asf[(size_t)mv * (i + MAB * ie)];
                if (nabi == 1) {
                    double ai0 = asf[(size_t)mv * (nabr + MAB * ie)];
                    for (int k = 0; k < nv; k++) {
                        double aex = ai[k] * ex;
                        tv[k] += aex;
                        tvi[k] += fabs(ai0 * aex);
                    }
                } else {
                    for (int k = 0; k < nv; k++) tv[k] += ai[k] * ex;
                }
            }
        }
        const zc *sj = &st[(size_t)nv * j];
        for (int k = 0; k < nv; k++) {
            v[k] += tv[k] * sj[k];
            vi[k] += (I * tvi[k]) * sj[k];
        }
    }
    free(tv); free(tvi);
}

/* comp_scpot: surfio.f90:29-41 */
void orc_comp_scpot(const orc_pot *p, double zi, zc *v, zc *vi)
{
    for (int k = 0; k < p->nv; k++) { v[k] = 0; vi[k] = 0; }
    scpot(p->nv, p->nv, v, vi, p->natms, p->ielm + p->natmb, p->z + p->natmb, zi - p->cc,
          p->nabr, p->nabi, p->asf, p->b, p->dchg, 1.0, p->ghk, p->ocr + p->natmb, p->st0);
    scpot(p->nv, p->nvb, v, vi, p->natmb, p->ielm, p->z, zi, p->nabr, p->nabi, p->asf, p->b,
          p->dchg, p->smesh, p->ghk, p->ocr, p->st1);
    scpot(p->nv, p->nvb, v, vi, p->natmb, p->ielm, p->z, zi + p->cc, p->nabr, p->nabi, p->asf,
          p->b, p->dchg, p->smesh, p->ghk, p->ocr, p->st2);
}

/* ------------------------------------------------------------------------- */
/* integrator coefficient tables (surf_prkn.f90:9-52; Blanes & Moan, JCAM 142 (2002) 313) */
static const double BAB4[13] = {
    0.0829844064174052, 0.245298957184271, 0.396309801498368, 0.604872665711080,
    -0.0390563049223486, -0.350171622895351, 0.1195241940131508, -0.350171622895351,
    -0.0390563049223486, 0.604872665711080, 0.396309801498368, 0.245298957184271,
    0.0829844064174052};
static const double BAB6[23] = {
    0.0414649985182624, 0.123229775946271, 0.198128671918067, 0.290553797799558,
    -0.0400061921041533, -0.127049212625417, 0.0752539843015807, -0.246331761062075,
    -0.0115113874206879, 0.357208872795928, 0.2366699247869311, 0.20477705429147,
    0.2366699247869311, 0.357208872795928, -0.0115113874206879, -0.246331761062075,
    0.0752539843015807, -0.127049212625417, -0.0400061921041533, 0.290553797799558,
    0.198128671918067, 0.123229775946271, 0.0414649985182624};

/* number of table columns used by int_method m (surflf.f90:28-41 and the *_int_init) */
int orc_table_cols(int method, int ns)
{
    switch (method) {
    case 1: case 4: return ns;
    case 2: case 3: return 3 * ns;
    case 5: return 6 * ns + 1;
    case 6: return 11 * ns + 1;
    case 7: case 8: return 2 * ns + 1;
    default: return -1;
    }
}

/* depth of table column c (0-based) for int_method m; the *_int_init routines:
 * hp_int_init surf_prkn.f90:174-197, m2_int_init surf_magnus.f90:4-22,
 * m4_int_init surf_magnus.f90:24-42, rk4_int_init surf_rk.f90:4-25 */
double orc_table_depth(int method, double dz, int ns, int c)
{
    if (method == 5 || method == 6) {
        int s = method == 5 ? 6 : 11;
        const double *bab = method == 5 ? BAB4 : BAB6;
        if (c == s * ns) return dz * ns;
        int l = c / s, k = c % s;
        double iz = dz * l;
        for (int j = 0; j < k; j++) iz = iz + dz * bab[2 * j + 1];
        return iz;
    }
    if (method == 1 || method == 4) return dz * (c + 0.5);
    if (method == 2 || method == 3) {
        const double g0 = (5.0 - sqrt(15.0)) / 10, g1 = 0.5, g2 = (5.0 + sqrt(15.0)) / 10;
        int l = c / 3, k = c % 3;
        return dz * (l + (k == 0 ? g0 : k == 1 ? g1 : g2));
    }
    if (method == 7 || method == 8) {
        if (c == 2 * ns) return dz * ns;
        int l = c / 2;
        return (c % 2 == 0) ? dz * l : dz * l + 0.5 * dz;
    }
    return 0.0;
}

/* table build: v(nv, K), vi(nv, K) */
int orc_table(const orc_pot *p, int method, double dz, int ns, zc *v, zc *vi)
{
    int K = orc_table_cols(method, ns);
    if (K < 0) return -1;
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < K; c++)
        orc_comp_scpot(p, orc_table_depth(method, dz, ns, c), v + (size_t)p->nv * c, vi + (size_t)p->nv * c);
    return K;
}

/* depth of table column c of an RKN table over a slice list (int_method 5/6): slice l starts
 * at z0 plus the thicknesses before it (summed in order), its stages follow hp_int_init
 * (surf_prkn.f90:187-195) with its own thickness, the last column ends the last slice.
 * No counterpart in the reference (uniform slices); the schedule the device builds. */
double orc_table_depth_steps(int method, const double *steps, int ns, double z0, int c)
{
    int s = method == 5 ? 6 : 11;
    const double *bab = method == 5 ? BAB4 : BAB6;
    int l = c / s, k = c % s;
    if (c == s * ns) { l = ns; k = 0; }
    double zs = z0;
    for (int j = 0; j < l; j++) zs = zs + steps[j];
    double iz = zs;
    for (int j = 0; j < k; j++) iz = iz + steps[l] * bab[2 * j + 1];
    return iz;
}

int orc_table_steps(const orc_pot *p, int method, const double *steps, int ns, double z0, zc *v, zc *vi)
{
    if (method != 5 && method != 6) return -1;
    int K = orc_table_cols(method, ns);
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < K; c++)
        orc_comp_scpot(p, orc_table_depth_steps(method, steps, ns, z0, c), v + (size_t)p->nv * c, vi + (size_t)p->nv * c);
    return K;
}

/* h211b_limit: h211b_control.f90:1-7 */
double orc_h211b_limit(double u)
{
    const double kappa = 1.0;
    return 1.0 + kappa * atan((u - 1.0) / kappa);
}

/* h211b_filter: h211b_control.f90:9-16 */
double orc_h211b_filter(double c1, double c0, double rho, int k)
{
    const int b = 4;
    return pow(c1, 1.0 / (b * k)) * pow(c0, 1.0 / (b * k)) * pow(rho, -1.0 / k);
}

/* ------------------------------------------------------------------------- */
/* dense helpers standing in for the BLAS/LAPACK calls of the reference */

/* C = alpha*A*B + beta*C, all n x n with leading dims lda, ldb, ldc (zgemm 'N','N') */
static void zgemm_nn(int n, zc alpha, const zc *A, int lda, const zc *B, int ldb, zc beta, zc *C, int ldc)
{
#pragma omp parallel for schedule(static) if (n >= 96)
    for (int j = 0; j < n; j++) {
        zc *cj = &C[(size_t)j * ldc];
        if (beta == 0) for (int i = 0; i < n; i++) cj[i] = 0;
        else if (beta != 1) for (int i = 0; i < n; i++) cj[i] *= beta;
        for (int l = 0; l < n; l++) {
            zc t = alpha * B[l + (size_t)j * ldb];
            if (t == 0) continue;
            const zc *al = &A[(size_t)l * lda];
            for (int i = 0; i < n; i++) cj[i] += t * al[i];
        }
    }
}

static double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }   /* izamax metric */

/* matAbdivB: matcomp.f90:53-69.  A is (lda, 2n) = [X | Y]; on exit Y <- X^{-1} Y.
 * zgetrf(n, 2n) with row partial pivoting chosen in the X block, then ztrsm(L,U,N,N). */
int orc_matAbdivB(int n, zc *A, int lda)
{
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k; double best = cabs1(AT(A, k, k, lda));
        for (int i = k + 1; i < n; i++) { double t = cabs1(AT(A, i, k, lda)); if (t > best) { best = t; p = i; } }
        if (best == 0.0) { if (!info) info = k + 1; continue; }
        if (p != k)
            for (int j = 0; j < 2 * n; j++) { zc t = AT(A, k, j, lda); AT(A, k, j, lda) = AT(A, p, j, lda); AT(A, p, j, lda) = t; }
        zc rp = 1.0 / AT(A, k, k, lda);
        for (int i = k + 1; i < n; i++) AT(A, i, k, lda) *= rp;
#pragma omp parallel for schedule(static) if (n >= 96)
        for (int j = k + 1; j < 2 * n; j++) {
            zc u = AT(A, k, j, lda);
            if (u == 0) continue;
            for (int i = k + 1; i < n; i++) AT(A, i, j, lda) -= AT(A, i, k, lda) * u;
        }
    }
    /* back substitution U X = Y' on the n right-hand columns */
#pragma omp parallel for schedule(static) if (n >= 96)
    for (int j = n; j < 2 * n; j++)
        for (int k = n - 1; k >= 0; k--) {
            zc x = AT(A, k, j, lda);
            if (x == 0) continue;
            x /= AT(A, k, k, lda);
            AT(A, k, j, lda) = x;
            for (int i = 0; i < k; i++) AT(A, i, j, lda) -= x * AT(A, i, k, lda);
        }
    return info;
}

/* cond_est: matcomp.f90:498-528 (Gershgorin bound on cond of the X^H block) */
double orc_cond_est(int n, const zc *A, int lda)
{
    double mn = cabs(AT(A, 0, 0, lda)), mx = mn;
    for (int j = 0; j < n; j++) {
        double t = 0.0;
        for (int i = 0; i < n; i++)
            if (i != j) { double re = creal(AT(A, i, j, lda)), im = cimag(AT(A, i, j, lda)); t += sqrt(re * re + im * im); }
        double d = cabs(AT(A, j, j, lda));
        if (d - t < mn) mn = d - t;
        if (d + t > mx) mx = d + t;
    }
    return mn < 0.0 ? 1e100 : mx / mn;
}

/* vvi2mate: matcomp.f90:253-279.  M(i,j): upper -conj(v+vi)[iv(j,i)], lower (-v+vi)[iv(i,j)],
 * diagonal -conj(v+vi)[1] - conj(g_j^2). */
void orc_vvi2mate(int n, const zc *v, const zc *vi, const int *iv, const zc *g, zc *m)
{
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) { int q = AT(iv, j, i, n) - 1; AT(m, i, j, n) = -conj(v[q] + vi[q]); }
        AT(m, j, j, n) = -conj(v[0] + vi[0]) - conj(g[j] * g[j]);
        for (int i = j + 1; i < n; i++) { int q = AT(iv, i, j, n) - 1; AT(m, i, j, n) = -v[q] + vi[q]; }
    }
}

/* vvi2matf: matcomp.f90:281-306 (used by rk4_intgs) */
static void vvi2matf(int n, const zc *v, const zc *vi, const int *iv, const zc *g, zc *m)
{
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) { int q = AT(iv, j, i, n) - 1; AT(m, i, j, n) = conj(vi[q] - v[q]); }
        AT(m, j, j, n) = -(v[0] + vi[0]) - g[j] * g[j];
        for (int i = j + 1; i < n; i++) { int q = AT(iv, i, j, n) - 1; AT(m, i, j, n) = -(v[q] + vi[q]); }
    }
}

/* invstransh: matcomp.f90:140-172.  t1 (n, 2n) <- [I | i diag(conj g) (I+F^H)^{-1}(I-F^H)] */
static void invstransh(int n, zc *t1, const zc *f, const zc *g)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (i == j) { AT(t1, i, j, n) = 1.0 + conj(AT(f, j, j, n)); AT(t1, i, n + j, n) = 1.0 - conj(AT(f, j, j, n)); }
            else { AT(t1, i, j, n) = conj(AT(f, j, i, n)); AT(t1, i, n + j, n) = -conj(AT(f, j, i, n)); }
        }
    orc_matAbdivB(n, t1, n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) AT(t1, i, n + j, n) = I * AT(t1, i, n + j, n) * conj(g[i]);
}

/* stransh: matcomp.f90:174-196, followed by f = conj(transpose(Y)) (surf_prkn.f90:329) */
static void stransh_out(int n, zc *t1, const zc *g, zc *f)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            zc y = AT(t1, i, j + n, n);
            if (i == j) { AT(t1, i, j, n) = conj(g[i]) - I * y; AT(t1, i, j + n, n) = conj(g[i]) + I * y; }
            else { AT(t1, i, j, n) = -I * y; AT(t1, i, j + n, n) = I * y; }
        }
    orc_matAbdivB(n, t1, n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) AT(f, i, j, n) = conj(AT(t1, j, i + n, n));
}

/* ------------------------------------------------------------------------- */
/* statistics a caller may read after an integration */
typedef struct { int n_renorm; int n_gemm; } orc_stats;

/* hp_int_spa: surf_prkn.f90:278-330 (symplectic RKN in ratio form).  steps (ns), when not
 * NULL, gives each slice its own thickness (a table built by orc_table_steps): the stage
 * weights take slice l's thickness, and the next slice's first kick, merged with this end
 * kick (psave), is recomputed with its own thickness where the two differ. */
static void hp_int_spa_steps(double dz, const double *steps, int ns, int nb, int nv, zc *f, const zc *g,
                             const int *iv, const zc *v, const zc *vi, int s, const double *bab, orc_stats *st);
static void hp_int_spa(double dz, int ns, int nb, int nv, zc *f, const zc *g, const int *iv,
                       const zc *v, const zc *vi, int s, const double *bab, orc_stats *st)
{
    hp_int_spa_steps(dz, NULL, ns, nb, nv, f, g, iv, v, vi, s, bab, st);
}
static void hp_int_spa_steps(double dz0, const double *steps, int ns, int nb, int nv, zc *f, const zc *g,
                             const int *iv, const zc *v, const zc *vi, int s, const double *bab, orc_stats *st)
{
    int n = nb;
    zc *t1 = malloc(sizeof(zc) * (size_t)n * 2 * n), *psave = malloc(sizeof(zc) * (size_t)n * n);
    zc *m = malloc(sizeof(zc) * (size_t)n * n);
    zc *Q = t1, *P = t1 + (size_t)n * n;
    invstransh(n, t1, f, g);
    int xid = 1;
    double dzp = dz0;   /* the previous slice's thickness */
    for (int l = 1; l <= ns; l++) {
        const double dz = steps ? steps[l - 1] : dz0;
        for (int k = 1; k <= s; k++) {
            double cb = bab[2 * k - 2], ca = bab[2 * k - 1];
            if (k > 1 || l == 1) orc_vvi2mate(n, v + (size_t)nv * (s * (l - 1) + k - 1), vi + (size_t)nv * (s * (l - 1) + k - 1), iv, g, m);
            if (k == 1 && !xid && l > 1 && dz != dzp) {
                /* the first kick with this slice's thickness: m still holds the previous
                 * slice's end matrix, at the depth this slice starts */
                zgemm_nn(n, cb * dz, Q, n, m, n, 1.0, P, n); st->n_gemm++;
            } else if (k == 1 && !xid && l > 1) {
                for (size_t e = 0; e < (size_t)n * n; e++) P[e] += psave[e];
            } else if (xid) {
                for (size_t e = 0; e < (size_t)n * n; e++) P[e] += (cb * dz) * m[e];
            } else {
                zgemm_nn(n, cb * dz, Q, n, m, n, 1.0, P, n); st->n_gemm++;
            }
            if (xid) {
                for (size_t e = 0; e < (size_t)n * n; e++) Q[e] = (ca * dz) * P[e];
                for (int i = 0; i < n; i++) AT(Q, i, i, n) += 1.0;
            } else {
                for (size_t e = 0; e < (size_t)n * n; e++) Q[e] += (ca * dz) * P[e];
            }
            xid = 0;
        }
        orc_vvi2mate(n, v + (size_t)nv * (s * l), vi + (size_t)nv * (s * l), iv, g, m);
        zgemm_nn(n, bab[0] * dz, Q, n, m, n, 0.0, psave, n); st->n_gemm++;
        for (size_t e = 0; e < (size_t)n * n; e++) P[e] += psave[e];
        double d = orc_cond_est(n, Q, n);
        if (d > 1e3 || l == ns) { orc_matAbdivB(n, t1, n); xid = 1; st->n_renorm++; }
        else xid = 0;
        dzp = dz;
    }
    stransh_out(n, t1, g, f);
    free(t1); free(psave); free(m);
}

/* hp_int_spa1d: surf_prkn.f90:485-524 (scalar N=1 form) */
static void hp_int_spa1d(double dz, int ns, zc *f, zc g, const zc *v, const zc *vi, int s, const double *bab)
{
    zc q = 1.0, p = I * (1.0 - conj(*f)) / (1.0 + conj(*f)) * conj(g);
    for (int l = 1; l <= ns; l++) {
        for (int k = 1; k <= s; k++) {
            double cb = (l != 1 && k == 1) ? 2 * bab[0] : bab[2 * k - 2], ca = bab[2 * k - 1];
            int c = s * (l - 1) + k - 1;
            zc mm = -conj(v[c] + vi[c]) - conj(g * g);
            p = p + cb * dz * mm * q;
            q = q + ca * dz * p;
        }
        p = p / q;
        q = 1.0;
    }
    zc mm = -conj(v[ns * s] + vi[ns * s]) - conj(g * g);
    p = p + bab[0] * dz * mm * q;
    *f = (conj(g) + I * p) / (conj(g) - I * p);
}

/* lf_int_spa: surf_magnus.f90:46-110 (Stormer-Verlet) */
static void lf_int_spa(double dz, int ns, int n, int nv, zc *f, const zc *g, const int *iv,
                       const zc *v, const zc *vi, orc_stats *st)
{
    zc *t1 = malloc(sizeof(zc) * (size_t)n * 2 * n), *m = malloc(sizeof(zc) * (size_t)n * n);
    zc *Q = t1, *P = t1 + (size_t)n * n;
    invstransh(n, t1, f, g);
    for (size_t e = 0; e < (size_t)n * n; e++) Q[e] = (dz / 2) * P[e];
    for (int i = 0; i < n; i++) AT(Q, i, i, n) += 1.0;
    int xid = 0;
    for (int l = 1; l <= ns; l++) {
        orc_vvi2mate(n, v + (size_t)nv * (l - 1), vi + (size_t)nv * (l - 1), iv, g, m);
        if (xid) for (size_t e = 0; e < (size_t)n * n; e++) P[e] += dz * m[e];
        else { zgemm_nn(n, dz, Q, n, m, n, 1.0, P, n); st->n_gemm++; }
        double dz2 = (l == ns) ? dz / 2 : dz;
        if (xid) {
            for (size_t e = 0; e < (size_t)n * n; e++) Q[e] = dz2 * P[e];
            for (int i = 0; i < n; i++) AT(Q, i, i, n) += 1.0;
        } else {
            for (size_t e = 0; e < (size_t)n * n; e++) Q[e] += dz2 * P[e];
        }
        double d = orc_cond_est(n, Q, n);
        if (d > 1e3 || l == ns) { orc_matAbdivB(n, t1, n); xid = 1; st->n_renorm++; }
        else xid = 0;
    }
    stransh_out(n, t1, g, f);
    free(t1); free(m);
}

/* hyperbolic_exp4r: matcomp.f90:388-405 */
static void hyp_exp4r(double h, int n, const zc *a, zc *q, zc *r, orc_stats *st)
{
    zgemm_nn(n, 1.0, a, n, a, n, 0.0, r, n); st->n_gemm++;
    double h3 = h * h * h, h5 = h3 * h * h;
    for (size_t e = 0; e < (size_t)n * n; e++) q[e] = (h3 / 6) * a[e] + (h5 / 120) * r[e];
    for (int i = 0; i < n; i++) AT(q, i, i, n) += h;
    for (size_t e = 0; e < (size_t)n * n; e++) r[e] = (h / 2) * a[e] - (h3 / 24) * r[e];
}

/* hyperbolic_exp6r: matcomp.f90:407-428 */
static void hyp_exp6r(double h, int n, const zc *a, zc *q, zc *r, zc *t, orc_stats *st)
{
    const double cq0 = 1.0 / 6, cq1 = 1.0 / 120, cq2 = 1.0 / 5040, cr0 = 0.5, cr1 = -1.0 / 24, cr2 = 1.0 / 240;
    zgemm_nn(n, 1.0, a, n, a, n, 0.0, r, n);
    zgemm_nn(n, 1.0, r, n, a, n, 0.0, t, n); st->n_gemm += 2;
    double h3 = h * h * h, h5 = h3 * h * h, h7 = h5 * h * h;
    for (size_t e = 0; e < (size_t)n * n; e++) q[e] = (h3 * cq0) * a[e] + (h5 * cq1) * r[e] + (h7 * cq2) * t[e];
    for (int i = 0; i < n; i++) AT(q, i, i, n) += h;
    for (size_t e = 0; e < (size_t)n * n; e++) r[e] = (h * cr0) * a[e] + (h3 * cr1) * r[e] + (h5 * cr2) * t[e];
}

/* m2_int_spa: surf_magnus.f90:114-178 */
static void m2_int_spa(double dz, int ns, int n, int nv, zc *f, const zc *g, const int *iv,
                       const zc *v, const zc *vi, orc_stats *st)
{
    size_t nn = (size_t)n * n;
    zc *t1 = malloc(sizeof(zc) * 2 * nn), *m = malloc(sizeof(zc) * nn), *r = malloc(sizeof(zc) * nn);
    zc *rp = calloc(nn, sizeof(zc)), *q = malloc(sizeof(zc) * nn);
    zc *Q = t1, *P = t1 + nn;
    invstransh(n, t1, f, g);
    int xid = 1;
    for (int l = 1; l <= ns; l++) {
        orc_vvi2mate(n, v + (size_t)nv * (l - 1), vi + (size_t)nv * (l - 1), iv, g, m);
        hyp_exp4r(dz, n, m, q, r, st);
        if (xid) { for (size_t e = 0; e < nn; e++) P[e] += r[e] + rp[e]; }
        else {
            for (size_t e = 0; e < nn; e++) rp[e] += r[e];
            zgemm_nn(n, 1.0, Q, n, rp, n, 1.0, P, n); st->n_gemm++;
        }
        memcpy(rp, r, sizeof(zc) * nn);
        if (xid) {
            zgemm_nn(n, 1.0, P, n, q, n, 0.0, Q, n); st->n_gemm++;
            for (int i = 0; i < n; i++) AT(Q, i, i, n) += 1.0;
        } else { zgemm_nn(n, 1.0, P, n, q, n, 1.0, Q, n); st->n_gemm++; }
        double d = orc_cond_est(n, Q, n);
        if (d > 1e3 || l == ns) { orc_matAbdivB(n, t1, n); xid = 1; st->n_renorm++; }
        else xid = 0;
    }
    for (size_t e = 0; e < nn; e++) P[e] += rp[e];
    stransh_out(n, t1, g, f);
    free(t1); free(m); free(r); free(rp); free(q);
}

/* m4_int_spa: surf_magnus.f90:180-251 */
static void m4_int_spa(double dz, int ns, int n, int nv, zc *f, const zc *g, const int *iv,
                       const zc *v, const zc *vi, orc_stats *st)
{
    size_t nn = (size_t)n * n;
    const double c0 = sqrt(15.0) / 36, c1 = 5.0 / 36;
    zc *t1 = malloc(sizeof(zc) * 2 * nn), *m1 = malloc(sizeof(zc) * nn), *m2 = malloc(sizeof(zc) * nn);
    zc *m3 = malloc(sizeof(zc) * nn), *rp = calloc(nn, sizeof(zc)), *r = malloc(sizeof(zc) * nn), *q = malloc(sizeof(zc) * nn);
    zc *Q = t1, *P = t1 + nn;
    invstransh(n, t1, f, g);
    int xid = 1;
    for (int l = 1; l <= ns; l++) {
        orc_vvi2mate(n, v + (size_t)nv * (3 * l - 3), vi + (size_t)nv * (3 * l - 3), iv, g, m1);
        orc_vvi2mate(n, v + (size_t)nv * (3 * l - 2), vi + (size_t)nv * (3 * l - 2), iv, g, m2);
        orc_vvi2mate(n, v + (size_t)nv * (3 * l - 1), vi + (size_t)nv * (3 * l - 1), iv, g, m3);
        for (size_t e = 0; e < nn; e++) { zc t = m3[e] - m1[e]; m3[e] = m1[e] - 2.0 * m2[e] + m3[e]; m1[e] = t; }
        hyp_exp4r(dz, n, m2, q, r, st);
        for (size_t e = 0; e < nn; e++) rp[e] = rp[e] + r[e] + dz * (-c0 * m1[e] + c1 * m3[e]);
        if (xid) { for (size_t e = 0; e < nn; e++) P[e] += rp[e]; }
        else { zgemm_nn(n, 1.0, Q, n, rp, n, 1.0, P, n); st->n_gemm++; }
        if (xid) {
            zgemm_nn(n, 1.0, P, n, q, n, 0.0, Q, n); st->n_gemm++;
            for (int i = 0; i < n; i++) AT(Q, i, i, n) += 1.0;
        } else { zgemm_nn(n, 1.0, P, n, q, n, 1.0, Q, n); st->n_gemm++; }
        for (size_t e = 0; e < nn; e++) rp[e] = r[e] + dz * (c0 * m1[e] + c1 * m3[e]);
        double d = orc_cond_est(n, Q, n);
        if (d > 1e3 || l == ns) { orc_matAbdivB(n, t1, n); xid = 1; st->n_renorm++; }
        else xid = 0;
    }
    for (size_t e = 0; e < nn; e++) P[e] += rp[e];
    stransh_out(n, t1, g, f);
    free(t1); free(m1); free(m2); free(m3); free(rp); free(r); free(q);
}

/* m6_int_spa: surf_magnus.f90:253-339 */
static void m6_int_spa(double dz, int ns, int n, int nv, zc *f, const zc *g, const int *iv,
                       const zc *v, const zc *vi, orc_stats *st)
{
    size_t nn = (size_t)n * n;
    const double c0 = sqrt(15.0) / 180, c1 = 1.0 / 18, c2 = 1.0 / 12960;
    const double d0 = 4.0 / (3.0 * sqrt(15.0)), d1 = 1.0 / 6;
    zc *t1 = malloc(sizeof(zc) * 2 * nn), *m1 = malloc(sizeof(zc) * nn), *m2 = malloc(sizeof(zc) * nn);
    zc *m3 = malloc(sizeof(zc) * nn), *tt = malloc(sizeof(zc) * nn), *rp = calloc(nn, sizeof(zc));
    zc *r = malloc(sizeof(zc) * nn), *q = malloc(sizeof(zc) * nn), *tw = malloc(sizeof(zc) * nn);
    zc *Q = t1, *P = t1 + nn;
    invstransh(n, t1, f, g);
    int xid = 1;
    for (int l = 1; l <= ns; l++) {
        orc_vvi2mate(n, v + (size_t)nv * (3 * l - 3), vi + (size_t)nv * (3 * l - 3), iv, g, m1);
        orc_vvi2mate(n, v + (size_t)nv * (3 * l - 2), vi + (size_t)nv * (3 * l - 2), iv, g, m2);
        orc_vvi2mate(n, v + (size_t)nv * (3 * l - 1), vi + (size_t)nv * (3 * l - 1), iv, g, m3);
        for (size_t e = 0; e < nn; e++) {
            r[e] = m3[e] - m1[e];
            q[e] = m1[e] - 2.0 * m2[e] + m3[e];
            m1[e] = r[e];
            m2[e] = m2[e] + d1 * q[e];
        }
        zgemm_nn(n, c2 * dz * dz * dz, m1, n, m1, n, 0.0, m3, n); st->n_gemm++;
        for (size_t e = 0; e < nn; e++) { m3[e] += c1 * dz * q[e]; tt[e] = m2[e] - d0 * m1[e]; }
        hyp_exp6r(dz / 2, n, tt, q, r, tw, st);
        for (size_t e = 0; e < nn; e++) rp[e] = rp[e] + r[e] - (c0 * dz) * m1[e] + m3[e];
        if (xid) { for (size_t e = 0; e < nn; e++) P[e] += rp[e]; }
        else { zgemm_nn(n, 1.0, Q, n, rp, n, 1.0, P, n); st->n_gemm++; }
        if (xid) {
            zgemm_nn(n, 1.0, P, n, q, n, 0.0, Q, n); st->n_gemm++;
            for (int i = 0; i < n; i++) AT(Q, i, i, n) += 1.0;
        } else { zgemm_nn(n, 1.0, P, n, q, n, 1.0, Q, n); st->n_gemm++; }
        memcpy(rp, r, sizeof(zc) * nn);
        for (size_t e = 0; e < nn; e++) tt[e] = m2[e] + d0 * m1[e];
        hyp_exp6r(dz / 2, n, tt, q, r, tw, st);
        for (size_t e = 0; e < nn; e++) rp[e] += r[e];
        zgemm_nn(n, 1.0, Q, n, rp, n, 1.0, P, n);
        zgemm_nn(n, 1.0, P, n, q, n, 1.0, Q, n); st->n_gemm += 2;
        for (size_t e = 0; e < nn; e++) rp[e] = r[e] + (dz * c0) * m1[e] + m3[e];
        double d = orc_cond_est(n, Q, n);
        if (d > 1e3 || l == ns) { orc_matAbdivB(n, t1, n); xid = 1; st->n_renorm++; }
        else xid = 0;
    }
    for (size_t e = 0; e < nn; e++) P[e] += rp[e];
    stransh_out(n, t1, g, f);
    free(t1); free(m1); free(m2); free(m3); free(tt); free(rp); free(r); free(q); free(tw);
}

/* rk4_int: surf_rk.f90:27-91 */
static void rk4_int(double dz, int ns, int n, int nv, zc *f, const zc *g, const int *iv,
                    const zc *v, const zc *vi, orc_stats *st)
{
    size_t nn = (size_t)n * n;
    zc *t1 = malloc(sizeof(zc) * 2 * nn), *t2 = malloc(sizeof(zc) * 2 * nn);
    zc *k1 = malloc(sizeof(zc) * 2 * nn), *k2 = malloc(sizeof(zc) * 2 * nn);
    zc *k3 = malloc(sizeof(zc) * 2 * nn), *k4 = malloc(sizeof(zc) * 2 * nn);
    zc *m0 = malloc(sizeof(zc) * nn), *m1 = malloc(sizeof(zc) * nn), *m2 = malloc(sizeof(zc) * nn);
    invstransh(n, t1, f, g);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) AT(t1, i, j, n) = (i == j) ? 1.0 : 0.0;
    int xid = 1;
    orc_vvi2mate(n, v, vi, iv, g, m2);
    for (int l = 1; l <= ns; l++) {
        memcpy(m0, m2, sizeof(zc) * nn);
        orc_vvi2mate(n, v + (size_t)nv * (2 * l - 1), vi + (size_t)nv * (2 * l - 1), iv, g, m1);
        orc_vvi2mate(n, v + (size_t)nv * (2 * l), vi + (size_t)nv * (2 * l), iv, g, m2);
        memcpy(k1, t1 + nn, sizeof(zc) * nn);
        if (xid) memcpy(k1 + nn, m0, sizeof(zc) * nn);
        else { zgemm_nn(n, 1.0, t1, n, m0, n, 0.0, k1 + nn, n); st->n_gemm++; }
        for (size_t e = 0; e < 2 * nn; e++) t2[e] = t1[e] + 0.5 * dz * k1[e];
        memcpy(k2, t2 + nn, sizeof(zc) * nn);
        zgemm_nn(n, 1.0, t2, n, m1, n, 0.0, k2 + nn, n);
        for (size_t e = 0; e < 2 * nn; e++) t2[e] = t1[e] + 0.5 * dz * k2[e];
        memcpy(k3, t2 + nn, sizeof(zc) * nn);
        zgemm_nn(n, 1.0, t2, n, m1, n, 0.0, k3 + nn, n);
        for (size_t e = 0; e < 2 * nn; e++) t2[e] = t1[e] + dz * k3[e];
        memcpy(k4, t2 + nn, sizeof(zc) * nn);
        zgemm_nn(n, 1.0, t2, n, m2, n, 0.0, k4 + nn, n);
        st->n_gemm += 3;
        for (size_t e = 0; e < 2 * nn; e++) t1[e] = t1[e] + (dz / 6) * (k1[e] + 2.0 * k2[e] + 2.0 * k3[e] + k4[e]);
        double d = orc_cond_est(n, t1, n);
        if (d > 1e3 || l == ns) {
            orc_matAbdivB(n, t1, n);
            for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) AT(t1, i, j, n) = (i == j) ? 1.0 : 0.0;
            xid = 1; st->n_renorm++;
        } else xid = 0;
    }
    stransh_out(n, t1, g, f);
    free(t1); free(t2); free(k1); free(k2); free(k3); free(k4); free(m0); free(m1); free(m2);
}

/* ---- rk4_intgs (surf_rk.f90:93-156) with bcgs/cqr (matgs.f90:4-37) and power_est
 *      (matcomp.f90:530-594).  Cholesky-QR of the stacked [q; p] (2n x n). ---- */
static void cqr_stack(int n, zc *q, zc *p)
{
    size_t nn = (size_t)n * n;
    zc *c = calloc(nn, sizeof(zc));
    /* c = [q;p]^H [q;p] (upper triangle) */
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) {
            zc acc = 0;
            for (int r = 0; r < n; r++) acc += conj(AT(q, r, i, n)) * AT(q, r, j, n) + conj(AT(p, r, i, n)) * AT(p, r, j, n);
            AT(c, i, j, n) = acc;
        }
    /* zpotrf 'U': c = U^H U */
    for (int j = 0; j < n; j++) {
        double d = creal(AT(c, j, j, n));
        for (int r = 0; r < j; r++) d -= creal(conj(AT(c, r, j, n)) * AT(c, r, j, n));
        d = sqrt(d);
        AT(c, j, j, n) = d;
        for (int k = j + 1; k < n; k++) {
            zc acc = AT(c, j, k, n);
            for (int r = 0; r < j; r++) acc -= conj(AT(c, r, j, n)) * AT(c, r, k, n);
            AT(c, j, k, n) = acc / d;
        }
    }
    /* ztrsm 'R','U','N','N': A <- A U^{-1} for both halves */
    zc *hal[2] = {q, p};
    for (int h = 0; h < 2; h++) {
        zc *A = hal[h];
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < j; k++) {
                zc u = AT(c, k, j, n);
                for (int r = 0; r < n; r++) AT(A, r, j, n) -= AT(A, r, k, n) * u;
            }
            zc dj = AT(c, j, j, n);
            for (int r = 0; r < n; r++) AT(A, r, j, n) /= dj;
        }
    }
    free(c);
}

static void zgemv_n(int n, const zc *A, const zc *x, zc *y)
{
    for (int i = 0; i < n; i++) y[i] = 0;
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) y[i] += AT(A, i, j, n) * x[j];
}
static void zgemv_c_acc(int n, const zc *A, const zc *x, zc *y)
{
    for (int j = 0; j < n; j++) { zc acc = 0; for (int i = 0; i < n; i++) acc += conj(AT(A, i, j, n)) * x[i]; y[j] += acc; }
}

static double power_est(int n, const zc *q, const zc *p, zc *x, int s)
{
    zc *y = malloc(sizeof(zc) * 2 * (size_t)n), *z = malloc(sizeof(zc) * (size_t)n);
    double d = -1.0, d0 = -1.0;
    for (int it = 1; it <= s; it++) {
        zgemv_n(n, q, x, y); zgemv_n(n, p, x, y + n);
        for (int i = 0; i < n; i++) z[i] = 0;
        zgemv_c_acc(n, q, y, z); zgemv_c_acc(n, p, y + n, z);
        zc cr = 0; double dn = 0;
        for (int i = 0; i < n; i++) { dn += creal(z[i]) * creal(z[i]) + cimag(z[i]) * cimag(z[i]); cr += conj(x[i]) * z[i]; }
        d = cabs(cr);
        dn = 1.0 / sqrt(dn);
        double dp = 0;
        for (int i = 0; i < n; i++) { z[i] *= dn; zc c = z[i] - x[i]; dp += creal(c) * creal(c) + cimag(c) * cimag(c); x[i] = z[i]; }
        dp = sqrt(dp);
        if (dp < 1e-2 || d - d0 < 1e-2 * d) break;
        d0 = d;
    }
    free(y); free(z);
    return d;
}

static void rk4_intgs(double dz, int ns, int n, int nv, zc *f, const zc *g, const int *iv,
                      const zc *v, const zc *vi, orc_stats *st)
{
    size_t nn = (size_t)n * n;
    zc *q = malloc(sizeof(zc) * nn), *p = malloc(sizeof(zc) * nn), *q2 = malloc(sizeof(zc) * nn);
    zc *r1 = malloc(sizeof(zc) * nn), *s1 = malloc(sizeof(zc) * nn), *r2 = malloc(sizeof(zc) * nn), *s2 = malloc(sizeof(zc) * nn);
    zc *r3 = malloc(sizeof(zc) * nn), *s3 = malloc(sizeof(zc) * nn), *r4 = malloc(sizeof(zc) * nn), *s4 = malloc(sizeof(zc) * nn);
    zc *m0 = malloc(sizeof(zc) * nn), *m1 = malloc(sizeof(zc) * nn), *m2 = malloc(sizeof(zc) * nn);
    zc *x = malloc(sizeof(zc) * (size_t)n);
    /* invstrans: matcomp.f90:89-115 */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            zc ig = 1.0 / g[i];
            if (i == j) { AT(q, i, j, n) = ig + ig * AT(f, i, j, n); AT(p, i, j, n) = -I + I * AT(f, i, j, n); }
            else { AT(q, i, j, n) = ig * AT(f, i, j, n); AT(p, i, j, n) = I * AT(f, i, j, n); }
        }
    cqr_stack(n, q, p);
    for (int i = 0; i < n; i++) x[i] = (1.0 + I) / sqrt(2.0 * n);
    vvi2matf(n, v, vi, iv, g, m2);
    for (int l = 1; l <= ns; l++) {
        memcpy(m0, m2, sizeof(zc) * nn);
        vvi2matf(n, v + (size_t)nv * (2 * l - 1), vi + (size_t)nv * (2 * l - 1), iv, g, m1);
        vvi2matf(n, v + (size_t)nv * (2 * l), vi + (size_t)nv * (2 * l), iv, g, m2);
        memcpy(r1, p, sizeof(zc) * nn);
        zgemm_nn(n, 1.0, m0, n, q, n, 0.0, s1, n);
        for (size_t e = 0; e < nn; e++) { q2[e] = q[e] + 0.5 * dz * r1[e]; r2[e] = p[e] + 0.5 * dz * s1[e]; }
        zgemm_nn(n, 1.0, m1, n, q2, n, 0.0, s2, n);
        for (size_t e = 0; e < nn; e++) { q2[e] = q[e] + 0.5 * dz * r2[e]; r3[e] = p[e] + 0.5 * dz * s2[e]; }
        zgemm_nn(n, 1.0, m1, n, q2, n, 0.0, s3, n);
        for (size_t e = 0; e < nn; e++) { q2[e] = q[e] + dz * r3[e]; r4[e] = p[e] + dz * s3[e]; }
        zgemm_nn(n, 1.0, m2, n, q2, n, 0.0, s4, n);
        st->n_gemm += 4;
        for (size_t e = 0; e < nn; e++) {
            q[e] = q[e] + (dz / 6) * (r1[e] + 2.0 * r2[e] + 2.0 * r3[e] + r4[e]);
            p[e] = p[e] + (dz / 6) * (s1[e] + 2.0 * s2[e] + 2.0 * s3[e] + s4[e]);
        }
        double d = power_est(n, q, p, x, n);
        if (d > 1e2) { cqr_stack(n, q, p); st->n_renorm++; }
    }
    /* strans: matcomp.f90:117-138, then f = p (p/q via matAdivB2) */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            zc xx = AT(q, i, j, n), yy = AT(p, i, j, n);
            AT(q, i, j, n) = g[i] * xx + I * yy;
            AT(p, i, j, n) = g[i] * xx - I * yy;
        }
    /* matAdivB2(n, p, q): p <- p q^{-1} = (q^T \ p^T)^T */
    zc *t = malloc(sizeof(zc) * 2 * nn);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) { AT(t, i, j, n) = AT(q, j, i, n); AT(t, i, j + n, n) = AT(p, j, i, n); }
    orc_matAbdivB(n, t, n);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) AT(f, i, j, n) = AT(t, j, i + n, n);
    free(t);
    free(q); free(p); free(q2); free(r1); free(s1); free(r2); free(s2); free(r3); free(s3); free(r4); free(s4);
    free(m0); free(m1); free(m2); free(x);
}

/* ------------------------------------------------------------------------- */
/* integrate one angle: the select case of surflf.f90:118-139.  f (nb, nb) in/out.
 * Returns 0, or -1 for an unknown method.  stats (2 ints) may be NULL. */
int orc_integrate(int method, double dz, int ns, int nb, int nv, zc *f, const zc *g,
                  const int *iv, const zc *v, const zc *vi, int *stats)
{
    orc_stats st = {0, 0};
    switch (method) {
    case 1: m2_int_spa(dz, ns, nb, nv, f, g, iv, v, vi, &st); break;
    case 2: m4_int_spa(dz, ns, nb, nv, f, g, iv, v, vi, &st); break;
    case 3: m6_int_spa(dz, ns, nb, nv, f, g, iv, v, vi, &st); break;
    case 4: lf_int_spa(dz, ns, nb, nv, f, g, iv, v, vi, &st); break;
    case 5: hp_int_spa(dz, ns, nb, nv, f, g, iv, v, vi, 6, BAB4, &st); break;
    case 6:
        if (nb == 1) hp_int_spa1d(dz, ns, f, g[0], v, vi, 11, BAB6);
        else hp_int_spa(dz, ns, nb, nv, f, g, iv, v, vi, 11, BAB6, &st);
        break;
    case 7: rk4_int(dz, ns, nb, nv, f, g, iv, v, vi, &st); break;
    case 8: rk4_intgs(dz, ns, nb, nv, f, g, iv, v, vi, &st); break;
    default: return -1;
    }
    if (stats) { stats[0] = st.n_renorm; stats[1] = st.n_gemm; }
    return 0;
}

/* the RKN integrators (int_method 5, 6; nb > 1) over a slice list (orc_table_steps) */
int orc_integrate_steps(int method, const double *steps, int ns, int nb, int nv, zc *f, const zc *g,
                        const int *iv, const zc *v, const zc *vi, int *stats)
{
    orc_stats st = {0, 0};
    if ((method != 5 && method != 6) || nb < 2) return -1;
    hp_int_spa_steps(steps[0], steps, ns, nb, nv, f, g, iv, v, vi, method == 5 ? 6 : 11, method == 5 ? BAB4 : BAB6, &st);
    if (stats) { stats[0] = st.n_renorm; stats[1] = st.n_gemm; }
    return 0;
}

/* gamma_i(theta, phi): surflf.f90:90-103 */
void orc_gamma(int nb, const int *ih, const int *ik, double wn, double ghx, double ghy,
               double gky, double ga, double az, zc *gma)
{
    double cga = cos(ga), sga = sin(ga), caz = cos(az), saz = sin(az);
    (void)sga;
    double wnx = wn * cga * caz, wny = wn * cga * saz;
    for (int i = 0; i < nb; i++) {
        double wgx = ghx * ih[i] + wnx, wgy = ghy * ih[i] + gky * ik[i] + wny;
        double s = wn * wn - wgx * wgx - wgy * wgy;
        gma[i] = (s >= 0.0) ? sqrt(s) : I * sqrt(-s);
    }
}

/* intensities of one angle: surflf.f90:143-151.  nb0 is 1-based. */
void orc_intensity(int nb, const zc *f, const zc *gma, int nb0, double wnsga2, double *f2)
{
    for (int i = 0; i < nb; i++) {
        double s = creal(gma[i]);
        zc a = AT(f, i, nb0 - 1, nb);
        f2[i] = (s > 1e-10) ? creal(a * conj(a)) * wnsga2 / s : 0.0;
    }
}
