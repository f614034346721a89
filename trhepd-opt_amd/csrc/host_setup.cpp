// Host-side per-structure setup: asfcrr (per-element scalars), srfghk (integer
// beam-pair map) and the plane-group orbit expansion of strfac.  These are the
// scalar / integer parts of the reference's setup; the per-component arithmetic
// (asfcef, the structure-factor phase sums, the potential table) runs on the device.
#include <cmath>

#include "internal.hpp"

namespace trhepd {

// Orbit of (x, y) under plane group nsg (strfac.f90:20-56 dispatching to the group
// routines p211 ... p6mm, strfac.f90:88-527).  Each group lists its special-position
// cases in the reference's order; a case is a predicate and the affine images
// (x, y) -> (m00 x + m01 y + t0, m10 x + m11 y + t1) that complete the orbit.
namespace {
struct Aff {
    signed char m00, m01, m10, m11;
    double t0, t1;
};
inline bool at(double v, double t) { return std::fabs(v - t) < 2e-4; }
inline bool zh(double v) { return at(v, 0.0) || at(v, 0.5); }

int emit(std::initializer_list<Aff> ops, double x, double y, double *xs, double *ys)
{
    int n = 0;
    xs[n] = x;
    ys[n] = y;
    ++n;
    for (const Aff &o : ops) {
        xs[n] = o.m00 * x + o.m01 * y + o.t0;
        ys[n] = o.m10 * x + o.m11 * y + o.t1;
        ++n;
    }
    return n;
}
}  // namespace

int orbit(int nsg, double x, double y, double *xs, double *ys)
{
    const double h = 0.5;
    const bool axis2 = zh(x) && zh(y);
    const bool third = (at(x, 1.0 / 3) && at(y, 2.0 / 3)) || (at(x, 2.0 / 3) && at(y, 1.0 / 3));
    const bool origin = at(x, 0) && at(y, 0);
    const bool c3 = origin || third;
    const bool four = origin || (at(x, h) && at(y, h));
    const bool two4 = (at(x, h) && at(y, 0)) || (at(x, 0) && at(y, h));
    const Aff R3a{0, -1, 1, -1, 0, 0}, R3b{-1, 1, -1, 0, 0, 0};   // threefold rotations
    const Aff SW{0, 1, 1, 0, 0, 0};                                // (y, x)
    switch (nsg) {
    case 1:
        return emit({}, x, y, xs, ys);
    case 2:
        return axis2 ? emit({}, x, y, xs, ys) : emit({{-1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
    case 3:
        return zh(y) ? emit({}, x, y, xs, ys) : emit({{1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
    case 4:
        return emit({{1, 0, 0, -1, h, 0}}, x, y, xs, ys);
    case 5:
        if (zh(y)) return emit({{1, 0, 0, -1, h, h}}, x, y, xs, ys);
        return emit({{1, 0, 0, -1, h, h}, {1, 0, 0, -1, 0, 0}, {1, 0, 0, 1, h, h}}, x, y, xs, ys);
    case 6:
        if (axis2) return emit({}, x, y, xs, ys);
        if (zh(x)) return emit({{1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
        if (zh(y)) return emit({{-1, 0, 0, 1, 0, 0}}, x, y, xs, ys);
        return emit({{1, 0, 0, -1, 0, 0}, {-1, 0, 0, 1, 0, 0}, {-1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
    case 7:
        if (axis2) return emit({{1, 0, 0, -1, 0, h}}, x, y, xs, ys);
        if (at(y, 0.25) || at(y, 0.75)) return emit({{-1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
        return emit({{1, 0, 0, -1, 0, h}, {-1, 0, 0, -1, 0, 0}, {-1, 0, 0, 1, 0, h}}, x, y, xs, ys);
    case 8:
        if (axis2) return emit({{-1, 0, 0, 1, h, h}}, x, y, xs, ys);
        return emit({{-1, 0, 0, 1, h, h}, {1, 0, 0, -1, h, h}, {-1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
    case 9:
        if (axis2) return emit({{-1, 0, 0, 1, h, h}}, x, y, xs, ys);
        if ((at(x, 0.25) || at(x, 0.75)) && (at(y, 0.25) || at(y, 0.75)))
            return emit({{1, 0, 0, -1, 0, 0}, {-1, 0, 0, 1, 0, 0}, {-1, 0, 0, -1, 0, 0}}, x, y, xs, ys);
        if (zh(x))
            return emit({{1, 0, 0, -1, 0, 0}, {-1, 0, 0, 1, h, h}, {-1, 0, 0, -1, h, h}}, x, y, xs, ys);
        if (zh(y))
            return emit({{-1, 0, 0, 1, 0, 0}, {1, 0, 0, -1, h, h}, {-1, 0, 0, -1, h, h}}, x, y, xs, ys);
        return emit({{1, 0, 0, -1, 0, 0}, {-1, 0, 0, 1, 0, 0}, {-1, 0, 0, -1, 0, 0}, {1, 0, 0, 1, h, h},
                     {1, 0, 0, -1, h, h}, {-1, 0, 0, 1, h, h}, {-1, 0, 0, -1, h, h}},
                    x, y, xs, ys);
    case 10:
        if (four) return emit({}, x, y, xs, ys);
        if (two4) return emit({{0, 1, -1, 0, 0, 0}}, x, y, xs, ys);
        return emit({{0, -1, 1, 0, 0, 0}, {-1, 0, 0, -1, 0, 0}, {0, 1, -1, 0, 0, 0}}, x, y, xs, ys);
    case 11:
        if (four) return emit({}, x, y, xs, ys);
        if (two4) return emit({SW}, x, y, xs, ys);
        if (at(y - x, 0) || at(x + y, 1) || zh(x) || zh(y))
            return emit({{0, -1, 1, 0, 0, 0}, {-1, 0, 0, -1, 0, 0}, {0, 1, -1, 0, 0, 0}}, x, y, xs, ys);
        return emit({SW, {0, -1, 1, 0, 0, 0}, {-1, 0, 0, 1, 0, 0}, {-1, 0, 0, -1, 0, 0},
                     {0, -1, -1, 0, 0, 0}, {0, 1, -1, 0, 0, 0}, {1, 0, 0, -1, 0, 0}},
                    x, y, xs, ys);
    case 12:
        if (four) return emit({{0, -1, -1, 0, h, h}}, x, y, xs, ys);
        if (two4) return emit({{0, 1, -1, 0, 0, 0}}, x, y, xs, ys);
        if (at(y - x, -h) || at(y - x, h) || at(x + y, h) || at(x + y, 1.5))
            return emit({{0, -1, 1, 0, 0, 0}, {-1, 0, 0, -1, 0, 0}, {0, 1, -1, 0, 0, 0}}, x, y, xs, ys);
        return emit({{0, -1, 1, 0, 0, 0}, {-1, 0, 0, -1, 0, 0}, {0, 1, -1, 0, 0, 0}, {0, -1, -1, 0, h, h},
                     {1, 0, 0, -1, h, h}, {0, 1, 1, 0, h, h}, {-1, 0, 0, 1, h, h}},
                    x, y, xs, ys);
    case 13:
        return c3 ? emit({}, x, y, xs, ys) : emit({R3a, R3b}, x, y, xs, ys);
    case 14:
        if (c3) return emit({}, x, y, xs, ys);
        if (at(2 * x - y, 0) || at(2 * y - x, 0) || at(2 * x - y, 1) || at(2 * y - x, 1) || at(x + y, 1))
            return emit({R3a, R3b}, x, y, xs, ys);
        return emit({{1, 0, 1, -1, 0, 0}, R3a, {-1, 1, 0, 1, 0, 0}, R3b, {0, -1, -1, 0, 0, 0}}, x, y, xs, ys);
    case 15:
        if (origin) return emit({}, x, y, xs, ys);
        if (third) return emit({SW}, x, y, xs, ys);
        if (at(x, 0) || at(y, 0) || at(x - y, 0)) return emit({R3a, R3b}, x, y, xs, ys);
        return emit({SW, R3a, {-1, 0, -1, 1, 0, 0}, R3b, {1, -1, 0, -1, 0, 0}}, x, y, xs, ys);
    case 16:
        if (origin) return emit({}, x, y, xs, ys);
        if (third) return emit({SW}, x, y, xs, ys);
        if ((at(x, h) && at(y, 0)) || (at(x, 0) && at(y, h)) || (at(x, h) && at(y, h)))
            return emit({R3a, R3b}, x, y, xs, ys);
        return emit({{1, -1, 1, 0, 0, 0}, R3a, {-1, 0, 0, -1, 0, 0}, R3b, {0, 1, -1, 1, 0, 0}}, x, y, xs, ys);
    case 17:
        if (origin) return emit({}, x, y, xs, ys);
        if (third) return emit({SW}, x, y, xs, ys);
        if ((at(x, h) && at(y, 0)) || (at(x, 0) && at(y, h)) || (at(x, h) && at(y, h)))
            return emit({R3a, R3b}, x, y, xs, ys);
        if (at(x, 0) || at(y, 0) || at(x - y, 0))
            return emit({R3a, R3b, {-1, 0, 0, -1, 0, 0}, {0, 1, -1, 1, 0, 0}, {1, -1, 1, 0, 0, 0}}, x, y, xs, ys);
        return emit({{1, 0, 1, -1, 0, 0}, {1, -1, 1, 0, 0, 0}, SW, R3a, {-1, 1, 0, 1, 0, 0},
                     {-1, 0, 0, -1, 0, 0}, {-1, 0, -1, 1, 0, 0}, R3b, {0, -1, -1, 0, 0, 0},
                     {0, 1, -1, 1, 0, 0}, {1, -1, 0, -1, 0, 0}},
                    x, y, xs, ys);
    default:
        return 0;
    }
}

}  // namespace trhepd

using namespace trhepd;

extern "C" int trhepd_asfcrr(double be, double *wn, int nelm, int nabr, int nabi, double *a,
                             double *b, double *c, double *dchg, double *dth, double *dtk,
                             const double *dtz, const double *elm, double tem, double smesh)
{
    if (!wn || !a || !b || !c || !dchg || !dth || !dtk || !dtz || !elm || nelm < 0 || nabr < 0 ||
        nabr + nabi > 6 || smesh == 0.0)
        return TRHEPD_E_ARG;
    const double pi = std::atan(1.0) * 4.0, c2m = 511.001, ek = .262466;
    *wn = std::sqrt(1e3 * be * ek * (1.0 + 0.5 * be / c2m));
    const double s = 4.0 * pi / smesh;
    const double rel = (1.0 + be / c2m) * s;
    const double rat = std::sqrt(100.0 / be * (1.0 + 50.0 / c2m) / (1.0 + be * 0.5 / c2m)) * s;
    const double vib = 11407.0 * tem;
    for (int e = 0; e < nelm; ++e) {
        double dwz;
        if (tem > 1e-5) {
            dwz = vib / (dtz[e] * dtz[e] * elm[e]);
            dth[e] = std::sqrt((vib + vib) / elm[e]) / (4.0 * pi * dth[e]);
            dtk[e] = std::sqrt((vib + vib) / elm[e]) / (4.0 * pi * dtk[e]);
        } else if (tem < -1e-5) {
            dwz = (8.0 * pi * pi) * dtz[e] * dtz[e];
        } else {
            dwz = dtz[e];
            dth[e] = std::sqrt(dth[e] * 0.5) / (pi + pi);
            dtk[e] = std::sqrt(dtk[e] * 0.5) / (pi + pi);
        }
        double *ae = a + 6 * e, *be_ = b + 6 * e, *ce = c + 6 * e;
        for (int j = 0; j < nabr; ++j) {
            ce[j] = be_[j] / (-16.0 * pi * pi);
            ae[j] = ae[j] * std::sqrt(4.0 * pi / (be_[j] + dwz)) * rel;
            be_[j] = -4.0 * pi * pi / (be_[j] + dwz);
        }
        if (std::fabs(dchg[e]) > 1e-4) dchg[e] = dchg[e] * 1.88972619 * rel;
        if (nabi >= 2) {
            for (int j = nabr; j < nabr + nabi; ++j) {
                ce[j] = be_[j] / (-16.0 * pi * pi);
                ae[j] = ae[j] * std::sqrt(4.0 * pi / be_[j]) * rat;
                be_[j] = -4.0 * pi * pi / be_[j];
            }
        }
    }
    return TRHEPD_OK;
}

extern "C" int trhepd_srfghk(int nvm, int nvb, int *igh, int *igk, int nb, const int *ih,
                             const int *ik, int *iv, int *nv_out)
{
    if (!igh || !igk || !ih || !ik || !iv || !nv_out || nb < 1 || nvb < 1 || nvb > nvm)
        return TRHEPD_E_ARG;
    int nv = nvb;
    for (int k = 0; k < nb - 1; ++k)
        for (int l = k + 1; l < nb; ++l) {
            const int h0 = ih[l] - ih[k], k0 = ik[l] - ik[k];
            int m = 1;
            for (; m < nv; ++m)
                if (igh[m] == h0 && igk[m] == k0) break;
            if (m == nv) {
                if (nv + 1 > nvm) return TRHEPD_E_ARG;
                igh[nv] = h0;
                igk[nv] = k0;
                ++nv;
            }
            iv[l + (size_t)k * nb] = m + 1;
        }
    *nv_out = nv;
    return TRHEPD_OK;
}
