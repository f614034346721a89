// Adaptive slice thickness for the RKN integrators (int_method 5, 6).
//
// The reference carries a step-size controller that nothing calls yet: h211b_limit and
// h211b_filter of sim-trhepd-rheed/src/h211b_control.f90:1-16 (linked into the surf_prkn and
// surf_magnus modules, Makefile:14,29-30).  They are Soderlind's H211b digital filter with the
// arctangent limiter (Soderlind, "Digital filters in adaptive time-stepping", ACM TOMS 29
// (2003) 1-26; Soderlind & Wang, JCAM 185 (2006) 225-243): with c = tol / err the ratio of
// the tolerance to the step's error estimate, rho the ratio of the last two steps and k the
// estimator's order, the next step is h * limit(filter(c_n, c_{n-1}, rho, k)),
//     filter = c_n^(1/(b k)) c_{n-1}^(1/(b k)) rho^(-1/k),  b = 4,
//     limit(u) = 1 + kappa atan((u - 1) / kappa),  kappa = 1,
// both exactly as the reference writes them (its rho exponent is -1/k, where Soderlind's
// H211b has -1/b).  The limiter keeps every step ratio inside (1 - pi/2, 1 + pi/2).
//
// trhepd_adaptive_steps puts the controller on the device integration: it advances a
// resident batch of pilot angles slice by slice from the top of the slab, and estimates each
// slice's error by step doubling: one slice of h and two of h/2 from the same state, both
// renormalised at their end (so that P alone is the state, with Q = I), compared as
// max over angles of max|P_h - P_h/2| / max|P_h/2|.  A slice whose estimate exceeds tol is
// retried with the controller's smaller step; an accepted one is kept (the state continues
// from the two half slices) and the controller sets the next step.  The list of accepted
// thicknesses is the output; trhepd_table_build_steps builds the table for it, and
// trhepd_propagate_batch integrates every angle of the scan over it.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

extern "C" double trhepd_h211b_limit(double u)
{
    const double kappa = 1.0;
    return 1.0 + kappa * std::atan((u - 1.0) / kappa);
}

extern "C" double trhepd_h211b_filter(double c1, double c0, double rho, int k)
{
    const int b = 4;
    return std::pow(c1, 1.0 / (b * k)) * std::pow(c0, 1.0 / (b * k)) * std::pow(rho, -1.0 / k);
}

using namespace trhepd;

namespace {

// the integration of ns slices of a slice list from depth z on a resident batch, each slice
// followed by its renormalisation decision (the last slice always renormalises)
int run_slices(trhepd_ctx *ctx, int batch_id, const trhepd_potential *pot, int method, int ns, const double *hs,
               double z)
{
    int id = 0;
    int rc = trhepd_table_build_steps(ctx, pot, method, ns, hs, z, &id);
    if (rc) return rc;
    Batch *b = find_batch(ctx, batch_id);
    const Table *t = find_table(ctx, id);
    for (int l = 1; l <= ns && !rc; ++l) {
        rc = runner_slice_enqueue(ctx, *b, *t, l);
        if (!rc) rc = runner_slice_finish(ctx, *b, *t);
    }
    const hipError_t e = hipStreamSynchronize(b->stream);   // before the table goes
    if (!rc && e) rc = hip_check(ctx, e, "adaptive slices");
    const int rc2 = trhepd_table_release(ctx, id);
    return rc ? rc : rc2;
}

}  // namespace

extern "C" int trhepd_adaptive_steps(trhepd_ctx *ctx, const trhepd_potential *pot, int method, double zmax,
                                     double tol, double h0, double hmin, double hmax, int nb, int nang,
                                     const int *iv, const double *gamma, const double *f, int nb0, int max_steps,
                                     double *steps, int *nsteps, int *nforced)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (method != TRHEPD_SP4 && method != TRHEPD_SP6)
        return fail(ctx, TRHEPD_E_ARG, "adaptive_steps: int_method 5 or 6, not %d", method);
    if (nb <= 48 || nb > 600)
        return fail(ctx, TRHEPD_E_ARG, "adaptive_steps: the pilot runs on the batched path, 48 < N <= 600 (N = %d)",
                    nb);
    if (!(zmax > 0.0) || !(tol > 0.0) || !(hmin > 0.0) || !(h0 >= hmin) || !(hmax >= h0) || !std::isfinite(zmax) ||
        !std::isfinite(hmax) || max_steps < 1 || !steps || !nsteps || nang < 1)
        return fail(ctx, TRHEPD_E_ARG, "adaptive_steps: need 0 < hmin <= h0 <= hmax, zmax > 0, tol > 0, max_steps >= 1");
    const int order = method == TRHEPD_SP4 ? 4 : 6, k = order + 1;
    TRY_HIP(ctx, hipSetDevice(ctx->device));

    // the resident batch of pilot angles (its table only serves the batch's creation)
    int tid0 = 0, bid = 0;
    double h = std::min(h0, zmax);
    int rc = trhepd_table_build_steps(ctx, pot, method, 1, &h, 0.0, &tid0);
    if (rc) return rc;
    rc = trhepd_batch_create(ctx, tid0, nb, nang, iv, gamma, f, nb0, nullptr, &bid);
    if (rc) {
        (void)trhepd_table_release(ctx, tid0);
        return rc;
    }
    Batch *b = find_batch(ctx, bid);
    const size_t sb = runner_state_bytes(*b);
    char *buf = nullptr;
    const size_t bytes = 2 * sb + (size_t)nang * 8;
    rc = dev_alloc(ctx, (void **)&buf, bytes);
    double2 *saved = (double2 *)buf, *single = (double2 *)(buf + sb);
    double *work = (double *)(buf + 2 * sb);
    if (!rc) rc = batch_prelude(ctx, *b);
    if (!rc) rc = runner_begin(ctx, *b, *find_table(ctx, tid0));

    double z = 0.0, h_prev = h, c_prev = -1.0;
    int n = 0, forced = 0;
    bool retry = false;   // the last attempt at depth z was rejected
    long trials = 0;
    while (!rc && zmax - z > 1e-12 * zmax) {
        if (++trials > 50L * max_steps) {
            rc = fail(ctx, TRHEPD_E_ARG, "adaptive_steps: no progress at depth %g of %g", z, zmax);
            break;
        }
        const double rest = zmax - z;
        h = std::min(h, rest);
        // no sliver below hmin at the bottom: take the rest on a first attempt, else half of it
        if (rest - h > 0.0 && rest - h < hmin) h = (!retry && rest <= hmax) ? rest : 0.5 * rest;
        rc = runner_save_p(ctx, *b, saved);
        if (!rc) rc = run_slices(ctx, bid, pot, method, 1, &h, z);
        if (!rc) rc = runner_save_p(ctx, *b, single);
        if (!rc) rc = runner_load_p(ctx, *b, saved);
        const double half[2] = {0.5 * h, 0.5 * h};
        if (!rc) rc = run_slices(ctx, bid, pot, method, 2, half, z);
        double err = 0.0;
        if (!rc) rc = runner_p_difference(ctx, *b, single, work, &err);
        if (rc) break;
        if (!std::isfinite(err)) {
            rc = fail(ctx, TRHEPD_E_NONFINITE, "adaptive_steps: the step error at depth %g (h = %g) is not finite", z, h);
            break;
        }
        const double c = err > 0.0 ? tol / err : HUGE_VAL;
        const bool at_min = h <= hmin * (1.0 + 1e-12);
        if (err <= tol || at_min) {
            if (n == max_steps) {
                rc = fail(ctx, TRHEPD_E_ARG, "adaptive_steps: more than max_steps = %d slices (depth %g of %g)",
                          max_steps, z, zmax);
                break;
            }
            if (!(err <= tol)) ++forced;
            steps[n++] = h;
            z += h;
            const double c0 = c_prev < 0.0 ? c : c_prev, rho = n > 1 ? h / h_prev : 1.0;
            const double r = trhepd_h211b_limit(trhepd_h211b_filter(c, c0, rho, k));
            h_prev = h;
            c_prev = c;
            h = std::min(hmax, std::max(hmin, h * r));
            retry = false;
        } else {
            rc = runner_load_p(ctx, *b, saved);   // back to the state at z
            // the controller's ratio, and at least a tenth less, so that rejections end
            const double r = trhepd_h211b_limit(trhepd_h211b_filter(c, c, 1.0, k));
            h = std::max(hmin, h * std::min(0.9, r));
            retry = true;
        }
    }
    if (b && b->stream) (void)hipStreamSynchronize(b->stream);
    if (buf) dev_free(ctx, buf, bytes);
    const int rc2 = trhepd_batch_release(ctx, bid);
    const int rc3 = trhepd_table_release(ctx, tid0);
    if (rc) return rc;
    if (rc2) return rc2;
    if (rc3) return rc3;
    *nsteps = n;
    if (nforced) *nforced = forced;
    return TRHEPD_OK;
}
