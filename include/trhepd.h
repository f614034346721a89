/*
 * trhepd.h -- C ABI of libtrhepd_hip.so, the gfx950 (MI355X) engine for the
 * sim-trhepd-rheed surface slice loop.
 *
 * Conventions (identical to what the Fortran driver passes through iso_c_binding):
 *   - complex(8) arrays are interleaved (re, im) doubles, passed as double*;
 *   - every matrix is column-major with the Fortran shape given in the comment;
 *   - index maps (iv, ielm, iord) are 1-based, exactly as the driver holds them;
 *   - every entry point returns 0 on success and a non-zero TRHEPD_E_* status on
 *     failure; trhepd_last_error() then describes it.  Nothing aborts inside the library.
 *   - a context is used from one host thread at a time; different contexts are independent.
 *
 * Citations name the reference routine each entry point replaces, relative to the
 * reference repository root (sim-trhepd-rheed/src unless stated).
 */
#ifndef TRHEPD_H
#define TRHEPD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TRHEPD_OK = 0,
    TRHEPD_E_ARG = 1,        /* invalid argument (shape, method, pointer) */
    TRHEPD_E_HIP = 2,        /* a HIP runtime call failed */
    TRHEPD_E_NOMEM = 3,      /* device allocation failed */
    TRHEPD_E_BUDGET = 4,     /* request exceeds the context's HBM budget */
    TRHEPD_E_SINGULAR = 5,   /* exact zero pivot in an LU factorisation */
    TRHEPD_E_NONFINITE = 6,  /* NaN/Inf reached an output */
    TRHEPD_E_NOTFOUND = 7,   /* unknown table / request id */
    TRHEPD_E_PENDING = 8     /* serving: the request has not finished yet */
};

/* integrator codes: the int_method values of surflf.f90:28-41 */
enum {
    TRHEPD_M2 = 1, TRHEPD_M4 = 2, TRHEPD_M6 = 3, TRHEPD_LF = 4,
    TRHEPD_SP4 = 5, TRHEPD_SP6 = 6, TRHEPD_RK4 = 7, TRHEPD_RK4GS = 8
};

/* arithmetic of the propagator products */
enum {
    TRHEPD_PREC_FP64 = 0,    /* all-FP64 complex */
    TRHEPD_PREC_MIXED = 1    /* large-N stage products as exact int8 digit-sliced MFMA
                                (base-254 digits per operand, int32 accumulation, FP64
                                recombination); FP64 state, table, LU and boundary
                                transforms.  N <= 48, the Magnus methods, RK4 and rk4gs run
                                the FP64 path. */
};

typedef struct trhepd_ctx trhepd_ctx;

typedef struct {
    int device;              /* HIP device ordinal */
    int precision;           /* TRHEPD_PREC_* */
    size_t hbm_budget;       /* bytes this context may hold on the device, live and kept for
                                reuse (0: no limit) */
} trhepd_opts;

/* Potential-table inputs: the module mod_scpot3 globals (surfio.f90:18-27) that
 * comp_scpot (surfio.f90:29-41) reads.  Host pointers. */
typedef struct {
    int nv, nvb, natms, natmb, nelm, nabr, nabi;
    double cc, smesh;
    const int *ielm;         /* (natmb+natms), bulk atoms first, 1-based element index */
    const double *z;         /* (natmb+natms), zoffset added */
    const double *dchg;      /* (nelm) */
    const double *ghk;       /* (nv) */
    const double *ocr;       /* (natmb+natms) */
    const double *b;         /* (6, nelm) */
    const double *asf;       /* (nv, 6, nelm) */
    const double *st0;       /* complex (nv, natms) surface structure factors */
    const double *st1;       /* complex (nvb, natmb) top bulk unit */
    const double *st2;       /* complex (nvb, natmb) second bulk unit */
} trhepd_potential;

/* ---- context ------------------------------------------------------------ */
int trhepd_ctx_create(const trhepd_opts *opts, trhepd_ctx **out);
int trhepd_ctx_destroy(trhepd_ctx *ctx);
const char *trhepd_last_error(const trhepd_ctx *ctx);
/* bytes currently held on the device by the context's live tables, batches and work areas.
 * Released blocks are kept for reuse (the budget counts them too; without a budget at most
 * 8 GiB) and returned to the device by trhepd_ctx_destroy. */
size_t trhepd_ctx_bytes(const trhepd_ctx *ctx);

/* ---- per-structure setup (replaces scpot.f90:11-110, surfsub.f90:9-41) -------- */

/* asfcrr (scpot.f90:11-58): energy / Debye-Waller correction of the element
 * coefficients, in place; a, b, c are (6, nelm).  Host arithmetic (per-element scalars). */
int trhepd_asfcrr(double be, double *wn, int nelm, int nabr, int nabi, double *a, double *b,
                  double *c, double *dchg, double *dth, double *dtk, const double *dtz,
                  const double *elm, double tem, double smesh);

/* srfghk (surfsub.f90:9-41): beam-pair -> Fourier-component map.  igh/igk hold nvb
 * entries on entry and room for nvm; iv is (nb, nb), strict lower triangle set.
 * *nv receives the extended component count.  Host integer work. */
int trhepd_srfghk(int nvm, int nvb, int *igh, int *igk, int nb, const int *ih, const int *ik,
                  int *iv, int *nv);

/* asfcef (scpot.f90:62-95) on the device: asf (nv, 6, nelm) and ghk (nv). */
int trhepd_asfcef(trhepd_ctx *ctx, int nv, int nelm, const int *igh, const int *igk, double ghx,
                  double ghy, double gky, int nabr, int nabi, const double *a, const double *c,
                  const double *dth, const double *dtk, double *asf, double *ghk);

/* scpot_stf (scpot.f90:100-110) with strfac (strfac.f90:5-76): plane-group orbits on
 * the host, phase sums on the device.  st is complex (nv, natm). */
int trhepd_scpot_stf(trhepd_ctx *ctx, int natm, int nv, int nsg, int ma, int mb, int na, int nb,
                     const double *gh, const double *gk, const double *x, const double *y,
                     double dx, double dy, double *st);

/* ---- potential table (replaces the *_int_init of surflf.f90:28-41) ------------- */

/* Builds v, vi (nv, K) on the device for int_method `method` (K = table columns of
 * that method: hp_int_init surf_prkn.f90:174-197, m2/m4_int_init surf_magnus.f90:4-42,
 * rk4_int_init surf_rk.f90:4-25).  The table stays resident; *table_id names it. */
int trhepd_table_build(trhepd_ctx *ctx, const trhepd_potential *pot, double dz, int ns,
                       int method, int *table_id);
/* The table of an RKN integrator (int_method 5 or 6) over a list of ns slices of thickness
 * steps[0..ns-1], the first starting at depth z0 (0 for a whole slab): slice l starts at z0
 * plus the thicknesses before it, its stage depths follow hp_int_init (surf_prkn.f90:187-195)
 * with that slice's thickness.  trhepd_propagate_batch and the batch entry points run such a
 * table like any other, each slice with its own thickness.  No counterpart in the reference,
 * whose slices are uniform (surfio.f90:197-210); for slice lists chosen by
 * trhepd_adaptive_steps. */
int trhepd_table_build_steps(trhepd_ctx *ctx, const trhepd_potential *pot, int method, int ns,
                             const double *steps, double z0, int *table_id);
/* number of columns K and components nv of a table */
int trhepd_table_shape(trhepd_ctx *ctx, int table_id, int *nv, int *ncols);
/* copies a table back to the host: v, vi complex (nv, K) */
int trhepd_table_fetch(trhepd_ctx *ctx, int table_id, double *v, double *vi);
int trhepd_table_release(trhepd_ctx *ctx, int table_id);

/* ---- bulk stage (replaces blkref, bulk.f90:7-124, with trmat trmatg.f90:6-59 and
 *      gcmi3 gcmi.f90:5-15) ---------------------------------------------------------
 * Reflection matrices of the semi-infinite bulk for nang angles and ngr groups of coupled
 * beams.  table_id: the bulk potential at the ns slice midpoints of one unit layer (built
 * with int_method 1 or 4, whose schedule is the midpoints, from the three-unit descriptor of
 * bulkio.f90:189-196).  nbg: group sizes (each <= 600); iv: blkghk's in-group map
 * iv(nbgm, nb) (bulksub.f90:297-336); gamma: complex (nb, nang) for the beams in group
 * order; phase: complex (nb) column phases of the layer shift, or NULL when there is none.
 * Every group's unit-layer transfer matrix is the product of the slice matrices (computed
 * as matrix cosines and sines of the slice Hamiltonian, the same matrix trmat builds from
 * its eigen-decomposition), then R_l = (T11 R + T12)(T21 R + T22)^-1 runs for up to ml
 * layers until max |d|R_ii|^2| < epsb after layer 20.  blocks receives the n_g x n_g
 * complex blocks, column-major, in (angle, group) order; layers (nang x ngr, may be NULL)
 * the layer count each used. */
int trhepd_bulk_reflect(trhepd_ctx *ctx, int table_id, int ngr, const int *nbg, int nbgm, const int *iv,
                        int nang, const double *gamma, const double *phase, int ml, double epsb,
                        double *blocks, int *layers);
/* Groups of up to n beams (default 64, at most 64) run one workgroup per (angle, group);
 * larger groups (up to 600 beams) run batched over the angles on the slice loop's per-angle
 * product, elementwise and LU kernels.  n = 0 sends every group to the batched path.  The two
 * agree to rounding.  No counterpart in the reference (blkref has one code path). */
int trhepd_set_bulk_batched_above(trhepd_ctx *ctx, int n);

/* ---- slice loop (replaces the angle-loop body of srfref, surflf.f90:79-153) ---- */

/* One batch of nang independent angles that share table `table_id`:
 *   iv      (nb, nb)          beam-pair map from srfghk
 *   gamma   complex (nb, nang) z wave numbers (surflf.f90:94-103)
 *   f       complex (nb, nb, nang) in: bulk reflection scattered by iord
 *                                   out: surface reflection matrix
 *   f2      (nb, nang) out intensities (surflf.f90:143-151), may be NULL
 *   nb0     1-based index of the (0,0) beam
 *   wnsga2  (nang) K sin^2(theta) per angle
 * The integrator is the table's int_method (1-8), at every N up to 600. */
int trhepd_propagate_batch(trhepd_ctx *ctx, int table_id, int nb, int nang, const int *iv,
                           const double *gamma, double *f, double *f2, int nb0,
                           const double *wnsga2);

/* trhepd_propagate_batch on device-resident arrays, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  iv, gamma, f, f2 and wnsga2 are
 * DEVICE pointers with the shapes of trhepd_propagate_batch; f is updated in place.  The
 * library's own work (signed gather map, counters, flags) lives in a fixed device array of
 * 9472 bytes per device, so N <= 48 (the small-N path) and int_method 4-7 only; larger N or
 * other methods return TRHEPD_E_ARG.  All contexts on one device share that array: calls that
 * overlap on one device race.  The call synchronises the context's stream before it returns;
 * statuses are those of trhepd_propagate_batch (an index of iv outside 1..nv is refused before
 * f is touched). */
int trhepd_propagate_batch_noalloc(trhepd_ctx *ctx, int table_id, int nb, int nang, const int *iv,
                                   const double *gamma, double *f, double *f2, int nb0,
                                   const double *wnsga2);

/* The same slice loop with device-resident inputs, for callers that run a batch more
 * than once or keep several batches on the device (serving).  batch_create uploads the
 * inputs (same arguments as trhepd_propagate_batch); batch_run integrates all angles,
 * leaving F and f2 on the device (the uploaded F is kept, so a batch can be re-run);
 * batch_fetch copies the results out (either pointer may be NULL). */
int trhepd_batch_create(trhepd_ctx *ctx, int table_id, int nb, int nang, const int *iv,
                        const double *gamma, const double *f, int nb0, const double *wnsga2,
                        int *batch_id);
int trhepd_batch_run(trhepd_ctx *ctx, int batch_id);
/* Runs n batches concurrently, each on its own HIP stream; their slices interleave so that
 * one batch's renormalisation solves overlap the others' stage products. */
int trhepd_batch_run_many(trhepd_ctx *ctx, int n, const int *batch_ids);
int trhepd_batch_fetch(trhepd_ctx *ctx, int batch_id, double *f, double *f2);
int trhepd_batch_release(trhepd_ctx *ctx, int batch_id);

/* ---- adaptive slice thickness (the unused controller of h211b_control.f90:1-16) ---- */

/* h211b_limit (h211b_control.f90:1-7): 1 + atan(u - 1), the limiter of a step ratio. */
double trhepd_h211b_limit(double u);
/* h211b_filter (h211b_control.f90:9-16): c1^(1/(4k)) c0^(1/(4k)) rho^(-1/k), the H211b
 * filter of the last two tolerance-to-error ratios c1 (latest), c0 and step ratio rho. */
double trhepd_h211b_filter(double c1, double c0, double rho, int k);
/* Chooses slice thicknesses for the slab [0, zmax] with the controller above: pilot angles
 * (nb, nang, iv, gamma, f, nb0 as for trhepd_propagate_batch; 48 < nb <= 600) advance slice
 * by slice on the device; each slice's error is the step-doubling estimate
 * max over angles of max|P_h - P_h/2| / max|P_h/2| of the ratio-form state after one slice of
 * h and after two of h/2 (both renormalised), and the slice is accepted when it is <= tol,
 * or when h is hmin (counted in *nforced, which may be NULL).  The next step is
 * h * h211b_limit(h211b_filter(tol/err, previous tol/err, h/h_previous, p + 1)) for the
 * method's order p (4 for sp4, 6 for sp6), kept within [hmin, hmax]; a rejected slice is
 * retried at that ratio, at most 0.9.  steps receives the *nsteps accepted thicknesses (at
 * most max_steps; they sum to zmax).  No counterpart in the reference (it defines the
 * controller but integrates uniform slices). */
int trhepd_adaptive_steps(trhepd_ctx *ctx, const trhepd_potential *pot, int method, double zmax,
                          double tol, double h0, double hmin, double hmax, int nb, int nang,
                          const int *iv, const double *gamma, const double *f, int nb0,
                          int max_steps, double *steps, int *nsteps, int *nforced);

/* ---- captured runs: batches of uneven beam counts in one launch ------------------
 * No counterpart in the reference (it runs one deck per process).  trhepd_graph_create
 * captures the complete integration of n resident batches (any mix of tables and beam
 * counts: exactly what trhepd_batch_run_many enqueues, every batch on its own stream) into
 * one HIP graph.  trhepd_graph_launch replays it as a single launch on the context's stream
 * and waits for it; every batch's F and f2 are then on the device (trhepd_batch_fetch), and
 * trhepd_last_stats reports the launch.  Batches with N > 48 need a periodic
 * renormalisation schedule (trhepd_set_renorm_period with K > 0 or -1) when the graph is
 * created, so that no decision needs the host.  A batch held by a graph cannot be released
 * before the graph.
 *
 * Stage-product tiles: every large-N batch picks its stage-product tile variant when it is
 * created, by timing the candidates on its own shape (FP64: 32, 48 or 64 columns per tile;
 * mixed: 2-4 column blocks x 4 or 8 row blocks); the choice is cached per shape. */
int trhepd_graph_create(trhepd_ctx *ctx, int n, const int *batch_ids, int *graph_id);
int trhepd_graph_launch(trhepd_ctx *ctx, int graph_id);
int trhepd_graph_release(trhepd_ctx *ctx, int graph_id);

/* statistics of the last propagate / batch_run call: renormalisations (LU solves) summed
 * over the batch, stage products executed, and the device time in milliseconds. */
int trhepd_last_stats(const trhepd_ctx *ctx, long long *n_renorm, long long *n_products,
                      double *device_ms);

/* Renormalisation schedule of the ratio-form integrators.  The reference renormalises an
 * angle when the Gershgorin estimate of cond(X^H) exceeds 1e3 (surf_prkn.f90:317-324);
 * that rule is the default (period = 0).  A period K > 0 renormalises every angle after
 * every K-th slice instead (and after the last): the change of variables is exact, so
 * results agree to rounding, and with no data-dependent branch the whole batch is
 * enqueued without host round trips.  period = -1 derives K per batch from the fastest
 * evanescent growth: K = max(1, floor(ln(1e3) / (max |Im gamma| * dz))), the number of
 * slices over which the most strongly growing mode alone gains a factor 1e3.  Applies to
 * later batch runs of the context. */
int trhepd_set_renorm_period(trhepd_ctx *ctx, int period);

/* Phase offsets between large-N batches run together (trhepd_batch_run_many).  No
 * counterpart in the reference (one deck per process).  With a periodic renormalisation
 * schedule every batch factorises at the same slices; with offsets on, the j-th of m
 * batches starts j/m of the first batch's period later (its stream waits for the first
 * batch's slice), so that its LU chains overlap the other batches' stage products instead
 * of each other.  Results do not depend on it; captured runs (trhepd_graph_create) and the
 * reference rule (period 0) start every batch at once.  mode: 1 on, 0 off, -1 the default
 * (on for TRHEPD_PREC_MIXED contexts, off for FP64). */
int trhepd_set_phase_offsets(trhepd_ctx *ctx, int mode);

/* Digits per operand of the mixed-precision stage products (5, 6 or 7; default 7).  The
 * digits are base 254 (int8 range): each adds 8 bits, the normwise error of one stage
 * product is about digits * 254^-digits, and the int8 matrix work grows as
 * digits (digits + 1) / 2.  Applies to batches created afterwards on a TRHEPD_PREC_MIXED
 * context. */
int trhepd_set_mixed_digits(trhepd_ctx *ctx, int digits);

/* ---- serving: continuous batching of rocking-curve requests --------------------
 * No counterpart in the reference (its driver runs one deck per process); each request
 * replaces one angle loop of srfref (surflf.f90:79-153) for one structure.
 *
 * trhepd_serve_submit queues a request: the potential descriptor of its structure (as for
 * trhepd_table_build), dz, ns and int_method (4, 5, 6 or 7), and the angle batch
 * (as for trhepd_batch_create; wnsga2 is required).  Every input is copied, so the caller's
 * arrays may be freed on return.  Nothing touches the device until admission.
 *
 * trhepd_serve_step runs one step of the loop: requests whose last kernels have finished
 * are copied back and their device memory is released; queued requests are admitted in
 * arrival order while their table and batch fit the context's HBM budget (a request that
 * cannot fit even with nothing else admitted fails with TRHEPD_E_BUDGET); the admitted
 * requests are advanced, each on its own stream, so their kernels overlap: one slice each
 * under the reference's renormalisation rule, up to 16 under a periodic schedule (at most
 * 4 slices ahead of the device).  A periodic-schedule request whose remaining stage work is
 * more than 8 times the smallest active request's yields (enqueues nothing new) until it has
 * yielded 2 s in total.  The counts of admitted, queued and finished-but-uncollected
 * requests are returned.
 *
 * trhepd_serve_result returns TRHEPD_E_PENDING until the request has finished; then it
 * copies f (nb, nb, nang) and f2 (nb, nang) (either may be NULL), the latency from submit
 * to completion in milliseconds, and forgets the request.  A failed request returns its
 * status and message once. */
int trhepd_serve_submit(trhepd_ctx *ctx, const trhepd_potential *pot, double dz, int ns, int method,
                        int nb, int nang, const int *iv, const double *gamma, const double *f,
                        int nb0, const double *wnsga2, int *request_id);
int trhepd_serve_step(trhepd_ctx *ctx, int *n_active, int *n_queued, int *n_done);
int trhepd_serve_result(trhepd_ctx *ctx, int request_id, double *f, double *f2, double *latency_ms);
/* The split of the last request trhepd_serve_result returned (finished or failed): time queued
 * (submit to admission) and time in service (admission to observed completion), ms.  Their sum
 * is its latency. */
int trhepd_serve_last_timing(trhepd_ctx *ctx, double *queued_ms, double *service_ms);
/* Merging of requests of one structure (no counterpart in the reference).  With max_angles > 0,
 * the queued requests that share the head of the queue's potential descriptor, dz, ns,
 * int_method, beam set (nb, iv) and nb0, and whose renormalisation schedule is the same (the
 * derived period depends on each request's largest |Im gamma|), are admitted with it as one
 * batch of at most max_angles angles; a merged batch that does not fit the HBM budget is
 * retried with half of them.  The head waits until window_ms after its submission before
 * admission, so that later requests can join it.  Results are bit-identical to those of the
 * requests served one by one; latency and timing are reported per request.  max_angles 0 (the
 * default) turns merging off. */
int trhepd_serve_set_merge(trhepd_ctx *ctx, int max_angles, double window_ms);

/* Kernel timing: when enabled, every stage-product kernel launch is bracketed by HIP
 * events on the context's stream; trhepd_kernel_stats returns the number of such
 * launches and their summed duration (ms) since the last enable. */
int trhepd_profile(trhepd_ctx *ctx, int enable);
int trhepd_kernel_stats(trhepd_ctx *ctx, long long *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif /* TRHEPD_H */
