// Internal declarations shared by the translation units of libtrhepd_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/trhepd.h"

namespace trhepd {

// Table entry: the two potential components of one Fourier coefficient at one depth.
struct VVi {
    double2 v;
    double2 vi;
};

// Coefficients of the symmetric RKN schemes (surf_prkn.f90:9-52, Blanes & Moan 2002):
// bab[2k-2] = c_b of stage k, bab[2k-1] = c_a of stage k, k = 1..s.
struct RknCoef {
    int s;
    double bab[23];
};

// The integrator stages the batched path runs, flattened for one slice.
RknCoef rkn_coef(int method);

struct Table {
    int id = 0;
    int method = 0;
    int nv = 0;
    int ncols = 0;
    int ns = 0;
    double dz = 0.0;            // the slice thickness (the largest one of a slice list)
    VVi *d = nullptr;   // [ncols][nv]
    size_t bytes = 0;
    // RKN tables built from a list of slice thicknesses (trhepd_table_build_steps): slice l
    // (1-based) is steps[l-1] thick and starts at z0 + steps[0] + ... + steps[l-2]; empty for
    // the uniform schedule.  d_steps is the same list on the device, behind the columns.
    std::vector<double> steps;
    double z0 = 0.0;
    const double *d_steps = nullptr;
};

// thickness of slice l (1-based; l past the last slice gives the last one's)
inline double slice_dz(const Table &t, int l)
{
    if (t.steps.empty()) return t.dz;
    return t.steps[(size_t)std::min(std::max(l, 1), t.ns) - 1];
}

struct Stats {
    long long n_renorm = 0;
    long long n_products = 0;
    double device_ms = 0.0;
};

struct Runner;   // working state of the batched large-N path (batched.hip)
struct Server;   // request loop state of the serving entry points (serve.cpp)

// a batch of angles with device-resident inputs, outputs and working memory; each batch
// runs on its own stream so that several batches overlap on the device
struct Batch {
    int id = 0, table_id = 0, nb = 0, nang = 0, nb0 = 1;
    bool has_f2 = false;
    char *d = nullptr;
    size_t bytes = 0;
    int *gmap = nullptr;
    double2 *gamma = nullptr, *fin = nullptr, *fout = nullptr;
    double *f2 = nullptr, *w = nullptr;
    int *cnt = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evs = nullptr;   // orders this batch's stream against a launch shared with others
    Runner *run = nullptr;
    Stats stats;
    double gim_max = 0.0;   // max |Im gamma| over beams and angles (evanescent growth rate)
};

// renormalisation period of a batch under the context's schedule (0: Gershgorin rule)
int batch_period(const trhepd_ctx *ctx, const Batch &b, const Table &t);

// a captured multi-batch run (trhepd_graph_create)
struct GraphRec {
    int id = 0;
    std::vector<int> batch_ids;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;   // fork / join events of the capture
};

// event pairs bracketing stage-product launches when profiling is on
struct Prof {
    bool on = false;
    std::vector<hipEvent_t> ev;
    int used = 0;
    long long launches = 0;
    double ms = 0.0;
};

// A device block the library has released: it may be handed out again once every stream
// that could still use it has passed the point of release (one event per stream).
struct PoolBlock {
    void *p = nullptr;
    size_t bytes = 0;
    std::vector<hipEvent_t> ev;
};

}  // namespace trhepd

struct trhepd_ctx {
    int device = 0;
    int precision = 0;
    size_t budget = 0;
    size_t held = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    int next_table = 1;
    std::vector<trhepd::Table> tables;
    // device address of the file-scope work area of trhepd_propagate_batch_noalloc on this
    // context's device (null if it could not be obtained; noalloc_err then holds the status)
    void *noalloc = nullptr;
    hipError_t noalloc_err = hipSuccess;
    trhepd::Stats stats;
    int next_batch = 1;
    int renorm_period = 0;   // 0: reference rule (cond_est > 1e3); K > 0: every K slices
    int mixed_digits = 7;    // digits per operand of the sliced-integer stage products
    int bulk_one_wg_max = 64;   // bulk groups of more beams run on the batched path
    std::vector<trhepd::Batch> batches;
    trhepd::Prof prof;
    // autotuned stage-product tiles: (precision, N, angles, digits) -> variant index
    std::map<std::array<int, 4>, int> tuned;
    trhepd::Server *serve = nullptr;
    std::vector<trhepd::GraphRec> graphs;
    int next_graph = 1;
    bool capturing = false;   // a stream capture is open: no host synchronisation, no timing events
    // batches run together start a fraction of the renormalisation period apart: 1 on, 0 off,
    // -1 on for the mixed precision only
    int phase_offsets = -1;
    // released device blocks kept for reuse (hipFree waits for the whole device): pending
    // until their events complete, then ready; pool_bytes counts both against the budget
    std::vector<trhepd::PoolBlock> pool_pending;
    std::multimap<size_t, void *> pool_ready;
    std::map<void *, size_t> pool_size;   // size of every block dev_alloc has handed out
    std::vector<hipEvent_t> pool_events;  // spare events
    size_t pool_bytes = 0;
    bool pool_off = false;                // the context is being destroyed: hipFree at once
};

namespace trhepd {

int fail(trhepd_ctx *ctx, int code, const char *fmt, ...);
int hip_check(trhepd_ctx *ctx, hipError_t e, const char *what);
#define TRY_HIP(ctx, call)                                               \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) return ::trhepd::hip_check(ctx, e_, #call); \
    } while (0)

// device allocation accounted against the context budget
int dev_alloc(trhepd_ctx *ctx, void **p, size_t bytes);
void dev_free(trhepd_ctx *ctx, void *p, size_t bytes);
// looks up this device's copy of the no-allocation entry point's work area (device current)
void noalloc_attach(trhepd_ctx *ctx);

Table *find_table(trhepd_ctx *ctx, int id);
Batch *find_batch(trhepd_ctx *ctx, int id);
// counters cleared, F reset to the uploaded input, start event (the first step of batch_start)
int batch_prelude(trhepd_ctx *ctx, Batch &b);
// starts a batch run on its stream: counters cleared, F reset to the uploaded input, then
// the boundary transform (large N) or the whole integration (small N)
int batch_start(trhepd_ctx *ctx, Batch &b, const Table &t);
// enqueues complete runs of n batches (start, interleaved slices, end) on their streams
int enqueue_batches(trhepd_ctx *ctx, const std::vector<Batch *> &bs, const std::vector<const Table *> &ts);

void server_destroy(trhepd_ctx *ctx);

// profiling: record an event on `stream` before / after a timed launch (no-op when off)
int prof_mark(trhepd_ctx *ctx, hipStream_t stream);
// accumulate the elapsed times of the recorded pairs (call after the stream is idle)
int prof_collect(trhepd_ctx *ctx);

// plane-group orbit of a fractional position (strfac.f90:15-65 and its 17 group routines)
int orbit(int nsg, double x, double y, double *xs, double *ys);

// kernels launchers ---------------------------------------------------------
int launch_asfcef(trhepd_ctx *ctx, int nv, int nelm, const int *igh, const int *igk, double ghx,
                  double ghy, double gky, int nabr, int nabi, const double *a, const double *c,
                  const double *dth, const double *dtk, double *asf, double *ghk);
int launch_strfac(trhepd_ctx *ctx, int natm, int nv, const double *gh, const double *gk,
                  const int *ndeg, const double *xb, const double *yb, double dx, double dy,
                  double *st);
int launch_table(trhepd_ctx *ctx, const trhepd_potential *pot, Table &t);

// Sliced-integer stage products (precision TRHEPD_PREC_MIXED) of the batched large-N path.
// Operands are split into five signed base-2^7 int8 digits under a power-of-two scale per
// row of Q and per column of M0; digit products are exact in int32 (v_mfma_i32_16x16x64_i8)
// and the complex product takes three real products (Qr Mr, Qi Mi, (Qr+Qi)(Mr+Mi)).
// Digit planes are stored pre-tiled: plane (m, p) holds, for each 16-row (16-column) block
// and 64-wide k chunk, the 64 lanes x 16 bytes of one MFMA operand fragment.
struct MixedGeo {
    int A, N, NP;
    size_t mat;   // NP*NP
    int KC;       // 64-wide k chunks, ceil(N / 64)
    int RB;       // 16-row blocks of [Q_1; ...; Q_A]: A*NP/16
    int RBP;      // RB rounded up to a multiple of 8 (row blocks per workgroup: WR)
    int CBP;      // 16-column blocks of M0 (NP/16) rounded up to a multiple of WC
    int WC;       // column blocks per workgroup tile (2..4)
    int WR;       // row blocks (waves) per workgroup tile (4 or 8)
    int ND;       // digits per operand (5..7)
    int PF;       // chunks of Q prefetch (1 or 2)
};
// tiles with room in 256 VGPRs for a third set of Q fragments
constexpr bool mixed_pf2_fits(int ND, int WC) { return WC * ND <= 18; }
// wc / wr / pf: column / row blocks per workgroup tile, Q prefetch depth (0: the default)
MixedGeo mixed_geo(int A, int N, int NP, int ND, int wc = 0, int wr = 0, int pf = 0);
size_t mixed_qd_bytes(const MixedGeo &g);   // digits of Q: 3 ND planes
size_t mixed_md_bytes(const MixedGeo &g);   // digits of one stage's M0: 3 ND planes
struct MixedStage {
    MixedGeo g;
    const int8_t *qd;
    const int *eq;       // [RBP*16] row scale exponents of Q
    const int8_t *md;
    const int *em;       // [CBP*16] column scale exponents of M0
    const double2 *Q;
    double2 *Qn;
    double2 *P;
    const double2 *g2c;
    const VVi *col;
    double alpha, beta;
    const double *coef;
    const int *xid;      // mode 0: angles whose drift is skipped (beta 0: Qn = Q), or null
    int mode;
    double *colsum, *diagabs;
    unsigned long long *rmax;   // mode 0/1: per-row maxima of Qn for the next split (bit patterns), or null
};
// digits of the off-diagonal slice matrices of `nst` consecutive table columns starting at
// `col0`, into md[st] / em[st]
void launch_msplit(hipStream_t st, const MixedGeo &g, const int *gmap, const VVi *col0, int nv, int nst,
                   int8_t *md, int *em);
// digits of Q (row scales into eq).  vmax: Q's per-row maxima when the kernel that wrote Q
// gathered them (one pass), else null (a pass of its own); clear: [RBP*16] zeroed for the
// next stage's maxima, or null
void launch_qsplit(hipStream_t st, const MixedGeo &g, const double2 *Q, int8_t *qd, int *eq,
                   const unsigned long long *vmax = nullptr, unsigned long long *clear = nullptr);
// P += alpha (Q M0 + Q diag(d)) from the digits, FP64 epilogue as the FP64 stage kernel
void launch_stage_i8(hipStream_t st, const MixedStage &S);

// slice-loop drivers
int run_small(trhepd_ctx *ctx, hipStream_t stream, int period, const Table &t, int nb, int nang, const int *d_gmap,
              const double2 *d_gamma, double2 *d_f, double *d_f2, int nb0, const double *d_wnsga2,
              int *d_counters);
// several small-N batches (same N, integrator, dz and nb0; any tables and slice counts) in
// one launch of the small-N kernel, one workgroup per (batch, angle)
struct SmallBatch {
    const Table *t;
    int nb, nang, nb0, period;
    const int *gmap;
    const double2 *gamma;
    double2 *f;
    double *f2;
    const double *w;
    int *cnt;
};
bool small_fusable(int nb, int method);
int run_small_multi(trhepd_ctx *ctx, hipStream_t stream, const std::vector<SmallBatch> &bs);
// blkref for one group of more than 64 beams, batched over angles (device pointers; see
// trhepd_bulk_reflect); out + a * out_stride receives angle a's n x n block
int bulk_large_group(trhepd_ctx *ctx, hipStream_t st, const Table &t, int n, int boff, int nang, const int *ivg,
                     int nbgm, const double2 *gamma, int nb, const double2 *phase, double gmax2, int ml, double epsb,
                     double2 *out, size_t out_stride, int *layers);
// batched large-N path, split into phases so that several batches interleave:
// runner_create allocates the batch's working memory; begin enqueues the boundary
// transform; slice_enqueue(l) the stages of slice l up to the end kick (the renormalisation
// flags are copied to the host behind an event); slice_finish waits for the flags and
// enqueues the LU solves; end enqueues the back transform and intensities.
int runner_create(trhepd_ctx *ctx, Batch &b, const Table &t);
// p[0..n) = v on a stream (a kernel, so that it can be captured into a graph)
void launch_fill_i32(hipStream_t st, int *p, int n, int v);
// picks the stage-product tile variant of a new batch (timed once per shape, then cached)
int autotune(trhepd_ctx *ctx, Batch &b, const Table &t);
void runner_destroy(trhepd_ctx *ctx, Batch &b);
// adaptive slice control: P of every angle to / from a device buffer of runner_state_bytes,
// and max over angles of max|P - ref| / max|P| (work: nang doubles on the device)
size_t runner_state_bytes(const Batch &b);
int runner_save_p(trhepd_ctx *ctx, Batch &b, double2 *dst);
int runner_load_p(trhepd_ctx *ctx, Batch &b, const double2 *src);
int runner_p_difference(trhepd_ctx *ctx, Batch &b, const double2 *ref, double *work, double *out);
int runner_begin(trhepd_ctx *ctx, Batch &b, const Table &t);
int runner_slice_enqueue(trhepd_ctx *ctx, Batch &b, const Table &t, int l);
int runner_slice_finish(trhepd_ctx *ctx, Batch &b, const Table &t);
int runner_end(trhepd_ctx *ctx, Batch &b, const Table &t);

}  // namespace trhepd
