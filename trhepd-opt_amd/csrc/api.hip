// C ABI of libtrhepd_hip.so: context, memory accounting and the entry points that
// replace the reference's per-structure setup, *_int_init and the angle loop of srfref.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>

#include "internal.hpp"

namespace trhepd {

int fail(trhepd_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int hip_check(trhepd_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess) return TRHEPD_OK;
    return fail(ctx, e == hipErrorOutOfMemory ? TRHEPD_E_NOMEM : TRHEPD_E_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

// Released device blocks are kept for reuse instead of being returned with hipFree, which
// waits for all work on the device: a context that runs several batches at once (serving)
// releases blocks while other batches' kernels are queued.  dev_free records an event on
// the context's stream and on every batch stream, the streams any kernel using the block
// was launched on, and the block is handed out again only after all of them have completed.

// pending blocks whose events have all completed become ready
static void pool_poll(trhepd_ctx *ctx)
{
    for (size_t i = 0; i < ctx->pool_pending.size();) {
        PoolBlock &b = ctx->pool_pending[i];
        bool done = true;
        for (hipEvent_t e : b.ev) {
            const hipError_t q = hipEventQuery(e);
            if (q == hipErrorNotReady) {
                done = false;
                break;
            }
        }
        if (!done) {
            ++i;
            continue;
        }
        for (hipEvent_t e : b.ev) ctx->pool_events.push_back(e);
        ctx->pool_ready.emplace(b.bytes, b.p);
        ctx->pool_pending.erase(ctx->pool_pending.begin() + i);
    }
}

// returns ready blocks to the device until the pool holds at most `keep` bytes
static void pool_trim(trhepd_ctx *ctx, size_t keep)
{
    while (ctx->pool_bytes > keep && !ctx->pool_ready.empty()) {
        auto it = std::prev(ctx->pool_ready.end());   // the largest first
        (void)hipFree(it->second);
        ctx->pool_bytes -= it->first;
        ctx->pool_ready.erase(it);
    }
}

// every pooled block back to the device (the device idle first)
static void pool_release_all(trhepd_ctx *ctx)
{
    (void)hipDeviceSynchronize();
    pool_poll(ctx);
    pool_trim(ctx, 0);
    for (PoolBlock &b : ctx->pool_pending) {   // events that failed to complete
        (void)hipFree(b.p);
        for (hipEvent_t e : b.ev) (void)hipEventDestroy(e);
    }
    ctx->pool_pending.clear();
    for (hipEvent_t e : ctx->pool_events) (void)hipEventDestroy(e);
    ctx->pool_events.clear();
    ctx->pool_bytes = 0;
}

int dev_alloc(trhepd_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return TRHEPD_OK;
    if (ctx->budget && ctx->held + bytes > ctx->budget)
        return fail(ctx, TRHEPD_E_BUDGET, "allocation of %zu bytes exceeds the HBM budget (%zu of %zu held)",
                    bytes, ctx->held, ctx->budget);
    pool_poll(ctx);
    // a released block of at least the size and at most half again (2 MiB more for small ones)
    auto it = ctx->pool_ready.lower_bound(bytes);
    if (it != ctx->pool_ready.end() && it->first - bytes <= std::max(bytes / 2, (size_t)2 << 20) &&
        (!ctx->budget || ctx->held + it->first <= ctx->budget)) {
        *p = it->second;
        ctx->pool_bytes -= it->first;
        ctx->held += it->first;
        ctx->pool_size[*p] = it->first;
        ctx->pool_ready.erase(it);
        return TRHEPD_OK;
    }
    // the pool counts against the budget beside the live blocks
    if (ctx->budget) {
        const size_t room = ctx->budget - ctx->held - bytes;
        if (ctx->pool_bytes > room) {
            pool_trim(ctx, room);
            if (ctx->pool_bytes > room) {   // pending blocks: wait for the device, then trim
                (void)hipDeviceSynchronize();
                pool_poll(ctx);
                pool_trim(ctx, room);
            }
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory && ctx->pool_bytes) {
        (void)hipGetLastError();
        (void)hipDeviceSynchronize();
        pool_poll(ctx);
        pool_trim(ctx, 0);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_check(ctx, e, "hipMalloc");
    }
    ctx->held += bytes;
    ctx->pool_size[*p] = bytes;
    return TRHEPD_OK;
}

void dev_free(trhepd_ctx *ctx, void *p, size_t bytes)
{
    if (!p) return;
    auto it = ctx->pool_size.find(p);
    const size_t sz = it != ctx->pool_size.end() ? it->second : bytes;
    if (it != ctx->pool_size.end()) ctx->pool_size.erase(it);
    ctx->held -= sz;
    // inside a stream capture an event record would become part of the graph
    bool pooled = !ctx->capturing && !ctx->pool_off;
    PoolBlock b;
    b.p = p;
    b.bytes = sz;
    auto mark = [&](hipStream_t st) {
        if (!pooled) return;
        hipEvent_t e = nullptr;
        if (!ctx->pool_events.empty()) {
            e = ctx->pool_events.back();
            ctx->pool_events.pop_back();
        } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            pooled = false;
            return;
        }
        b.ev.push_back(e);
        if (hipEventRecord(e, st) != hipSuccess) pooled = false;
    };
    mark(ctx->stream);
    for (const Batch &bt : ctx->batches) mark(bt.stream);
    if (!pooled) {
        (void)hipGetLastError();
        for (hipEvent_t e : b.ev) ctx->pool_events.push_back(e);
        (void)hipDeviceSynchronize();
        (void)hipFree(p);
        return;
    }
    ctx->pool_pending.push_back(std::move(b));
    ctx->pool_bytes += sz;
    // without a budget the pool is held to a fixed size, so that it does not keep memory
    // other users of the device need
    constexpr size_t kPoolCap = (size_t)8 << 30;
    if (!ctx->budget && ctx->pool_bytes > kPoolCap) {
        pool_poll(ctx);
        pool_trim(ctx, kPoolCap);
    }
}

int prof_mark(trhepd_ctx *ctx, hipStream_t stream)
{
    if (ctx->capturing) return TRHEPD_OK;
    Prof &P = ctx->prof;
    if (!P.on) return TRHEPD_OK;
    if (P.used == (int)P.ev.size()) {
        hipEvent_t e;
        TRY_HIP(ctx, hipEventCreate(&e));
        P.ev.push_back(e);
    }
    TRY_HIP(ctx, hipEventRecord(P.ev[P.used++], stream));
    return TRHEPD_OK;
}

int prof_collect(trhepd_ctx *ctx)
{
    Prof &P = ctx->prof;
    for (int i = 0; i + 1 < P.used; i += 2) {
        float ms = 0;
        TRY_HIP(ctx, hipEventElapsedTime(&ms, P.ev[i], P.ev[i + 1]));
        P.ms += ms;
        P.launches++;
    }
    P.used = 0;
    return TRHEPD_OK;
}

int batch_period(const trhepd_ctx *ctx, const Batch &b, const Table &t)
{
    if (ctx->renorm_period >= 0) return ctx->renorm_period;
    const double rate = b.gim_max * t.dz;
    if (!(rate > 0.0)) return t.ns;
    return std::max(1, (int)std::floor(std::log(1e3) / rate));
}

Table *find_table(trhepd_ctx *ctx, int id)
{
    for (Table &t : ctx->tables)
        if (t.id == id) return &t;
    return nullptr;
}

// surf_prkn.f90:11-52 (Blanes & Moan, JCAM 142 (2002) 313-330)
RknCoef rkn_coef(int method)
{
    static const double b4[13] = {
        0.0829844064174052, 0.245298957184271, 0.396309801498368, 0.604872665711080,
        -0.0390563049223486, -0.350171622895351, 0.1195241940131508, -0.350171622895351,
        -0.0390563049223486, 0.604872665711080, 0.396309801498368, 0.245298957184271,
        0.0829844064174052};
    static const double b6[23] = {
        0.0414649985182624, 0.123229775946271, 0.198128671918067, 0.290553797799558,
        -0.0400061921041533, -0.127049212625417, 0.0752539843015807, -0.246331761062075,
        -0.0115113874206879, 0.357208872795928, 0.2366699247869311, 0.20477705429147,
        0.2366699247869311, 0.357208872795928, -0.0115113874206879, -0.246331761062075,
        0.0752539843015807, -0.127049212625417, -0.0400061921041533, 0.290553797799558,
        0.198128671918067, 0.123229775946271, 0.0414649985182624};
    RknCoef c{};
    if (method == TRHEPD_SP4) {
        c.s = 6;
        std::memcpy(c.bab, b4, sizeof b4);
    } else if (method == TRHEPD_SP6) {
        c.s = 11;
        std::memcpy(c.bab, b6, sizeof b6);
    } else {   // Stormer-Verlet: one kick, one drift per slice
        c.s = 1;
        c.bab[0] = 1.0;
        c.bab[1] = 1.0;
        c.bab[2] = 1.0;
    }
    return c;
}

// counters cleared, F reset to the uploaded input, start event
int batch_prelude(trhepd_ctx *ctx, Batch &b)
{
    const size_t nn = (size_t)b.nb * b.nb;
    b.stats = Stats();
    launch_fill_i32(b.stream, b.cnt, 16, 0);
    TRY_HIP(ctx, hipEventRecord(b.ev0, b.stream));
    TRY_HIP(ctx, hipMemcpyAsync(b.fout, b.fin, nn * b.nang * 16, hipMemcpyDeviceToDevice, b.stream));
    return TRHEPD_OK;
}

int batch_start(trhepd_ctx *ctx, Batch &b, const Table &t)
{
    int rc = batch_prelude(ctx, b);
    if (rc) return rc;
    if (b.run) return runner_begin(ctx, b, t);
    return run_small(ctx, b.stream, batch_period(ctx, b, t), t, b.nb, b.nang, b.gmap, b.gamma, b.fout,
                     b.has_f2 ? b.f2 : nullptr, b.nb0, b.w, b.cnt);
}

// Small-N batches that share N, integrator, dz and nb0 (structures with the same beams)
// run as one launch on the first one's stream; the others' streams wait for their prelude
// before it and resume after it, so every batch's own events still bracket its work.
static int start_small_group(trhepd_ctx *ctx, const std::vector<Batch *> &g, const std::vector<const Table *> &gt)
{
    if (g.size() == 1) return batch_start(ctx, *g[0], *gt[0]);
    std::vector<SmallBatch> sb;
    for (size_t i = 0; i < g.size(); ++i) {
        Batch &b = *g[i];
        int rc = batch_prelude(ctx, b);
        if (rc) return rc;
        sb.push_back(SmallBatch{gt[i], b.nb, b.nang, b.nb0, batch_period(ctx, b, *gt[i]), b.gmap, b.gamma, b.fout,
                                b.has_f2 ? b.f2 : nullptr, b.w, b.cnt});
        if (i) {
            TRY_HIP(ctx, hipEventRecord(b.evs, b.stream));
            TRY_HIP(ctx, hipStreamWaitEvent(g[0]->stream, b.evs, 0));
        }
    }
    int rc = run_small_multi(ctx, g[0]->stream, sb);
    if (rc) return rc;
    TRY_HIP(ctx, hipEventRecord(g[0]->evs, g[0]->stream));
    for (size_t i = 1; i < g.size(); ++i) TRY_HIP(ctx, hipStreamWaitEvent(g[i]->stream, g[0]->evs, 0));
    return TRHEPD_OK;
}

int enqueue_batches(trhepd_ctx *ctx, const std::vector<Batch *> &bs, const std::vector<const Table *> &ts)
{
    const int n = (int)bs.size();
    int rc = TRHEPD_OK, L = 0;
    std::vector<bool> done(n, false);
    for (int i = 0; i < n && !rc; ++i) {
        if (done[i]) continue;
        if (bs[i]->run || !small_fusable(bs[i]->nb, ts[i]->method)) {
            rc = batch_start(ctx, *bs[i], *ts[i]);
            if (bs[i]->run) L = std::max(L, ts[i]->ns);
            done[i] = true;
            continue;
        }
        std::vector<Batch *> g;
        std::vector<const Table *> gt;
        for (int j = i; j < n; ++j)
            if (!done[j] && !bs[j]->run && bs[j]->nb == bs[i]->nb && ts[j]->method == ts[i]->method &&
                ts[j]->dz == ts[i]->dz && ts[j]->steps == ts[i]->steps && bs[j]->nb0 == bs[i]->nb0 &&
                small_fusable(bs[j]->nb, ts[j]->method)) {
                g.push_back(bs[j]);
                gt.push_back(ts[j]);
                done[j] = true;
            }
        rc = start_small_group(ctx, g, gt);
    }
    // Phase offsets: with a fixed renormalisation period every batch factorises at the same
    // slices, and batches that start together reach their LU chains (a few workgroups per
    // launch, bound by the panel latency) together, leaving most CUs idle.  The j-th of m
    // large-N batches therefore starts j/m of the first one's period later: its stream waits
    // for the first batch's slice off[j], so that its LU chains fall among the other batches'
    // stage products.  The batches' results do not depend on it.
    std::vector<int> off(n, 0);
    {
        int lead = -1, m = 0;
        for (int i = 0; i < n; ++i)
            if (bs[i]->run) {
                if (lead < 0) lead = i;
                ++m;
            }
        const int period = lead >= 0 ? batch_period(ctx, *bs[lead], *ts[lead]) : 0;
        const bool on = ctx->phase_offsets < 0 ? ctx->precision == TRHEPD_PREC_MIXED : ctx->phase_offsets > 0;
        if (m > 1 && period > 1 && period < ts[lead]->ns && on && !ctx->capturing)
            for (int i = lead + 1, j = 1; i < n; ++i)
                if (bs[i]->run) off[i] = std::min((j++ * period + m / 2) / m, ts[lead]->ns - 1);
        for (int i = 0; i < n; ++i) L = std::max(L, bs[i]->run ? ts[i]->ns + off[i] : 0);
        for (int l = 1; l <= L && !rc; ++l) {
            // slices of all large-N batches interleaved: while the host waits for one batch's
            // renormalisation flags the others' kernels are already queued on their streams
            for (int i = 0; i < n && !rc; ++i)
                if (bs[i]->run && off[i] > 0 && l == off[i] + 1) {
                    // the lead batch has queued slices 1 .. off[i] and their renormalisations
                    rc = hip_check(ctx, hipEventRecord(bs[i]->evs, bs[lead]->stream), "phase event");
                    if (!rc) rc = hip_check(ctx, hipStreamWaitEvent(bs[i]->stream, bs[i]->evs, 0), "phase wait");
                }
            for (int i = 0; i < n && !rc; ++i) {
                const int li = l - off[i];
                if (bs[i]->run && li >= 1 && li <= ts[i]->ns) rc = runner_slice_enqueue(ctx, *bs[i], *ts[i], li);
            }
            for (int i = 0; i < n && !rc; ++i) {
                const int li = l - off[i];
                if (bs[i]->run && li >= 1 && li <= ts[i]->ns) rc = runner_slice_finish(ctx, *bs[i], *ts[i]);
            }
        }
    }
    for (int i = 0; i < n && !rc; ++i)
        if (bs[i]->run) rc = runner_end(ctx, *bs[i], *ts[i]);
    return rc;
}

Batch *find_batch(trhepd_ctx *ctx, int id)
{
    for (Batch &b : ctx->batches)
        if (b.id == id) return &b;
    return nullptr;
}

}  // namespace trhepd

using namespace trhepd;

extern "C" {

int trhepd_ctx_create(const trhepd_opts *opts, trhepd_ctx **out)
{
    if (!out) return TRHEPD_E_ARG;
    *out = nullptr;
    trhepd_ctx *c = new (std::nothrow) trhepd_ctx();
    if (!c) return TRHEPD_E_NOMEM;
    if (opts) {
        c->device = opts->device;
        c->precision = opts->precision;
        c->budget = opts->hbm_budget;
    }
    if (c->precision != TRHEPD_PREC_FP64 && c->precision != TRHEPD_PREC_MIXED) {
        delete c;
        return TRHEPD_E_ARG;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0 || c->device < 0 || c->device >= ndev) {
        delete c;
        return e != hipSuccess ? TRHEPD_E_HIP : TRHEPD_E_ARG;
    }
    e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return TRHEPD_E_HIP;
    }
    // device address of the fixed work area of trhepd_propagate_batch_noalloc on this device
    noalloc_attach(c);
    *out = c;
    return TRHEPD_OK;
}

int trhepd_ctx_destroy(trhepd_ctx *ctx)
{
    if (!ctx) return TRHEPD_E_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    server_destroy(ctx);
    (void)hipDeviceSynchronize();
    ctx->pool_off = true;   // the streams go away below
    for (GraphRec &gr : ctx->graphs) {
        (void)hipGraphExecDestroy(gr.exec);
        (void)hipGraphDestroy(gr.graph);
        for (hipEvent_t e : gr.events) (void)hipEventDestroy(e);
    }
    for (Table &t : ctx->tables) dev_free(ctx, t.d, t.bytes);
    for (Batch &b : ctx->batches) {
        runner_destroy(ctx, b);
        dev_free(ctx, b.d, b.bytes);
        (void)hipEventDestroy(b.ev0);
        (void)hipEventDestroy(b.ev1);
        (void)hipEventDestroy(b.evs);
        (void)hipStreamDestroy(b.stream);
    }
    for (hipEvent_t e : ctx->prof.ev) (void)hipEventDestroy(e);
    pool_release_all(ctx);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TRHEPD_OK;
}

const char *trhepd_last_error(const trhepd_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

size_t trhepd_ctx_bytes(const trhepd_ctx *ctx) { return ctx ? ctx->held : 0; }

int trhepd_asfcef(trhepd_ctx *ctx, int nv, int nelm, const int *igh, const int *igk, double ghx,
                  double ghy, double gky, int nabr, int nabi, const double *a, const double *c,
                  const double *dth, const double *dtk, double *asf, double *ghk)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (nv < 1 || nelm < 1 || nabr < 0 || nabr + nabi > 6 || !igh || !igk || !a || !c || !dth || !dtk ||
        !asf || !ghk)
        return fail(ctx, TRHEPD_E_ARG, "asfcef: bad arguments");
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~size_t(255); return o; };
    const size_t oh = take(nv * 4), ok = take(nv * 4), oa = take(6 * nelm * 8), oc = take(6 * nelm * 8),
                 o1 = take(nelm * 8), o2 = take(nelm * 8), oasf = take((size_t)nv * 6 * nelm * 8),
                 og = take(nv * 8);
    char *d;
    int rc = dev_alloc(ctx, (void **)&d, off);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(d + oh, igh, nv * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + ok, igk, nv * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + oa, a, 6 * nelm * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + oc, c, 6 * nelm * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o1, dth, nelm * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o2, dtk, nelm * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) {
        rc = launch_asfcef(ctx, nv, nelm, (int *)(d + oh), (int *)(d + ok), ghx, ghy, gky, nabr, nabi,
                           (double *)(d + oa), (double *)(d + oc), (double *)(d + o1), (double *)(d + o2),
                           (double *)(d + oasf), (double *)(d + og));
        if (!rc) e = hipMemcpyAsync(asf, d + oasf, (size_t)nv * 6 * nelm * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (!rc && !e) e = hipMemcpyAsync(ghk, d + og, nv * 8, hipMemcpyDeviceToHost, ctx->stream);
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, d, off);
    if (rc) return rc;
    if (e) return hip_check(ctx, e, "asfcef");
    return hip_check(ctx, e2, "asfcef sync");
}

int trhepd_scpot_stf(trhepd_ctx *ctx, int natm, int nv, int nsg, int ma, int mb, int na, int nb,
                     const double *gh, const double *gk, const double *x, const double *y, double dx,
                     double dy, double *st)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (natm < 0 || nv < 1 || nsg < 1 || nsg > 17 || ma * nb - mb * na == 0 || !gh || !gk || !st ||
        (natm > 0 && (!x || !y)))
        return fail(ctx, TRHEPD_E_ARG, "scpot_stf: bad arguments (nsg must be 1..17)");
    if (natm == 0) return TRHEPD_OK;
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    // orbit expansion (strfac.f90:15-65): host, <= 12 positions per atom
    std::vector<int> ndeg(natm);
    std::vector<double> xb(12 * (size_t)natm), yb(12 * (size_t)natm);
    const double s = (double)(ma * nb - mb * na);
    auto red = [](double v) { return v - std::floor(v + 1e-4); };
    for (int j = 0; j < natm; ++j) {
        double xs[12], ys[12];
        const double x0 = red((nb * x[j] - na * y[j]) / s), y0 = red((ma * y[j] - mb * x[j]) / s);
        const int n = orbit(nsg, x0, y0, xs, ys);
        ndeg[j] = n;
        for (int i = 0; i < n; ++i) {
            const double u = red(xs[i]), w = red(ys[i]);
            xb[12 * j + i] = ma * u + na * w;
            yb[12 * j + i] = mb * u + nb * w;
        }
    }
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~size_t(255); return o; };
    const size_t o_gh = take(nv * 8), o_gk = take(nv * 8), o_n = take(natm * 4), o_x = take(12 * natm * 8),
                 o_y = take(12 * natm * 8), o_st = take((size_t)nv * natm * 16);
    char *d;
    int rc = dev_alloc(ctx, (void **)&d, off);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(d + o_gh, gh, nv * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o_gk, gk, nv * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o_n, ndeg.data(), natm * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o_x, xb.data(), 12 * natm * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) e = hipMemcpyAsync(d + o_y, yb.data(), 12 * natm * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!e) {
        rc = launch_strfac(ctx, natm, nv, (double *)(d + o_gh), (double *)(d + o_gk), (int *)(d + o_n),
                           (double *)(d + o_x), (double *)(d + o_y), dx, dy, (double *)(d + o_st));
        if (!rc) e = hipMemcpyAsync(st, d + o_st, (size_t)nv * natm * 16, hipMemcpyDeviceToHost, ctx->stream);
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, d, off);
    if (rc) return rc;
    if (e) return hip_check(ctx, e, "scpot_stf");
    return hip_check(ctx, e2, "scpot_stf sync");
}

int trhepd_table_build(trhepd_ctx *ctx, const trhepd_potential *pot, double dz, int ns, int method,
                       int *table_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (!pot || !table_id || ns < 1 || !(dz > 0.0) || pot->nv < 1 || pot->nvb < 1 || pot->nvb > pot->nv ||
        pot->nelm < 1 || pot->nabr < 0 || pot->nabr + pot->nabi > 6 || pot->natms < 0 || pot->natmb < 0 ||
        !pot->ielm || !pot->z || !pot->dchg || !pot->ghk || !pot->ocr || !pot->b || !pot->asf ||
        (pot->natms && !pot->st0) || (pot->natmb && (!pot->st1 || !pot->st2)))
        return fail(ctx, TRHEPD_E_ARG, "table_build: bad potential descriptor");
    if (method < TRHEPD_M2 || method > TRHEPD_RK4GS) return fail(ctx, TRHEPD_E_ARG, "unknown int_method %d", method);
    for (int i = 0; i < pot->natmb + pot->natms; ++i)
        if (pot->ielm[i] < 1 || pot->ielm[i] > pot->nelm)
            return fail(ctx, TRHEPD_E_ARG, "ielm(%d) = %d out of 1..%d", i + 1, pot->ielm[i], pot->nelm);
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    Table t;
    t.method = method;
    t.ns = ns;
    t.dz = dz;
    int rc = launch_table(ctx, pot, t);
    if (rc) return rc;
    t.id = ctx->next_table++;
    ctx->tables.push_back(t);
    *table_id = t.id;
    return TRHEPD_OK;
}

int trhepd_table_build_steps(trhepd_ctx *ctx, const trhepd_potential *pot, int method, int ns, const double *steps,
                             double z0, int *table_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (method != TRHEPD_SP4 && method != TRHEPD_SP6)
        return fail(ctx, TRHEPD_E_ARG, "table_build_steps: slice lists are for int_method 5 and 6, not %d", method);
    if (ns < 1 || !steps || !table_id) return fail(ctx, TRHEPD_E_ARG, "table_build_steps: no slices");
    double hmax = 0.0;
    for (int l = 0; l < ns; ++l) {
        if (!(steps[l] > 0.0) || !std::isfinite(steps[l]))
            return fail(ctx, TRHEPD_E_ARG, "table_build_steps: slice %d has thickness %g", l + 1, steps[l]);
        hmax = std::max(hmax, steps[l]);
    }
    if (!std::isfinite(z0)) return fail(ctx, TRHEPD_E_ARG, "table_build_steps: z0 is not finite");
    if (!pot || pot->nv < 1 || pot->nvb < 1 || pot->nvb > pot->nv || pot->nelm < 1 || pot->nabr < 0 ||
        pot->nabr + pot->nabi > 6 || pot->natms < 0 || pot->natmb < 0 || !pot->ielm || !pot->z || !pot->dchg ||
        !pot->ghk || !pot->ocr || !pot->b || !pot->asf || (pot->natms && !pot->st0) ||
        (pot->natmb && (!pot->st1 || !pot->st2)))
        return fail(ctx, TRHEPD_E_ARG, "table_build_steps: bad potential descriptor");
    for (int i = 0; i < pot->natmb + pot->natms; ++i)
        if (pot->ielm[i] < 1 || pot->ielm[i] > pot->nelm)
            return fail(ctx, TRHEPD_E_ARG, "ielm(%d) = %d out of 1..%d", i + 1, pot->ielm[i], pot->nelm);
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    Table t;
    t.method = method;
    t.ns = ns;
    t.dz = hmax;
    t.steps.assign(steps, steps + ns);
    t.z0 = z0;
    int rc = launch_table(ctx, pot, t);
    if (rc) return rc;
    t.id = ctx->next_table++;
    ctx->tables.push_back(t);
    *table_id = t.id;
    return TRHEPD_OK;
}

int trhepd_table_shape(trhepd_ctx *ctx, int table_id, int *nv, int *ncols)
{
    if (!ctx) return TRHEPD_E_ARG;
    Table *t = find_table(ctx, table_id);
    if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "no table %d", table_id);
    if (nv) *nv = t->nv;
    if (ncols) *ncols = t->ncols;
    return TRHEPD_OK;
}

int trhepd_table_fetch(trhepd_ctx *ctx, int table_id, double *v, double *vi)
{
    if (!ctx) return TRHEPD_E_ARG;
    Table *t = find_table(ctx, table_id);
    if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "no table %d", table_id);
    if (!v || !vi) return fail(ctx, TRHEPD_E_ARG, "table_fetch: null output");
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<VVi> h((size_t)t->nv * t->ncols);
    TRY_HIP(ctx, hipMemcpyAsync(h.data(), t->d, h.size() * sizeof(VVi), hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t e = 0; e < h.size(); ++e) {
        v[2 * e] = h[e].v.x;
        v[2 * e + 1] = h[e].v.y;
        vi[2 * e] = h[e].vi.x;
        vi[2 * e + 1] = h[e].vi.y;
    }
    return TRHEPD_OK;
}

int trhepd_table_release(trhepd_ctx *ctx, int table_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    // a resident batch (and so any graph that captured it) reads the table on every run
    for (const Batch &b : ctx->batches)
        if (b.table_id == table_id)
            return fail(ctx, TRHEPD_E_ARG, "table %d is used by batch %d: release the batch first", table_id, b.id);
    for (size_t i = 0; i < ctx->tables.size(); ++i)
        if (ctx->tables[i].id == table_id) {
            TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            dev_free(ctx, ctx->tables[i].d, ctx->tables[i].bytes);
            ctx->tables.erase(ctx->tables.begin() + i);
            return TRHEPD_OK;
        }
    return fail(ctx, TRHEPD_E_NOTFOUND, "no table %d", table_id);
}

int trhepd_batch_create(trhepd_ctx *ctx, int table_id, int nb, int nang, const int *iv, const double *gamma,
                        const double *f, int nb0, const double *wnsga2, int *batch_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    Table *t = find_table(ctx, table_id);
    if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "no table %d", table_id);
    if (nb < 1 || nang < 1 || !iv || !gamma || !f || nb0 < 1 || nb0 > nb || !batch_id)
        return fail(ctx, TRHEPD_E_ARG, "batch_create: bad arguments");
    // the Magnus methods and rk4gs run on the batched path at every N
    const bool magnus = t->method == TRHEPD_M2 || t->method == TRHEPD_M4 || t->method == TRHEPD_M6 ||
                        t->method == TRHEPD_RK4GS;
    // signed gather map (row-major): upper i<j -> +iv(j,i), lower i>j -> -iv(i,j), diagonal 0
    std::vector<int> gm((size_t)nb * nb);
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            int g = 0;
            if (i < j) g = iv[j + (size_t)i * nb];
            else if (i > j) g = -iv[i + (size_t)j * nb];
            if ((i != j) && (g == 0 || std::abs(g) > t->nv))
                return fail(ctx, TRHEPD_E_ARG, "iv(%d,%d) = %d outside 1..nv=%d", i > j ? i + 1 : j + 1,
                            i > j ? j + 1 : i + 1, std::abs(g), t->nv);
            gm[(size_t)i * nb + j] = g;
        }
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)nb * nb;
    Batch b;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_gm = take(nn * 4), o_g = take((size_t)nb * nang * 16), o_fi = take(nn * nang * 16),
                 o_fo = take(nn * nang * 16), o_f2 = take((size_t)nb * nang * 8), o_w = take((size_t)nang * 8),
                 o_c = take(64);
    int rc = dev_alloc(ctx, (void **)&b.d, off);
    if (rc) return rc;
    b.bytes = off;
    b.gmap = (int *)(b.d + o_gm);
    b.gamma = (double2 *)(b.d + o_g);
    b.fin = (double2 *)(b.d + o_fi);
    b.fout = (double2 *)(b.d + o_fo);
    b.f2 = (double *)(b.d + o_f2);
    b.w = (double *)(b.d + o_w);
    b.cnt = (int *)(b.d + o_c);
    b.table_id = table_id;
    b.nb = nb;
    b.nang = nang;
    b.nb0 = nb0;
    b.has_f2 = wnsga2 != nullptr;
    for (size_t e = 0; e < (size_t)nb * nang; ++e) b.gim_max = std::max(b.gim_max, std::fabs(gamma[2 * e + 1]));
    hipError_t e = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
    if (!e) e = hipEventCreate(&b.ev0);
    if (!e) e = hipEventCreate(&b.ev1);
    if (!e) e = hipEventCreateWithFlags(&b.evs, hipEventDisableTiming);
    if (!e) e = hipMemcpyAsync(b.gmap, gm.data(), nn * 4, hipMemcpyHostToDevice, b.stream);
    if (!e) e = hipMemcpyAsync(b.gamma, gamma, (size_t)nb * nang * 16, hipMemcpyHostToDevice, b.stream);
    if (!e) e = hipMemcpyAsync(b.fin, f, nn * nang * 16, hipMemcpyHostToDevice, b.stream);
    if (!e && wnsga2) e = hipMemcpyAsync(b.w, wnsga2, (size_t)nang * 8, hipMemcpyHostToDevice, b.stream);
    if (!e) e = hipStreamSynchronize(b.stream);   // gm is a host temporary
    int rc2 = e ? hip_check(ctx, e, "batch upload") : TRHEPD_OK;
    if (!rc2 && (nb > 48 || magnus)) rc2 = runner_create(ctx, b, *t);
    if (rc2) {
        if (b.stream) (void)hipStreamSynchronize(b.stream);   // b is not among the batches dev_free marks
        runner_destroy(ctx, b);
        dev_free(ctx, b.d, b.bytes);
        if (b.ev0) (void)hipEventDestroy(b.ev0);
        if (b.ev1) (void)hipEventDestroy(b.ev1);
        if (b.evs) (void)hipEventDestroy(b.evs);
        if (b.stream) (void)hipStreamDestroy(b.stream);
        return rc2;
    }
    b.id = ctx->next_batch++;
    ctx->batches.push_back(b);
    *batch_id = b.id;
    return TRHEPD_OK;
}

int trhepd_batch_run_many(trhepd_ctx *ctx, int n, const int *batch_ids)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (n < 1 || !batch_ids) return fail(ctx, TRHEPD_E_ARG, "batch_run_many: no batches");
    std::vector<Batch *> bs;
    std::vector<const Table *> ts;
    for (int i = 0; i < n; ++i) {
        Batch *b = find_batch(ctx, batch_ids[i]);
        if (!b) return fail(ctx, TRHEPD_E_NOTFOUND, "no batch %d", batch_ids[i]);
        for (Batch *o : bs)
            if (o == b) return fail(ctx, TRHEPD_E_ARG, "batch %d listed twice", batch_ids[i]);
        Table *t = find_table(ctx, b->table_id);
        if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "table %d of batch %d was released", b->table_id, b->id);
        bs.push_back(b);
        ts.push_back(t);
    }
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    ctx->stats = Stats();
    int rc = enqueue_batches(ctx, bs, ts);
    for (int i = 0; i < n; ++i) {
        if (!rc) rc = hip_check(ctx, hipEventRecord(bs[i]->ev1, bs[i]->stream), "event");
    }
    int singular = 0;
    for (int i = 0; i < n; ++i) {
        Batch &b = *bs[i];
        const hipError_t e = hipStreamSynchronize(b.stream);
        if (!rc && e) rc = hip_check(ctx, e, "batch stream");
        if (rc) continue;
        int cnt[16];
        TRY_HIP(ctx, hipMemcpy(cnt, b.cnt, 64, hipMemcpyDeviceToHost));
        float ms = 0;
        TRY_HIP(ctx, hipEventElapsedTime(&ms, b.ev0, b.ev1));
        b.stats.n_renorm = cnt[0];
        b.stats.n_products = (long long)b.nang * ts[i]->ns * rkn_coef(ts[i]->method).s;
        b.stats.device_ms = ms;
        singular |= cnt[1];
        ctx->stats.n_renorm += b.stats.n_renorm;
        ctx->stats.n_products += b.stats.n_products;
    }
    if (rc) return rc;
    // device span: first start to last end
    float span = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, bs[i]->ev0, bs[j]->ev1) == hipSuccess) span = std::max(span, ms);
        }
    ctx->stats.device_ms = span;
    if ((rc = prof_collect(ctx))) return rc;
    if (singular) return fail(ctx, TRHEPD_E_SINGULAR, "exact zero pivot in a renormalisation LU");
    return TRHEPD_OK;
}

int trhepd_batch_run(trhepd_ctx *ctx, int batch_id) { return trhepd_batch_run_many(ctx, 1, &batch_id); }

int trhepd_batch_fetch(trhepd_ctx *ctx, int batch_id, double *f, double *f2)
{
    if (!ctx) return TRHEPD_E_ARG;
    Batch *b = find_batch(ctx, batch_id);
    if (!b) return fail(ctx, TRHEPD_E_NOTFOUND, "no batch %d", batch_id);
    if (f2 && !b->has_f2) return fail(ctx, TRHEPD_E_ARG, "batch %d was created without wnsga2", batch_id);
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)b->nb * b->nb;
    if (f) TRY_HIP(ctx, hipMemcpyAsync(f, b->fout, nn * b->nang * 16, hipMemcpyDeviceToHost, b->stream));
    if (f2) TRY_HIP(ctx, hipMemcpyAsync(f2, b->f2, (size_t)b->nb * b->nang * 8, hipMemcpyDeviceToHost, b->stream));
    TRY_HIP(ctx, hipStreamSynchronize(b->stream));
    if (f2)
        for (size_t e = 0; e < (size_t)b->nb * b->nang; ++e)
            if (!std::isfinite(f2[e])) return fail(ctx, TRHEPD_E_NONFINITE, "non-finite intensity at %zu", e);
    if (f)
        for (size_t e = 0; e < 2 * nn * b->nang; ++e)
            if (!std::isfinite(f[e])) return fail(ctx, TRHEPD_E_NONFINITE, "non-finite reflection at %zu", e / 2);
    return TRHEPD_OK;
}

int trhepd_batch_release(trhepd_ctx *ctx, int batch_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    for (const GraphRec &gr : ctx->graphs)
        for (int id : gr.batch_ids)
            if (id == batch_id)
                return fail(ctx, TRHEPD_E_ARG, "batch %d is part of graph %d: release the graph first", batch_id, gr.id);
    for (size_t i = 0; i < ctx->batches.size(); ++i)
        if (ctx->batches[i].id == batch_id) {
            Batch &b = ctx->batches[i];
            TRY_HIP(ctx, hipStreamSynchronize(b.stream));
            runner_destroy(ctx, b);
            dev_free(ctx, b.d, b.bytes);
            (void)hipEventDestroy(b.ev0);
            (void)hipEventDestroy(b.ev1);
            (void)hipEventDestroy(b.evs);
            (void)hipStreamDestroy(b.stream);
            ctx->batches.erase(ctx->batches.begin() + i);
            return TRHEPD_OK;
        }
    return fail(ctx, TRHEPD_E_NOTFOUND, "no batch %d", batch_id);
}

int trhepd_propagate_batch(trhepd_ctx *ctx, int table_id, int nb, int nang, const int *iv, const double *gamma,
                           double *f, double *f2, int nb0, const double *wnsga2)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (f2 && !wnsga2) return fail(ctx, TRHEPD_E_ARG, "propagate_batch: f2 needs wnsga2");
    if (nang == 0 && nb >= 1) {
        ctx->stats = Stats();
        return TRHEPD_OK;
    }
    int id = 0;
    int rc = trhepd_batch_create(ctx, table_id, nb, nang, iv, gamma, f, nb0, f2 ? wnsga2 : nullptr, &id);
    if (rc) return rc;
    rc = trhepd_batch_run(ctx, id);
    if (!rc) rc = trhepd_batch_fetch(ctx, id, f, f2);
    const int rc2 = trhepd_batch_release(ctx, id);
    return rc ? rc : rc2;
}

int trhepd_set_renorm_period(trhepd_ctx *ctx, int period)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (period < -1) return fail(ctx, TRHEPD_E_ARG, "renormalisation period must be >= -1 (got %d)", period);
    ctx->renorm_period = period;
    return TRHEPD_OK;
}

int trhepd_set_phase_offsets(trhepd_ctx *ctx, int mode)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (mode < -1 || mode > 1) return fail(ctx, TRHEPD_E_ARG, "phase offsets: mode -1, 0 or 1 (got %d)", mode);
    ctx->phase_offsets = mode;
    return TRHEPD_OK;
}

int trhepd_set_mixed_digits(trhepd_ctx *ctx, int digits)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (digits < 5 || digits > 7) return fail(ctx, TRHEPD_E_ARG, "mixed digits must be 5, 6 or 7 (got %d)", digits);
    ctx->mixed_digits = digits;
    return TRHEPD_OK;
}

int trhepd_profile(trhepd_ctx *ctx, int enable)
{
    if (!ctx) return TRHEPD_E_ARG;
    ctx->prof.on = enable != 0;
    ctx->prof.used = 0;
    ctx->prof.launches = 0;
    ctx->prof.ms = 0.0;
    return TRHEPD_OK;
}

int trhepd_kernel_stats(trhepd_ctx *ctx, long long *launches, double *total_ms)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (launches) *launches = ctx->prof.launches;
    if (total_ms) *total_ms = ctx->prof.ms;
    return TRHEPD_OK;
}

int trhepd_last_stats(const trhepd_ctx *ctx, long long *n_renorm, long long *n_products, double *device_ms)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (n_renorm) *n_renorm = ctx->stats.n_renorm;
    if (n_products) *n_products = ctx->stats.n_products;
    if (device_ms) *device_ms = ctx->stats.device_ms;
    return TRHEPD_OK;
}

int trhepd_graph_create(trhepd_ctx *ctx, int n, const int *batch_ids, int *graph_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (n < 1 || !batch_ids || !graph_id) return fail(ctx, TRHEPD_E_ARG, "graph_create: no batches");
    std::vector<Batch *> bs;
    std::vector<const Table *> ts;
    for (int i = 0; i < n; ++i) {
        Batch *b = find_batch(ctx, batch_ids[i]);
        if (!b) return fail(ctx, TRHEPD_E_NOTFOUND, "no batch %d", batch_ids[i]);
        for (Batch *o : bs)
            if (o == b) return fail(ctx, TRHEPD_E_ARG, "batch %d listed twice", batch_ids[i]);
        const Table *t = find_table(ctx, b->table_id);
        if (!t) return fail(ctx, TRHEPD_E_NOTFOUND, "table %d of batch %d was released", b->table_id, b->id);
        if (b->run && batch_period(ctx, *b, *t) <= 0)
            return fail(ctx, TRHEPD_E_ARG, "graph capture needs a periodic renormalisation schedule "
                                           "(trhepd_set_renorm_period with K > 0 or -1)");
        bs.push_back(b);
        ts.push_back(t);
    }
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GraphRec gr;
    gr.events.resize(n + 1, nullptr);
    for (hipEvent_t &e : gr.events) TRY_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int rc = hip_check(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "begin capture");
    if (!rc) {
        ctx->capturing = true;
        rc = hip_check(ctx, hipEventRecord(gr.events[0], ctx->stream), "fork");
        for (int i = 0; i < n && !rc; ++i) rc = hip_check(ctx, hipStreamWaitEvent(bs[i]->stream, gr.events[0], 0), "fork");
        if (!rc) rc = enqueue_batches(ctx, bs, ts);
        for (int i = 0; i < n && !rc; ++i) {
            rc = hip_check(ctx, hipEventRecord(gr.events[i + 1], bs[i]->stream), "join");
            if (!rc) rc = hip_check(ctx, hipStreamWaitEvent(ctx->stream, gr.events[i + 1], 0), "join");
        }
        ctx->capturing = false;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &gr.graph);
        if (!rc && e) rc = hip_check(ctx, e, "end capture");
    }
    if (!rc) rc = hip_check(ctx, hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0), "instantiate");
    if (rc) {
        if (gr.graph) (void)hipGraphDestroy(gr.graph);
        for (hipEvent_t e : gr.events)
            if (e) (void)hipEventDestroy(e);
        return rc;
    }
    gr.id = ctx->next_graph++;
    gr.batch_ids.assign(batch_ids, batch_ids + n);
    ctx->graphs.push_back(gr);
    *graph_id = gr.id;
    return TRHEPD_OK;
}

int trhepd_graph_launch(trhepd_ctx *ctx, int graph_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    GraphRec *gr = nullptr;
    for (GraphRec &g : ctx->graphs)
        if (g.id == graph_id) gr = &g;
    if (!gr) return fail(ctx, TRHEPD_E_NOTFOUND, "no graph %d", graph_id);
    for (int id : gr->batch_ids) {
        const Batch *b = find_batch(ctx, id);
        if (!b || !find_table(ctx, b->table_id))
            return fail(ctx, TRHEPD_E_NOTFOUND, "graph %d lost batch %d or its table", graph_id, id);
    }
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    ctx->stats = Stats();
    TRY_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    TRY_HIP(ctx, hipGraphLaunch(gr->exec, ctx->stream));
    TRY_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    TRY_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.device_ms = ms;
    int singular = 0;
    for (int id : gr->batch_ids) {
        Batch *b = find_batch(ctx, id);
        const Table *t = b ? find_table(ctx, b->table_id) : nullptr;
        if (!b || !t) return fail(ctx, TRHEPD_E_NOTFOUND, "graph %d lost batch %d or its table", graph_id, id);
        int cnt[16];
        TRY_HIP(ctx, hipMemcpy(cnt, b->cnt, 64, hipMemcpyDeviceToHost));
        ctx->stats.n_renorm += cnt[0];
        ctx->stats.n_products += (long long)b->nang * t->ns * rkn_coef(t->method).s;
        singular |= cnt[1];
    }
    if (singular) return fail(ctx, TRHEPD_E_SINGULAR, "exact zero pivot in a renormalisation LU");
    return TRHEPD_OK;
}

int trhepd_graph_release(trhepd_ctx *ctx, int graph_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    for (size_t i = 0; i < ctx->graphs.size(); ++i)
        if (ctx->graphs[i].id == graph_id) {
            GraphRec &gr = ctx->graphs[i];
            TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipGraphExecDestroy(gr.exec);
            (void)hipGraphDestroy(gr.graph);
            for (hipEvent_t e : gr.events) (void)hipEventDestroy(e);
            ctx->graphs.erase(ctx->graphs.begin() + i);
            return TRHEPD_OK;
        }
    return fail(ctx, TRHEPD_E_NOTFOUND, "no graph %d", graph_id);
}

}  // extern "C"
