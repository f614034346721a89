// Request loop for serving rocking curves: continuous batching at slice granularity.
//
// A request is one structure (its potential table inputs, dz, ns, integrator, beam set) and
// a set of glancing angles; it replaces one srfref angle loop (surflf.f90:79-153).  Requests
// wait in arrival order until their device memory fits the context's HBM budget; an admitted
// request owns a table, a batch and a stream, and every step of the loop advances all
// admitted requests by one slice, so the kernels of different requests overlap on the device
// and requests that arrive later join while earlier ones are in flight.  A request that
// reaches its last slice is copied back and its device memory released at once, which lets
// the next queued request in.  Latency is measured from submit to the step that observed
// completion (host steady clock).
#include <chrono>
#include <cstring>
#include <deque>
#include <map>

#include "internal.hpp"

namespace trhepd {

namespace {

using Clock = std::chrono::steady_clock;

// deep copy of a potential descriptor (the caller's arrays may go away after submit)
struct PotCopy {
    trhepd_potential p{};
    std::vector<int> ielm;
    std::vector<double> z, dchg, ghk, ocr, b, asf, st0, st1, st2;

    void take(const trhepd_potential &s)
    {
        p = s;
        const size_t na = (size_t)s.natmb + s.natms;
        ielm.assign(s.ielm, s.ielm + na);
        z.assign(s.z, s.z + na);
        ocr.assign(s.ocr, s.ocr + na);
        dchg.assign(s.dchg, s.dchg + s.nelm);
        ghk.assign(s.ghk, s.ghk + s.nv);
        b.assign(s.b, s.b + (size_t)6 * s.nelm);
        asf.assign(s.asf, s.asf + (size_t)s.nv * 6 * s.nelm);
        if (s.natms) st0.assign(s.st0, s.st0 + (size_t)2 * s.nv * s.natms);
        if (s.natmb) {
            st1.assign(s.st1, s.st1 + (size_t)2 * s.nvb * s.natmb);
            st2.assign(s.st2, s.st2 + (size_t)2 * s.nvb * s.natmb);
        }
        p.ielm = ielm.data();
        p.z = z.data();
        p.ocr = ocr.data();
        p.dchg = dchg.data();
        p.ghk = ghk.data();
        p.b = b.data();
        p.asf = asf.data();
        p.st0 = st0.empty() ? nullptr : st0.data();
        p.st1 = st1.empty() ? nullptr : st1.data();
        p.st2 = st2.empty() ? nullptr : st2.data();
    }
};

enum { QUEUED = 0, ACTIVE = 1, FINISHING = 2, DONE = 3, FAILED = 4 };
constexpr int RUN_AHEAD = 4;        // slices a request may be enqueued ahead of the device
constexpr int SLICES_PER_STEP = 16;  // slices one step may enqueue for a periodic-schedule request

struct Request {
    int id = 0, state = QUEUED;
    PotCopy pot;
    double dz = 0;
    int ns = 0, method = 0, nb = 0, nang = 0, nb0 = 1;
    std::vector<int> iv;
    std::vector<double> gamma, f, w, f2;
    int table_id = 0, batch_id = 0;
    int slice = 0;                       // next slice to enqueue (1-based)
    hipEvent_t ev[RUN_AHEAD] = {};       // end of the last RUN_AHEAD slices
    Clock::time_point t_submit, t_admit, t_done;
    int rc = TRHEPD_OK;
    long long n_renorm = 0;              // renormalisations summed over the request's angles
    std::string err;
};

}  // namespace

struct Server {
    double last_queued_ms = 0.0, last_service_ms = 0.0;   // of the last collected request
    std::deque<int> queue;
    std::map<int, Request> reqs;
    int next_id = 1;
    size_t held_at_refusal = 0;   // device bytes held when the head of the queue last did not fit
};

static Server &server(trhepd_ctx *ctx)
{
    if (!ctx->serve) ctx->serve = new Server();
    return *ctx->serve;
}

static void drop_device_state(trhepd_ctx *ctx, Request &r)
{
    if (r.batch_id) (void)trhepd_batch_release(ctx, r.batch_id);
    if (r.table_id) (void)trhepd_table_release(ctx, r.table_id);
    r.batch_id = r.table_id = 0;
    for (hipEvent_t &e : r.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

void server_destroy(trhepd_ctx *ctx)
{
    if (!ctx->serve) return;
    for (auto &kv : ctx->serve->reqs) drop_device_state(ctx, kv.second);
    delete ctx->serve;
    ctx->serve = nullptr;
}

// admits one request: builds its table and batch; E_BUDGET / E_NOMEM leave it queued
static int admit(trhepd_ctx *ctx, Request &r)
{
    r.t_admit = Clock::now();
    int rc = trhepd_table_build(ctx, &r.pot.p, r.dz, r.ns, r.method, &r.table_id);
    if (!rc)
        rc = trhepd_batch_create(ctx, r.table_id, r.nb, r.nang, r.iv.data(), r.gamma.data(), r.f.data(), r.nb0,
                                 r.w.data(), &r.batch_id);
    for (int k = 0; k < RUN_AHEAD && !rc; ++k) rc = hip_check(ctx, hipEventCreateWithFlags(&r.ev[k], hipEventDisableTiming), "event");
    if (!rc) {
        Batch *b = find_batch(ctx, r.batch_id);
        const Table *t = find_table(ctx, r.table_id);
        rc = batch_start(ctx, *b, *t);
        r.slice = 1;
        if (!rc && !b->run) {   // small N: the whole integration is one launch
            rc = hip_check(ctx, hipEventRecord(b->ev1, b->stream), "event");
            r.state = FINISHING;
        } else if (!rc) {
            r.state = ACTIVE;
        }
    }
    if (rc) drop_device_state(ctx, r);
    return rc;
}

static int complete(trhepd_ctx *ctx, Request &r)
{
    r.f.assign((size_t)2 * r.nb * r.nb * r.nang, 0.0);
    r.f2.assign((size_t)r.nb * r.nang, 0.0);
    int rc = trhepd_batch_fetch(ctx, r.batch_id, r.f.data(), r.f2.data());
    // counters of the run: [0] renormalisations, [1] exact zero pivots seen by the LU
    int cnt[16] = {};
    if (!rc) {
        const Batch *b = find_batch(ctx, r.batch_id);
        rc = hip_check(ctx, hipMemcpy(cnt, b->cnt, sizeof cnt, hipMemcpyDeviceToHost), "request counters");
    }
    if (!rc) r.n_renorm = cnt[0];
    if (!rc && cnt[1]) rc = fail(ctx, TRHEPD_E_SINGULAR, "request %d: exact zero pivot in a renormalisation LU", r.id);
    r.t_done = Clock::now();
    drop_device_state(ctx, r);
    return rc;
}

}  // namespace trhepd

using namespace trhepd;

extern "C" {

int trhepd_serve_submit(trhepd_ctx *ctx, const trhepd_potential *pot, double dz, int ns, int method, int nb,
                        int nang, const int *iv, const double *gamma, const double *f, int nb0,
                        const double *wnsga2, int *request_id)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (!pot || !iv || !gamma || !f || !wnsga2 || !request_id || nb < 1 || nang < 1 || ns < 1 || nb0 < 1 ||
        nb0 > nb || !(dz > 0.0))
        return fail(ctx, TRHEPD_E_ARG, "serve_submit: bad arguments");
    if (method != TRHEPD_SP4 && method != TRHEPD_SP6 && method != TRHEPD_LF && method != TRHEPD_RK4)
        return fail(ctx, TRHEPD_E_ARG, "int_method %d is not available on the device path at N = %d", method, nb);
    Server &S = server(ctx);
    Request r;
    r.id = S.next_id++;
    r.pot.take(*pot);
    r.dz = dz;
    r.ns = ns;
    r.method = method;
    r.nb = nb;
    r.nang = nang;
    r.nb0 = nb0;
    r.iv.assign(iv, iv + (size_t)nb * nb);
    r.gamma.assign(gamma, gamma + (size_t)2 * nb * nang);
    r.f.assign(f, f + (size_t)2 * nb * nb * nang);
    r.w.assign(wnsga2, wnsga2 + nang);
    r.t_submit = Clock::now();
    const int id = r.id;
    S.reqs.emplace(id, std::move(r));
    S.queue.push_back(id);
    *request_id = id;
    return TRHEPD_OK;
}

int trhepd_serve_step(trhepd_ctx *ctx, int *n_active, int *n_queued, int *n_done)
{
    if (!ctx) return TRHEPD_E_ARG;
    TRY_HIP(ctx, hipSetDevice(ctx->device));
    Server &S = server(ctx);
    // 1. completion of requests whose last kernels were enqueued by an earlier step
    for (auto &kv : S.reqs) {
        Request &r = kv.second;
        if (r.state != FINISHING) continue;
        Batch *b = find_batch(ctx, r.batch_id);
        const hipError_t q = hipEventQuery(b->ev1);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return hip_check(ctx, q, "request completion");
        r.rc = complete(ctx, r);
        if (r.rc) r.err = ctx->err;
        r.state = r.rc ? FAILED : DONE;
    }
    // 2. admission in arrival order while the HBM budget allows; after a refusal the head of
    //    the queue is tried again only once memory has been released
    while (!S.queue.empty() && !(S.held_at_refusal && ctx->held >= S.held_at_refusal)) {
        Request &r = S.reqs.at(S.queue.front());
        const size_t held = ctx->held;
        const int rc = admit(ctx, r);
        S.held_at_refusal = 0;
        if (rc == TRHEPD_E_BUDGET || rc == TRHEPD_E_NOMEM) {
            bool any = false;
            for (auto &kv : S.reqs) any |= kv.second.state == ACTIVE || kv.second.state == FINISHING;
            if (any) {   // waits for memory to free up
                S.held_at_refusal = held;
                break;
            }
        }
        if (rc) {
            r.rc = rc;
            r.err = ctx->err;
            r.state = FAILED;
            r.t_done = r.t_admit = Clock::now();
        }
        S.queue.pop_front();
    }
    // 3. advance the admitted requests.  Requests under the reference's renormalisation rule
    //    need the host after every slice (the flags of the Gershgorin test): one slice each,
    //    interleaved across their streams.  Requests under a periodic schedule need no host
    //    decision: each is given as many slices as its run-ahead events allow (up to
    //    SLICES_PER_STEP), without blocking, so that cheap requests are not held to the pace of
    //    the most expensive one.  The host blocks only when no request could advance.
    std::vector<Request *> act, gated;
    for (auto &kv : S.reqs)
        if (kv.second.state == ACTIVE) act.push_back(&kv.second);
    std::vector<Request *> host_rule;
    bool moved = false;
    for (Request *r : act) {
        Batch *b = find_batch(ctx, r->batch_id);
        const Table *t = find_table(ctx, r->table_id);
        if (!b->run || batch_period(ctx, *b, *t) <= 0) {
            host_rule.push_back(r);
            continue;
        }
        for (int k = 0; k < SLICES_PER_STEP && r->state == ACTIVE; ++k) {
            // slice l may be enqueued once slice l - RUN_AHEAD has finished
            if (r->slice > RUN_AHEAD) {
                const hipError_t q = hipEventQuery(r->ev[r->slice % RUN_AHEAD]);
                if (q == hipErrorNotReady) {
                    if (k == 0) gated.push_back(r);
                    break;
                }
                if (q != hipSuccess) return hip_check(ctx, q, "request run-ahead");
            }
            int rc = runner_slice_enqueue(ctx, *b, *t, r->slice);
            if (!rc) rc = runner_slice_finish(ctx, *b, *t);
            if (rc) return rc;
            TRY_HIP(ctx, hipEventRecord(r->ev[r->slice % RUN_AHEAD], b->stream));
            if (r->slice == t->ns) {
                if ((rc = runner_end(ctx, *b, *t))) return rc;
                TRY_HIP(ctx, hipEventRecord(b->ev1, b->stream));
                r->state = FINISHING;
            }
            ++r->slice;
            moved = true;
        }
    }
    for (Request *r : host_rule) {
        Batch *b = find_batch(ctx, r->batch_id);
        const Table *t = find_table(ctx, r->table_id);
        if (r->slice > RUN_AHEAD) TRY_HIP(ctx, hipEventSynchronize(r->ev[r->slice % RUN_AHEAD]));
        int rc = runner_slice_enqueue(ctx, *b, *t, r->slice);
        if (rc) return rc;
    }
    for (Request *r : host_rule) {
        Batch *b = find_batch(ctx, r->batch_id);
        const Table *t = find_table(ctx, r->table_id);
        int rc = runner_slice_finish(ctx, *b, *t);
        if (rc) return rc;
        TRY_HIP(ctx, hipEventRecord(r->ev[r->slice % RUN_AHEAD], b->stream));
        if (r->slice == t->ns) {
            if ((rc = runner_end(ctx, *b, *t))) return rc;
            TRY_HIP(ctx, hipEventRecord(b->ev1, b->stream));
            r->state = FINISHING;
        }
        ++r->slice;
        moved = true;
    }
    // nothing could advance: wait for the oldest gating slice of one request instead of spinning
    if (!moved && !gated.empty())
        TRY_HIP(ctx, hipEventSynchronize(gated[0]->ev[gated[0]->slice % RUN_AHEAD]));
    int na = 0, nd = 0;
    for (auto &kv : S.reqs) {
        na += kv.second.state == ACTIVE || kv.second.state == FINISHING;
        nd += kv.second.state == DONE || kv.second.state == FAILED;
    }
    if (n_active) *n_active = na;
    if (n_queued) *n_queued = (int)S.queue.size();
    if (n_done) *n_done = nd;
    return TRHEPD_OK;
}

int trhepd_serve_result(trhepd_ctx *ctx, int request_id, double *f, double *f2, double *latency_ms)
{
    if (!ctx) return TRHEPD_E_ARG;
    Server &S = server(ctx);
    auto it = S.reqs.find(request_id);
    if (it == S.reqs.end()) return fail(ctx, TRHEPD_E_NOTFOUND, "no request %d", request_id);
    Request &r = it->second;
    if (r.state != DONE && r.state != FAILED) return TRHEPD_E_PENDING;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
    S.last_queued_ms = std::chrono::duration<double, std::milli>(r.t_admit - r.t_submit).count();
    S.last_service_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_admit).count();
    const int rc = r.rc;
    if (rc) {
        ctx->err = r.err;
    } else {
        if (f) std::memcpy(f, r.f.data(), r.f.size() * sizeof(double));
        if (f2) std::memcpy(f2, r.f2.data(), r.f2.size() * sizeof(double));
    }
    S.reqs.erase(it);
    return rc;
}

}  // extern "C"

extern "C" int trhepd_serve_last_timing(trhepd_ctx *ctx, double *queued_ms, double *service_ms)
{
    if (!ctx) return TRHEPD_E_ARG;
    Server &S = server(ctx);
    if (queued_ms) *queued_ms = S.last_queued_ms;
    if (service_ms) *service_ms = S.last_service_ms;
    return TRHEPD_OK;
}
