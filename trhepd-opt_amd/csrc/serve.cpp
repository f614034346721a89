// Request loop for serving rocking curves: continuous batching at slice granularity.
//
// A request is one structure (its potential table inputs, dz, ns, integrator, beam set) and
// a set of glancing angles; it replaces one srfref angle loop (surflf.f90:79-153).  Requests
// wait in arrival order until their device memory fits the context's HBM budget; an admitted
// request owns a table, a batch and a stream, and every step of the loop advances the
// admitted requests by one or more slices, so the kernels of different requests overlap on
// the device and requests that arrive later join while earlier ones are in flight (a request
// with far more remaining work than the smallest yields to it for a bounded time).  A request that
// reaches its last slice is copied back and its device memory released at once, which lets
// the next queued request in.  Latency is measured from submit to the step that observed
// completion (host steady clock).
//
// Merging (off unless trhepd_serve_set_merge enables it): queued requests of one structure
// (identical potential descriptor, dz, ns, integrator, beam set and nb0) that would run
// under the same renormalisation schedule are admitted together as one batch of all their
// angles, up to a limit of angles, and split again at completion.  Every angle's
// arithmetic is the same in a batch of any composition, so a merged request's results are
// bit-identical to its results alone; merging only saves tables, launches and streams.  A
// request may wait up to a window after its submission so that others can join it.
#include <algorithm>
#include <chrono>
#include <cmath>
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
constexpr double HOLD_RATIO = 8.0;     // remaining work, relative to the smallest, above which a request yields
constexpr double HOLD_MAX_MS = 2000.0; // longest a request yields in total

struct Request {
    int id = 0, state = QUEUED;
    PotCopy pot;
    double gim_max = 0.0;                // max |Im gamma| over its beams and angles
    double dz = 0;
    int ns = 0, method = 0, nb = 0, nang = 0, nb0 = 1;
    std::vector<int> iv;
    std::vector<double> gamma, f, w, f2;
    int table_id = 0, batch_id = 0;
    int lead = 0;                        // merged: the request that holds the shared batch
    int off = 0;                         // merged: first angle of this request in that batch
    std::vector<int> members;            // the lead: every request of its batch, in angle order
    int slice = 0;                       // next slice to enqueue (1-based)
    hipEvent_t ev[RUN_AHEAD] = {};       // end of the last RUN_AHEAD slices
    bool holding = false;                // yielding to requests with far less remaining work
    Clock::time_point held_since;
    double held_ms = 0.0;                // time yielded in earlier spells
    Clock::time_point t_submit, t_admit, t_done;
    int rc = TRHEPD_OK;
    long long n_renorm = 0;              // renormalisations summed over the angles of its batch
    std::string err;
};

}  // namespace

struct Server {
    double last_queued_ms = 0.0, last_service_ms = 0.0;   // of the last collected request
    std::deque<int> queue;
    std::map<int, Request> reqs;
    int next_id = 1;
    size_t held_at_refusal = 0;   // device bytes held when the head of the queue last did not fit
    int merge_angles = 0;         // merging: most angles of one merged batch (0: no merging)
    double merge_window_ms = 0.0; // merging: how long a request may wait for others to join
};

// whether two requests run the same structure under the same renormalisation schedule
static bool same_structure(const trhepd_ctx *ctx, const Request &a, const Request &b)
{
    const trhepd_potential &p = a.pot.p, &q = b.pot.p;
    if (a.dz != b.dz || a.ns != b.ns || a.method != b.method || a.nb != b.nb || a.nb0 != b.nb0 || a.iv != b.iv)
        return false;
    if (p.nv != q.nv || p.nvb != q.nvb || p.natms != q.natms || p.natmb != q.natmb || p.nelm != q.nelm ||
        p.nabr != q.nabr || p.nabi != q.nabi || p.cc != q.cc || p.smesh != q.smesh)
        return false;
    const PotCopy &x = a.pot, &y = b.pot;
    if (x.ielm != y.ielm || x.z != y.z || x.dchg != y.dchg || x.ghk != y.ghk || x.ocr != y.ocr || x.b != y.b ||
        x.asf != y.asf || x.st0 != y.st0 || x.st1 != y.st1 || x.st2 != y.st2)
        return false;
    // the derived periodic schedule depends on the batch's largest |Im gamma|: equal periods only
    if (ctx->renorm_period < 0) {
        auto period = [&](const Request &r) {
            const double rate = r.gim_max * r.dz;
            return rate > 0.0 ? std::max(1, (int)std::floor(std::log(1e3) / rate)) : r.ns;
        };
        if (period(a) != period(b)) return false;
    }
    return true;
}

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

// admits a request with the requests merged into it (`group`: the lead first, then the others
// in queue order; each holds only its own angles): builds one table and one batch of all their
// angles; E_BUDGET / E_NOMEM leave them queued
static int admit(trhepd_ctx *ctx, Server &S, const std::vector<int> &group)
{
    Request &r = S.reqs.at(group[0]);
    const Clock::time_point now = Clock::now();
    int nang = 0;
    for (int id : group) nang += S.reqs.at(id).nang;
    std::vector<double> gamma, f, w;
    const double *pg = r.gamma.data(), *pf = r.f.data(), *pw = r.w.data();
    if (group.size() > 1) {
        gamma.reserve((size_t)2 * r.nb * nang);
        f.reserve((size_t)2 * r.nb * r.nb * nang);
        w.reserve(nang);
        for (int id : group) {
            const Request &m = S.reqs.at(id);
            gamma.insert(gamma.end(), m.gamma.begin(), m.gamma.end());
            f.insert(f.end(), m.f.begin(), m.f.end());
            w.insert(w.end(), m.w.begin(), m.w.end());
        }
        pg = gamma.data();
        pf = f.data();
        pw = w.data();
    }
    int rc = trhepd_table_build(ctx, &r.pot.p, r.dz, r.ns, r.method, &r.table_id);
    if (!rc) rc = trhepd_batch_create(ctx, r.table_id, r.nb, nang, r.iv.data(), pg, pf, r.nb0, pw, &r.batch_id);
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
    if (rc) {
        drop_device_state(ctx, r);
        return rc;
    }
    r.members = group;
    int off = 0;
    for (int id : group) {
        Request &m = S.reqs.at(id);
        m.t_admit = now;
        m.off = off;
        off += m.nang;
        if (id != r.id) {
            m.lead = r.id;
            m.state = ACTIVE;
        }
    }
    return TRHEPD_OK;
}

// copies a finished batch back and hands every request of it its own angles
static void complete(trhepd_ctx *ctx, Server &S, Request &r)
{
    int nang = 0;
    for (int id : r.members) nang += S.reqs.at(id).nang;
    std::vector<double> f((size_t)2 * r.nb * r.nb * nang), f2((size_t)r.nb * nang);
    int rc = trhepd_batch_fetch(ctx, r.batch_id, f.data(), f2.data());
    // counters of the run: [0] renormalisations, [1] exact zero pivots seen by the LU
    int cnt[16] = {};
    if (!rc) {
        const Batch *b = find_batch(ctx, r.batch_id);
        rc = hip_check(ctx, hipMemcpy(cnt, b->cnt, sizeof cnt, hipMemcpyDeviceToHost), "request counters");
    }
    if (!rc && cnt[1]) rc = fail(ctx, TRHEPD_E_SINGULAR, "request %d: exact zero pivot in a renormalisation LU", r.id);
    const std::string err = rc ? ctx->err : std::string();
    const Clock::time_point now = Clock::now();
    drop_device_state(ctx, r);
    for (int id : r.members) {
        Request &m = S.reqs.at(id);
        m.t_done = now;
        m.n_renorm = cnt[0];
        m.rc = rc;
        m.err = err;
        m.state = rc ? FAILED : DONE;
        if (rc) continue;
        const size_t fa = (size_t)2 * r.nb * r.nb, f2a = (size_t)r.nb;
        m.f.assign(f.begin() + fa * m.off, f.begin() + fa * (m.off + m.nang));
        m.f2.assign(f2.begin() + f2a * m.off, f2.begin() + f2a * (m.off + m.nang));
    }
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
    for (size_t e = 0; e < (size_t)nb * nang; ++e) r.gim_max = std::max(r.gim_max, std::fabs(gamma[2 * e + 1]));
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
        complete(ctx, S, r);
    }
    // 2. admission in arrival order while the HBM budget allows; after a refusal the head of
    //    the queue is tried again only once memory has been released
    while (!S.queue.empty() && !(S.held_at_refusal && ctx->held >= S.held_at_refusal)) {
        Request &r = S.reqs.at(S.queue.front());
        // merging: the head may wait out its window for requests of its structure to join
        std::vector<int> group{r.id};
        if (S.merge_angles > 0) {
            const double waited = std::chrono::duration<double, std::milli>(Clock::now() - r.t_submit).count();
            if (waited < S.merge_window_ms) break;
            int nang = r.nang;
            for (size_t k = 1; k < S.queue.size(); ++k) {
                const Request &m = S.reqs.at(S.queue[k]);
                if (nang + m.nang > S.merge_angles || !same_structure(ctx, r, m)) continue;
                group.push_back(m.id);
                nang += m.nang;
            }
        }
        const size_t held = ctx->held;
        int rc = admit(ctx, S, group);
        // a merged batch that does not fit is retried with fewer of the requests
        while ((rc == TRHEPD_E_BUDGET || rc == TRHEPD_E_NOMEM) && group.size() > 1) {
            group.resize(group.size() / 2);
            rc = admit(ctx, S, group);
        }
        if (rc && rc != TRHEPD_E_BUDGET && rc != TRHEPD_E_NOMEM && group.size() > 1) {   // the head on its own
            group.resize(1);
            rc = admit(ctx, S, group);
        }
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
            S.queue.pop_front();
            continue;
        }
        for (int id : group) S.queue.erase(std::find(S.queue.begin(), S.queue.end(), id));
    }
    // 3. advance the admitted requests.  Requests under the reference's renormalisation rule
    //    need the host after every slice (the flags of the Gershgorin test): one slice each,
    //    interleaved across their streams.  Requests under a periodic schedule need no host
    //    decision: each is given as many slices as its run-ahead events allow (up to
    //    SLICES_PER_STEP), without blocking, so that cheap requests are not held to the pace of
    //    the most expensive one.  The host blocks only when no request could advance.
    //    Shortest remaining work first, within a bound: the kernels of all streams share a few
    //    hardware queues, so one slice of a small request would wait behind several launches
    //    of every large one.  A periodic-schedule request whose remaining stage work (angles x
    //    N^3 x stages x slices left) exceeds HOLD_RATIO times the smallest active one's
    //    enqueues nothing new (what it has queued still runs) until HOLD_MAX_MS of such
    //    yielding in total, after which it is never held again.  Results do not depend on it.
    std::vector<Request *> act, gated;
    for (auto &kv : S.reqs)
        if (kv.second.state == ACTIVE && !kv.second.lead) act.push_back(&kv.second);
    std::vector<Request *> host_rule;
    bool moved = false;
    const Clock::time_point now = Clock::now();
    std::vector<double> rem(act.size(), -1.0);
    double mrem = HUGE_VAL;
    for (size_t i = 0; i < act.size(); ++i) {
        const Batch *b = find_batch(ctx, act[i]->batch_id);
        const Table *t = find_table(ctx, act[i]->table_id);
        if (!b->run || batch_period(ctx, *b, *t) <= 0) continue;
        const double n = b->nb;
        rem[i] = (double)b->nang * n * n * n * rkn_coef(t->method).s * (t->ns - act[i]->slice + 1);
        mrem = std::min(mrem, rem[i]);
    }
    for (size_t i = 0; i < act.size(); ++i) {
        Request *r = act[i];
        Batch *b = find_batch(ctx, r->batch_id);
        const Table *t = find_table(ctx, r->table_id);
        if (rem[i] < 0.0) {
            host_rule.push_back(r);
            continue;
        }
        const double held = r->held_ms + (r->holding ? std::chrono::duration<double, std::milli>(now - r->held_since).count() : 0.0);
        if (rem[i] > HOLD_RATIO * mrem && held < HOLD_MAX_MS) {
            if (!r->holding) {
                r->holding = true;
                r->held_since = now;
            }
            continue;
        }
        if (r->holding) {
            r->held_ms = held;
            r->holding = false;
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

extern "C" int trhepd_serve_set_merge(trhepd_ctx *ctx, int max_angles, double window_ms)
{
    if (!ctx) return TRHEPD_E_ARG;
    if (max_angles < 0 || !(window_ms >= 0.0))
        return fail(ctx, TRHEPD_E_ARG, "serve_set_merge: max_angles %d, window %g ms", max_angles, window_ms);
    Server &S = server(ctx);
    S.merge_angles = max_angles;
    S.merge_window_ms = window_ms;
    return TRHEPD_OK;
}

extern "C" int trhepd_serve_last_timing(trhepd_ctx *ctx, double *queued_ms, double *service_ms)
{
    if (!ctx) return TRHEPD_E_ARG;
    Server &S = server(ctx);
    if (queued_ms) *queued_ms = S.last_queued_ms;
    if (service_ms) *service_ms = S.last_service_ms;
    return TRHEPD_OK;
}
