// serve.cpp -- continuous batching of covariance and field solves on one GPU.
//
// Requests (MCcov or MCfield problems of any size, precision and path count)
// arrive with scmc_submit_mccov / scmc_submit_mcfield.  The scheduler runs on
// the host: `max_active` slot threads, each driving its own HIP streams, so
// the kernels of all admitted requests are in flight on the device together
// (they share the GPU through its hardware queues).  A slot takes the
// first pending request whose device footprint fits the remaining HBM budget
// (FIFO with backfill), so requests are admitted as slots and memory free up.
// Each admitted request runs in its slot's session (its own HIP streams and
// accumulators, reused from request to request) and advances in slices of
// paths, so the kernels of all admitted requests overlap on the GPU; a cancel takes effect at the next slice boundary and
// frees the slot.  Results are normalised exactly as scmc_mccov and depend
// only on the request (Philox paths, fixed slice schedule), not on the mix.
//
// Update modes for the mixed covariance updates: each chunk's update launched
// on its slot's stream (default), the small requests' updates in shared
// launches (shared_launches), or every update as a job of one resident loop
// kernel that stays on the GPU and takes work items from a job table in host
// memory (resident_loop, ResidentLoop in resident.cpp).  The results are the
// same in all three (bit-identical for the resident loop); on the request
// workload separate launches measured fastest, so they are the default.
#include "scmc_internal.h"

#include <cstdlib>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace scmc {

using clk = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int kind = 0;   // 0 MCcov, 1 MCfield
    ProblemCopy pc;
    int64_t max_itr = 0;
    int precision = 0;
    int64_t footprint = 0;
    int state = 0;   // 0 pending, 1 running, 2 done
    int rc = SCMC_OK;
    std::string err;
    std::atomic<bool> cancel{false};
    std::vector<double> ms_re, ms_im, mm_re, mm_im;
    scmc_stats st{};
    clk::time_point t_submit, t_start, t_done;
};

struct Server {
    scmc_server_config cfg{};
    int64_t budget = 0, used = 0, slice = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::map<int64_t, std::unique_ptr<Request>> all;
    std::vector<std::thread> slots;
    std::vector<int64_t> held;   // device bytes each slot's session keeps (at its last request's end)
    int64_t next_id = 1;
    bool stop = false;
    bool shared_updates = false;   // mixed updates of concurrent small requests in shared launches
    UpdateBatcher batcher;
    std::unique_ptr<ResidentLoop> loop;   // mixed updates in one resident kernel (resident_loop)
    // A server still running when the process exits (no scmc_server_stop)
    // is stopped here: joinable threads would otherwise end the process in
    // std::terminate.  Running requests are cancelled at their next slice.
    ~Server()
    {
        bool joined = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
            for (auto &kv : all) kv.second->cancel = true;
        }
        cv_work.notify_all();
        for (auto &t : slots)
            if (t.joinable()) {
                t.join();
                joined = true;
            }
        loop.reset();
        if (joined) reaper_stop();
    }
};

// The running server.  Callers of scmc_wait/scmc_poll hold a reference while
// they use it, so scmc_server_stop can drop the global one while a waiter is
// still inside: the Server is destroyed by whoever leaves it last.
static std::mutex g_srv_mu;
static std::shared_ptr<Server> g_srv;

int64_t request_footprint(int nv, int nl, int precision, bool near)
{
    const double n = (double)nv * nl;
    const bool mixed = precision == SCMC_MIXED;
    const double tile = mixed ? kTileMixed : kTileF64;
    const double nt = std::ceil(n / tile);
    const double tiles = nt * (nt + 1) / 2;
    // accumulator copies (multiple + single scattering), at most 32 + 8
    const double ks = std::min(32.0, std::max(1.0, std::ceil(8192.0 / tiles)));
    const double acc = (ks + std::min(ks, 8.0)) * n * n * 16.0;
    // chunk buffers: expanded rows, factors, vertex records
    const double cap = mixed ? 4.0e9 / (8.0 * n) : 2.5e9 / (16.0 * n);
    const double kchunk = std::max(65536.0, std::min(4.0e6, cap));
    const double cs = mixed ? 8.0 : 16.0;
    double chunk = kchunk * ((mixed ? 8.0 : 16.0) * n + 2.0 * (nv + nl) * cs * 2.0 + 80.0);
    // near field: whole single-scatter rows, one per path (<= one per vertex)
    if (near) chunk += kchunk * n * cs;
    return (int64_t)(acc + chunk + 64.0 * 1024 * 1024);
}

// Device memory of an admitted MCfield request: accumulator, per-view work
// (64 entry splits) and the records of its largest chunk (field_chunk_paths).
int64_t request_footprint_field(int nv, int nl, int precision)
{
    (void)precision;   // both modes keep FP64 accumulators and records
    const double n = (double)nv * nl;
    const double kchunk = std::max(2097152.0, std::min(8388608.0, 67108864.0 / std::max(nl, 1)));   // vertices
    const double paths = std::min(1048576.0, kchunk / 10.0);
    const double work = 16.0 * (kchunk * nl + paths * nl + n + 64.0 * n);
    const double records = kchunk * (64.0 + 4.0) + paths * (80.0 + 12.0);
    return (int64_t)(16.0 * n + work + records + 64.0 * 1024 * 1024);
}

static void run_request(Server *S, Request *r, Session &sess)
{
    // device blocks come from and go back to the session's pool while the
    // request runs; they become reusable once it has synchronised (below)
    tls_block_pool = &sess.pool;
    int rc = sess.init(&r->pc.pr, r->kind, r->precision);
    const int64_t slice = S->slice > 0 ? S->slice : std::max<int64_t>(1000, r->max_itr / 8);
    for (int64_t p0 = 0; !rc && p0 < r->max_itr; p0 += slice) {
        if (r->cancel.load()) {
            rc = set_error(SCMC_E_CANCELLED, "request %lld cancelled", (long long)r->id);
            break;
        }
        // only the last slice waits for the device: the chunks of one slice
        // stay in flight while the next one is traced
        const int64_t len = std::min(slice, r->max_itr - p0);
        rc = sess.run(p0, len, p0 + len >= r->max_itr);
    }
    if (!rc && r->kind == 0) {
        const size_t ne = (size_t)r->pc.pr.nv * r->pc.pr.nv * r->pc.pr.nl * r->pc.pr.nl;
        r->ms_re.resize(ne);
        r->ms_im.resize(ne);
        r->mm_re.resize(ne);
        r->mm_im.resize(ne);
        rc = sess.result(r->max_itr, r->ms_re.data(), r->ms_im.data(), r->mm_re.data(), r->mm_im.data());
    } else if (!rc) {
        // the field u (nv x nl) in the first output pair
        const size_t ne = (size_t)r->pc.pr.nv * r->pc.pr.nl;
        r->ms_re.resize(ne);
        r->ms_im.resize(ne);
        rc = sess.result(r->max_itr, r->ms_re.data(), r->ms_im.data(), nullptr, nullptr);
    }
    sess.fill_stats(&r->st);
    r->rc = rc;
    if (rc) r->err = scmc_last_error();
    // every kernel of the request has finished: result() synchronised the
    // session's stream after draining both (on an error path, wait here,
    // resident-loop jobs included)
    if (rc) {
        for (auto &sl : sess.ce.slot) (void)sess.ce.wait_loop(sl);
        if (sess.stream2) (void)stream_wait(sess.stream2);
        if (sess.stream) (void)stream_wait(sess.stream);
    }
    sess.pool.settle();
    tls_block_pool = nullptr;
}

// A slot keeps one session between requests: its streams and device buffers
// are reused (they only grow), so admitting a request allocates nothing in
// the steady state and finishing one frees nothing.  Releasing a session
// costs the whole server: hipHostFree and hipFree wait for the whole device,
// and with many slots that is seconds of queued work, during which the
// runtime holds every other slot's launches and stream creations.  So a
// session is released only when its request failed or when the sessions of
// all slots together keep more than half the budget (the bytes they hold,
// not the footprint estimates, which assume full chunks): admitted requests
// are charged against the budget, so device memory stays within 1.5 budgets.
static void slot_loop(Server *S, int slot)
{
    tls_sleepy_waits = true;   // wait without keeping a core spinning
    std::unique_ptr<Session> sess;
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] {
                if (S->stop) return true;
                for (Request *q : S->pending)
                    if (S->used + q->footprint <= S->budget || S->used == 0) return true;
                return false;
            });
            if (S->stop) return;
            for (auto it = S->pending.begin(); it != S->pending.end(); ++it)
                if (S->used + (*it)->footprint <= S->budget || S->used == 0) {
                    r = *it;
                    S->pending.erase(it);
                    break;
                }
            if (!r) continue;
            S->used += r->footprint;
            r->state = 1;
            r->t_start = clk::now();
        }
        if (!sess) sess.reset(new Session());
        sess->ce.batcher = S->shared_updates ? &S->batcher : nullptr;
        sess->ce.loop = S->loop.get();
        run_request(S, r, *sess);
        bool release = r->rc != 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->held[(size_t)slot] = (int64_t)sess->pool.held;
            int64_t kept = 0;
            for (int64_t h : S->held) kept += h;
            if (2 * kept > S->budget) release = true;
            if (release) S->held[(size_t)slot] = 0;
        }
        if (release) sess.reset();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->used -= r->footprint;
            r->state = 2;
            r->t_done = clk::now();
        }
        S->cv_work.notify_all();
        S->cv_done.notify_all();
    }
}

static int copy_out(Request *r, double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats)
{
    if (stats) {
        *stats = r->st;
        stats->ms_total = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        // a request cancelled before admission was never started
        const clk::time_point adm = r->t_start.time_since_epoch().count() ? r->t_start : r->t_done;
        stats->ms_queued = std::chrono::duration<double, std::milli>(adm - r->t_submit).count();
    }
    if (r->rc) return set_error(r->rc, "%s", r->err.c_str());
    const size_t ne = r->ms_re.size();
    auto put = [&](double *re, double *im, const std::vector<double> &a, const std::vector<double> &b) {
        if (!re || a.size() != ne) return;
        if (im) {
            memcpy(re, a.data(), ne * sizeof(double));
            memcpy(im, b.data(), ne * sizeof(double));
        } else {
            for (size_t i = 0; i < ne; ++i) {
                re[2 * i] = a[i];
                re[2 * i + 1] = b[i];
            }
        }
    };
    put(ms_re, ms_im, r->ms_re, r->ms_im);
    put(mm_re, mm_im, r->mm_re, r->mm_im);
    return SCMC_OK;
}

} // namespace scmc

using namespace scmc;

extern "C" {

int64_t scmc_request_footprint(int32_t nv, int32_t nl, int32_t precision)
{
    if (nv < 1 || nl < 1) return -1;
    return request_footprint(nv, nl, precision, false);
}

int scmc_server_start(const scmc_server_config *cfg)
{
    std::lock_guard<std::mutex> g(g_srv_mu);
    if (g_srv) return set_error(SCMC_E_BUSY, "server already running");
    int rc = ensure_device(nullptr);
    if (rc) return rc;
    std::shared_ptr<Server> S(new Server());
    if (cfg) S->cfg = *cfg;
    if (S->cfg.max_active <= 0) S->cfg.max_active = 4;
    size_t freeb = 0, total = 0;
    (void)hipMemGetInfo(&freeb, &total);
    S->budget = S->cfg.hbm_budget_gb > 0 ? (int64_t)(S->cfg.hbm_budget_gb * 1e9) : (int64_t)(0.8 * (double)freeb);
    S->slice = S->cfg.paths_per_slice;
    // Stop a server the caller leaves running before the HIP runtime shuts
    // down at exit: HIP registers its teardown when it initialises (above),
    // and atexit handlers run last-registered first.
    static std::once_flag at_exit;
    std::call_once(at_exit, [] { std::atexit([] { (void)scmc_server_stop(); }); });
    S->shared_updates = S->cfg.shared_launches != 0;
    if (S->shared_updates && (rc = S->batcher.init())) return rc;
    if (S->cfg.resident_loop) {
        if (S->shared_updates) return set_error(SCMC_E_ARG, "shared_launches and resident_loop exclude each other");
        S->loop.reset(new ResidentLoop());
        if ((rc = S->loop->init(S->cfg.resident_workgroups))) return rc;
    }
    Server *raw = S.get();
    reaper_start();
    S->held.assign((size_t)S->cfg.max_active, 0);
    for (int i = 0; i < S->cfg.max_active; ++i) S->slots.emplace_back(slot_loop, raw, i);
    g_srv = std::move(S);
    return SCMC_OK;
}

int scmc_server_stop(void)
{
    std::shared_ptr<Server> S;
    {
        std::lock_guard<std::mutex> g(g_srv_mu);
        if (!g_srv) return SCMC_OK;
        S = std::move(g_srv);
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (auto &kv : S->all) kv.second->cancel = true;
        // requests never admitted finish now, as cancelled; running ones
        // stop at their next slice boundary and finish in their slot
        for (Request *r : S->pending) {
            r->rc = SCMC_E_CANCELLED;
            r->err = "server stopped before the request was admitted";
            r->state = 2;
            r->t_done = clk::now();
        }
        S->pending.clear();
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
    for (auto &t : S->slots) t.join();   // their sessions' buffers go to the reaper
    if (S->loop) {
        // every job is finished (the slots have drained): end the instance
        (void)S->loop->stop();
        S->loop.reset();
    }
    reaper_stop();
    // every request is finished now: wake any waiter that is still asleep
    S->cv_done.notify_all();
    return SCMC_OK;
}

static int submit(const scmc_problem *pr, int kind, int64_t max_itr, int32_t precision, int64_t *request_id)
{
    int rc = validate_problem(pr, max_itr, true);
    if (rc) return rc;
    if (precision != SCMC_FP64 && precision != SCMC_MIXED) return set_error(SCMC_E_ARG, "invalid precision %d", precision);
    std::lock_guard<std::mutex> g(g_srv_mu);
    if (!g_srv) return set_error(SCMC_E_ARG, "server not started");
    Server *S = g_srv.get();
    std::unique_ptr<Request> r(new Request());
    r->kind = kind;
    r->pc.assign(pr);
    r->max_itr = max_itr;
    r->precision = precision;
    r->footprint = kind == 0 ? request_footprint(pr->nv, pr->nl, precision, !(pr->is_ff_l && pr->is_ff_v))
                             : request_footprint_field(pr->nv, pr->nl, precision);
    r->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        if (request_id) *request_id = r->id;
        S->pending.push_back(r.get());
        S->all[r->id] = std::move(r);
    }
    S->cv_work.notify_all();
    return SCMC_OK;
}

static int finish(int64_t id, bool block, double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats)
{
    std::shared_ptr<Server> S;   // keeps the server alive across a concurrent stop
    {
        std::lock_guard<std::mutex> g(g_srv_mu);
        if (!g_srv) return set_error(SCMC_E_ARG, "server not started");
        S = g_srv;
    }
    std::unique_ptr<Request> done;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) return set_error(SCMC_E_NOTFOUND, "unknown request %lld", (long long)id);
        Request *r = it->second.get();
        if (block)
            S->cv_done.wait(lk, [&] { return r->state == 2; });
        else if (r->state != 2)
            return SCMC_E_BUSY;
        done = std::move(it->second);
        S->all.erase(it);
    }
    return copy_out(done.get(), ms_re, ms_im, mm_re, mm_im, stats);
}

int scmc_submit_mccov(const scmc_problem *pr, int64_t max_itr, int32_t precision, int64_t *request_id)
{
    return submit(pr, 0, max_itr, precision, request_id);
}

int scmc_submit_mcfield(const scmc_problem *pr, int64_t max_itr, int32_t precision, int64_t *request_id)
{
    return submit(pr, 1, max_itr, precision, request_id);
}

int scmc_poll(int64_t id, double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats)
{
    return finish(id, false, ms_re, ms_im, mm_re, mm_im, stats);
}

int scmc_wait(int64_t id, double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats)
{
    return finish(id, true, ms_re, ms_im, mm_re, mm_im, stats);
}

int scmc_cancel(int64_t id)
{
    std::lock_guard<std::mutex> g(g_srv_mu);
    if (!g_srv) return set_error(SCMC_E_ARG, "server not started");
    Server *S = g_srv.get();
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return set_error(SCMC_E_NOTFOUND, "unknown request %lld", (long long)id);
    Request *r = it->second.get();
    r->cancel = true;
    if (r->state == 0) {
        // never admitted: drop it from the queue and finish it as cancelled
        S->pending.erase(std::remove(S->pending.begin(), S->pending.end(), r), S->pending.end());
        r->rc = SCMC_E_CANCELLED;
        r->err = "request cancelled before admission";
        r->state = 2;
        r->t_done = clk::now();
        S->cv_done.notify_all();
    }
    return SCMC_OK;
}

} // extern "C"
