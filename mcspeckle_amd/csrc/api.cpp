// api.cpp -- C-ABI of libscmc.so (include/scmc.h): argument checks, device
// buffers and the chunked solve pipeline
//     trace(count) -> scan -> trace(emit) -> NEE factors -> MFMA update
// for MCcov, and trace -> coefficients -> field accumulation for MCfield.
#include "kernels.h"
#include "scmc.h"
#include "scmc_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace scmc {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

bool hip_ok(hipError_t e, const char *what, int *rc)
{
    if (e == hipSuccess) return true;
    *rc = set_error(SCMC_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return false;
}

// ------------------------------------------------------------ buffers
// hipFree synchronises the whole device: it returns once every kernel queued
// so far, on any stream, has finished.  That is what makes releasing a
// buffer safe while kernels queued on another stream may still read it (the
// path records of the previous chunk).  While a server runs, the frees go to
// a reaper thread instead, so a slot thread that grows a buffer does not
// wait for the other slots' work; the block is still freed only after that
// device-wide wait.
namespace {
struct Reaper {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<void *> q;
    bool stop = false;
    std::thread th;
};
std::mutex g_reaper_mu;
Reaper *g_reaper = nullptr;   // guarded by g_reaper_mu
} // namespace

static void device_free(void *p)
{
    {
        std::lock_guard<std::mutex> g(g_reaper_mu);
        if (g_reaper) {
            {
                std::lock_guard<std::mutex> l(g_reaper->mu);
                g_reaper->q.push_back(p);
            }
            g_reaper->cv.notify_one();
            return;
        }
    }
    (void)hipFree(p);
}

void reaper_start()
{
    std::lock_guard<std::mutex> g(g_reaper_mu);
    if (g_reaper) return;
    Reaper *R = new Reaper();
    R->th = std::thread([R] {
        (void)ensure_device(nullptr);
        std::unique_lock<std::mutex> l(R->mu);
        for (;;) {
            R->cv.wait(l, [R] { return R->stop || !R->q.empty(); });
            if (R->q.empty()) return;   // stop requested and drained
            void *p = R->q.front();
            R->q.pop_front();
            l.unlock();
            (void)hipFree(p);
            l.lock();
        }
    });
    g_reaper = R;
}

void reaper_stop()
{
    Reaper *R;
    {
        std::lock_guard<std::mutex> g(g_reaper_mu);
        R = g_reaper;
        g_reaper = nullptr;   // later frees are synchronous again
    }
    if (!R) return;
    {
        std::lock_guard<std::mutex> l(R->mu);
        R->stop = true;
    }
    R->cv.notify_all();
    R->th.join();   // frees what was queued
    delete R;
}

// Idle device blocks kept across sessions and servers.  A session's pool
// gives blocks up only once its request has synchronised (settle(), or the
// session's end, after its streams drained), so no queued kernel can still
// use them: they are kept here, up to kIdleCap bytes, for the next session or
// call instead of going back to the driver.  Without this, a server started
// after another one had freed its sessions' memory allocated all of it anew
// while its other slots kept work queued, and with 16 slots those
// allocations held every slot's HIP calls for up to 2.3 s.
namespace {
struct IdleBlock {
    int dev;   // the device the block was allocated on
    void *p;
};
struct IdleCache {
    std::mutex mu;
    std::multimap<size_t, IdleBlock> blocks;   // by size
    size_t bytes = 0;
};
IdleCache g_idle;
constexpr size_t kIdleCap = (size_t)64 << 30;
} // namespace

// blocks go back on the device that is current, as they were allocated
static void idle_put(void *p, size_t bytes)
{
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        std::lock_guard<std::mutex> g(g_idle.mu);
        if (g_idle.bytes + bytes <= kIdleCap) {
            g_idle.blocks.emplace(bytes, IdleBlock{dev, p});
            g_idle.bytes += bytes;
            return;
        }
    }
    device_free(p);
}

// smallest idle block of the current device of at least want bytes, at most
// twice that
static void *idle_take(size_t want, size_t *got)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(g_idle.mu);
    for (auto it = g_idle.blocks.lower_bound(want); it != g_idle.blocks.end() && it->first <= 2 * want; ++it) {
        if (it->second.dev != dev) continue;
        void *p = it->second.p;
        *got = it->first;
        g_idle.bytes -= it->first;
        g_idle.blocks.erase(it);
        return p;
    }
    return nullptr;
}

// every idle block back to the driver (an allocation failed), each freed on
// its own device
static void idle_trim()
{
    std::multimap<size_t, IdleBlock> b;
    {
        std::lock_guard<std::mutex> g(g_idle.mu);
        b.swap(g_idle.blocks);
        g_idle.bytes = 0;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (const auto &kv : b) {
        if (kv.second.dev != cur) (void)hipSetDevice(kv.second.dev);
        (void)hipFree(kv.second.p);
        if (kv.second.dev != cur) (void)hipSetDevice(cur);
    }
}

thread_local BlockPool *tls_block_pool = nullptr;

void *BlockPool::take(size_t want, size_t *got)
{
    // smallest ready block of at least want bytes, at most twice that
    size_t best = ready.size();
    for (size_t i = 0; i < ready.size(); ++i)
        if (ready[i].bytes >= want && ready[i].bytes <= 2 * want &&
            (best == ready.size() || ready[i].bytes < ready[best].bytes))
            best = i;
    if (best == ready.size()) return nullptr;
    void *q = ready[best].p;
    *got = ready[best].bytes;
    ready.erase(ready.begin() + (long)best);
    return q;
}

void BlockPool::settle()
{
    ready.insert(ready.end(), retired.begin(), retired.end());
    retired.clear();
    size_t total = 0;
    for (const Blk &b : ready) total += b.bytes;
    // over the cap: give the largest blocks back
    std::sort(ready.begin(), ready.end(), [](const Blk &a, const Blk &b) { return a.bytes < b.bytes; });
    while (total > cap && !ready.empty()) {
        total -= ready.back().bytes;
        held -= ready.back().bytes;
        idle_put(ready.back().p, ready.back().bytes);
        ready.pop_back();
    }
}

void BlockPool::trim(bool now)
{
    for (const std::vector<Blk> *v : {&ready, &retired})
        for (const Blk &b : *v) {
            held -= b.bytes;
            if (now)
                (void)hipFree(b.p);
            else
                idle_put(b.p, b.bytes);
        }
    ready.clear();
    retired.clear();
}

thread_local Arena *tls_arena = nullptr;

DevBuf::~DevBuf() { release(); }
void DevBuf::release()
{
    if (p && !borrowed) {
        if (pool)
            pool->put(p, bytes);
        else
            device_free(p);
    }
    p = nullptr;
    bytes = 0;
    pool = nullptr;
    borrowed = false;
}
bool DevBuf::ensure(size_t b, int *rc)
{
    if (b <= bytes && p) return true;
    release();
    if (Arena *A = tls_arena) {
        const size_t want = ((b ? b : 16) + 255) & ~(size_t)255;
        if (want > A->cap - A->used) {
            *rc = set_error(SCMC_E_NOMEM,
                            "allocation-free call: %zu more bytes needed, %zu of the %zu-byte scratch are left",
                            want, A->cap - A->used, A->cap);
            return false;
        }
        p = A->base + A->used;
        A->used += want;
        bytes = want;
        borrowed = true;
        return true;
    }
    // 1/8 headroom: the next chunk, a few percent larger, fits without a
    // reallocation (and its device-wide wait)
    size_t want = b ? b + b / 8 : 16;
    want = (want + 255) & ~(size_t)255;
    BlockPool *bp = tls_block_pool;
    size_t got = 0;
    void *q = bp ? bp->take(want, &got) : nullptr;
    if (!q && (q = idle_take(want, &got)) && bp) bp->held += got;
    if (q) {
        p = q;
        bytes = got;
        pool = bp;
        return true;
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        // kept blocks may be what the driver is missing
        (void)hipGetLastError();
        if (bp) bp->trim(true);
        idle_trim();
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        p = nullptr;
        *rc = set_error(SCMC_E_NOMEM, "device allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
        return false;
    }
    bytes = want;
    pool = bp;
    if (bp) bp->held += want;
    return true;
}

thread_local bool tls_sleepy_waits = false;

hipError_t event_wait(hipEvent_t e)
{
    if (!tls_sleepy_waits) return hipEventSynchronize(e);
    // a few yields first (short waits stay short), then sleeps growing to 0.2 ms
    for (int i = 0;; ++i) {
        const hipError_t q = hipEventQuery(e);
        if (q != hipErrorNotReady) return q;
        if (i < 32)
            std::this_thread::yield();
        else
            std::this_thread::sleep_for(std::chrono::microseconds(std::min(200, 10 * (i - 31))));
    }
}

hipError_t stream_wait(hipStream_t s)
{
    if (!tls_sleepy_waits) return hipStreamSynchronize(s);
    // one event per thread, destroyed when the thread ends (slot threads end
    // in scmc_server_stop, before the runtime shuts down)
    struct Holder {
        hipEvent_t ev = nullptr;
        ~Holder() { if (ev) (void)hipEventDestroy(ev); }
    };
    static thread_local Holder h;
    hipError_t e = hipSuccess;
    if (!h.ev && (e = hipEventCreateWithFlags(&h.ev, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventRecord(h.ev, s)) != hipSuccess) return e;
    return event_wait(h.ev);
}

int ensure_device(int *dev_out)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return set_error(SCMC_E_DEVICE, "no HIP device visible");
    int dev = 0;
    const char *env = getenv("SCMC_DEVICE");
    if (env) dev = atoi(env);
    if (dev < 0 || dev >= n) dev = 0;
    e = hipSetDevice(dev);
    if (e != hipSuccess) return set_error(SCMC_E_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    if (dev_out) *dev_out = dev;
    return SCMC_OK;
}

// ------------------------------------------------------ validation
static int check_phase(const scmc_phase &ph, const char *name, bool needs_type)
{
    if (needs_type && (ph.type < 1 || ph.type > 3)) return set_error(SCMC_E_ARG, "Invalid %s scattering type %d", name, ph.type);
    if (ph.type == SCMC_TABULATED) {
        if (!ph.amp_re || ph.n_amp < 2) return set_error(SCMC_E_ARG, "%s: tabulated evalAmp missing", name);
        if (!ph.icdf || ph.n_icdf < 2) return set_error(SCMC_E_ARG, "%s: tabulated sampleIcdf missing", name);
    }
    if (ph.type == SCMC_HG && !(ph.g >= -1 && ph.g <= 1)) return set_error(SCMC_E_ARG, "Invalid %s g input", name);
    return SCMC_OK;
}

int validate_problem(const scmc_problem *pr, int64_t max_itr, bool need_views)
{
    if (!pr) return set_error(SCMC_E_ARG, "null problem");
    if (pr->dim != 2 && pr->dim != 3) return set_error(SCMC_E_ARG, "Invalid box size");   // MCcov.m:48-51
    for (int j = 0; j < pr->dim; ++j)
        if (!(pr->box_max[j] >= pr->box_min[j])) return set_error(SCMC_E_ARG, "Invalid box size");
    if (pr->nl < 1 || !pr->l) return set_error(SCMC_E_ARG, "Invalid light source input");   // MCcov.m:60
    if (need_views && (pr->nv < 1 || !pr->v)) return set_error(SCMC_E_ARG, "Invalid view source input");   // MCcov.m:68
    if (!(pr->sigt > 0)) return set_error(SCMC_E_ARG, "sigt must be positive");
    if (!(pr->albedo >= 0 && pr->albedo <= 1)) return set_error(SCMC_E_ARG, "Invalid albedo input");
    if (!(pr->lambda > 0)) return set_error(SCMC_E_ARG, "lambda must be positive");
    if (pr->smp_flg != 1 && pr->smp_flg != 2) return set_error(SCMC_E_ARG, "smpFlg must be 1 or 2");
    if (max_itr < 1) return set_error(SCMC_E_ARG, "Invalid photons number");   // scmc.m:82-84
    int rc = check_phase(pr->amp, "ampfunc", true);
    if (rc) return rc;
    if (pr->has_amp0) {
        scmc_phase a0 = pr->amp0;
        a0.type = pr->amp.type;
        rc = check_phase(a0, "ampfunc0", false);
        if (rc) return rc;
    }
    return SCMC_OK;
}

// ---------------------------------------------------- device problem
int ProblemState::setup(const scmc_problem *pr, hipStream_t s)
{
    int rc = SCMC_OK;
    const int dim = pr->dim;
    DevProblem &P = dp;
    memset(&P, 0, sizeof P);
    het = false;   // setup_medium sets it (a reused state may have held a medium)
    P.dim = dim;
    P.sigt = pr->sigt;
    P.albedo = pr->albedo;
    P.lambda = pr->lambda;
    P.k = M_PI * 2.0 / pr->lambda;
    P.inv_lambda = 1.0 / pr->lambda;
    for (int j = 0; j < 3; ++j) {
        P.bmin[j] = j < dim ? pr->box_min[j] : 0.0;
        P.bmax[j] = j < dim ? pr->box_max[j] : 0.0;
    }
    P.nl = pr->nl;
    P.nv = pr->nv;
    P.is_ff_l = pr->is_ff_l ? 1 : 0;
    P.is_ff_v = pr->is_ff_v ? 1 : 0;
    P.do_cbs = pr->do_cbs ? 1 : 0;
    P.smp_flg = pr->smp_flg;
    P.seed = pr->seed;
    P.path_offset = pr->path_offset;
    P.w0p_iso = 1.0 / sqrt(pow(2.0, dim - 1) * M_PI);
    P.ff_sign = pr->is_ff_v ? -1.0 : 1.0;
    // a view's exit distance through the last axis's faces is at most
    // W_last / max(|v_last|, 1e-7); through another axis j at least the
    // point's distance to its nearer face j (|v_j| <= 1)
    P.far_last_bound = INFINITY;
    if (pr->is_ff_v && pr->nv > 0 && pr->v) {
        double vmin = INFINITY;
        for (int i = 0; i < pr->nv; ++i) vmin = std::min(vmin, fabs(pr->v[(size_t)i * dim + dim - 1]));
        const double b = (P.bmax[dim - 1] - P.bmin[dim - 1]) / std::max(vmin, 1.0000000000000004e-07);
        if (std::isfinite(b)) P.far_last_bound = (float)(b * (1 + 1e-5));
    }
    // unmeanl / meanl: MCcov.m:82-83, MCfield.m:77-82
    double um[3] = {0, 0, 0};
    if (pr->has_lmean0) {
        for (int j = 0; j < dim; ++j) um[j] = pr->lmean0[j];
    } else {
        for (int i = 0; i < pr->nl; ++i)
            for (int j = 0; j < dim; ++j) um[j] += pr->l[(size_t)i * dim + j];
        for (int j = 0; j < dim; ++j) um[j] /= pr->nl;
    }
    double ss = 0;
    for (int j = 0; j < dim; ++j) ss += um[j] * um[j];
    double nrm = sqrt(ss);
    for (int j = 0; j < 3; ++j) {
        P.unmean[j] = um[j];
        P.meanl[j] = j < dim ? um[j] / nrm : 0.0;
    }
    // directions / positions
    size_t lb = sizeof(double) * (size_t)dim * pr->nl, vb = sizeof(double) * (size_t)dim * std::max(pr->nv, 1);
    if (!dl.ensure(lb, &rc) || !dv.ensure(vb, &rc)) return rc;
    if (!hip_ok(hipMemcpyAsync(dl.p, pr->l, lb, hipMemcpyHostToDevice, s), "copy l", &rc)) return rc;
    if (pr->nv > 0 && pr->v)
        if (!hip_ok(hipMemcpyAsync(dv.p, pr->v, sizeof(double) * (size_t)dim * pr->nv, hipMemcpyHostToDevice, s),
                    "copy v", &rc))
            return rc;
    P.l = (const double *)dl.p;
    P.v = (const double *)dv.p;
    rc = upload_phase(pr->amp, pr->amp.type, tab_amp, tab_icdf, P.amp, s);
    if (rc) return rc;
    P.amp.iso = P.w0p_iso;
    P.has_amp0 = pr->has_amp0 ? 1 : 0;
    if (pr->has_amp0) {
        rc = upload_phase(pr->amp0, pr->amp.type, tab_amp0, tab_icdf0, P.amp0, s);
        if (rc) return rc;
        P.amp0.iso = P.w0p_iso;
    }
    // hold host copies alive until the async copies complete
    return hip_ok(stream_wait(s), "setup sync", &rc) ? SCMC_OK : rc;
}

int ProblemState::upload_phase(const scmc_phase &ph, int type, DevBuf &amp, DevBuf &icdf, DevPhase &out, hipStream_t s)
{
    int rc = SCMC_OK;
    out.type = type;
    out.g = ph.g;
    out.n_amp = 0;
    out.amp = nullptr;
    out.n_icdf = 0;
    out.cos_icdf = nullptr;
    if (type != SCMC_TABULATED) return SCMC_OK;
    std::vector<double> a((size_t)ph.n_amp * 2);
    for (int i = 0; i < ph.n_amp; ++i) {
        a[2 * i] = ph.amp_re[i];
        a[2 * i + 1] = ph.amp_im ? ph.amp_im[i] : 0.0;
    }
    if (!amp.ensure(a.size() * sizeof(double), &rc) || !icdf.ensure(sizeof(double) * ph.n_icdf, &rc)) return rc;
    // on the session's stream (a null-stream copy would wait for every other
    // session's work); a is local, so wait for the copies here
    if (!hip_ok(hipMemcpyAsync(amp.p, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy evalAmp",
                &rc) ||
        !hip_ok(hipMemcpyAsync(icdf.p, ph.icdf, sizeof(double) * ph.n_icdf, hipMemcpyHostToDevice, s),
                "copy sampleIcdf", &rc))
        return rc;
    launch_cos_table((double *)icdf.p, ph.n_icdf, s);
    if (!hip_ok(hipGetLastError(), "cos table", &rc) || !hip_ok(stream_wait(s), "table sync", &rc))
        return rc;
    out.n_amp = ph.n_amp;
    out.amp = (const double2 *)amp.p;
    out.n_icdf = ph.n_icdf;
    out.cos_icdf = (const double *)icdf.p;
    return SCMC_OK;
}

double box_volume(const scmc_problem *pr)
{
    double V = 1;
    for (int j = 0; j < pr->dim; ++j) V *= pr->box_max[j] - pr->box_min[j];
    return V;
}

// ------------------------------------------------------ traced chunk
int TraceChunk::run(const DevProblem &P, uint64_t p0, int64_t n, hipStream_t s, const HetMedium *M)
{
    int rc = SCMC_OK;
    np = n;
    if (!counts.ensure(sizeof(int32_t) * (size_t)n, &rc) || !offsets.ensure(sizeof(int64_t) * (size_t)(n + 1), &rc) ||
        !paths.ensure(sizeof(PathRec) * (size_t)n, &rc) ||
        !tsum.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(scan_tiles(n), 1), &rc) ||
        !ctr.ensure(sizeof(unsigned long long), &rc))
        return rc;
    if (M)
        launch_hetero_trace(P, *M, p0, n, (int32_t *)counts.p, nullptr, nullptr, nullptr, nullptr, false,
                            (unsigned long long *)ctr.p, s);
    else
        launch_trace_count(P, p0, n, (int32_t *)counts.p, (unsigned long long *)ctr.p, s);
    launch_scan((const int32_t *)counts.p, n, (int64_t *)offsets.p, (int64_t *)tsum.p, s);
    int64_t total = 0;
    if (!hip_ok(hipMemcpyAsync(&total, (int64_t *)offsets.p + n, sizeof total, hipMemcpyDeviceToHost, s), "copy total",
                &rc))
        return rc;
    if (!hip_ok(stream_wait(s), "trace sync", &rc)) return rc;
    nvert = total;
    if (!verts.ensure(sizeof(VertexRec) * (size_t)std::max<int64_t>(total, 1), &rc) ||
        !vpath.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1), &rc))
        return rc;
    if (M)
        launch_hetero_trace(P, *M, p0, n, nullptr, (int64_t *)offsets.p, (PathRec *)paths.p, (VertexRec *)verts.p,
                            (int32_t *)vpath.p, true, (unsigned long long *)ctr.p, s);
    else
        launch_trace_emit(P, p0, n, (const int64_t *)offsets.p, (PathRec *)paths.p, (VertexRec *)verts.p,
                          (int32_t *)vpath.p, (unsigned long long *)ctr.p, s);
    return hip_ok(hipGetLastError(), "trace launch", &rc) ? SCMC_OK : rc;
}

// Upload a host path set (SoA) as device records.
int TraceChunk::load(const scmc_pathset *ps, hipStream_t s)
{
    int rc = SCMC_OK;
    np = ps->n_paths;
    nvert = ps->n_vertices;
    std::vector<PathRec> hp((size_t)np);
    std::vector<VertexRec> hv((size_t)std::max<int64_t>(nvert, 1));
    std::vector<int32_t> hvp((size_t)std::max<int64_t>(nvert, 1));
    for (int64_t i = 0; i < np; ++i) {
        PathRec &r = hp[(size_t)i];
        for (int j = 0; j < 3; ++j) {
            r.x1[j] = ps->x1[3 * i + j];
            r.w1[j] = ps->w1[3 * i + j];
        }
        r.px = ps->px[i];
        r.pad = 0;
        r.w0p = make_double2(ps->w0p_re[i], ps->w0p_im ? ps->w0p_im[i] : 0.0);
        for (int64_t k = ps->offset[i]; k < ps->offset[i + 1]; ++k) {
            VertexRec &v = hv[(size_t)k];
            for (int j = 0; j < 3; ++j) {
                v.x[j] = ps->x[3 * k + j];
                v.ow[j] = ps->ow[3 * k + j];
            }
            v.weight = ps->weight[k];
            v.xi = ps->xi ? ps->xi[k] : 0.0;
            hvp[(size_t)k] = (int32_t)i;
        }
    }
    if (!offsets.ensure(sizeof(int64_t) * (size_t)(np + 1), &rc) || !paths.ensure(sizeof(PathRec) * (size_t)np, &rc) ||
        !verts.ensure(sizeof(VertexRec) * hv.size(), &rc) || !vpath.ensure(sizeof(int32_t) * hvp.size(), &rc))
        return rc;
    if (!hip_ok(hipMemcpy(offsets.p, ps->offset, sizeof(int64_t) * (size_t)(np + 1), hipMemcpyHostToDevice), "copy offsets", &rc) ||
        !hip_ok(hipMemcpy(paths.p, hp.data(), sizeof(PathRec) * (size_t)np, hipMemcpyHostToDevice), "copy paths", &rc) ||
        !hip_ok(hipMemcpy(verts.p, hv.data(), sizeof(VertexRec) * hv.size(), hipMemcpyHostToDevice), "copy vertices", &rc) ||
        !hip_ok(hipMemcpy(vpath.p, hvp.data(), sizeof(int32_t) * hvp.size(), hipMemcpyHostToDevice), "copy vpath", &rc))
        return rc;
    (void)s;
    return SCMC_OK;
}

// ---------------------------------------------------- covariance engine
// (Re)start a solve.  A reused engine keeps its device buffers (they only
// grow) and resets everything that belongs to the solve: accumulator copies
// (zeroed on first use), the tile order when N changes, statistics.
int CovEngine::init(const ProblemState &ps, bool mixed_, hipStream_t s)
{
    int rc = SCMC_OK;
    if ((rc = drain())) return rc;
    mixed = mixed_;
    nv = ps.dp.nv;
    nl = ps.dp.nl;
    n = nv * nl;
    tile = mixed ? kTileMixed : kTileF64;
    ks_m = ks_s = 0;
    last_scale_log2 = 0;
    ms_herk = herk_flops = ms_expand = ms_accum = ms_nee = 0;
    herk_launches = 0;
    nee_done = nullptr;
    if (order_n != n || order_tile != tile) order.release();
    if (!nonfinite.ensure(sizeof(unsigned long long), &rc)) return rc;
    launch_zero(nonfinite.p, sizeof(unsigned long long), s);
    if (ps.het) {
        if (!mean.ensure(sizeof(double2) * (size_t)n, &rc) ||
            !mean_part.ensure(sizeof(double2) * (size_t)n * kHetMeanSplits, &rc))
            return rc;
        launch_zero(mean.p, sizeof(double2) * (size_t)n, s);
    }
    return SCMC_OK;
}

static int pick_ksplit(int64_t k, int n, int tile)
{
    const int nt = (n + tile - 1) / tile;
    const int64_t tiles = (int64_t)nt * (nt + 1) / 2;
    // aim for >= 16 rounds of workgroups over 256 CUs x 2 (load balance) with
    // >= 2048 vertices per work item (FP32 sums stay short, launch amortised)
    int64_t want = (8192 + tiles - 1) / tiles;
    int64_t maxs = std::max<int64_t>(1, k / 2048);
    int64_t ks = std::min<int64_t>(want, maxs);
    if (ks >= 8) ks = (ks / 8) * 8;   // keep splits aligned with the 8 XCDs
    return (int)std::max<int64_t>(1, std::min<int64_t>(ks, 32));
}

int CovEngine::ensure_acc(DevBuf &acc, int &ks_have, int ks, hipStream_t s)
{
    int rc = SCMC_OK;
    if (acc.p && ks <= ks_have) return SCMC_OK;
    size_t one = sizeof(double2) * (size_t)n * n;
    if (ks_have == 0) {
        // first use in this solve: reuse the buffer when it is large enough
        if (!acc.ensure(one * ks, &rc)) return rc;
        launch_zero(acc.p, one * ks, s);
        ks_have = ks;
        return SCMC_OK;
    }
    // grow: allocate a larger set of copies, copy the existing sums over
    DevBuf nb;
    if (!nb.ensure(one * ks, &rc)) return rc;
    launch_zero(nb.p, one * ks, s);
    if (acc.p && ks_have > 0)
        if (!hip_ok(hipMemcpyAsync(nb.p, acc.p, one * ks_have, hipMemcpyDeviceToDevice, s), "grow acc", &rc)) return rc;
    if (!hip_ok(stream_wait(s), "acc sync", &rc)) return rc;
    std::swap(acc.p, nb.p);
    std::swap(acc.bytes, nb.bytes);
    std::swap(acc.pool, nb.pool);   // each block goes back where it came from
    ks_have = ks;
    return SCMC_OK;
}

// Producer half of a chunk (stream sp): NEE factors of the traced chunk into
// the current slot and, in mixed precision, the FP16 hi/lo row expansion.
// The slot's previous update must be finished before its buffers are
// overwritten: its events are harvested first (host wait, at most two chunks
// behind).
int CovEngine::produce(const ProblemState &ps, TraceChunk &tc, hipStream_t sp)
{
    int rc = SCMC_OK;
    CovSlot &S = slot[cur];
    if (S.used && (rc = harvest(S))) return rc;
    const DevProblem &P = ps.dp;
    const int64_t np = tc.np, nvert = tc.nvert;
    const int64_t k2 = nvert - np;
    const size_t cs = mixed ? sizeof(float2) : sizeof(double2);
    const bool cbs = P.do_cbs != 0;
    const size_t kk = (size_t)std::max<int64_t>(k2, 1);
    if (!S.A1.ensure(cs * kk * nv, &rc) || !S.B1.ensure(cs * kk * nl, &rc) || !S.SA.ensure(cs * (size_t)np * nv, &rc) ||
        !S.SB.ensure(cs * (size_t)np * nl, &rc) || !S.Gp.ensure(cs * (size_t)np * nl, &rc) ||
        !S.maxG.ensure(sizeof(float) * (size_t)np, &rc) || !S.maxF2.ensure(sizeof(float) * (size_t)np, &rc) ||
        !S.bound.ensure(16, &rc))
        return rc;
    if (cbs)
        if (!S.A2.ensure(cs * kk * nv, &rc) || !S.B2.ensure(cs * kk * nl, &rc) || !S.Fp2.ensure(cs * (size_t)np * nv, &rc))
            return rc;
    const bool near = !(P.is_ff_l && P.is_ff_v);
    if (near)
        if (!S.U.ensure(cs * (size_t)np * n, &rc) || !S.ones.ensure(cs * (size_t)np, &rc)) return rc;
    if (!S.ev[0])
        for (auto &e : S.ev)
            if (!hip_ok(hipEventCreate(&e), "event", &rc)) return rc;
    if (!S.bound_host && !hip_ok(hipHostMalloc((void **)&S.bound_host, sizeof(unsigned)), "pinned bound", &rc))
        return rc;
    launch_zero(S.bound.p, 16, sp);
    CovFactors f;
    f.A1 = S.A1.p;
    f.B1 = S.B1.p;
    f.A2 = cbs ? S.A2.p : nullptr;
    f.B2 = cbs ? S.B2.p : nullptr;
    f.SA = S.SA.p;
    f.SB = S.SB.p;
    f.Gp = S.Gp.p;
    f.Fp2 = cbs ? S.Fp2.p : nullptr;
    f.U = near ? S.U.p : nullptr;
    f.ones = near ? S.ones.p : nullptr;
    f.maxG = (float *)S.maxG.p;
    f.maxF2 = (float *)S.maxF2.p;
    f.bound = (unsigned *)S.bound.p;
    f.nonfinite = (unsigned long long *)nonfinite.p;
    if (ps.het) {
        if (!S.EV0.ensure(sizeof(double2) * (size_t)np * nv, &rc) || !S.UL0.ensure(sizeof(double2) * (size_t)np * nl, &rc))
            return rc;
        f.EV0 = (double2 *)S.EV0.p;
        f.UL0 = (double2 *)S.UL0.p;
    }
    (void)hipEventRecord(S.ev[0], sp);
    launch_nee_cov(P, ps.het ? &ps.hm : nullptr, np, nvert, (const PathRec *)tc.paths.p, (const VertexRec *)tc.verts.p,
                   (const int32_t *)tc.vpath.p, (const int64_t *)tc.offsets.p, f, mixed, sp);
    if (ps.het)
        launch_hetero_mean(P, np, (const PathRec *)tc.paths.p, f.EV0, f.UL0, (double2 *)mean.p,
                           (double2 *)mean_part.p, sp);
    (void)hipEventRecord(S.ev[1], sp);
    nee_done = S.ev[1];
    // the two rank-K problems of the chunk: multiple scattering, and single
    // scattering as Kronecker factors (far field) or whole rows with a unit
    // second factor (near field)
    HerkProblem &hm = S.hp[0], &hs = S.hp[1];
    hm = HerkProblem();
    hm.nv = nv;
    hm.nl = nl;
    hm.n = n;
    hm.k = k2;
    hm.A1 = S.A1.p;
    hm.B1 = S.B1.p;
    hm.A2 = cbs ? S.A2.p : nullptr;
    hm.B2 = cbs ? S.B2.p : nullptr;
    hm.bound = (const unsigned *)S.bound.p;
    hs = HerkProblem();
    hs.nv = near ? n : nv;
    hs.nl = near ? 1 : nl;
    hs.n = n;
    hs.k = np;
    hs.A1 = near ? S.U.p : S.SA.p;
    hs.B1 = near ? S.ones.p : S.SB.p;
    hs.bound = (const unsigned *)S.bound.p + 1;
    if (mixed) {
        if (!S.zbuf_m.ensure(mixed_zbuf_bytes(std::max<int64_t>(k2, 1), n), &rc) ||
            !S.zbuf_s.ensure(mixed_zbuf_bytes(std::max<int64_t>(np, 1), n), &rc))
            return rc;
        hm.zbuf = S.zbuf_m.p;
        hs.zbuf = S.zbuf_s.p;
        launch_expand_mixed(hm, sp);
        launch_expand_mixed(hs, sp);
    } else {
        if (!S.zbuf_m.ensure(fp64_zbuf_bytes(std::max<int64_t>(k2, 1), n), &rc) ||
            !S.zbuf_s.ensure(fp64_zbuf_bytes(std::max<int64_t>(np, 1), n), &rc))
            return rc;
        hm.zbuf = S.zbuf_m.p;
        hs.zbuf = S.zbuf_s.p;
        launch_expand_fp64(hm, sp);
        launch_expand_fp64(hs, sp);
    }
    (void)hipEventRecord(S.ev[2], sp);
    S.np = np;
    S.k2 = k2;
    return hip_ok(hipGetLastError(), "covariance factors", &rc) ? SCMC_OK : rc;
}

// Consumer half (stream sc): waits for the slot's factors, runs the
// matrix-core updates into the accumulators (only ever written on sc, so
// chunks accumulate in order), then moves to the other slot.
// ------------------------------------------------------ shared updates
int UpdateBatcher::init()
{
    int rc = SCMC_OK;
    if (!stream && !hip_ok(hipStreamCreate(&stream), "batcher stream", &rc)) return rc;
    return SCMC_OK;
}

UpdateBatcher::~UpdateBatcher()
{
    if (stream) (void)stream_wait(stream);
    for (hipEvent_t e : inflight) (void)hipEventDestroy(e);
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

int UpdateBatcher::run(Job &j)
{
    int rc = SCMC_OK;
    std::unique_lock<std::mutex> lk(mu);
    pending.push_back(&j);
    while (!j.launched) {
        // retire finished groups (non-blocking queries; an error status also
        // retires the group, so a failed device cannot leave waiters spinning)
        for (size_t i = 0; i < inflight.size();) {
            if (hipEventQuery(inflight[i]) != hipErrorNotReady) {
                spare.push_back(inflight[i]);
                inflight.erase(inflight.begin() + (long)i);
            } else {
                ++i;
            }
        }
        if (launching || (int)inflight.size() >= max_inflight) {
            cv.wait_for(lk, std::chrono::microseconds(100));
            continue;
        }
        // launch everything pending as one group (more than kMaxUpdateGroup
        // problems go out as several launches in the same stream order)
        launching = true;
        std::vector<Job *> take;
        take.swap(pending);
        hipEvent_t end = nullptr;
        if (!spare.empty()) {
            end = spare.back();
            spare.pop_back();
        }
        lk.unlock();
        std::vector<const HerkProblem *> list;
        bool ok = end || hip_ok(hipEventCreateWithFlags(&end, hipEventDisableTiming), "batcher event", &rc);
        for (Job *q : take) {
            ok = ok && hip_ok(hipStreamWaitEvent(stream, q->ready, 0), "batcher wait", &rc);
            if (ok) (void)hipEventRecord(q->t0, stream);
            list.push_back(q->hm);
            list.push_back(q->hs);
        }
        if (ok) {
            launch_herk_mixed_group(list.data(), (int)list.size(), stream);
            for (Job *q : take) (void)hipEventRecord(q->t1, stream);
            (void)hipEventRecord(end, stream);
            ok = hip_ok(hipGetLastError(), "shared update", &rc);
        }
        lk.lock();
        launching = false;
        if (end) inflight.push_back(end);
        groups += 1;
        jobs += (int64_t)take.size();
        // a failed launch releases its waiters with the error
        for (Job *q : take) {
            q->launched = true;
            if (!ok) {
                q->rc = rc;
                q->err = scmc_last_error();
            }
        }
        cv.notify_all();
    }
    if (j.rc) return set_error(j.rc, "%s", j.err.c_str());
    return SCMC_OK;
}

int CovEngine::consume(hipStream_t sc)
{
    int rc = SCMC_OK;
    CovSlot &S = slot[cur];
    // resident loop: the previous chunk's update runs outside this stream's
    // order and adds into the accumulators set up (and possibly grown) below
    if (mixed && loop && (rc = wait_loop(slot[cur ^ 1]))) return rc;
    if (!hip_ok(hipStreamWaitEvent(sc, S.ev[2], 0), "wait factors", &rc)) return rc;
    const int ksm = pick_ksplit(S.k2, n, tile), kss = pick_ksplit(S.np, n, tile);
    if ((rc = ensure_acc(acc_m, ks_m, std::max(ksm, ks_m), sc))) return rc;
    if ((rc = ensure_acc(acc_s, ks_s, std::max(kss, ks_s), sc))) return rc;
    if (!order.p) {
        // kept in order_host until the next N: the copy runs on the solve's
        // stream (a null-stream copy would wait for every other session)
        order_n = n;
        order_tile = tile;
        order_host = blocked_tile_order((n + tile - 1) / tile);
        if (!order.ensure(sizeof(int) * order_host.size(), &rc)) return rc;
        if (!hip_ok(hipMemcpyAsync(order.p, order_host.data(), sizeof(int) * order_host.size(), hipMemcpyHostToDevice,
                                   sc),
                    "tile order", &rc))
            return rc;
    }
    HerkProblem &hm = S.hp[0], &hs = S.hp[1];
    hm.acc = (double2 *)acc_m.p;
    hm.ksplit = ksm;
    hm.order = order.p;
    hs.acc = (double2 *)acc_s.p;
    hs.ksplit = kss;
    hs.order = order.p;
    // one launch for both problems, mixed or FP64 (ev[3]..ev[4])
    if (mixed && loop) {
        // serving with the resident loop: the chunk's two updates become jobs
        // once their inputs (factors, expansion, zeroed accumulators, tile
        // order) are complete, so the loop never waits on a kernel; the
        // previous chunk's jobs finished above
        (void)hipEventRecord(S.ev[6], sc);
        if (!hip_ok(event_wait(S.ev[6]), "update inputs", &rc)) return rc;
        S.loop_t0 = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
        S.looped = true;
        if ((rc = loop->publish(hm, &S.tk[0])) || (rc = loop->publish(hs, &S.tk[1]))) return rc;
        // extra workgroups where the hardware finds room, as a launch would
        // have (they leave when no item is left, the resident ones stay)
        if (loop->helpers && (rc = loop->help(S.tk[0], S.tk[1], sc))) return rc;
    } else if (mixed && batcher && n <= kSharedUpdateMaxN) {
        // serving: one launch shared with the other small requests' chunks
        // (a larger problem fills the GPU alone and launches on its own
        // stream, so small updates do not queue behind it)
        // the accumulators were zeroed / grown and the tile order uploaded on
        // sc after ev[2]: the shared stream waits for all of it
        (void)hipEventRecord(S.ev[6], sc);
        UpdateBatcher::Job j;
        j.hm = &hm;
        j.hs = &hs;
        j.ready = S.ev[6];
        j.t0 = S.ev[3];
        j.t1 = S.ev[4];
        if ((rc = batcher->run(j))) return rc;
        if (!hip_ok(hipStreamWaitEvent(sc, S.ev[4], 0), "wait shared update", &rc)) return rc;
    } else {
        (void)hipEventRecord(S.ev[3], sc);
        if (mixed)
            launch_herk_mixed_update(hm, &hs, sc);
        else
            launch_herk_fp64(hm, &hs, sc);
        (void)hipEventRecord(S.ev[4], sc);
    }
    // the scale bound goes to pinned memory on the stream: a blocking copy on
    // the null stream would wait for the update in flight on the other stream
    if (mixed && !hip_ok(hipMemcpyAsync(S.bound_host, S.bound.p, sizeof(unsigned), hipMemcpyDeviceToHost, sc),
                         "copy bound", &rc))
        return rc;
    (void)hipEventRecord(S.ev[5], sc);
    if (!hip_ok(hipGetLastError(), "covariance update", &rc)) return rc;
    S.used = true;
    cur ^= 1;
    return SCMC_OK;
}

// Wait for a slot's events and add its timings to the statistics.
int CovEngine::wait_loop(CovSlot &S)
{
    int rc = SCMC_OK;
    for (LoopTicket &t : S.tk) {
        if (t.idx >= 0 && loop && (rc = loop->wait(t))) return rc;
        t.idx = -1;
    }
    return SCMC_OK;
}

int CovEngine::harvest(CovSlot &S)
{
    int rc = SCMC_OK;
    if (!S.used) return SCMC_OK;
    const bool looped = S.looped;
    S.looped = false;
    if ((rc = wait_loop(S))) return rc;
    if (!hip_ok(event_wait(S.ev[5]), "covariance sync", &rc)) return rc;
    float tn = 0, tx = 0, th = 0, ts = 0;
    (void)hipEventElapsedTime(&tn, S.ev[0], S.ev[1]);
    (void)hipEventElapsedTime(&tx, S.ev[1], S.ev[2]);
    if (looped) {
        // resident loop: publish to done as the host saw it (queueing included)
        th = (float)(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
                         .count() -
                     S.loop_t0);
    } else {
        (void)hipEventElapsedTime(&th, S.ev[3], S.ev[4]);
        (void)hipEventElapsedTime(&ts, S.ev[4], S.ev[5]);
    }
    ms_nee += tn;
    ms_accum += tx + th + ts;
    // the timed launch's rank-1 terms: both problems
    const int64_t kt = S.k2 + S.np;
    if (kt > 0) {
        ms_herk += th;
        ms_expand += tx;
        herk_launches += 1;
        herk_flops += 8.0 * (double)n * (double)n * (double)kt;
    }
    if (mixed) {
        const unsigned b = *S.bound_host;
        float bf;
        memcpy(&bf, &b, sizeof bf);
        int ex = 0;
        if (bf > 0) frexpf(bf, &ex);
        last_scale_log2 = 15 - ex;
    }
    S.used = false;
    return SCMC_OK;
}

// The solve's count of vertices with NaN/Inf factors (after the stream's work).
int CovEngine::read_nonfinite(int64_t *out, hipStream_t s)
{
    int rc = SCMC_OK;
    unsigned long long h = 0;
    if (nonfinite.p &&
        (!hip_ok(hipMemcpyAsync(&h, nonfinite.p, sizeof h, hipMemcpyDeviceToHost, s), "copy nonfinite", &rc) ||
         !hip_ok(stream_wait(s), "nonfinite sync", &rc)))
        return rc;
    *out = (int64_t)h;
    return SCMC_OK;
}

int CovEngine::drain()
{
    int rc = SCMC_OK;
    for (auto &S : slot)
        if ((rc = harvest(S))) return rc;
    return SCMC_OK;
}

CovEngine::~CovEngine()
{
    // resident-loop jobs still read this engine's buffers
    for (auto &S : slot) (void)wait_loop(S);
    for (auto &S : slot) {
        for (auto &e : S.ev)
            if (e) (void)hipEventDestroy(e);
        if (S.bound_host) (void)hipHostFree(S.bound_host);
    }
}

// One chunk on one stream (path-set replay).
int CovEngine::chunk(const ProblemState &ps, TraceChunk &tc, hipStream_t s)
{
    int rc = produce(ps, tc, s);
    if (!rc) rc = consume(s);
    if (!rc) rc = drain();
    return rc;
}

int CovEngine::finalize(const ProblemState &ps, const scmc_problem *pr, int64_t max_itr, double *ms_re, double *ms_im,
                        double *mm_re, double *mm_im, hipStream_t s)
{
    int rc = SCMC_OK;
    const size_t ne = (size_t)n * n;
    DevBuf &out = fin_out, &af = fin_af;
    if (!out.ensure(sizeof(double) * 2 * ne, &rc) || !af.ensure(sizeof(double2) * (size_t)n, &rc)) return rc;
    // MCcov.m:293-299: x V sigt / maxItr; MCCovHetro.m:258-265: / maxItr only
    // (its first-vertex density already carries sum(sigt) * cell volume)
    const double scale = ps.het ? 1.0 / (double)max_itr : 1.0 / (double)max_itr * box_volume(pr) * pr->sigt;
    // far field: the single-scatter amplitude table is a per-row factor; near
    // field it is already inside the rows
    const bool far = ps.dp.is_ff_l && ps.dp.is_ff_v;
    if (far) launch_af_table(ps.dp, (double2 *)af.p, s);
    double *o = (double *)out.p;
    struct Dst {
        double *re, *im;
    };
    auto emit = [&](const DevBuf &acc, int ks, double sc, const double2 *rowf, Dst d) -> int {
        if (!d.re) return SCMC_OK;
        if (!acc.p) {
            launch_zero(out.p, sizeof(double) * 2 * ne, s);
        } else {
            launch_cov_finalize((const double2 *)acc.p, ks, nv, nl, sc, rowf, o, d.im ? o + ne : nullptr, s);
        }
        int r2 = SCMC_OK;
        if (d.im) {
            if (!hip_ok(hipMemcpyAsync(d.re, o, sizeof(double) * ne, hipMemcpyDeviceToHost, s), "copy re", &r2) ||
                !hip_ok(hipMemcpyAsync(d.im, o + ne, sizeof(double) * ne, hipMemcpyDeviceToHost, s), "copy im", &r2))
                return r2;
        } else {
            if (!hip_ok(hipMemcpyAsync(d.re, o, sizeof(double) * 2 * ne, hipMemcpyDeviceToHost, s), "copy out", &r2))
                return r2;
        }
        return hip_ok(stream_wait(s), "finalize sync", &r2) ? SCMC_OK : r2;
    };
    // MCcov.m:293-299: Mm/2 with CBS, then /maxItr*V*sigt; Ms /maxItr*V*sigt
    if ((rc = emit(acc_s, ks_s, scale, far ? (const double2 *)af.p : nullptr, Dst{ms_re, ms_im}))) return rc;
    if ((rc = emit(acc_m, ks_m, pr->do_cbs ? scale / 2 : scale, nullptr, Dst{mm_re, mm_im}))) return rc;
    return SCMC_OK;
}

int CovEngine::finalize_mean(int64_t max_itr, double *re, double *im, hipStream_t s)
{
    int rc = SCMC_OK;
    if (!re) return SCMC_OK;
    std::vector<double> h((size_t)n * 2);
    if (!hip_ok(hipMemcpyAsync(h.data(), mean.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s), "copy mean", &rc) ||
        !hip_ok(stream_wait(s), "mean sync", &rc))
        return rc;
    const double sc = 1.0 / (double)max_itr;   // MCCovHetro.m:276 with cs = 1
    for (int i = 0; i < n; ++i) {
        if (im) {
            re[i] = h[2 * (size_t)i] * sc;
            im[i] = h[2 * (size_t)i + 1] * sc;
        } else {
            re[2 * i] = h[2 * (size_t)i] * sc;
            re[2 * i + 1] = h[2 * (size_t)i + 1] * sc;
        }
    }
    return SCMC_OK;
}

int64_t chunk_paths(int n, int64_t remaining, double mean_len, bool mixed)
{
    // vertices per chunk: large enough to fill the GPU, bounded by the
    // expanded-row buffer (8 B per row and vertex mixed, 16 B FP64: ~4 and
    // ~2.5 GB) and factor memory
    double cap = mixed ? 4.0e9 / (8.0 * n) : 2.5e9 / (16.0 * n);
    int64_t target_vertices = (int64_t)std::max(65536.0, std::min(4.0e6, cap));
    int64_t p = (int64_t)(target_vertices / std::max(mean_len, 1.0));
    p = std::max<int64_t>(256, std::min<int64_t>(p, 1 << 20));
    return std::min(p, remaining);
}

} // namespace scmc

using namespace scmc;

// ======================================================= C entry points
extern "C" {

const char *scmc_last_error(void) { return g_err.c_str(); }

const char *scmc_version(void) { return "scmc-mi355x 0.3 (abi 3, gfx950)"; }

int scmc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}


} // extern "C"

namespace scmc {

void ProblemCopy::assign(const scmc_problem *src)
{
    pr = *src;
    const int dim = src->dim;
    l.assign(src->l, src->l + (size_t)dim * src->nl);
    pr.l = l.data();
    if (src->v && src->nv > 0) {
        v.assign(src->v, src->v + (size_t)dim * src->nv);
        pr.v = v.data();
    }
    auto cp = [](const scmc_phase &ph, scmc_phase &out, std::vector<double> &re, std::vector<double> &im,
                 std::vector<double> &ic) {
        if (ph.type != SCMC_TABULATED) return;
        re.assign(ph.amp_re, ph.amp_re + ph.n_amp);
        if (ph.amp_im)
            im.assign(ph.amp_im, ph.amp_im + ph.n_amp);
        else
            im.assign((size_t)ph.n_amp, 0.0);
        ic.assign(ph.icdf, ph.icdf + ph.n_icdf);
        out.amp_re = re.data();
        out.amp_im = im.data();
        out.icdf = ic.data();
    };
    cp(src->amp, pr.amp, a_re, a_im, a_ic);
    if (src->has_amp0) {
        scmc_phase a0 = src->amp0;
        a0.type = src->amp.type;
        cp(a0, pr.amp0, b_re, b_im, b_ic);
    }
}

int Session::init(const scmc_problem *pr, int kind_, int precision, const double *wl_re, const double *wl_im,
                  const scmc_medium *med)
{
    int rc = validate_problem(pr, 1, true);
    if (rc) return rc;
    if (precision != SCMC_FP64 && precision != SCMC_MIXED) return set_error(SCMC_E_ARG, "invalid precision %d", precision);
    if (kind_ != 0 && kind_ != 1) return set_error(SCMC_E_ARG, "invalid session kind %d", kind_);
    if ((rc = ensure_device(nullptr))) return rc;
    kind = kind_;
    mixed = precision == SCMC_MIXED;
    pc.assign(pr);
    // a reused session keeps its streams and device buffers
    if (!stream && !hip_ok(hipStreamCreate(&stream), "stream", &rc)) return rc;
    if (kind == 0 && !stream2 && !hip_ok(hipStreamCreate(&stream2), "stream", &rc)) return rc;
    mean_len = 20;
    if ((rc = ps.setup(&pc.pr, stream))) return rc;
    if (med && (rc = ps.setup_medium(&pc.pr, med, stream))) return rc;
    memset(&st, 0, sizeof st);
    return kind == 0 ? ce.init(ps, mixed, stream) : fe.init(ps, stream, wl_re, wl_im);
}

Session::~Session()
{
    if (stream) (void)stream_wait(stream);
    if (stream2) (void)stream_wait(stream2);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
}

int Session::run(int64_t first, int64_t n, bool settle)
{
    int rc = SCMC_OK;
    if (first < 0 || n < 0) return set_error(SCMC_E_ARG, "invalid path range");
    for (int64_t p0 = first; p0 < first + n;) {
        const int64_t rem = first + n - p0;
        const int64_t cap = kind == 0 ? chunk_paths(ce.n, rem, mean_len, mixed)
                                      : field_chunk_paths(pc.pr.nv, pc.pr.nl, rem, mean_len);
        // equal chunks rather than full ones and a short tail
        const int64_t nch = (rem + cap - 1) / cap, np = (rem + nch - 1) / nch;
        auto a = std::chrono::steady_clock::now();
        hipStream_t ts = kind == 0 ? stream2 : stream;
        // the trace overwrites the path records the previous chunk's factors read
        if (kind == 0 && ce.nee_done && !hip_ok(hipStreamWaitEvent(ts, ce.nee_done, 0), "wait factors", &rc)) return rc;
        if ((rc = tc.run(ps.dp, (uint64_t)p0, np, ts, ps.het ? &ps.hm : nullptr))) return rc;
        if (!hip_ok(stream_wait(ts), "trace", &rc)) return rc;
        st.ms_trace += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
        if (kind == 0) {
            // the tracer (a narrow kernel: one thread per path) runs on stream2
            // beside the matrix-core update of the previous chunk; the factors,
            // the expansion and the update fill the whole GPU and follow each
            // other on stream
            if ((rc = ce.produce(ps, tc, stream)) || (rc = ce.consume(stream))) return rc;
        } else {
            rc = fe.chunk(ps, tc, mixed, stream, &st.ms_accum);
        }
        if (rc) return rc;
        st.paths += np;
        st.vertices += tc.nvert;
        st.chunks += 1;
        mean_len = (double)st.vertices / (double)st.paths;
        p0 += np;
    }
    if (!settle) return SCMC_OK;
    if (kind == 0) {
        if ((rc = ce.drain())) return rc;
        st.ms_accum = ce.ms_accum;
        st.ms_nee = ce.ms_nee;
        return ce.read_nonfinite(&st.nonfinite, stream);
    }
    return fe.read_nonfinite(&st.nonfinite, stream);
}

int Session::result(int64_t max_itr, double *a_re, double *a_im, double *b_re, double *b_im)
{
    if (max_itr < 1) return set_error(SCMC_E_ARG, "Invalid photons number");
    if (kind == 0) return ce.finalize(ps, &pc.pr, max_itr, a_re, a_im, b_re, b_im, stream);
    return fe.finalize(&pc.pr, max_itr, a_re, a_im, stream);
}

void Session::fill_stats(scmc_stats *out) const
{
    *out = st;
    if (kind == 0) {
        out->fp16_scale_log2 = ce.last_scale_log2;
        out->ms_herk = ce.ms_herk;
        out->herk_launches = ce.herk_launches;
        out->herk_flops = ce.herk_flops;
        out->ms_expand = ce.ms_expand;
    }
}

// Session handles.  A call holds a reference to its session and its call
// lock for the whole call: calls on one session from several threads run one
// after another, and scmc_session_destroy during another thread's call only
// drops the handle -- the session is freed when that call returns.
struct SessionHandle {
    std::mutex call_mu;
    std::unique_ptr<Session> s;
};
static std::mutex g_sess_mu;
static std::map<int64_t, std::shared_ptr<SessionHandle>> g_sessions;
static int64_t g_next_session = 1;

static std::shared_ptr<SessionHandle> find_session(int64_t h)
{
    std::lock_guard<std::mutex> g(g_sess_mu);
    auto it = g_sessions.find(h);
    return it == g_sessions.end() ? nullptr : it->second;
}

} // namespace scmc

extern "C" {

int scmc_session_create(const scmc_problem *pr, int32_t kind, int32_t precision, int64_t *session)
{
    if (!session) return set_error(SCMC_E_ARG, "null session handle");
    std::unique_ptr<Session> s(new Session());
    int rc = s->init(pr, kind, precision);
    if (rc) return rc;
    // sessions the caller never destroys are released before the HIP runtime
    // shuts down at exit (init has initialised HIP; atexit runs LIFO)
    static std::once_flag at_exit;
    std::call_once(at_exit, [] {
        std::atexit([] {
            std::map<int64_t, std::shared_ptr<SessionHandle>> left;
            {
                std::lock_guard<std::mutex> g(g_sess_mu);
                left.swap(g_sessions);
            }
        });
    });
    std::shared_ptr<SessionHandle> h(new SessionHandle());
    h->s = std::move(s);
    std::lock_guard<std::mutex> g(g_sess_mu);
    *session = g_next_session++;
    g_sessions[*session] = std::move(h);
    return SCMC_OK;
}

int scmc_session_run(int64_t session, int64_t first, int64_t n_paths)
{
    auto h = find_session(session);
    if (!h) return set_error(SCMC_E_NOTFOUND, "unknown session %lld", (long long)session);
    std::lock_guard<std::mutex> g(h->call_mu);
    return h->s->run(first, n_paths);
}

int scmc_session_result(int64_t session, int64_t max_itr, double *a_re, double *a_im, double *b_re, double *b_im)
{
    auto h = find_session(session);
    if (!h) return set_error(SCMC_E_NOTFOUND, "unknown session %lld", (long long)session);
    std::lock_guard<std::mutex> g(h->call_mu);
    return h->s->result(max_itr, a_re, a_im, b_re, b_im);
}

int scmc_session_stats(int64_t session, scmc_stats *stats)
{
    auto h = find_session(session);
    if (!h) return set_error(SCMC_E_NOTFOUND, "unknown session %lld", (long long)session);
    std::lock_guard<std::mutex> g(h->call_mu);
    if (stats) h->s->fill_stats(stats);
    return SCMC_OK;
}

int scmc_session_destroy(int64_t session)
{
    std::shared_ptr<SessionHandle> victim;
    {
        std::lock_guard<std::mutex> g(g_sess_mu);
        auto it = g_sessions.find(session);
        if (it == g_sessions.end()) return set_error(SCMC_E_NOTFOUND, "unknown session %lld", (long long)session);
        victim = std::move(it->second);
        g_sessions.erase(it);
    }
    return SCMC_OK;
}

static int one_shot(const scmc_problem *pr, int kind, int64_t max_itr, int32_t precision, double *a_re, double *a_im,
                    double *b_re, double *b_im, scmc_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    int rc = validate_problem(pr, max_itr, true);
    if (rc) return rc;
    Session s;
    if ((rc = s.init(pr, kind, precision))) return rc;
    if ((rc = s.run(0, max_itr))) return rc;
    rc = s.result(max_itr, a_re, a_im, b_re, b_im);
    if (stats) {
        s.fill_stats(stats);
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return rc;
}

int scmc_mccov(const scmc_problem *pr, int64_t max_itr, int32_t precision, double *ms_re, double *ms_im,
               double *mm_re, double *mm_im, scmc_stats *stats)
{
    return one_shot(pr, 0, max_itr, precision, ms_re, ms_im, mm_re, mm_im, stats);
}

int scmc_mcfield(const scmc_problem *pr, int64_t max_itr, int32_t precision, double *u_re, double *u_im,
                 scmc_stats *stats)
{
    return one_shot(pr, 1, max_itr, precision, u_re, u_im, nullptr, nullptr, stats);
}


int scmc_mcfield_wave(const scmc_problem *pr, const double *wl_re, const double *wl_im, int64_t max_itr,
                      int32_t precision, double *u_re, double *u_im, scmc_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    int rc = validate_problem(pr, max_itr, true);
    if (rc) return rc;
    if (!wl_re) return set_error(SCMC_E_ARG, "Wl must have one weight per light");
    Session s;
    if ((rc = s.init(pr, 1, precision, wl_re, wl_im))) return rc;
    if ((rc = s.run(0, max_itr))) return rc;
    rc = s.result(max_itr, u_re, u_im, nullptr, nullptr);
    if (stats) {
        s.fill_stats(stats);
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return rc;
}

int scmc_trace(const scmc_problem *pr, int64_t n_paths, int64_t *n_vertices, int64_t *offset, double *x1, double *px,
               double *w1, double *w0p_re, double *w0p_im, double *x, double *ow, double *weight, double *xi)
{
    int rc = validate_problem(pr, n_paths, false);
    if (rc) return rc;
    if ((rc = ensure_device(nullptr))) return rc;
    hipStream_t s;
    if (!hip_ok(hipStreamCreate(&s), "stream", &rc)) return rc;
    {
        ProblemState ps;
        TraceChunk tc;
        if (!(rc = ps.setup(pr, s)) && !(rc = tc.run(ps.dp, 0, n_paths, s))) {
            if (n_vertices) *n_vertices = tc.nvert;
            if (offset && x) {
                std::vector<int64_t> off((size_t)n_paths + 1);
                std::vector<PathRec> hp((size_t)n_paths);
                std::vector<VertexRec> hv((size_t)std::max<int64_t>(tc.nvert, 1));
                if (hip_ok(hipMemcpy(off.data(), tc.offsets.p, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost), "copy", &rc) &&
                    hip_ok(hipMemcpy(hp.data(), tc.paths.p, sizeof(PathRec) * hp.size(), hipMemcpyDeviceToHost), "copy", &rc) &&
                    hip_ok(hipMemcpy(hv.data(), tc.verts.p, sizeof(VertexRec) * (size_t)tc.nvert, hipMemcpyDeviceToHost), "copy", &rc)) {
                    for (int64_t i = 0; i <= n_paths; ++i) offset[i] = off[(size_t)i];
                    for (int64_t i = 0; i < n_paths; ++i) {
                        for (int j = 0; j < 3; ++j) {
                            x1[3 * i + j] = hp[(size_t)i].x1[j];
                            w1[3 * i + j] = hp[(size_t)i].w1[j];
                        }
                        px[i] = hp[(size_t)i].px;
                        w0p_re[i] = hp[(size_t)i].w0p.x;
                        w0p_im[i] = hp[(size_t)i].w0p.y;
                    }
                    for (int64_t k = 0; k < tc.nvert; ++k) {
                        for (int j = 0; j < 3; ++j) {
                            x[3 * k + j] = hv[(size_t)k].x[j];
                            ow[3 * k + j] = hv[(size_t)k].ow[j];
                        }
                        weight[k] = hv[(size_t)k].weight;
                        xi[k] = hv[(size_t)k].xi;
                    }
                }
            }
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

int scmc_mccov_replay(const scmc_problem *pr, const scmc_pathset *pset, int64_t max_itr, int32_t precision,
                      double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats)
{
    int rc = validate_problem(pr, max_itr, true);
    if (rc) return rc;
    if (!pset || pset->n_paths < 1) return set_error(SCMC_E_ARG, "empty path set");
    if ((rc = ensure_device(nullptr))) return rc;
    hipStream_t s;
    if (!hip_ok(hipStreamCreate(&s), "stream", &rc)) return rc;
    {
        ProblemState ps;
        TraceChunk tc;
        CovEngine ce;
        if (!(rc = ps.setup(pr, s)) && !(rc = ce.init(ps, precision == SCMC_MIXED, s)) && !(rc = tc.load(pset, s)) &&
            !(rc = ce.chunk(ps, tc, s)))
            rc = ce.finalize(ps, pr, max_itr, ms_re, ms_im, mm_re, mm_im, s);
        if (!rc && stats) {
            memset(stats, 0, sizeof *stats);
            stats->paths = tc.np;
            stats->vertices = tc.nvert;
            stats->chunks = 1;
            rc = ce.read_nonfinite(&stats->nonfinite, s);
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

int scmc_mcfield_replay(const scmc_problem *pr, const scmc_pathset *pset, int64_t max_itr, int32_t precision,
                        double *u_re, double *u_im, scmc_stats *stats)
{
    int rc = validate_problem(pr, max_itr, true);
    if (rc) return rc;
    if (!pset || pset->n_paths < 1) return set_error(SCMC_E_ARG, "empty path set");
    if ((rc = ensure_device(nullptr))) return rc;
    hipStream_t s;
    if (!hip_ok(hipStreamCreate(&s), "stream", &rc)) return rc;
    {
        ProblemState ps;
        TraceChunk tc;
        FieldEngine fe;
        double a = 0;
        if (!(rc = ps.setup(pr, s)) && !(rc = fe.init(ps, s)) && !(rc = tc.load(pset, s)) &&
            !(rc = fe.chunk(ps, tc, precision == SCMC_MIXED, s, &a)))
            rc = fe.finalize(pr, max_itr, u_re, u_im, s);
        if (!rc && stats) {
            memset(stats, 0, sizeof *stats);
            stats->paths = tc.np;
            stats->vertices = tc.nvert;
            stats->chunks = 1;
            stats->ms_accum = a;
            rc = fe.read_nonfinite(&stats->nonfinite, s);
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

} // extern "C"
