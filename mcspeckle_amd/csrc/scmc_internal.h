// scmc_internal.h -- host-side state shared by the C-ABI entry points.
#pragma once
#include "hetero.h"
#include "kernels.h"
#include "scmc.h"

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace scmc {

int set_error(int code, const char *fmt, ...);
bool hip_ok(hipError_t e, const char *what, int *rc);
// Host waits for a stream or an event.  On a thread with tls_sleepy_waits
// set (a server slot thread) they poll with short sleeps instead of HIP's
// busy wait, so a server with many slots does not keep a host core spinning
// per slot; measured with 4 and 16 slots, latency and throughput are the
// same either way.
extern thread_local bool tls_sleepy_waits;
hipError_t stream_wait(hipStream_t s);
hipError_t event_wait(hipEvent_t e);
int ensure_device(int *dev_out);
int validate_problem(const scmc_problem *pr, int64_t max_itr, bool need_views);
double box_volume(const scmc_problem *pr);
int64_t chunk_paths(int n, int64_t remaining, double mean_len, bool mixed);
int64_t field_chunk_paths(int nv, int nl, int64_t remaining, double mean_len);

// Owning device allocation that only grows.
// Serving: device blocks a session frees are kept and handed out again
// instead of going back to the driver (hipFree waits for the whole device
// and can stall every other slot).  A block freed during a request may still
// be read by that request's queued kernels, so it becomes reusable only at
// settle(), which the server calls once the request has fully synchronised.
// DevBuf allocates from the pool that tls_block_pool names (the slot
// thread's session while it runs a request) and returns a block to the pool
// it came from.  Blocks a pool gives up (over its cap, or at the session's
// end) go to a process-wide cache of idle blocks, which any later session
// or call takes from before it asks the driver.
struct BlockPool {
    struct Blk {
        void *p;
        size_t bytes;
    };
    std::vector<Blk> ready, retired;
    size_t cap = (size_t)16 << 30;   // ready bytes kept past settle()
    size_t held = 0;                 // bytes taken from the driver through this pool and not given back
    void *take(size_t want, size_t *got);
    void put(void *p, size_t bytes) { retired.push_back({p, bytes}); }
    void settle();
    void trim(bool now = false);   // every block back to the driver (now: hipFree on this thread)
    ~BlockPool() { trim(); }
};
extern thread_local BlockPool *tls_block_pool;

// Allocation-free calls: while a thread's tls_arena is set, DevBuf::ensure
// carves its blocks from that arena -- device memory that outlives the call
// and that the library does not allocate per call -- instead of asking the
// driver, and release() gives nothing back.  A block is carved at its exact
// size (256-byte aligned); running out is an error, never an allocation.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};
extern thread_local Arena *tls_arena;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    BlockPool *pool = nullptr;   // where p goes back to (null: the driver)
    bool borrowed = false;       // p lies in an arena (nothing to free)
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    void release();
    bool ensure(size_t b, int *rc);
};

// While started, DevBuf frees are handed to a background thread (hipFree's
// device-wide wait moves off the calling thread); reaper_stop frees what is
// queued and makes frees synchronous again.  The server brackets its slot
// threads with them.
void reaper_start();
void reaper_stop();

// Device copy of the problem description (directions, tables, constants).
struct ProblemState {
    DevProblem dp;
    DevBuf dl, dv, tab_amp, tab_icdf, tab_amp0, tab_icdf0;
    // voxel medium (MCCovHetro)
    bool het = false;
    HetMedium hm{};
    DevBuf dsig, dalb, dcell_icdf;
    int setup(const scmc_problem *pr, hipStream_t s);
    int setup_medium(const scmc_problem *pr, const scmc_medium *m, hipStream_t s);
    int upload_phase(const scmc_phase &ph, int type, DevBuf &amp, DevBuf &icdf, DevPhase &out, hipStream_t s);
};

// One traced chunk of paths: offsets, path records and vertex records.
struct TraceChunk {
    int64_t np = 0, nvert = 0;
    DevBuf counts, offsets, paths, verts, vpath, tsum;   // tsum: scan scratch
    DevBuf ctr;                                         // the walks' path counter
    int run(const DevProblem &P, uint64_t p0, int64_t n, hipStream_t s, const HetMedium *M = nullptr);
    int load(const scmc_pathset *ps, hipStream_t s);
};

// Serving: the resident update loop (herk_mixed_loop_kernel) and its host
// side.  publish() hands one mixed update to the running loop (launching an
// instance when none runs, or queueing the next one behind an instance that
// has begun to leave) and returns a ticket; wait() returns once that job's
// items are all in its accumulators.  Jobs are numbered within a generation
// of kLoopJobs; when the table is full, publish() waits for every job,
// stops the instance, clears the counters and starts a new generation (a
// ticket of an older generation is done).  No call waits on the device
// without a bound: the loop leaves after idle_ms without work.
struct LoopTicket {
    uint64_t gen = 0;
    int idx = -1;   // -1: nothing published
};
struct ResidentLoop {
    LoopCtl *H = nullptr;        // host view (pinned, coherent)
    LoopCtl *Hd = nullptr;       // the same memory as the device sees it
    LoopDev *D = nullptr;
    hipStream_t stream = nullptr;
    int grid = 0;
    uint64_t idle_ticks = 0;
    std::mutex mu;
    unsigned njobs = 0;
    uint64_t gen = 0;
    int inst = -1;               // the newest instance queued (-1: none this generation)
    int64_t instances = 0, jobs = 0;   // statistics
    int init(int workgroups);    // workgroups <= 0: one per CU
    int publish(const HerkProblem &hp, LoopTicket *t);
    // A helper instance on s (the publishing request's stream) that takes
    // items of the published jobs a and b only (an empty ticket is skipped)
    // until none is left: one workgroup per item, at most two per CU.
    int help(const LoopTicket &a, const LoopTicket &b, hipStream_t s);
    int cus = 0;
    bool helpers = true;         // publish callers add helper instances
    int wait(const LoopTicket &t);
    int stop();
    ~ResidentLoop();

  private:
    int launch_next();           // with mu held
    int recycle();               // with mu held
    int wait_index(unsigned j, uint64_t g);
};

// Factor buffers of one in-flight chunk (two alternate, so the factors of
// the next chunk are built while the current one updates the accumulators).
struct CovSlot {
    DevBuf A1, B1, A2, B2, SA, SB, Gp, Fp2, maxG, maxF2, U, ones, EV0, UL0, bound, zbuf_m, zbuf_s;
    hipEvent_t ev[7] = {};   // factors start / NEE done / expansion done; update start / update done /
                             // bound copied; update inputs ready (the accumulators and tile order too)
    HerkProblem hp[2];       // multiple and single scattering
    bool used = false;       // events recorded and not yet harvested
    int64_t np = 0, k2 = 0;
    unsigned *bound_host = nullptr;   // pinned copy of bound[0], written before ev[5]
    LoopTicket tk[2];                 // resident loop: the chunk's two jobs
    bool looped = false;              // the update went to the resident loop
    double loop_t0 = 0;               // resident loop: host time of publishing (ms)
};

// Covariance accumulation state across chunks.
// Serving: the mixed-precision updates of the requests running at the same
// time go out as shared launches (launch_herk_mixed_group).  A slot thread
// submits its chunk's two problems; whichever thread finds no launch in
// progress and fewer than max_inflight groups still running on the
// batcher's stream launches everything pending as one group, after making
// the stream wait on each job's factors.  Each job's t1 event is recorded
// after its group, and the submitting session's stream waits on it.  No
// GPU-side waits are involved: a thread that cannot launch re-checks every
// 100 us.
// Largest N = Nv*Nl whose updates go through the shared launches (36 tiles
// of 128 or fewer: a single such problem leaves most of the GPU idle).
constexpr int kSharedUpdateMaxN = 1024;
struct UpdateBatcher {
    struct Job {
        const HerkProblem *hm = nullptr, *hs = nullptr;
        hipEvent_t ready = nullptr, t0 = nullptr, t1 = nullptr;   // t1: after the group
        bool launched = false;
        int rc = SCMC_OK;   // the group's launch status
        std::string err;
    };
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Job *> pending;
    std::vector<hipEvent_t> inflight;   // end events of launched groups
    std::vector<hipEvent_t> spare;
    bool launching = false;
    int max_inflight = 2;
    int64_t groups = 0, jobs = 0;       // launches and problems pairs (statistics)
    int init();
    ~UpdateBatcher();
    int run(Job &j);   // returns once j's group is launched
};

struct CovEngine {
    UpdateBatcher *batcher = nullptr;   // serving: shared update launches
    ResidentLoop *loop = nullptr;       // serving: updates go to the resident loop
    bool mixed = false;
    int nv = 0, nl = 0, n = 0, tile = 0;
    int ks_m = 0, ks_s = 0;
    int last_scale_log2 = 0;
    double ms_herk = 0, herk_flops = 0, ms_expand = 0, ms_accum = 0, ms_nee = 0;
    int64_t herk_launches = 0;
    DevBuf acc_m, acc_s, order;
    int order_n = 0, order_tile = 0;   // N and tile edge the tile order was built for
    std::vector<int> order_host;   // its host copy (source of the async upload)
    DevBuf mean, mean_part;   // voxel medium: mean field (MCCovHetro's mean1), per-range partials
    DevBuf nonfinite;         // vertices with NaN/Inf factors in this solve (device counter)
    DevBuf fin_out, fin_af;   // finalize: output staging and amplitude table (kept: hipFree
                              // synchronises the device, which would stall other sessions)
    CovSlot slot[2];
    int cur = 0;
    hipEvent_t nee_done = nullptr;   // the last produced chunk has read its path records
    ~CovEngine();
    int init(const ProblemState &ps, bool mixed, hipStream_t s);
    int ensure_acc(DevBuf &acc, int &ks_have, int ks, hipStream_t s);
    int produce(const ProblemState &ps, TraceChunk &tc, hipStream_t sp);
    int consume(hipStream_t sc);
    int harvest(CovSlot &S);
    int drain();
    int wait_loop(CovSlot &S);          // the slot's resident-loop jobs are done
    int chunk(const ProblemState &ps, TraceChunk &tc, hipStream_t s);
    int finalize(const ProblemState &ps, const scmc_problem *pr, int64_t max_itr, double *ms_re, double *ms_im,
                 double *mm_re, double *mm_im, hipStream_t s);
    int finalize_mean(int64_t max_itr, double *re, double *im, hipStream_t s);
    int read_nonfinite(int64_t *out, hipStream_t s);
};

// Field accumulation state across chunks.
struct FieldEngine {
    int nv = 0, nl = 0, ncol = 0;
    bool wave = false;   // lights combined with weights (MCfieldOnWave)
    bool het = false;    // voxel medium (MCsampleHetro)
    DevBuf acc, work, out, wl;
    DevBuf nonfinite;    // paths with NaN/Inf coefficients in this solve (device counter)
    int init(const ProblemState &ps, hipStream_t s, const double *wl_re = nullptr, const double *wl_im = nullptr);
    int chunk(const ProblemState &ps, TraceChunk &tc, bool mixed, hipStream_t s, double *ms_accum);
    int finalize(const scmc_problem *pr, int64_t max_itr, double *u_re, double *u_im, hipStream_t s);
    int read_nonfinite(int64_t *out, hipStream_t s);
};

// Deep copy of a problem description (the session owns its arrays).
struct ProblemCopy {
    scmc_problem pr;
    std::vector<double> l, v, a_re, a_im, a_ic, b_re, b_im, b_ic;
    void assign(const scmc_problem *src);
};

// A solve in progress: device accumulators plus the traced-chunk buffers.
struct Session {
    BlockPool pool;   // serving: reused device blocks (first member: destroyed last)
    int kind = 0;   // 0 MCcov, 1 MCfield
    bool mixed = false;
    ProblemCopy pc;
    hipStream_t stream = nullptr, stream2 = nullptr;   // stream2: covariance tracer
    ProblemState ps;
    TraceChunk tc;
    CovEngine ce;
    FieldEngine fe;
    scmc_stats st{};
    double mean_len = 20;
    // kind 0 MCcov, 1 MCfield; wl_re (kind 1 only) makes it MCfieldOnWave
    int init(const scmc_problem *pr, int kind, int precision, const double *wl_re = nullptr,
             const double *wl_im = nullptr, const scmc_medium *med = nullptr);
    // settle = false leaves the last chunks in flight (no device sync, the
    // statistics lag behind); a later run or result() continues the stream
    int run(int64_t first, int64_t n, bool settle = true);
    int result(int64_t max_itr, double *a_re, double *a_im, double *b_re, double *b_im);
    void fill_stats(scmc_stats *out) const;
    ~Session();
};

} // namespace scmc
