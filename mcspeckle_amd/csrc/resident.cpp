// resident.cpp -- host side of the resident update loop (serving).
//
// The loop kernel (herk_mixed_loop_kernel) stays on the GPU while there is
// work: the slot threads of the server publish their chunks' mixed updates as
// jobs in a table in pinned, coherent host memory, and the kernel's
// workgroups take the jobs' work items as they appear, so the updates of all
// requests in flight run in one resident launch.  Jobs are published only
// once their inputs (expanded rows, zeroed accumulators, bound) are complete,
// so the loop never waits on another kernel; it waits only for the host, and
// with a bound: a workgroup leaves after idle_ticks without work.
#include "scmc_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

namespace scmc {

namespace {
constexpr double kIdleMs = 2.0;          // a workgroup leaves after this long without work
constexpr double kWaitLimitS = 600.0;    // a job not finished after this long is an error
}

int ResidentLoop::init(int workgroups)
{
    int rc = SCMC_OK, dev = 0;
    if ((rc = ensure_device(&dev))) return rc;
    int khz = 0;
    if (!hip_ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count", &rc)) return rc;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
    // one workgroup per CU: the other half of every CU's LDS and registers
    // stays free for the kernels that prepare the jobs (tracer, factors,
    // expansion), which then run beside the loop at nearly the speed they
    // have alone.  Measured per request (mixed, CBS, one slot): N = 4096,
    // 1e5 paths: 0.76 s with 256 workgroups, 1.24 / 1.64 / 1.29 s with 384 /
    // 480 / 512 (the factors and expansion of the next chunk starve), 0.68 s
    // with launches; N = 1024: 0.111 s against 0.116 s with launches.
    grid = workgroups > 0 ? workgroups : std::max(8, cus);
    idle_ticks = (uint64_t)(kIdleMs * khz);
    if (!hip_ok(hipHostMalloc((void **)&H, sizeof(LoopCtl), hipHostMallocCoherent | hipHostMallocMapped),
                "loop control", &rc))
        return rc;
    memset(H, 0, sizeof(LoopCtl));
    if (!hip_ok(hipHostGetDevicePointer((void **)&Hd, H, 0), "loop control pointer", &rc)) return rc;
    if (!hip_ok(hipMalloc((void **)&D, sizeof(LoopDev)), "loop counters", &rc)) return rc;
    // The loop's stream gets a hardware queue of its own: HIP maps a
    // process's streams onto a few in-order hardware queues, and a request
    // stream sharing the queue of the resident kernel would wait behind it
    // until it leaves.  A stream with a CU mask (here every CU) is given its
    // own queue.
    std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0xFFFFFFFFu);
    if (cus % 32) mask.back() = (1u << (cus % 32)) - 1;
    if (!hip_ok(hipExtStreamCreateWithCUMask(&stream, (uint32_t)mask.size(), mask.data()), "loop stream", &rc))
        return rc;
    if (!hip_ok(hipMemsetAsync(D, 0, sizeof(LoopDev), stream), "loop counters", &rc) ||
        !hip_ok(hipStreamSynchronize(stream), "loop counters", &rc))
        return rc;
    return SCMC_OK;
}

ResidentLoop::~ResidentLoop()
{
    (void)stop();
    if (stream) (void)hipStreamDestroy(stream);
    if (D) (void)hipFree(D);
    if (H) (void)hipHostFree(H);
}

int ResidentLoop::launch_next()
{
    int rc = SCMC_OK;
    const int next = inst + 1;
    // instance next - 2 used this flag, and it has ended: instance next - 1
    // has started (its announcement is why we are here), or none runs
    __atomic_store_n(&H->closing[next & 1], 0u, __ATOMIC_SEQ_CST);
    launch_herk_mixed_loop(Hd, D, next, idle_ticks, grid, stream);
    if (!hip_ok(hipGetLastError(), "resident loop launch", &rc)) return rc;
    inst = next;
    instances += 1;
    return SCMC_OK;
}

int ResidentLoop::publish(const HerkProblem &hp, LoopTicket *t)
{
    int rc = SCMC_OK;
    LoopJob job;
    loop_job_from(hp, &job);
    t->idx = -1;
    if (job.nitems8 == 0) return SCMC_OK;   // nothing to add
    std::lock_guard<std::mutex> lk(mu);
    if (njobs == (unsigned)kLoopJobs && (rc = recycle())) return rc;
    const unsigned j = njobs;
    H->jobs[j] = job;
    __atomic_store_n(&H->done[j], 0u, __ATOMIC_RELAXED);
    njobs = j + 1;
    // publish (a full barrier on the host), then look for an instance that
    // has begun to leave: the kernel's announce-then-look mirrors this
    __atomic_store_n(&H->published, njobs, __ATOMIC_SEQ_CST);
    if (inst < 0 || __atomic_load_n(&H->closing[inst & 1], __ATOMIC_SEQ_CST))
        if ((rc = launch_next())) return rc;
    t->gen = gen;
    t->idx = (int)j;
    jobs += 1;
    return SCMC_OK;
}

int ResidentLoop::help(const LoopTicket &a, const LoopTicket &b, hipStream_t s)
{
    int rc = SCMC_OK;
    // the two tickets of one publish are consecutive jobs of one generation
    int lo = -1, hi = -1, items = 0;
    for (const LoopTicket *t : {&a, &b}) {
        if (t->idx < 0) continue;
        lo = lo < 0 ? t->idx : std::min(lo, t->idx);
        hi = std::max(hi, t->idx);
        items += H->jobs[t->idx].nitems8;
    }
    const int g = std::min(items, 2 * std::max(cus, 1));
    if (lo < 0 || g <= 0) return SCMC_OK;
    launch_herk_mixed_loop(Hd, D, -1, 0, g, s, (unsigned)lo, (unsigned)hi);
    return hip_ok(hipGetLastError(), "resident loop helper", &rc) ? SCMC_OK : rc;
}

int ResidentLoop::wait_index(unsigned j, uint64_t g)
{
    int rc = SCMC_OK;
    const auto t0 = std::chrono::steady_clock::now();
    for (int spin = 0;; ++spin) {
        // a newer generation started only after every job of this one finished
        if (__atomic_load_n(&gen, __ATOMIC_ACQUIRE) != g) return SCMC_OK;
        if (__atomic_load_n(&H->done[j], __ATOMIC_ACQUIRE)) return SCMC_OK;
        const hipError_t e = hipStreamQuery(stream);
        if (e != hipSuccess && e != hipErrorNotReady) {
            hip_ok(e, "resident loop", &rc);
            return rc;
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (s > kWaitLimitS) return set_error(SCMC_E_DEVICE, "resident loop: job %u not finished after %.0f s", j, s);
        std::this_thread::sleep_for(std::chrono::microseconds(spin < 100 ? 10 : 50));
    }
}

int ResidentLoop::wait(const LoopTicket &t)
{
    if (t.idx < 0) return SCMC_OK;
    return wait_index((unsigned)t.idx, t.gen);
}

// Every job of the generation finished, the instances stopped, the counters
// and the table cleared: the next job is job 0 of a new generation.
int ResidentLoop::recycle()
{
    int rc = SCMC_OK;
    for (unsigned j = 0; j < njobs; ++j)
        if ((rc = wait_index(j, gen))) return rc;
    __atomic_store_n(&H->stop, 1u, __ATOMIC_SEQ_CST);
    if (!hip_ok(stream_wait(stream), "resident loop stop", &rc)) return rc;
    // helper instances run on the requests' own streams: one still queued
    // there would start after the reset below and take an item of an old
    // job (its job index is reused by the next generation), so wait for the
    // whole device; all their jobs are done, so they end at once
    if (!hip_ok(hipDeviceSynchronize(), "resident loop helpers", &rc)) return rc;
    if (!hip_ok(hipMemsetAsync(D, 0, sizeof(LoopDev), stream), "loop counters", &rc) ||
        !hip_ok(stream_wait(stream), "loop counters", &rc))
        return rc;
    for (unsigned j = 0; j < njobs; ++j) H->done[j] = 0;
    H->closing[0] = H->closing[1] = 0;
    __atomic_store_n(&H->published, 0u, __ATOMIC_SEQ_CST);
    __atomic_store_n(&H->stop, 0u, __ATOMIC_SEQ_CST);
    njobs = 0;
    inst = -1;
    __atomic_store_n(&gen, gen + 1, __ATOMIC_RELEASE);
    return SCMC_OK;
}

int ResidentLoop::stop()
{
    int rc = SCMC_OK;
    if (!H) return SCMC_OK;
    std::lock_guard<std::mutex> lk(mu);
    if (inst < 0) return SCMC_OK;
    for (unsigned j = 0; j < njobs; ++j)
        if ((rc = wait_index(j, gen))) break;
    __atomic_store_n(&H->stop, 1u, __ATOMIC_SEQ_CST);
    int r2 = SCMC_OK;
    if (!hip_ok(stream_wait(stream), "resident loop stop", &r2) && !rc) rc = r2;
    __atomic_store_n(&H->stop, 0u, __ATOMIC_SEQ_CST);
    inst = -1;
    // the counters stay: a later publish continues this generation
    return rc;
}

} // namespace scmc
