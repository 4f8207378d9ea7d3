// noalloc.hip -- the MCfield solve without device allocation.
//
// scmc_mcfield_noalloc computes exactly what scmc_mcfield computes (same
// Philox walks, same kernels), but every device buffer of the call -- the
// problem copy, the traced chunk's path and vertex records, the field work
// area and the accumulators -- lies in one file-scope device array,
// `scratch`, of kScratchBytes.  The call allocates and frees no device
// memory and creates no HIP object after the first call on a device: that
// first call looks up the array's device address (hipGetSymbolAddress) and
// creates the device's stream; both are kept until the process ends.
//
// Chunks are sized so that the records of a chunk fit beside the fixed part
// (problem copy, accumulators, per-view work); when a chunk's paths turn out
// to hold more vertices than its share, only the longest prefix of paths that
// fits is emitted and the rest is traced again in the next chunk, so a call
// never runs out of scratch part-way for any path count.
#include "scmc_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace scmc {

constexpr size_t kScratchBytes = (size_t)128 << 20;   // per device, held until the process ends
constexpr int kMaxDevices = 64;
constexpr int64_t kVertsPerPathShare = 48;   // vertex records reserved per path of a chunk
constexpr int64_t kMinChunkPaths = 64;

// The array's name is a host-side stand-in: host code uses the address that
// hipGetSymbolAddress returns for the current device, never `scratch` itself.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoAllocDevice {
    std::once_flag once;
    int rc = SCMC_OK;
    std::string err;
    char *base = nullptr;          // device address of `scratch` on this device
    hipStream_t stream = nullptr;
    std::mutex call_mu;            // calls in one process run one after another
};
static NoAllocDevice g_noalloc[kMaxDevices];

static int noalloc_setup(int dev, NoAllocDevice **out)
{
    if (dev < 0 || dev >= kMaxDevices) return set_error(SCMC_E_DEVICE, "device %d beyond %d", dev, kMaxDevices);
    NoAllocDevice &D = g_noalloc[dev];
    std::call_once(D.once, [&D] {
        // the device is current (ensure_device)
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            D.rc = SCMC_E_DEVICE;
            D.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return;
        }
        D.base = (char *)p;
        e = hipStreamCreate(&D.stream);
        if (e != hipSuccess) {
            D.rc = SCMC_E_DEVICE;
            D.err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        }
    });
    if (D.rc) return set_error(D.rc, "%s", D.err.c_str());
    *out = &D;
    return SCMC_OK;
}

static size_t al(size_t b) { return ((b ? b : 16) + 255) & ~(size_t)255; }

// Scratch the call holds whatever its chunk size: the problem copy
// (ProblemState::setup), the accumulator and the non-finite counter
// (FieldEngine::init).
static size_t fixed_bytes(const scmc_problem *pr)
{
    const size_t dim = (size_t)pr->dim;
    size_t b = al(8 * dim * (size_t)pr->nl) + al(8 * dim * (size_t)std::max(pr->nv, 1));
    auto tab = [](const scmc_phase &ph, int type) {
        return type == SCMC_TABULATED ? al(16 * (size_t)ph.n_amp) + al(8 * (size_t)ph.n_icdf) : (size_t)0;
    };
    b += tab(pr->amp, pr->amp.type);
    if (pr->has_amp0) b += tab(pr->amp0, pr->amp.type);
    b += al(16 * (size_t)pr->nv * (size_t)pr->nl) + al(sizeof(unsigned long long));
    b += al(sizeof(unsigned long long));   // the walks' path counter
    return b;
}

// Scratch of one chunk of np paths and nvert vertices: TraceChunk's records
// and scan tiles, and the field work area (FieldEngine::chunk).
static size_t chunk_bytes(int64_t np, int64_t nvert, int nv, int nl)
{
    return al(4 * (size_t)np) + al(8 * (size_t)(np + 1)) + al(sizeof(PathRec) * (size_t)np) +
           al(8 * (size_t)std::max<int64_t>(scan_tiles(np), 1)) + al(sizeof(VertexRec) * (size_t)std::max<int64_t>(nvert, 1)) +
           al(4 * (size_t)std::max<int64_t>(nvert, 1)) + al(16 * field_work_elems(np, nvert, nv, nl, false));
}

// Largest chunk (paths) that fits beside `fixed`, at kVertsPerPathShare
// vertex records per path; 0 if not even kMinChunkPaths fit.
static int64_t chunk_paths_that_fit(size_t fixed, int nv, int nl)
{
    auto fits = [&](int64_t np) { return fixed + chunk_bytes(np, np * kVertsPerPathShare, nv, nl) <= kScratchBytes; };
    if (!fits(kMinChunkPaths)) return 0;
    int64_t lo = kMinChunkPaths, hi = kMinChunkPaths;
    while (hi < (1 << 20) && fits(hi * 2)) hi *= 2;
    if (hi < (1 << 20)) {
        int64_t a = hi, b = hi * 2;   // fits(a), !fits(b)
        while (b - a > 1) {
            const int64_t m = (a + b) / 2;
            (fits(m) ? a : b) = m;
        }
        hi = a;
    }
    return std::max(lo, hi);
}

// Sets the calling thread's arena for the lifetime of the guard.
struct ArenaGuard {
    Arena a;
    explicit ArenaGuard(char *base) { a.base = base, a.cap = kScratchBytes, a.used = 0, tls_arena = &a; }
    ~ArenaGuard() { tls_arena = nullptr; }
};

static int field_noalloc(const scmc_problem *pr, int64_t max_itr, int32_t precision, double *u_re, double *u_im,
                         scmc_stats *stats)
{
    const auto t_call = std::chrono::steady_clock::now();
    int rc = validate_problem(pr, max_itr, true);
    if (rc) return rc;
    if (precision != SCMC_FP64 && precision != SCMC_MIXED) return set_error(SCMC_E_ARG, "invalid precision %d", precision);
    if (!u_re) return set_error(SCMC_E_ARG, "null output");
    // the size limit is known before any device work
    const size_t fixed = fixed_bytes(pr);
    const int64_t np_max = chunk_paths_that_fit(fixed, pr->nv, pr->nl);
    if (np_max == 0)
        return set_error(SCMC_E_ARG,
                         "MCfield without allocation: nv x nl = %d x %d needs more than its %zu-byte scratch "
                         "(%zu bytes before any path); use scmc_mcfield",
                         pr->nv, pr->nl, kScratchBytes, fixed);
    const int64_t nvert_max = np_max * kVertsPerPathShare;
    int dev = 0;
    if ((rc = ensure_device(&dev))) return rc;
    NoAllocDevice *D = nullptr;
    if ((rc = noalloc_setup(dev, &D))) return rc;
    std::lock_guard<std::mutex> lock(D->call_mu);
    hipStream_t s = D->stream;
    ArenaGuard arena(D->base);
    const bool mixed = precision == SCMC_MIXED;

    ProblemState ps;
    FieldEngine fe;
    TraceChunk tc;
    // problem copy and accumulators (FieldEngine::init clears the accumulator
    // and the counter: the scratch holds the previous call's values)
    if ((rc = ps.setup(pr, s)) || (rc = fe.init(ps, s))) return rc;
    // the chunk buffers at their largest, so later chunks reuse them
    if (!tc.ctr.ensure(sizeof(unsigned long long), &rc) || !tc.counts.ensure(4 * (size_t)np_max, &rc) ||
        !tc.offsets.ensure(8 * (size_t)(np_max + 1), &rc) ||
        !tc.paths.ensure(sizeof(PathRec) * (size_t)np_max, &rc) ||
        !tc.tsum.ensure(8 * (size_t)std::max<int64_t>(scan_tiles(np_max), 1), &rc) ||
        !tc.verts.ensure(sizeof(VertexRec) * (size_t)nvert_max, &rc) ||
        !tc.vpath.ensure(4 * (size_t)nvert_max, &rc) ||
        !fe.work.ensure(16 * field_work_elems(np_max, nvert_max, pr->nv, pr->nl, false), &rc))
        return rc;

    scmc_stats st{};
    std::vector<int64_t> off;
    for (int64_t done = 0; done < max_itr;) {
        const auto t0 = std::chrono::steady_clock::now();
        const int64_t np = std::min(np_max, max_itr - done);
        launch_trace_count(ps.dp, (uint64_t)done, np, (int32_t *)tc.counts.p, (unsigned long long *)tc.ctr.p, s);
        launch_scan((const int32_t *)tc.counts.p, np, (int64_t *)tc.offsets.p, (int64_t *)tc.tsum.p, s);
        int64_t total = 0;
        if (!hip_ok(hipMemcpyAsync(&total, (int64_t *)tc.offsets.p + np, sizeof total, hipMemcpyDeviceToHost, s),
                    "copy total", &rc) ||
            !hip_ok(hipStreamSynchronize(s), "trace sync", &rc))
            return rc;
        int64_t use = np;
        if (total > nvert_max) {
            // the longest prefix of paths whose records fit; the rest go into
            // the next chunk (the offsets of a prefix are its own scan)
            off.resize((size_t)np + 1);
            if (!hip_ok(hipMemcpyAsync(off.data(), tc.offsets.p, sizeof(int64_t) * (size_t)(np + 1),
                                       hipMemcpyDeviceToHost, s),
                        "copy offsets", &rc) ||
                !hip_ok(hipStreamSynchronize(s), "offsets sync", &rc))
                return rc;
            use = (int64_t)(std::upper_bound(off.begin(), off.end(), nvert_max) - off.begin()) - 1;
            if (use < 1)
                return set_error(SCMC_E_NOMEM, "MCfield without allocation: path %lld has %lld vertices, beyond %lld",
                                 (long long)done, (long long)off[1], (long long)nvert_max);
            total = off[(size_t)use];
        }
        launch_trace_emit(ps.dp, (uint64_t)done, use, (const int64_t *)tc.offsets.p, (PathRec *)tc.paths.p,
                          (VertexRec *)tc.verts.p, (int32_t *)tc.vpath.p, (unsigned long long *)tc.ctr.p, s);
        if (!hip_ok(hipGetLastError(), "trace launch", &rc)) return rc;
        tc.np = use;
        tc.nvert = total;
        st.ms_trace += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = fe.chunk(ps, tc, mixed, s, &st.ms_accum))) return rc;
        st.paths += use;
        st.vertices += total;
        st.chunks += 1;
        done += use;
    }
    if ((rc = fe.read_nonfinite(&st.nonfinite, s))) return rc;
    if ((rc = fe.finalize(pr, max_itr, u_re, u_im, s))) return rc;
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call).count();
    if (stats) *stats = st;
    return SCMC_OK;
}

} // namespace scmc

extern "C" {

int scmc_mcfield_noalloc(const scmc_problem *pr, int64_t max_itr, int32_t precision, double *u_re, double *u_im,
                         scmc_stats *stats)
{
    try {
        return scmc::field_noalloc(pr, max_itr, precision, u_re, u_im, stats);
    } catch (const std::exception &e) {
        return scmc::set_error(SCMC_E_DEVICE, "scmc_mcfield_noalloc: %s", e.what());
    }
}

} // extern "C"
