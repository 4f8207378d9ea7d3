// kernels.h -- host-side launch interfaces of the speckle Monte Carlo kernels.
#pragma once
#include "device_common.h"

#include <vector>

namespace scmc {

// One scattering vertex (64 B): position, incoming direction, albedo weight
// W_n and the random-phase uniform xi_n of MCfield.m:220.
struct VertexRec {
    double x[3];
    double ow[3];
    double weight;
    double xi;
};

// Per-path quantities of the first vertex (MCcov.m:114-153).
struct PathRec {
    double x1[3];
    double px;
    double w1[3];
    double pad;
    double2 w0p;
};

// ---------------------------------------------------------------- tracer
// Workgroups of 256 lanes for a kernel whose lanes stay resident and take
// paths from a counter (take_path): as many as fit on the device at once
// (occupancy query), but no more than np paths need.
inline unsigned resident_blocks(const void *kernel, int64_t np)
{
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0);
    const int64_t b = (int64_t)(cus > 1 ? cus : 1) * (per > 1 ? per : 1);
    const int64_t need = (np + 255) / 256;
    return (unsigned)(b < need ? b : need);
}

// Vertex counts of paths [p0, p0 + np) (ids relative to P.path_offset).
// ctr: one device counter the launch clears and the walk hands paths out from.
void launch_trace_count(const DevProblem &P, uint64_t p0, int64_t np, int32_t *counts, unsigned long long *ctr,
                        hipStream_t s);
// offsets[np + 1] = exclusive scan of counts (offsets[np] = total).
// Exclusive scan of counts[np] into offsets[np + 1]; tsum: scratch of
// scan_tiles(np) int64 (at least one).
int64_t scan_tiles(int64_t np);
void launch_scan(const int32_t *counts, int64_t np, int64_t *offsets, int64_t *tsum, hipStream_t s);
void launch_trace_emit(const DevProblem &P, uint64_t p0, int64_t np, const int64_t *offsets, PathRec *paths,
                       VertexRec *verts, int32_t *vpath, unsigned long long *ctr, hipStream_t s);

// -------------------------------------------------- covariance factors
// Factor arrays for the rank-K update C += sum_k u_k u_k^H with
//   u_k(v,l) = A1[k][v] B1[k][l] + A2[k][v] B2[k][l]   (A2/B2 null without CBS)
// Single scattering uses A1 = sqrt(albedo/px) e_v0, B1 = e_l0 (the
// amplitude product a(l.v) is a per-row factor applied when finalising) in
// the far field, and whole rows U with B = 1 in the near field.
struct CovFactors {
    void *A1, *B1, *A2, *B2;   // float2* (mixed) or double2* (fp64)
    void *SA, *SB;             // single-scatter factors per path
    void *Gp, *Fp2;            // path-level e_l0_ms, e_v0_ms
    void *U, *ones;            // near field: single-scatter rows [np][nv*nl], ones[np]
    double2 *EV0 = nullptr;    // voxel medium: W/px e_v0 per path (mean field)
    double2 *UL0 = nullptr;    // voxel medium: unattenuated u_l0 per path
    float *maxG, *maxF2;       // path-level maxima (mixed scaling)
    unsigned *bound;           // [0] multi, [1] single: float bits of max |u|
    unsigned long long *nonfinite = nullptr;   // += vertices whose factors are NaN/Inf
};
struct HetMedium;   // hetero.h
// M is null for a homogeneous medium.
void launch_nee_cov(const DevProblem &P, const HetMedium *M, int64_t np, int64_t nvert, const PathRec *paths,
                    const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, const CovFactors &f,
                    bool mixed, hipStream_t s);

// ------------------------------------------------------ voxel media
// Walks of MCCovHetro.m: count pass (emit = false, fills counts) or emit pass;
// ctr: one device counter the launch clears and the walk hands paths out from.
void launch_hetero_trace(const DevProblem &P, const HetMedium &M, uint64_t p0, int64_t np, int32_t *counts,
                         int64_t *offsets, PathRec *paths, VertexRec *verts, int32_t *vpath, bool emit,
                         unsigned long long *ctr, hipStream_t s);
// mean1 += sum_p af_ang_vl EV0[p] UL0[p]^T (MCCovHetro.m:210-211); part holds
// kHetMeanSplits * nv * nl double2.
constexpr int kHetMeanSplits = 64;
void launch_hetero_mean(const DevProblem &P, int64_t np, const PathRec *paths, const double2 *EV0,
                        const double2 *UL0, double2 *mean, double2 *part, hipStream_t s);

// ------------------------------------------------------------ HERK
struct HerkProblem {
    int nv, nl, n;
    int64_t k;             // number of rank-1 terms
    const void *A1, *B1, *A2, *B2;
    double2 *acc;          // [ksplit][n][n] row-major FP64 accumulators
    const unsigned *bound; // mixed: float bits of max |u| (may be null)
    int ksplit;
    void *zbuf = nullptr;          // expanded rows: FP16 hi/lo (mixed_zbuf_bytes) or FP64 (fp64_zbuf_bytes)
    const void *order = nullptr;   // int2 tile order (blocked_tile_order) over the update's tile edge
};
// (ti, tj) pairs of the lower-triangle tiles in the blocked order used by
// the mixed update, flattened.
std::vector<int> blocked_tile_order(int nt);
// Bytes of the expanded-row buffer of the mixed update for k terms.
size_t mixed_zbuf_bytes(int64_t k, int n);
// Tile edge of the mixed and FP64 kernels (the finaliser needs it).
constexpr int kTileMixed = 128;
constexpr int kTileF64 = 64;
// Mixed precision in two steps: FP16 hi/lo row expansion into hp.zbuf, then
// the matrix-core update from hp.zbuf into hp.acc.  The update launch takes
// a second problem of the same N and tile order (b, may be null), whose work
// items follow a's in the same grid.
void launch_expand_mixed(const HerkProblem &hp, hipStream_t s);
void launch_herk_mixed_update(const HerkProblem &a, const HerkProblem *b, hipStream_t s);
// Up to kMaxUpdateGroup problems (each with its own N, tile order and
// accumulators; null or empty entries skipped) per launch, more in several.
constexpr int kMaxUpdateGroup = 8;
void launch_herk_mixed_group(const HerkProblem *const *hp, int count, hipStream_t s);
// FP64 in two steps as well: row expansion into hp.zbuf, then the update,
// which takes a second problem of the same N and tile order (b, may be null)
// whose work items follow a's in the same grid.
size_t fp64_zbuf_bytes(int64_t k, int n);
void launch_expand_fp64(const HerkProblem &hp, hipStream_t s);
void launch_herk_fp64(const HerkProblem &a, const HerkProblem *b, hipStream_t s);

// Resident update loop (serving).  One job is one mixed update (one problem
// of a chunk, as launch_herk_mixed_update would run it); the host writes its
// descriptor into LoopCtl (pinned, coherent host memory), then raises
// `published`; the loop kernel sets done[j] once every item of job j has been
// added into its accumulators.  LoopDev holds the kernel's own counters in
// device memory (zeroed before the first instance).
constexpr int kLoopJobs = 4096;
struct LoopJob {
    int64_t nks;              // K-steps of the expanded rows
    const void *Z;            // expanded rows
    double2 *acc;             // accumulator copies
    const unsigned *bound;    // scale bound (may be null)
    const void *order;        // tile order
    int32_t n, ntile, ksplit, nitems;
    int32_t nitems8;          // nitems rounded up to a multiple of 8
    int32_t pad;
};
struct LoopCtl {
    unsigned published;       // jobs published (host)
    unsigned stop;            // host: idle workgroups leave
    unsigned closing[2];      // kernel: a workgroup of instance i left (closing[i & 1])
    unsigned pad[12];
    unsigned done[kLoopJobs];
    LoopJob jobs[kLoopJobs];
};
struct LoopDev {
    unsigned active[8];                  // per XCD label: first job not yet exhausted
    unsigned next[kLoopJobs * 8];        // per job and label: items taken / 8
    unsigned finished[kLoopJobs];        // per job: items finished
};
void loop_job_from(const HerkProblem &hp, LoopJob *out);
// One instance of the loop kernel (grid workgroups of 256 threads) on s; it
// returns when idle for idle_ticks of the device wall clock, or on stop.  A
// helper instance (inst < 0) takes items of the published jobs
// help_lo..help_hi only and returns when they are all taken.
void launch_herk_mixed_loop(LoopCtl *ctl_dev, LoopDev *dev, int inst, uint64_t idle_ticks, int grid, hipStream_t s,
                            unsigned help_lo = 0, unsigned help_hi = 0);

// Sum the ksplit accumulators, restore the upper triangle, scale, apply the
// optional per-row factor rowf[(v,l)] (single scattering amplitude), and write
// the MATLAB nv x nv x nl x nl column-major layout (split or interleaved).
void launch_cov_finalize(const double2 *acc, int ksplit, int nv, int nl, double scale,
                         const double2 *rowf, double *out_re, double *out_im, hipStream_t s);
// Single-scatter amplitude table af_ang_vl(v,l) = a(l_l . v_v) (far field).
void launch_af_table(const DevProblem &P, double2 *af, hipStream_t s);

// --------------------------------------------------------------- field
// u(v,l) += sum over vertices of the MCfield.m:207-222 contributions,
// FP64 accumulators acc[nv*nl] (column-major, v fastest).  With wl (device,
// nl weights) the lights form one wave (MCfieldOnWave.m) and acc has nv
// entries.  With M (a voxel medium, MCsampleHetro.m) E is the ray march of
// evalphaseattHetro (FP64 march; the mixed mode evaluates sin/cos/exp in
// FP32 in the far field).  work holds
// field_work_elems(...) double2.
// nonfinite (may be null) += paths whose coefficients are NaN/Inf.
void launch_field(const DevProblem &P, const HetMedium *M, int64_t np, int64_t nvert, const PathRec *paths,
                  const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, double2 *work, double2 *acc,
                  bool mixed, const double2 *wl, unsigned long long *nonfinite, hipStream_t s);
size_t field_work_elems(int64_t np, int64_t nvert, int nv, int nl, bool wave);

// Device-wide helpers.
void launch_zero(void *p, size_t bytes, hipStream_t s);
// In place: t[i] = cos(t[i]) (the tabulated sampleIcdf angles -> cosines).
void launch_cos_table(double *t, int n, hipStream_t s);

} // namespace scmc
