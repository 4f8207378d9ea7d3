// tracer.hip -- counter-based random-walk tracer (FP64), one lane per path.
//
// Restates the path sampling of code/MCcov.m:112-289 (identical in
// code/MCfield.m:112-248 and the vertex store of MCfieldFourierQuad.m:154-213):
// first vertex by expSmpX / uniform, first direction isotropic or from
// ampfunc0, then free flight, escape test, albedo roulette and a new direction
// from the phase function.  Every random number is Philox4x32-10 keyed by the
// seed with counter (vertex, group, path), so a path depends only on its id:
// the count and emit passes below retrace identical walks, and the vertex
// order in HBM (path-major) is independent of scheduling.
#include "kernels.h"

namespace scmc {

// Both passes keep their lanes resident: a lane whose walk has ended takes the
// next path index (take_path) at the wave's next refill (refill_due: once a
// quarter of the wave is idle), so a wave does not idle while its longest walk
// finishes (for the slab workloads the longest of 64 walks is ~6x the mean).
// Outputs are indexed by path, so they do not depend on which lane walked
// which path.
template <int DIM>
__global__ __launch_bounds__(256) void trace_count_kernel(DevProblem P, uint64_t p0, int64_t np, int32_t *counts,
                                                          unsigned long long *ctr)
{
    __shared__ double lcos[kIcdfLds];
    const double *tab = stage_icdf(P.amp, lcos);
    WavePaths wp;
    bool active = false;
    int64_t id = 0;
    uint64_t path = 0;
    double x[3] = {0, 0, 0}, w[3] = {0, 0, 0}, ow[3] = {0, 0, 0}, px = 1, weight = 0;
    uint32_t n = 0;
    for (;;) {
        if (refill_due(wp, active)) {
            const int64_t t = take_path(wp, !active, np, ctr);
            if (t >= 0) {
                id = t;
                path = P.path_offset + p0 + (uint64_t)t;
                double2 w0p;
                path_start<DIM>(P, path, x, px, w, w0p);
                ow[0] = ow[1] = ow[2] = 0;   // a new path has no incoming direction
                weight = P.albedo;
                n = 0;
                active = true;
            }
        }
        if (!__ballot(active)) break;
        if (active) {
            ++n;
            const double2 a = uniform_pair(P.seed, path, n, 0);
            if (!path_step<DIM>(P, path, n, x, w, ow, weight, a.y, tab)) {
                counts[id] = (int32_t)n;
                active = false;
            }
        }
    }
}

template <int DIM>
__global__ __launch_bounds__(256) void trace_emit_kernel(DevProblem P, uint64_t p0, int64_t np, const int64_t *offsets,
                                                         PathRec *paths, VertexRec *verts, int32_t *vpath,
                                                         unsigned long long *ctr)
{
    __shared__ double lcos[kIcdfLds];
    const double *tab = stage_icdf(P.amp, lcos);
    WavePaths wp;
    bool active = false;
    int64_t id = 0, o = 0, end = 0;
    uint64_t path = 0;
    double x[3] = {0, 0, 0}, w[3] = {0, 0, 0}, ow[3] = {0, 0, 0}, px = 1, weight = 0;
    uint32_t n = 0;
    for (;;) {
        if (refill_due(wp, active)) {
            const int64_t t = take_path(wp, !active, np, ctr);
            if (t >= 0) {
                id = t;
                path = P.path_offset + p0 + (uint64_t)t;
                double2 w0p;
                path_start<DIM>(P, path, x, px, w, w0p);
                ow[0] = ow[1] = ow[2] = 0;   // a new path has no incoming direction
                PathRec pr;
                for (int j = 0; j < 3; ++j) {
                    pr.x1[j] = x[j];
                    pr.w1[j] = w[j];
                }
                pr.px = px;
                pr.pad = 0;
                pr.w0p = w0p;
                paths[t] = pr;
                o = offsets[t];
                end = offsets[t + 1];
                weight = P.albedo;
                n = 0;
                active = true;
            }
        }
        if (!__ballot(active)) break;
        if (active) {
            ++n;
            const double2 a = uniform_pair(P.seed, path, n, 0);
            if (o < end) {
                VertexRec r;
                for (int j = 0; j < 3; ++j) {
                    r.x[j] = x[j];
                    r.ow[j] = ow[j];
                }
                r.weight = weight;
                r.xi = a.x;
                verts[o] = r;
                vpath[o] = (int32_t)id;
                ++o;
            }
            if (!path_step<DIM>(P, path, n, x, w, ow, weight, a.y, tab)) active = false;
        }
    }
}

// Exclusive scan of the int32 vertex counts into int64 offsets (offsets[np]
// = total), in three launches: per-tile sums, a scan of the tile sums, and a
// rescan of each tile from its base.  A tile is 256 threads x 8 consecutive
// counts, so every load and store is contiguous across the lanes.
constexpr int kScanThreads = 256, kScanPer = 8, kScanTile = kScanThreads * kScanPer;

// Exclusive scan of one value per thread across the workgroup; returns the
// thread's prefix and sets *total to the workgroup sum.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t *sh, int64_t *total)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int64_t u = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int64_t incl = sh[t];
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sum_kernel(const int32_t *counts, int64_t np, int64_t *tsum)
{
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int64_t v = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j)
        if (base + j < np) v += counts[base + j];
    int64_t total;
    (void)block_exclusive_scan(v, sh, &total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// In place: tsum[b] becomes the sum of the tiles before b; offsets[np] the total.
__global__ __launch_bounds__(kScanThreads) void scan_tile_base_kernel(int64_t *tsum, int64_t nt, int64_t np,
                                                                      int64_t *offsets)
{
    __shared__ int64_t sh[kScanThreads];
    const int64_t per = (nt + kScanThreads - 1) / kScanThreads;
    const int64_t a = (int64_t)threadIdx.x * per, b = a + per < nt ? a + per : nt;
    int64_t v = 0;
    for (int64_t i = a; i < b; ++i) v += tsum[i];
    int64_t total;
    int64_t run = block_exclusive_scan(v, sh, &total);
    for (int64_t i = a; i < b; ++i) {
        const int64_t c = tsum[i];
        tsum[i] = run;
        run += c;
    }
    if (threadIdx.x == 0) offsets[np] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_apply_kernel(const int32_t *counts, int64_t np,
                                                                       const int64_t *tbase, int64_t *offsets)
{
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int32_t c[kScanPer];
    int64_t v = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        c[j] = base + j < np ? counts[base + j] : 0;
        v += c[j];
    }
    int64_t total;
    int64_t run = tbase[blockIdx.x] + block_exclusive_scan(v, sh, &total);
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        if (base + j < np) offsets[base + j] = run;
        run += c[j];
    }
}

int64_t scan_tiles(int64_t np) { return (np + kScanTile - 1) / kScanTile; }

void launch_trace_count(const DevProblem &P, uint64_t p0, int64_t np, int32_t *counts, unsigned long long *ctr,
                        hipStream_t s)
{
    if (np <= 0) return;
    launch_zero(ctr, sizeof(unsigned long long), s);
    if (P.dim == 3)
        hipLaunchKernelGGL(trace_count_kernel<3>, dim3(resident_blocks((const void *)trace_count_kernel<3>, np)),
                           dim3(256), 0, s, P, p0, np, counts, ctr);
    else
        hipLaunchKernelGGL(trace_count_kernel<2>, dim3(resident_blocks((const void *)trace_count_kernel<2>, np)),
                           dim3(256), 0, s, P, p0, np, counts, ctr);
}

void launch_scan(const int32_t *counts, int64_t np, int64_t *offsets, int64_t *tsum, hipStream_t s)
{
    const int64_t nt = scan_tiles(np);
    if (nt > 0) hipLaunchKernelGGL(scan_tile_sum_kernel, dim3((unsigned)nt), dim3(kScanThreads), 0, s, counts, np, tsum);
    hipLaunchKernelGGL(scan_tile_base_kernel, dim3(1), dim3(kScanThreads), 0, s, tsum, nt, np, offsets);
    if (nt > 0)
        hipLaunchKernelGGL(scan_tile_apply_kernel, dim3((unsigned)nt), dim3(kScanThreads), 0, s, counts, np,
                           (const int64_t *)tsum, offsets);
}

void launch_trace_emit(const DevProblem &P, uint64_t p0, int64_t np, const int64_t *offsets, PathRec *paths,
                       VertexRec *verts, int32_t *vpath, unsigned long long *ctr, hipStream_t s)
{
    if (np <= 0) return;
    launch_zero(ctr, sizeof(unsigned long long), s);
    if (P.dim == 3)
        hipLaunchKernelGGL(trace_emit_kernel<3>, dim3(resident_blocks((const void *)trace_emit_kernel<3>, np)),
                           dim3(256), 0, s, P, p0, np, offsets, paths, verts, vpath, ctr);
    else
        hipLaunchKernelGGL(trace_emit_kernel<2>, dim3(resident_blocks((const void *)trace_emit_kernel<2>, np)),
                           dim3(256), 0, s, P, p0, np, offsets, paths, verts, vpath, ctr);
}

// The tabulated direction table: each sampleIcdf angle replaced by its
// cosine, with the same cos the walk used per draw, so a walk is unchanged.
__global__ void cos_table_kernel(double *t, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) t[i] = cos(t[i]);
}

void launch_cos_table(double *t, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(cos_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, n);
}

__global__ void zero_kernel(uint4 *p, size_t n16)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

// Clears with a fill kernel rather than hipMemsetAsync so the same code is
// safe inside captured graphs.  bytes must be a multiple of 16.
// Zeroes `bytes` rounded up to whole 16-byte words: the buffers are device
// allocations padded to 256 bytes (DevBuf), so the rounding stays inside them.
void launch_zero(void *p, size_t bytes, hipStream_t s)
{
    size_t n16 = (bytes + 15) / 16;
    if (!n16) return;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)p, n16);
}

} // namespace scmc
