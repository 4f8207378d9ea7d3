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

template <int DIM>
__global__ __launch_bounds__(256) void trace_count_kernel(DevProblem P, uint64_t p0, int64_t np, int32_t *counts)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    uint64_t path = P.path_offset + p0 + (uint64_t)i;
    double x[3] = {0, 0, 0}, w[3] = {0, 0, 0}, ow[3] = {0, 0, 0}, px;
    double2 w0p;
    path_start<DIM>(P, path, x, px, w, w0p);
    double weight = P.albedo;
    uint32_t n = 0;
    for (;;) {
        ++n;
        double2 a = uniform_pair(P.seed, path, n, 0);
        if (!path_step<DIM>(P, path, n, x, w, ow, weight, a.y)) break;
    }
    counts[i] = (int32_t)n;
}

template <int DIM>
__global__ __launch_bounds__(256) void trace_emit_kernel(DevProblem P, uint64_t p0, int64_t np, const int64_t *offsets,
                                                         PathRec *paths, VertexRec *verts, int32_t *vpath)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    uint64_t path = P.path_offset + p0 + (uint64_t)i;
    double x[3] = {0, 0, 0}, w[3] = {0, 0, 0}, ow[3] = {0, 0, 0}, px;
    double2 w0p;
    path_start<DIM>(P, path, x, px, w, w0p);
    PathRec pr;
    for (int j = 0; j < 3; ++j) {
        pr.x1[j] = x[j];
        pr.w1[j] = w[j];
    }
    pr.px = px;
    pr.pad = 0;
    pr.w0p = w0p;
    paths[i] = pr;
    double weight = P.albedo;
    int64_t o = offsets[i];
    const int64_t end = offsets[i + 1];
    uint32_t n = 0;
    for (;;) {
        ++n;
        double2 a = uniform_pair(P.seed, path, n, 0);
        if (o < end) {
            VertexRec r;
            for (int j = 0; j < 3; ++j) {
                r.x[j] = x[j];
                r.ow[j] = ow[j];
            }
            r.weight = weight;
            r.xi = a.x;
            verts[o] = r;
            vpath[o] = (int32_t)i;
            ++o;
        }
        if (!path_step<DIM>(P, path, n, x, w, ow, weight, a.y)) break;
    }
}

// Single-workgroup exclusive scan of int32 counts into int64 offsets.  The
// chunk sizes here are at most a few hundred thousand paths.
__global__ __launch_bounds__(1024) void scan_kernel(const int32_t *counts, int64_t np, int64_t *offsets)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (np + 1023) / 1024;
    const int64_t a = t * per, b = (a + per < np) ? a + per : np;
    int64_t s = 0;
    for (int64_t i = a; i < b; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;   // exclusive prefix of this thread's slice
    for (int64_t i = a; i < b; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (t == 1023) offsets[np] = part[1023];
}

void launch_trace_count(const DevProblem &P, uint64_t p0, int64_t np, int32_t *counts, hipStream_t s)
{
    if (np <= 0) return;
    dim3 g((unsigned)((np + 255) / 256));
    if (P.dim == 3)
        hipLaunchKernelGGL(trace_count_kernel<3>, g, dim3(256), 0, s, P, p0, np, counts);
    else
        hipLaunchKernelGGL(trace_count_kernel<2>, g, dim3(256), 0, s, P, p0, np, counts);
}

void launch_scan(const int32_t *counts, int64_t np, int64_t *offsets, hipStream_t s)
{
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, counts, np, offsets);
}

void launch_trace_emit(const DevProblem &P, uint64_t p0, int64_t np, const int64_t *offsets, PathRec *paths,
                       VertexRec *verts, int32_t *vpath, hipStream_t s)
{
    if (np <= 0) return;
    dim3 g((unsigned)((np + 255) / 256));
    if (P.dim == 3)
        hipLaunchKernelGGL(trace_emit_kernel<3>, g, dim3(256), 0, s, P, p0, np, offsets, paths, verts, vpath);
    else
        hipLaunchKernelGGL(trace_emit_kernel<2>, g, dim3(256), 0, s, P, p0, np, offsets, paths, verts, vpath);
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
