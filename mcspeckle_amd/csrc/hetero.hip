// hetero.hip -- walks and the mean field of code/MCCovHetro.m (voxel media).
//
// The walks fill the same path and vertex records as the homogeneous tracer,
// so the NEE factors (with the cell-by-cell attenuation of hetero.h) and the
// matrix-core covariance update are shared.  A path whose first vertex falls
// outside the box (MCCovHetro.m:121-123, the iteration is spent) is stored as
// one vertex of weight 0 at the box centre: every factor it feeds is zero and
// the path-major layout (one first vertex per path) is kept.
#include "hetero.h"
#include "kernels.h"

#include <algorithm>

namespace scmc {

// Both passes keep their lanes resident and refill them from a path counter
// (take_path, refill_due), as the homogeneous tracer does: Woodcock walks in
// a voxel medium vary in length as much as free flights in a slab.  Outputs
// are indexed by path, so they do not depend on which lane walked a path.
template <int DIM>
__global__ __launch_bounds__(256) void hetero_count_kernel(DevProblem P, HetMedium M, uint64_t p0, int64_t np,
                                                           int32_t *counts, unsigned long long *ctr)
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
                ow[0] = ow[1] = ow[2] = 0;   // a new path has no incoming direction
                n = 0;
                if (hetero_path_start<DIM>(P, M, path, x, px, w, w0p)) {
                    weight = med_albedo<DIM>(P, M, x);
                    active = true;
                } else {
                    counts[t] = 1;   // stored as one zero-weight vertex
                }
            }
        }
        if (!__ballot(active)) {
            // a path whose first vertex misses the box ends at its start, so
            // an idle wave ends only once the counter is exhausted
            if (wp.exhausted) break;
            continue;
        }
        if (active) {
            ++n;
            const double2 a = uniform_pair(P.seed, path, n, 0);
            const double2 e = uniform_pair(P.seed, path, n, 3);
            if (!hetero_path_step<DIM>(P, M, path, n, x, w, ow, weight, a.y, e.y, tab)) {
                counts[id] = (int32_t)n;
                active = false;
            }
        }
    }
}

template <int DIM>
__global__ __launch_bounds__(256) void hetero_emit_kernel(DevProblem P, HetMedium M, uint64_t p0, int64_t np,
                                                          const int64_t *offsets, PathRec *paths, VertexRec *verts,
                                                          int32_t *vpath, unsigned long long *ctr)
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
                ow[0] = ow[1] = ow[2] = 0;   // a new path has no incoming direction
                o = offsets[t];
                end = offsets[t + 1];
                n = 0;
                PathRec pr;
                pr.pad = 0;
                if (hetero_path_start<DIM>(P, M, path, x, px, w, w0p)) {
                    for (int j = 0; j < 3; ++j) {
                        pr.x1[j] = x[j];
                        pr.w1[j] = w[j];
                    }
                    pr.px = px;
                    pr.w0p = w0p;
                    weight = med_albedo<DIM>(P, M, x);
                    active = true;
                } else {
                    // one vertex of weight 0 at the box centre
                    for (int j = 0; j < 3; ++j) {
                        pr.x1[j] = j < DIM ? 0.5 * (P.bmin[j] + P.bmax[j]) : 0.0;
                        pr.w1[j] = P.meanl[j];
                    }
                    pr.px = 1;
                    pr.w0p = make_double2(P.w0p_iso, 0.0);
                    VertexRec r;
                    for (int j = 0; j < 3; ++j) {
                        r.x[j] = pr.x1[j];
                        r.ow[j] = 0;
                    }
                    r.weight = 0;
                    r.xi = 0;
                    verts[o] = r;
                    vpath[o] = (int32_t)t;
                }
                paths[t] = pr;
            }
        }
        if (!__ballot(active)) {
            // a path whose first vertex misses the box ends at its start, so
            // an idle wave ends only once the counter is exhausted
            if (wp.exhausted) break;
            continue;
        }
        if (active) {
            ++n;
            const double2 a = uniform_pair(P.seed, path, n, 0);
            const double2 e = uniform_pair(P.seed, path, n, 3);
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
            if (!hetero_path_step<DIM>(P, M, path, n, x, w, ow, weight, a.y, e.y, tab)) active = false;
        }
    }
}

// MCCovHetro.m:152, 210-211: mean1(v,l) += W_1/px af_ang_vl(v,l) e_v0(v) u_l0(l)
// with u_l0 the unattenuated light term (evalphaseatt with sigt = 0).  EV0
// holds W_1/px * e_v0 per path, UL0 u_l0 per path (both written by the NEE
// path kernel).  Thread per (v, l), workgroup column per range of paths
// (summed in order); the ranges' partials are added in order by
// hetero_mean_reduce_kernel.
template <int DIM>
__global__ __launch_bounds__(256) void hetero_mean_kernel(DevProblem P, int64_t np, int nsplit, const PathRec *paths,
                                                          const double2 *EV0, const double2 *UL0, double2 *part)
{
    const int64_t n = (int64_t)P.nv * P.nl;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int v = (int)(idx % P.nv), l = (int)(idx / P.nv);
    const int split = blockIdx.y;
    const int64_t p0 = np * split / nsplit, p1 = np * (split + 1) / nsplit;
    const bool far = P.is_ff_l && P.is_ff_v;
    double2 af = far ? eval_amp<DIM>(dot<DIM>(P.l + (size_t)l * DIM, P.v + (size_t)v * DIM), P.amp)
                     : make_double2(0.0, 0.0);
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t p = p0; p < p1; ++p) {
        if (!far) {
            double rl[3], rv[3];
            light_dir<DIM>(P, paths[p].x1, l, rl);
            view_dir<DIM>(P, paths[p].x1, v, rv);
            af = eval_amp<DIM>(dot<DIM>(rl, rv), P.amp);
        }
        const double2 t = cmul(af, cmul(EV0[p * P.nv + v], UL0[p * P.nl + l]));
        acc.x += t.x;
        acc.y += t.y;
    }
    part[(size_t)split * n + idx] = acc;
}

__global__ void hetero_mean_reduce_kernel(int64_t n, int nsplit, const double2 *part, double2 *mean)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    double2 s = mean[idx];
    for (int k = 0; k < nsplit; ++k) {
        const double2 t = part[(size_t)k * n + idx];
        s.x += t.x;
        s.y += t.y;
    }
    mean[idx] = s;
}

void launch_hetero_trace(const DevProblem &P, const HetMedium &M, uint64_t p0, int64_t np, int32_t *counts,
                         int64_t *offsets, PathRec *paths, VertexRec *verts, int32_t *vpath, bool emit,
                         unsigned long long *ctr, hipStream_t s)
{
    if (np <= 0) return;
    launch_zero(ctr, sizeof(unsigned long long), s);
    auto go = [&](auto kernel, auto... args) {
        hipLaunchKernelGGL(kernel, dim3(resident_blocks((const void *)kernel, np)), dim3(256), 0, s, P, M, p0, np,
                           args..., ctr);
    };
    if (!emit) {
        if (P.dim == 3)
            go(hetero_count_kernel<3>, counts);
        else
            go(hetero_count_kernel<2>, counts);
        return;
    }
    if (P.dim == 3)
        go(hetero_emit_kernel<3>, (const int64_t *)offsets, paths, verts, vpath);
    else
        go(hetero_emit_kernel<2>, (const int64_t *)offsets, paths, verts, vpath);
}

void launch_hetero_mean(const DevProblem &P, int64_t np, const PathRec *paths, const double2 *EV0,
                        const double2 *UL0, double2 *mean, double2 *part, hipStream_t s)
{
    const int64_t n = (int64_t)P.nv * P.nl;
    if (np <= 0 || n <= 0) return;
    int nsplit = (int)std::min<int64_t>(kHetMeanSplits, std::max<int64_t>(1, np / 128));
    dim3 g((unsigned)((n + 255) / 256), (unsigned)nsplit);
    if (P.dim == 3)
        hipLaunchKernelGGL(hetero_mean_kernel<3>, g, dim3(256), 0, s, P, np, nsplit, paths, EV0, UL0, part);
    else
        hipLaunchKernelGGL(hetero_mean_kernel<2>, g, dim3(256), 0, s, P, np, nsplit, paths, EV0, UL0, part);
    hipLaunchKernelGGL(hetero_mean_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, nsplit, part,
                       mean);
}

} // namespace scmc
