// field_engine.cpp -- host side of the MCfield accumulation (MCfield.m:190-254).
#include "scmc_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace scmc {

int64_t field_chunk_paths(int nv, int nl, int64_t remaining, double mean_len)
{
    (void)nv;
    // vertices per chunk: enough paths to keep the tracer's waves busy, with
    // the per-vertex coefficients (nl columns of 16 B) near 1 GB
    int64_t target_vertices = std::max<int64_t>(1 << 21, std::min<int64_t>(1 << 23, (1LL << 26) / std::max(nl, 1)));
    int64_t p = (int64_t)(target_vertices / std::max(mean_len, 1.0));
    p = std::max<int64_t>(256, std::min<int64_t>(p, 1 << 20));
    return std::min(p, remaining);
}

int FieldEngine::init(const ProblemState &ps, hipStream_t s, const double *wl_re, const double *wl_im)
{
    int rc = SCMC_OK;
    nv = ps.dp.nv;
    nl = ps.dp.nl;
    wave = wl_re != nullptr;
    het = ps.het;
    ncol = wave ? 1 : nl;
    if (!acc.ensure(sizeof(double2) * (size_t)nv * ncol, &rc) || !nonfinite.ensure(sizeof(unsigned long long), &rc))
        return rc;
    launch_zero(acc.p, sizeof(double2) * (size_t)nv * ncol, s);
    launch_zero(nonfinite.p, sizeof(unsigned long long), s);
    if (wave) {
        // the light weights Wl of MCfieldOnWave.m (nl complex)
        std::vector<double> h((size_t)nl * 2);
        for (int l = 0; l < nl; ++l) {
            h[2 * (size_t)l] = wl_re[l];
            h[2 * (size_t)l + 1] = wl_im ? wl_im[l] : 0.0;
        }
        if (!wl.ensure(sizeof(double2) * (size_t)nl, &rc)) return rc;
        if (!hip_ok(hipMemcpyAsync(wl.p, h.data(), sizeof(double2) * (size_t)nl, hipMemcpyHostToDevice, s), "wl", &rc) ||
            !hip_ok(stream_wait(s), "wl sync", &rc))
            return rc;
    }
    return SCMC_OK;
}

int FieldEngine::chunk(const ProblemState &ps, TraceChunk &tc, bool mixed, hipStream_t s, double *ms_accum)
{
    int rc = SCMC_OK;
    const size_t need = field_work_elems(tc.np, tc.nvert, nv, nl, wave);
    if (!work.ensure(sizeof(double2) * need, &rc)) return rc;
    if (tls_arena) {
        // allocation-free calls create no events: host time around the
        // launch and its completion
        auto t0 = std::chrono::steady_clock::now();
        launch_field(ps.dp, ps.het ? &ps.hm : nullptr, tc.np, tc.nvert, (const PathRec *)tc.paths.p,
                     (const VertexRec *)tc.verts.p, (const int32_t *)tc.vpath.p, (const int64_t *)tc.offsets.p,
                     (double2 *)work.p, (double2 *)acc.p, mixed, wave ? (const double2 *)wl.p : nullptr,
                     (unsigned long long *)nonfinite.p, s);
        if (!hip_ok(hipGetLastError(), "field launch", &rc) || !hip_ok(stream_wait(s), "field sync", &rc))
            return rc;
        *ms_accum += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return SCMC_OK;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    launch_field(ps.dp, ps.het ? &ps.hm : nullptr, tc.np, tc.nvert, (const PathRec *)tc.paths.p,
                 (const VertexRec *)tc.verts.p,
                 (const int32_t *)tc.vpath.p, (const int64_t *)tc.offsets.p, (double2 *)work.p, (double2 *)acc.p, mixed,
                 wave ? (const double2 *)wl.p : nullptr, (unsigned long long *)nonfinite.p, s);
    (void)hipEventRecord(e1, s);
    if (!hip_ok(hipGetLastError(), "field launch", &rc)) return rc;
    if (!hip_ok(event_wait(e1), "field sync", &rc)) return rc;
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms_accum += t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SCMC_OK;
}

int FieldEngine::read_nonfinite(int64_t *out, hipStream_t s)
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

int FieldEngine::finalize(const scmc_problem *pr, int64_t max_itr, double *u_re, double *u_im, hipStream_t s)
{
    int rc = SCMC_OK;
    const size_t ne = (size_t)nv * ncol;
    std::vector<double> h(ne * 2);
    if (!hip_ok(hipMemcpyAsync(h.data(), acc.p, sizeof(double) * 2 * ne, hipMemcpyDeviceToHost, s), "copy field", &rc) ||
        !hip_ok(stream_wait(s), "field sync", &rc))
        return rc;
    // MCfield.m:253-254 (MCfieldOnWave.m:265-267); a voxel medium carries the
    // volume in px and scales by sqrt(1/maxItr) (MCsampleHetro.m:222)
    const double sc = het ? sqrt(1.0 / (double)max_itr) : sqrt(1.0 / (double)max_itr * box_volume(pr) * pr->sigt);
    for (size_t i = 0; i < ne; ++i) {
        if (u_im) {
            u_re[i] = h[2 * i] * sc;
            u_im[i] = h[2 * i + 1] * sc;
        } else {
            u_re[2 * i] = h[2 * i] * sc;
            u_re[2 * i + 1] = h[2 * i + 1] * sc;
        }
    }
    return SCMC_OK;
}

} // namespace scmc
