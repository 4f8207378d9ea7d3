// field_engine.cpp -- host side of the MCfield accumulation (MCfield.m:190-254).
#include "scmc_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace scmc {

int64_t field_chunk_paths(int nv, int nl, int64_t remaining, double mean_len)
{
    (void)nv;
    (void)nl;
    int64_t target_vertices = 1 << 21;
    int64_t p = (int64_t)(target_vertices / std::max(mean_len, 1.0));
    p = std::max<int64_t>(256, std::min<int64_t>(p, 1 << 20));
    return std::min(p, remaining);
}

int FieldEngine::init(const ProblemState &ps, hipStream_t s)
{
    int rc = SCMC_OK;
    nv = ps.dp.nv;
    nl = ps.dp.nl;
    if (!acc.ensure(sizeof(double2) * (size_t)nv * nl, &rc)) return rc;
    launch_zero(acc.p, sizeof(double2) * (size_t)nv * nl, s);
    return SCMC_OK;
}

int FieldEngine::chunk(const ProblemState &ps, TraceChunk &tc, bool mixed, hipStream_t s, double *ms_accum)
{
    int rc = SCMC_OK;
    // coef[nvert*nl] | coefR[np*nl] | aftab[nv*nl] | part[64*nv*nl]
    size_t need = (size_t)tc.nvert * nl + (size_t)tc.np * nl + (size_t)nv * nl + (size_t)64 * nv * nl;
    if (!work.ensure(sizeof(double2) * need, &rc)) return rc;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    launch_field(ps.dp, tc.np, tc.nvert, (const PathRec *)tc.paths.p, (const VertexRec *)tc.verts.p,
                 (const int32_t *)tc.vpath.p, (const int64_t *)tc.offsets.p, (double2 *)work.p, (double2 *)acc.p, mixed,
                 s);
    (void)hipEventRecord(e1, s);
    if (!hip_ok(hipGetLastError(), "field launch", &rc)) return rc;
    if (!hip_ok(hipEventSynchronize(e1), "field sync", &rc)) return rc;
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms_accum += t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SCMC_OK;
}

int FieldEngine::finalize(const scmc_problem *pr, int64_t max_itr, double *u_re, double *u_im, hipStream_t s)
{
    int rc = SCMC_OK;
    const size_t ne = (size_t)nv * nl;
    std::vector<double> h(ne * 2);
    if (!hip_ok(hipMemcpyAsync(h.data(), acc.p, sizeof(double) * 2 * ne, hipMemcpyDeviceToHost, s), "copy field", &rc) ||
        !hip_ok(hipStreamSynchronize(s), "field sync", &rc))
        return rc;
    // MCfield.m:253-254
    const double sc = sqrt(1.0 / (double)max_itr * box_volume(pr) * pr->sigt);
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
