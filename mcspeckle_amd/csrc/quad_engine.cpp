// quad_engine.cpp -- host side of MCfieldFourierQuad (code/MCfieldFourierQuad.m:1-303)
// and MCfieldFourierQuadOnWave (code/MCfieldFourierQuadOnWave.m:1-337).
#include "quad.h"
#include "scmc_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <vector>

namespace scmc {

static int colon_len(double a, double s, double b)
{
    // element count of MATLAB's a:s:b (with its relative tolerance)
    double q = (b - a) / s;
    return (int)floor(q + q * 1e-12 + 1e-10) + 1;
}

int quad_grid_setup(const double *box_min, const double *box_max, double v_max, double v_stp, double lambda,
                    bool parity_fix, QuadGeom &g, double *bmin, double *bmax, std::vector<double> *gz)
{
    if (!(v_stp > 0) || !(v_max > 0) || !(lambda > 0))
        return set_error(SCMC_E_ARG, "v_max, v_stp and lambda must be positive");
    // MCfieldFourierQuad.m:80-98
    g.nv0 = colon_len(-v_max, v_stp, v_max);
    g.box_stp = lambda / (v_max * 2 + v_stp);
    const double bw0 = box_max[0] - box_min[0];
    g.nx = (int)ceil(bw0 / g.box_stp / 2) * 2 + 1;
    // MCfieldFourierQuadOnWave.m:100-102: the cell grid takes the parity of
    // the view grid (the transform size may then be even)
    if (parity_fix && (g.nv0 - g.nx) % 2 != 0) g.nx += 1;
    // the padding / crop of MCfieldFourierQuad.m:103-106 must be whole cells
    if ((g.nv0 - g.nx) % 2 != 0)
        return set_error(SCMC_E_ARG, "view grid (%d) and cell grid (%d) differ by an odd number of cells", g.nv0, g.nx);
    const double bwn = g.nx * g.box_stp;
    for (int j = 0; j < 3; ++j) {
        bmin[j] = box_min[j];
        bmax[j] = box_max[j];
    }
    for (int j = 0; j < 2; ++j) {
        bmin[j] = box_min[j] - (bwn - bw0) / 2;
        bmax[j] = box_max[j] + (bwn - bw0) / 2;
    }
    g.gx0 = bmin[0] + g.box_stp / 2;
    g.gy0 = bmin[1] + g.box_stp / 2;
    // MCfieldFourierQuad.m:103-106
    g.nx_pad = std::max((g.nv0 - g.nx) / 2, 0);
    g.nva_pad = std::max((g.nx - g.nv0) / 2, 0);
    g.nfft = g.nx + 2 * g.nx_pad;
    // MCfieldFourierQuad.m:108-121: z_bin from the widest view of the grid,
    // max(1 - |vz|) over the grid.  |vx| = |-v_max + i v_stp| is largest at
    // i = 0 or nv0 - 1 (a rounded sum is monotone in each term), and so are
    // vx^2 + vy^2 and 1 - |vz|: the four corners give the maximum exactly.
    double worst = 0;
    for (int i : {0, g.nv0 - 1})
        for (int j : {0, g.nv0 - 1}) {
            double vx = -(-v_max + i * v_stp), vy = -(-v_max + j * v_stp);
            double q = 1 - fabs(sqrt(1 - vx * vx - vy * vy));
            worst = std::max(worst, q);
        }
    if (!(worst > 0)) return set_error(SCMC_E_ARG, "view grid has no oblique direction (v_max too small)");
    g.z_bin = lambda / worst / 5;
    g.z_bin = g.z_bin / 2;
    const int nb = colon_len(bmin[2], g.z_bin, bmax[2]);
    g.nz = nb + 1;
    if (g.nz > kQuadMaxSlabs) return set_error(SCMC_E_ARG, "too many z slabs (%d)", g.nz);
    if (gz) {
        gz->resize((size_t)g.nz);
        for (int i = 0; i < nb; ++i) (*gz)[(size_t)i] = bmin[2] + i * g.z_bin;
        (*gz)[(size_t)g.nz - 1] = bmax[2];
    }
    return SCMC_OK;
}

} // namespace scmc

using namespace scmc;

extern "C" {

static int quad_grid_query(const double box_min[3], const double box_max[3], double v_max, double v_stp,
                           double lambda, bool wave, int32_t *nv0, int32_t *nx, int32_t *nz)
{
    QuadGeom g;
    double bmin[3], bmax[3];
    int rc = quad_grid_setup(box_min, box_max, v_max, v_stp, lambda, wave, g, bmin, bmax, nullptr);
    if (rc) return rc;
    if (nv0) *nv0 = g.nv0;
    if (nx) *nx = g.nx;
    if (nz) *nz = g.nz;
    return SCMC_OK;
}

int scmc_quad_grid(const double box_min[3], const double box_max[3], double v_max, double v_stp, double lambda,
                   int32_t *nv0, int32_t *nx, int32_t *nz)
{
    return quad_grid_query(box_min, box_max, v_max, v_stp, lambda, false, nv0, nx, nz);
}

int scmc_quad_grid_wave(const double box_min[3], const double box_max[3], double v_max, double v_stp, double lambda,
                        int32_t *nv0, int32_t *nx, int32_t *nz)
{
    return quad_grid_query(box_min, box_max, v_max, v_stp, lambda, true, nv0, nx, nz);
}

// Device buffers and the stream of the Fourier solves, kept from call to
// call (they only grow): the slab grids and transforms of a 511^2 problem are
// ~0.4 GB, and allocating and freeing them per call cost ~2 ms of device-wide
// waits.  Calls in one process take turns (mu).  Never destroyed: the HIP
// runtime may be gone by the time static destructors run.
struct QuadWorkspace {
    std::mutex mu;
    hipStream_t s = nullptr;
    ProblemState ps;
    DevBuf dgz, grid, W, tw, T1, H, u, dwl, dwlk, dcount;
    double *host_u = nullptr;   // pinned staging of the result (a pageable copy ran at ~2 GB/s)
    size_t host_u_bytes = 0;
};
static QuadWorkspace *quad_workspace()
{
    static QuadWorkspace *w = new QuadWorkspace();
    return w;
}

// MCfieldFourierQuad.m:1-303, or with wl_re MCfieldFourierQuadOnWave.m:1-337.
static int fourier_quad_run(const scmc_problem *pr_in, const double *wl_re, const double *wl_im, double v_max,
                            double v_stp, int64_t max_itr, double vsign, int32_t precision, double *u_re, double *u_im,
                            scmc_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    const bool wave = wl_re != nullptr;
    if (!pr_in) return set_error(SCMC_E_ARG, "null problem");
    if (pr_in->dim != 3) return set_error(SCMC_E_ARG, "MCfieldFourierQuad is 3-D only");
    if (precision != SCMC_FP64 && precision != SCMC_MIXED) return set_error(SCMC_E_ARG, "invalid precision %d", precision);
    int rc = validate_problem(pr_in, max_itr, false);
    if (rc) return rc;
    QuadGeom g;
    double bmin[3], bmax[3];
    std::vector<double> gz;
    if ((rc = quad_grid_setup(pr_in->box_min, pr_in->box_max, v_max, v_stp, pr_in->lambda, wave, g, bmin, bmax, &gz)))
        return rc;
    if ((rc = ensure_device(nullptr))) return rc;
    const bool fp32 = precision == SCMC_MIXED;
    // the walk runs in the enlarged box with lmean0 (default [0;0;1]) and no
    // ampfunc0 (MCfieldFourierQuad.m:59-64, 96-98, 166-172)
    scmc_problem pr = *pr_in;
    for (int j = 0; j < 3; ++j) {
        pr.box_min[j] = bmin[j];
        pr.box_max[j] = bmax[j];
    }
    pr.is_ff_l = 1;
    pr.is_ff_v = 1;
    pr.nv = 0;
    pr.v = nullptr;
    pr.has_amp0 = 0;
    if (!pr.has_lmean0) {
        pr.has_lmean0 = 1;
        pr.lmean0[0] = 0;
        pr.lmean0[1] = 0;
        pr.lmean0[2] = 1;
    }
    const double lz = pr.l[3 * (pr.nl - 1) + 2];   // l(end)
    g.l_diratt = lz > 0 ? 1 / fabs(lz) : 0;
    g.l_idiratt = lz < 0 ? 1 / fabs(lz) : 0;
    QuadWorkspace &ws = *quad_workspace();
    std::lock_guard<std::mutex> lock(ws.mu);
    if (!ws.s && !hip_ok(hipStreamCreate(&ws.s), "stream", &rc)) return rc;
    hipStream_t s = ws.s;
    scmc_stats st;
    memset(&st, 0, sizeof st);
    {
        ProblemState &ps = ws.ps;
        DevBuf &dgz = ws.dgz, &grid = ws.grid, &W = ws.W, &tw = ws.tw, &T1 = ws.T1, &H = ws.H, &u = ws.u,
               &dwl = ws.dwl, &dwlk = ws.dwlk, &dcount = ws.dcount;
        const int nx = g.nx, nv0 = g.nv0, nsl = g.nz - 1;
        const int nl = wave ? 1 : pr.nl;   // output planes
        const size_t cs = fp32 ? sizeof(float2) : sizeof(double2);
        const size_t ngrid = (size_t)nsl * nl * nx * nx;
        do {
            if ((rc = ps.setup(&pr, s))) break;
            if (!dgz.ensure(sizeof(double) * gz.size(), &rc) || !grid.ensure(cs * ngrid, &rc)) break;
            if (!hip_ok(hipMemcpyAsync(dgz.p, gz.data(), sizeof(double) * gz.size(), hipMemcpyHostToDevice, s), "gz",
                        &rc))
                break;
            g.gz = (const double *)dgz.p;
            if (wave) {
                std::vector<double> hw((size_t)pr.nl * 2);
                // the FP32-grid kernels' form: phase gradients in cycles per
                // unit length, weights in FP32
                std::vector<WaveLight> hk((size_t)pr.nl);
                const double kc = vsign * (1.0 / pr.lambda);
                for (int l = 0; l < pr.nl; ++l) {
                    hw[2 * (size_t)l] = wl_re[l];
                    hw[2 * (size_t)l + 1] = wl_im ? wl_im[l] : 0.0;
                    hk[(size_t)l] = WaveLight{kc * pr.l[3 * (size_t)l], kc * pr.l[3 * (size_t)l + 1],
                                              kc * pr.l[3 * (size_t)l + 2], (float)hw[2 * (size_t)l],
                                              (float)hw[2 * (size_t)l + 1]};
                }
                if (!dwl.ensure(sizeof(double) * hw.size(), &rc) || !dwlk.ensure(sizeof(WaveLight) * hk.size(), &rc) ||
                    !hip_ok(hipMemcpyAsync(dwl.p, hw.data(), sizeof(double) * hw.size(), hipMemcpyHostToDevice, s),
                            "wl", &rc) ||
                    !hip_ok(hipMemcpyAsync(dwlk.p, hk.data(), sizeof(WaveLight) * hk.size(), hipMemcpyHostToDevice, s),
                            "wl", &rc) ||
                    !hip_ok(stream_wait(s), "wl sync", &rc))
                    break;
                g.wave = 1;
                g.vsign = vsign;
                g.wl = (const double2 *)dwl.p;
                g.wlk = (const WaveLight *)dwlk.p;
            }
            launch_zero(grid.p, (cs * ngrid + 15) / 16 * 16 > grid.bytes ? grid.bytes : (cs * ngrid + 15) / 16 * 16, s);
            // At most kQuadFusedLights lights (a wave: kWaveLightsLds, its one
            // output plane and its lights staged in LDS): walk and scatter in
            // one pass (no vertex records, no counting pass).  More: trace
            // into records, then a vertex-parallel scatter.
            const bool fused = wave ? pr.nl <= kWaveLightsLds : pr.nl <= kQuadFusedLights;
            if (fused) {
                // [0]: vertex count; [1 + launch]: the launch's path counter
                const int64_t kLaunch = (int64_t)1 << 24;   // paths per launch
                const int64_t nlaunch = (max_itr + kLaunch - 1) / kLaunch;
                if (!dcount.ensure(sizeof(unsigned long long) * (size_t)(1 + nlaunch), &rc)) break;
                launch_zero(dcount.p, sizeof(unsigned long long) * (size_t)(1 + nlaunch), s);
                auto a = std::chrono::steady_clock::now();
                for (int64_t p0 = 0, li = 0; p0 < max_itr; p0 += kLaunch, ++li) {
                    const int64_t np = std::min<int64_t>(max_itr - p0, kLaunch);
                    if (!launch_quad_trace_scatter(ps.dp, g, (uint64_t)p0, np, grid.p, fp32,
                                                   (unsigned long long *)dcount.p,
                                                   (unsigned long long *)dcount.p + 1 + li, s)) {
                        rc = set_error(SCMC_E_UNSUPPORTED, "fused Fourier pass for %d lights", pr.nl);
                        break;
                    }
                    st.paths += np;
                    st.chunks += 1;
                }
                if (rc) break;
                unsigned long long hv = 0;
                if (!hip_ok(hipMemcpyAsync(&hv, dcount.p, sizeof hv, hipMemcpyDeviceToHost, s), "vertex count", &rc) ||
                    !hip_ok(stream_wait(s), "quad scatter", &rc))
                    break;
                st.ms_accum += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
                st.vertices = (int64_t)hv;
            } else {
                TraceChunk tc;
                double mean_len = 20;
                for (int64_t p0 = 0; p0 < max_itr;) {
                    // ~3.2e7 vertices (2 GB of records) per chunk: a trace pass
                    // takes about the longest walk's step count times the step
                    // latency whatever the chunk size, so few large chunks
                    // amortise it
                    int64_t np = std::min<int64_t>(max_itr - p0, std::max<int64_t>(256, (int64_t)(3.2e7 / mean_len)));
                    auto a = std::chrono::steady_clock::now();
                    if ((rc = tc.run(ps.dp, (uint64_t)p0, np, s))) break;
                    launch_quad_scatter(ps.dp, g, tc.nvert, (const PathRec *)tc.paths.p, (const VertexRec *)tc.verts.p,
                                        (const int32_t *)tc.vpath.p, (const int64_t *)tc.offsets.p, grid.p, fp32, s);
                    if (!hip_ok(stream_wait(s), "quad scatter", &rc)) break;
                    st.ms_accum += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
                    st.paths += np;
                    st.vertices += tc.nvert;
                    st.chunks += 1;
                    mean_len = (double)st.vertices / (double)st.paths;
                    p0 += np;
                }
                if (rc) break;
            }
            const int64_t nstore = st.vertices;
            if (rc) break;
            // DFT matrix of fliplr(flipud(fftshift(fft2(ifftshift(.))))) and the crop:
            // W[o][ix] = exp(+2 pi i ((o + nva_pad - a)(ix + nx_pad - h) mod n)/n),
            // h = floor(n/2) (the shift of fftshift/ifftshift), a = n - 1 - h;
            // built on the device from the n twiddles
            const int n = g.nfft;
            std::vector<double> htw((size_t)n * 2);
            for (int r = 0; r < n; ++r) {
                const double ang = 2 * M_PI * (double)r / n;
                htw[2 * (size_t)r] = cos(ang);
                htw[2 * (size_t)r + 1] = sin(ang);
            }
            if (!W.ensure(cs * (size_t)nv0 * nx, &rc) || !tw.ensure(sizeof(double) * htw.size(), &rc) ||
                !T1.ensure(cs * (size_t)nsl * nl * nv0 * nx, &rc) || !H.ensure(cs * (size_t)nsl * nl * nv0 * nv0, &rc) ||
                !u.ensure(sizeof(double2) * (size_t)nv0 * nv0 * nl, &rc))
                break;
            if (!hip_ok(hipMemcpyAsync(tw.p, htw.data(), sizeof(double) * htw.size(), hipMemcpyHostToDevice, s), "tw",
                        &rc))
                break;
            auto a = std::chrono::steady_clock::now();
            // H[b] = W X[b] W^T: factored DFT passes when n factors (513 =
            // 27 x 19), else T1[b] = W (nv0 x nx) X[b] (nx x nx), H[b] = T1[b] W^T
            if (!launch_quad_dft2(g, nsl * nl, grid.p, T1.p, H.p, (const double2 *)tw.p, fp32, s)) {
                launch_quad_dft_matrix(g, (const double2 *)tw.p, W.p, fp32, s);
                launch_cgemm(nv0, nx, nx, W.p, 0, grid.p, (int64_t)nx * nx, T1.p, (int64_t)nv0 * nx, nsl * nl, false,
                             fp32, s);
                launch_cgemm(nv0, nv0, nx, T1.p, (int64_t)nv0 * nx, W.p, 0, H.p, (int64_t)nv0 * nv0, nsl * nl, true,
                             fp32, s);
            }
            // slab phases exp(zexp0 tz), boxexp0 (MCfieldFourierQuad.m:124-138,
            // 281-282) and MCfieldFourierQuad.m:302-303 (maxItr = stored
            // vertices), per output pixel on the device
            QuadView view;
            view.v_max = v_max;
            view.v_stp = v_stp;
            view.vsign = vsign;
            view.sigt = pr.sigt;
            view.lambda = pr.lambda;
            view.zmin = bmin[2];
            view.zmax = bmax[2];
            view.wave = wave ? 1 : 0;
            const double bwn = g.nx * g.box_stp;
            const double V = bwn * bwn * (bmax[2] - bmin[2]);
            view.scale = nstore ? sqrt(1.0 / (double)nstore * V * pr.sigt) : 0.0;
            launch_quad_combine(nv0, nl, g.nz, H.p, g.gz, view, (double2 *)u.p, fp32, s);
            const size_t ne = (size_t)nv0 * nv0 * nl;
            if (ws.host_u_bytes < sizeof(double) * 2 * ne) {
                if (ws.host_u) (void)hipHostFree(ws.host_u);
                ws.host_u = nullptr;
                ws.host_u_bytes = 0;
                if (!hip_ok(hipHostMalloc((void **)&ws.host_u, sizeof(double) * 2 * ne, hipHostMallocDefault),
                            "pinned result", &rc))
                    break;
                ws.host_u_bytes = sizeof(double) * 2 * ne;
            }
            const double *hu = ws.host_u;
            if (!hip_ok(hipMemcpyAsync(ws.host_u, u.p, sizeof(double) * 2 * ne, hipMemcpyDeviceToHost, s), "copy u",
                        &rc) ||
                !hip_ok(stream_wait(s), "quad sync", &rc))
                break;
            st.ms_nee += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
            if (u_im) {
                for (size_t k = 0; k < ne; ++k) {
                    u_re[k] = hu[2 * k];
                    u_im[k] = hu[2 * k + 1];
                }
            } else {
                memcpy(u_re, hu, sizeof(double) * 2 * ne);
            }
        } while (0);
        // an error leaves queued work behind: finish it before the next call
        if (rc) (void)stream_wait(s);
    }
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return rc;
}

int scmc_mcfield_fourier_quad(const scmc_problem *pr, double v_max, double v_stp, int64_t max_itr, double vsign,
                              int32_t precision, double *u_re, double *u_im, scmc_stats *stats)
{
    return fourier_quad_run(pr, nullptr, nullptr, v_max, v_stp, max_itr, vsign, precision, u_re, u_im, stats);
}

int scmc_mcfield_fourier_quad_wave(const scmc_problem *pr, const double *wl_re, const double *wl_im, double v_max,
                                   double v_stp, int64_t max_itr, double vsign, int32_t precision, double *u_re,
                                   double *u_im, scmc_stats *stats)
{
    if (!wl_re) return set_error(SCMC_E_ARG, "Wl must have one weight per light");
    return fourier_quad_run(pr, wl_re, wl_im, v_max, v_stp, max_itr, vsign, precision, u_re, u_im, stats);
}

} // extern "C"
