// hetero_engine.cpp -- host side of MCCovHetro (code/MCCovHetro.m:1-276):
// the per-medium constants of MCCovHetro.m:37-66 and expSmpXHetro.m:4-10,
// and the C entry point.
#include "scmc_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace scmc {

// invCDFvec.m:1-8 on cumsum(sigt(:)'/sum(sigt(:))): MATLAB's hist with n_out
// equal bins over [min, max] (inner edges belong to the upper bin), then the
// cumulative counts.
static void cell_icdf(const std::vector<double> &sig, int n_out, std::vector<int32_t> &icdf)
{
    double tot = 0;
    for (double v : sig) tot += v;
    const int n = (int)sig.size();
    std::vector<double> cdf((size_t)n);
    double c = 0;
    for (int i = 0; i < n; ++i) {
        c += sig[(size_t)i] / tot;
        cdf[(size_t)i] = c;
    }
    double lo = cdf[0], hi = cdf[0];
    for (double v : cdf) {
        lo = std::min(lo, v);
        hi = std::max(hi, v);
    }
    std::vector<int32_t> cnt((size_t)n_out, 0);
    for (double v : cdf) {
        int b;
        if (hi == lo) {
            b = n_out / 2;
        } else {
            const double bw = (hi - lo) / n_out;
            b = (int)floor((v - lo) / bw);
            b = std::max(0, std::min(b, n_out - 1));
            while (b > 0 && v < lo + bw * b) --b;
            while (b < n_out - 1 && v >= lo + bw * (b + 1)) ++b;
        }
        cnt[(size_t)b] += 1;
    }
    icdf.resize((size_t)n_out);
    int32_t run = 0;
    for (int b = 0; b < n_out; ++b) {
        run += cnt[(size_t)b];
        icdf[(size_t)b] = run;
    }
}

// An orthonormal basis of the complement of lmean (null(lmean') up to an
// in-plane rotation): 2-D the normal, 3-D two cross products with the axis
// least aligned with lmean.
static void null_basis(const double *lm, int dim, double u[3][2])
{
    for (int j = 0; j < 3; ++j) u[j][0] = u[j][1] = 0;
    if (dim == 2) {
        u[0][0] = -lm[1];
        u[1][0] = lm[0];
        return;
    }
    int a = 0;
    for (int j = 1; j < 3; ++j)
        if (fabs(lm[j]) < fabs(lm[a])) a = j;
    double e[3] = {0, 0, 0};
    e[a] = 1;
    double c1[3] = {lm[1] * e[2] - lm[2] * e[1], lm[2] * e[0] - lm[0] * e[2], lm[0] * e[1] - lm[1] * e[0]};
    const double n1 = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
    for (int j = 0; j < 3; ++j) c1[j] /= n1;
    const double c2[3] = {lm[1] * c1[2] - lm[2] * c1[1], lm[2] * c1[0] - lm[0] * c1[2], lm[0] * c1[1] - lm[1] * c1[0]};
    for (int j = 0; j < 3; ++j) {
        u[j][0] = c1[j];
        u[j][1] = c2[j];
    }
}

int ProblemState::setup_medium(const scmc_problem *pr, const scmc_medium *m, hipStream_t s)
{
    int rc = SCMC_OK;
    const int dim = pr->dim;
    if (!m || !m->sigt) return set_error(SCMC_E_ARG, "missing sigt grid");
    HetMedium &M = hm;
    memset(&M, 0, sizeof M);
    int nvox = 1;
    for (int j = 0; j < 3; ++j) {
        M.nb[j] = 1;
        M.bin[j] = 1.0;
        M.inv_bin[j] = 1.0;
    }
    for (int j = 0; j < dim; ++j) {
        if (!(m->box_bin[j] > 0)) return set_error(SCMC_E_ARG, "box_bin must be positive");
        // findBin.m:4: Nb = round((box_max - box_min) ./ box_bin)
        const int nb = (int)lround((pr->box_max[j] - pr->box_min[j]) / m->box_bin[j]);
        if (nb < 1 || nb != m->nb[j])
            return set_error(SCMC_E_ARG, "sigt grid (%d cells on axis %d) does not match (box_max-box_min)./box_bin (%d)",
                             m->nb[j], j + 1, nb);
        M.nb[j] = nb;
        M.bin[j] = m->box_bin[j];
        M.inv_bin[j] = 1.0 / m->box_bin[j];
        nvox *= nb;
    }
    std::vector<double> sig(m->sigt, m->sigt + nvox);
    double tot = 0, smax = 0;
    for (double v : sig) {
        if (!(v >= 0)) return set_error(SCMC_E_ARG, "sigt must be non-negative");
        tot += v;
        smax = std::max(smax, v);
    }
    if (!(tot > 0)) return set_error(SCMC_E_ARG, "sigt grid is zero everywhere");
    double vol = 1;
    for (int j = 0; j < dim; ++j) vol *= m->box_bin[j];
    M.V = tot * vol;   // MCCovHetro.m:66
    M.smax = smax;
    std::vector<int32_t> icdf;
    M.n_icdf = nvox * 100;   // MCCovHetro.m:65
    cell_icdf(sig, M.n_icdf, icdf);
    if (!dsig.ensure(sizeof(double) * sig.size(), &rc) || !dcell_icdf.ensure(sizeof(int32_t) * icdf.size(), &rc)) return rc;
    if (!hip_ok(hipMemcpyAsync(dsig.p, sig.data(), sizeof(double) * sig.size(), hipMemcpyHostToDevice, s), "sigt", &rc) ||
        !hip_ok(hipMemcpyAsync(dcell_icdf.p, icdf.data(), sizeof(int32_t) * icdf.size(), hipMemcpyHostToDevice, s),
                "cell icdf", &rc))
        return rc;
    M.sigt = (const double *)dsig.p;
    M.icdf = (const int32_t *)dcell_icdf.p;
    std::vector<double> alb;
    if (m->albedo) {
        alb.assign(m->albedo, m->albedo + nvox);
        if (!dalb.ensure(sizeof(double) * alb.size(), &rc) ||
            !hip_ok(hipMemcpyAsync(dalb.p, alb.data(), sizeof(double) * alb.size(), hipMemcpyHostToDevice, s), "albedo",
                    &rc))
            return rc;
        M.albedo = (const double *)dalb.p;
    }
    // expSmpXHetro.m:4-10 and MCCovHetro.m:45-48
    M.has_lrad = m->has_lrad ? 1 : 0;
    if (m->has_lrad) {
        if (!(m->lrad > 0)) return set_error(SCMC_E_ARG, "lrad must be positive");
        if (pr->smp_flg == 1 && pr->nl != 1) return set_error(SCMC_E_ARG, "lrad with smpFlg 1 needs one light");
        M.lrad = m->lrad;
    } else {
        double d2 = 0;
        for (int j = 0; j < dim; ++j) d2 += (pr->box_max[j] - pr->box_min[j]) * (pr->box_max[j] - pr->box_min[j]);
        M.lrad = sqrt(d2);
    }
    double c = 0;
    for (int j = 0; j < dim; ++j) c = std::max(c, std::max(fabs(pr->box_max[j]), fabs(pr->box_min[j])));
    M.pushc = c * dim;
    null_basis(dp.meanl, dim, M.u);
    M.cap = 4 * (M.nb[0] + M.nb[1] + (dim == 3 ? M.nb[2] : 0)) + 16;
    het = true;
    return hip_ok(stream_wait(s), "medium sync", &rc) ? SCMC_OK : rc;
}

} // namespace scmc

using namespace scmc;

extern "C" int scmc_mccov_hetro(const scmc_problem *pr, const scmc_medium *med, int64_t max_itr, int32_t precision,
                                double *ms_re, double *ms_im, double *mm_re, double *mm_im, double *mean_re,
                                double *mean_im, scmc_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!pr) return set_error(SCMC_E_ARG, "null problem");
    // MCCovHetro has no lmean0: meanl is the mean light (MCCovHetro.m:40-41)
    scmc_problem q = *pr;
    q.has_lmean0 = 0;
    // the scalar sigt is not used with a grid; the scalar albedo only without one
    if (!(q.sigt > 0)) q.sigt = 1.0;
    if (med && med->albedo) q.albedo = 1.0;
    int rc = validate_problem(&q, max_itr, true);
    if (rc) return rc;
    Session s;
    if ((rc = s.init(&q, 0, precision, nullptr, nullptr, med))) return rc;
    if ((rc = s.run(0, max_itr))) return rc;
    if ((rc = s.result(max_itr, ms_re, ms_im, mm_re, mm_im))) return rc;
    if ((rc = s.ce.finalize_mean(max_itr, mean_re, mean_im, s.stream))) return rc;
    if (stats) {
        s.fill_stats(stats);
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return SCMC_OK;
}

// MCsampleHetro.m:1-227: the field sample u (Nv x Nl) of a voxel medium.  The
// walks are MCCovHetro's; lmean0 is honoured (MCsampleHetro.m:35-40); there is
// no lrad argument (expSmpXHetro with lrad = [], MCsampleHetro.m:97).
extern "C" int scmc_mcfield_hetro(const scmc_problem *pr, const scmc_medium *med, int64_t max_itr, int32_t precision,
                                  double *u_re, double *u_im, scmc_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!pr) return set_error(SCMC_E_ARG, "null problem");
    if (med && med->has_lrad) return set_error(SCMC_E_ARG, "MCsampleHetro has no lrad argument");
    scmc_problem q = *pr;
    if (!(q.sigt > 0)) q.sigt = 1.0;
    if (med && med->albedo) q.albedo = 1.0;
    int rc = validate_problem(&q, max_itr, true);
    if (rc) return rc;
    Session s;
    if ((rc = s.init(&q, 1, precision, nullptr, nullptr, med))) return rc;
    if ((rc = s.run(0, max_itr))) return rc;
    if ((rc = s.result(max_itr, u_re, u_im, nullptr, nullptr))) return rc;
    if (stats) {
        s.fill_stats(stats);
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return SCMC_OK;
}
