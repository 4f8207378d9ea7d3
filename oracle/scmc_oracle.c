/*
 * scmc_oracle.c -- FP64 CPU restatement of the SCMC Monte Carlo speckle core.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / timed CPU baseline.  The
 * product library never links or calls it.
 *
 * Reference files restated (paths relative to the reference repository):
 *   code/MCcov.m, code/MCfield.m, code/MCfieldFourierQuad.m,
 *   code/coreFunctions/{expSmpX,evalphaseatt,cubeDist,evalampfunc_general,
 *     evaluateHG,evalampfunc_amplitude,smpampfunc_general,sampleHG,smpampfunc,
 *     rotateBytheta,invCDF}.m,
 *   interface/scatter/tabulatedAmplitudeScatter.m.
 *
 * Deviations from MATLAB, all deliberate and shared with the product:
 *   - rand/randn come from Philox4x32-10 with a fixed slot map (below), not
 *     MATLAB's Mersenne twister / ziggurat; randn uses Box-Muller.
 *   - acos arguments are clamped to [-1,1] before the tabulated lookup
 *     (MATLAB returns a complex acos for |c|>1 by one ulp and then fails to
 *     index); Fourier grid indices are clamped to the grid (MATLAB errors on
 *     the measure-zero x == box edge case).
 *
 * RNG slot map.  counter = (vertex, group, path_lo, path_hi), key = seed.
 * Each call yields two 53-bit uniforms (u0, u1):
 *   vertex 0 (prologue): g0 = (u_z, u_x)  g1 = (u_y, d0)  g2 = (d1, d2)  g3 = (d3, -)
 *   vertex n >= 1      : g0 = (xi, u_flight)  g1 = (u_roulette, d0)  g2 = (d1, d2)  g3 = (d3, -)
 * d0..d3 drive a direction sample: Box-Muller pairs (d0,d1), (d2,d3) for the
 * isotropic randn; d0 = inversion uniform and d1 = azimuth/sign uniform for
 * HG and tabulated sampling.
 */
#include "scmc_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* ------------------------------------------------------------------ RNG */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u53(uint32_t a, uint32_t b)
{
    uint64_t v = (((uint64_t)a << 32) | b) >> 11;
    return (double)v * 0x1.0p-53;
}

void orc_uniform_pair(uint64_t seed, uint64_t path, uint32_t vertex, uint32_t group, double out[2])
{
    uint32_t ctr[4] = {vertex, group, (uint32_t)path, (uint32_t)(path >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    out[0] = u53(o[0], o[1]);
    out[1] = u53(o[2], o[3]);
}

/* ------------------------------------------------------ core functions */

static double dot3(const double *a, const double *b, int dim)
{
    /* explicit fused chain; the GPU kernels use the same association */
    if (dim == 3) return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
    return fma(a[1], b[1], a[0] * b[0]);
}

/* cubeDist.m:4-11 */
double orc_cube_dist(const double *x, const double *bmin, const double *bmax, const double *d, int dim)
{
    const double thr = pow(0.1, 7);
    double best = INFINITY;
    for (int j = 0; j < dim; ++j) {
        double ad = fmax(fabs(d[j]), thr);
        double bd = (d[j] <= 0) ? (x[j] - bmin[j]) / ad : (-x[j] + bmax[j]) / ad;
        if (bd < best) best = bd;
    }
    return best;
}

/* evalphaseatt.m:1-20.  Directions (far field) or positions (near field) are
 * sign * d[:, i]. */
void orc_eval_phase_att(const double *x, const double *d, int m, int dim, int is_ff, double sign,
                        double sigt, double lambda, const double *bmin, const double *bmax,
                        double *out_re, double *out_im)
{
    const double k = M_PI * 2.0 / lambda;
    for (int i = 0; i < m; ++i) {
        double dv[3] = {0, 0, 0};
        for (int j = 0; j < dim; ++j) dv[j] = sign * d[i * dim + j];
        double pv, rv, bdv;
        if (is_ff) {
            double nd[3] = {-dv[0], -dv[1], -dv[2]};
            pv = dot3(x, dv, dim);
            rv = 1.0;
            bdv = orc_cube_dist(x, bmin, bmax, nd, dim);
        } else {
            double dd[3] = {0, 0, 0};
            double s = 0;
            for (int j = 0; j < dim; ++j) { dd[j] = x[j] - dv[j]; s += dd[j] * dd[j]; }
            double ndist = sqrt(s);
            double ndir[3] = {0, 0, 0};
            for (int j = 0; j < dim; ++j) ndir[j] = -(dd[j] / ndist);
            pv = ndist;
            rv = 1.0 / pow(ndist + 0.01 * lambda, (dim - 1) / 2.0);
            bdv = fmin(orc_cube_dist(x, bmin, bmax, ndir, dim), ndist);
        }
        double mag = exp(-sigt / 2 * bdv) * rv;
        double ph = k * pv;
        out_re[i] = mag * cos(ph);
        out_im[i] = mag * sin(ph);
    }
}

/* evaluateHG.m:25-39 with cosineFlag = 1 */
double orc_evaluate_hg(double c, double g, int dim)
{
    if (dim == 3) return 1.0 / 4.0 / M_PI * ((1 - g * g) / pow(1 + (g * g - 2 * (g * c)), 1.5));
    return 1.0 / 2.0 / M_PI * ((1 - g * g) / (1 + (g * g - 2 * (g * c))));
}

/* evalampfunc_general.m:5-18, evalampfunc_amplitude.m:5-9 */
void orc_eval_amp(double c, const orc_phase *ph, int dim, double *re, double *im)
{
    switch (ph->type) {
    case 1:
        *re = 1.0 / sqrt(pow(2.0, dim - 1) * M_PI);
        *im = 0;
        return;
    case 2: {
        double cc = c > 1 ? 1 : (c < -1 ? -1 : c);
        double ang = acos(cc);
        int n = ph->n_amp;
        double scale = (dim == 2) ? (n - 1) / (2 * M_PI) : (n - 1) / M_PI;
        long idx = lround(ang * scale);
        if (idx < 0) idx = 0;
        if (idx > n - 1) idx = n - 1;
        *re = ph->amp_re[idx];
        *im = ph->amp_im ? ph->amp_im[idx] : 0.0;
        return;
    }
    default:
        *re = sqrt(orc_evaluate_hg(c, ph->g, dim));
        *im = 0;
        return;
    }
}

static cplx eval_amp_c(double c, const orc_phase *ph, int dim)
{
    double re, im;
    orc_eval_amp(c, ph, dim, &re, &im);
    return CMPLX(re, im);
}

/* sampleHG.m:3-17 */
double orc_sample_hg(double g, int dim, double u)
{
    if (fabs(g) < 0.001) return 2 * u - 1;
    if (dim == 3) {
        double t = (1 - g * g) / (1 - g + 2 * g * u);
        return (1 / (2 * g)) * ((1 + g * g) - t * t);
    }
    double th = 2 * atan((1 - g) / (1 + g) * tan(M_PI / 2 * (1 - 2 * u)));
    return cos(th);
}

/* rotateBytheta.m:4-27 */
void orc_rotate_by_theta(const double *ow, double costheta, int dim, double u, double *w)
{
    double s2 = 1 - costheta * costheta;
    double sintheta = s2 > 0 ? sqrt(s2) : 0.0;
    if (dim == 2) {
        sintheta = ((u > 0.5) * 2 - 1) * sintheta;
        w[0] = costheta * ow[0] + sintheta * ow[1];
        w[1] = -sintheta * ow[0] + costheta * ow[1];
        return;
    }
    double phi = M_PI * 2 * u;
    double sinphi = sin(phi), cosphi = cos(phi);
    if (fabs(ow[2]) > 0.999999999) {
        w[0] = sintheta * cosphi;
        w[1] = sintheta * sinphi;
        w[2] = costheta * ow[2] / fabs(ow[2]);
    } else {
        double temp = sqrt(1 - ow[2] * ow[2]);
        w[0] = sintheta * (ow[0] * ow[2] * cosphi - ow[1] * sinphi) / temp + ow[0] * costheta;
        w[1] = sintheta * (ow[1] * ow[2] * cosphi + ow[0] * sinphi) / temp + ow[1] * costheta;
        w[2] = -sintheta * cosphi * temp + ow[2] * costheta;
    }
}

static void box_muller(double ua, double ub, double *n0, double *n1)
{
    double r = sqrt(-2.0 * log(1.0 - ua));
    double t = 2.0 * M_PI * ub;
    *n0 = r * cos(t);
    *n1 = r * sin(t);
}

/* randn(dim,1) normalised: smpampfunc_general.m:7-8, MCcov.m:148 */
static void iso_dir(const double u[4], int dim, double *w)
{
    double n0, n1, n2, n3;
    box_muller(u[0], u[1], &n0, &n1);
    if (dim == 3) {
        box_muller(u[2], u[3], &n2, &n3);
        double nrm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        w[0] = n0 / nrm; w[1] = n1 / nrm; w[2] = n2 / nrm;
    } else {
        double nrm = sqrt(n0 * n0 + n1 * n1);
        w[0] = n0 / nrm; w[1] = n1 / nrm;
    }
}

/* smpampfunc_general.m:5-20, smpampfunc.m:3-7 */
void orc_sample_dir(const double *ow, const orc_phase *ph, int dim, const double u[4], double *w)
{
    if (ph->type == 1) { iso_dir(u, dim, w); return; }
    double cosa;
    if (ph->type == 2) {
        int n = ph->n_icdf;
        double stp = 1.0 / (n - 1);
        long i = lround(u[0] / stp);
        if (i > n - 1) i = n - 1;
        cosa = cos(ph->icdf[i]);
    } else {
        cosa = orc_sample_hg(ph->g, dim, u[0]);
    }
    orc_rotate_by_theta(ow, cosa, dim, u[1], w);
}

/* expSmpX.m:1-12 ; u = (u_z, u_x, u_y) */
void orc_exp_smp_x(const double *bmin, const double *bmax, const double *lmean, double sigt, int dim,
                   const double u[3], double *x, double *px)
{
    int e = dim - 1;
    double lz = fabs(lmean[e]);
    double sigs = sigt / lz;
    double dmax = bmax[e] - bmin[e];
    double rmax = 1 - exp(-dmax * sigs);
    double d = -log(-u[0] * rmax + 1) / sigs;
    x[0] = u[1] * (bmax[0] - bmin[0]) + bmin[0];
    if (dim == 3) x[1] = u[2] * (bmax[1] - bmin[1]) + bmin[1];
    x[e] = (d + bmin[e]) * (lmean[e] > 0) + (-d + bmax[e]) * (lmean[e] <= 0);
    *px = exp(-sigs * d) / rmax * sigs * dmax;
}

/* ---------------------------------------------------------- path walk */

typedef struct {
    double x1[3], px, w1[3];
    cplx w0p;
} path_head;

typedef struct {
    double x[3], ow[3], weight, xi;
} vtx;

static void mean_dir(const orc_problem *pr, double *unmean, double *meanl)
{
    int dim = pr->dim;
    for (int j = 0; j < 3; ++j) unmean[j] = 0;
    if (pr->has_lmean0) {
        for (int j = 0; j < dim; ++j) unmean[j] = pr->lmean0[j];
    } else {
        for (int i = 0; i < pr->nl; ++i)
            for (int j = 0; j < dim; ++j) unmean[j] += pr->l[i * dim + j];
        for (int j = 0; j < dim; ++j) unmean[j] /= pr->nl;
    }
    double s = 0;
    for (int j = 0; j < dim; ++j) s += unmean[j] * unmean[j];
    double nrm = sqrt(s);
    for (int j = 0; j < 3; ++j) meanl[j] = j < dim ? unmean[j] / nrm : 0;
}

static int outside(const double *x, const double *bmin, const double *bmax, int dim)
{
    for (int j = 0; j < dim; ++j)
        if (x[j] > bmax[j] || x[j] < bmin[j]) return 1;
    return 0;
}

/* One path: MCcov.m:114-153 (first vertex and direction) and the vertex loop
 * MCcov.m:199-289 without the contributions.  Calls emit for each vertex;
 * returns the vertex count. buf grows as needed. */
static int64_t walk(const orc_problem *pr, uint64_t path, const double *unmean, const double *meanl,
                    path_head *h, vtx **buf, int64_t *cap)
{
    const int dim = pr->dim;
    const double kill_thr = 0.2;
    double g0[2], g1[2], g2[2], g3[2];
    orc_uniform_pair(pr->seed, path, 0, 0, g0);
    orc_uniform_pair(pr->seed, path, 0, 1, g1);
    orc_uniform_pair(pr->seed, path, 0, 2, g2);
    orc_uniform_pair(pr->seed, path, 0, 3, g3);
    double x[3] = {0, 0, 0}, px;
    if (pr->smp_flg == 1) {
        /* MCcov.m:121 */
        x[0] = g0[1] * (pr->box_max[0] - pr->box_min[0]) + pr->box_min[0];
        if (dim == 3) x[1] = g1[0] * (pr->box_max[1] - pr->box_min[1]) + pr->box_min[1];
        x[dim - 1] = g0[0] * (pr->box_max[dim - 1] - pr->box_min[dim - 1]) + pr->box_min[dim - 1];
        px = 1;
    } else {
        double u[3] = {g0[0], g0[1], g1[0]};
        orc_exp_smp_x(pr->box_min, pr->box_max, unmean, pr->sigt, dim, u, x, &px);
    }
    double du[4] = {g1[1], g2[0], g2[1], g3[0]};
    double w[3] = {0, 0, 0};
    cplx w0p;
    if (!pr->has_amp0) {
        iso_dir(du, dim, w);
        w0p = 1.0 / sqrt(pow(2.0, dim - 1) * M_PI);
    } else {
        orc_phase p0 = pr->amp0;
        p0.type = pr->amp.type;
        orc_sample_dir(meanl, &p0, dim, du, w);
        w0p = eval_amp_c(dot3(meanl, w, dim), &p0, dim);
    }
    memcpy(h->x1, x, sizeof x);
    h->px = px;
    memcpy(h->w1, w, sizeof w);
    h->w0p = w0p;

    double weight = pr->albedo;
    double ow[3] = {0, 0, 0};
    int64_t n = 0;
    for (;;) {
        ++n;
        double a[2];
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 0, a);
        if (n > *cap) {
            *cap = *cap ? *cap * 2 : 64;
            *buf = (vtx *)realloc(*buf, sizeof(vtx) * (size_t)*cap);
        }
        vtx *vv = &(*buf)[n - 1];
        memcpy(vv->x, x, sizeof x);
        memcpy(vv->ow, ow, sizeof ow);
        vv->weight = weight;
        vv->xi = a[0];
        /* MCcov.m:267-268 */
        double d = -log(-a[1] + 1) / pr->sigt;
        for (int j = 0; j < dim; ++j) x[j] = x[j] + d * w[j];
        /* MCcov.m:272-274 */
        if (outside(x, pr->box_min, pr->box_max, dim)) break;
        double b[2];
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 1, b);
        /* MCcov.m:278-284 */
        if (weight < kill_thr) {
            if (b[0] > pr->albedo) break;
        } else {
            weight = weight * pr->albedo;
        }
        /* MCcov.m:287-288 */
        double c[2], e[2];
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 2, c);
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 3, e);
        double u4[4] = {b[1], c[0], c[1], e[0]};
        memcpy(ow, w, sizeof w);
        orc_sample_dir(ow, &pr->amp, dim, u4, w);
    }
    return n;
}

int64_t orc_trace_count(const orc_problem *pr, uint64_t p0, int64_t np, int64_t *counts)
{
    double unmean[3], meanl[3];
    mean_dir(pr, unmean, meanl);
    vtx *buf = NULL;
    int64_t cap = 0, total = 0;
    path_head h;
    for (int64_t i = 0; i < np; ++i) {
        int64_t n = walk(pr, p0 + (uint64_t)i, unmean, meanl, &h, &buf, &cap);
        counts[i] = n;
        total += n;
    }
    free(buf);
    return total;
}

void orc_trace_fill(const orc_problem *pr, uint64_t p0, int64_t np, orc_pathset *ps)
{
    double unmean[3], meanl[3];
    mean_dir(pr, unmean, meanl);
    vtx *buf = NULL;
    int64_t cap = 0;
    path_head h;
    for (int64_t i = 0; i < np; ++i) {
        int64_t n = walk(pr, p0 + (uint64_t)i, unmean, meanl, &h, &buf, &cap);
        int64_t o = ps->offset[i];
        for (int j = 0; j < 3; ++j) {
            ps->x1[3 * i + j] = h.x1[j];
            ps->w1[3 * i + j] = h.w1[j];
        }
        ps->px[i] = h.px;
        ps->w0p_re[i] = creal(h.w0p);
        ps->w0p_im[i] = cimag(h.w0p);
        for (int64_t k = 0; k < n; ++k) {
            for (int j = 0; j < 3; ++j) {
                ps->x[3 * (o + k) + j] = buf[k].x[j];
                ps->ow[3 * (o + k) + j] = buf[k].ow[j];
            }
            ps->weight[o + k] = buf[k].weight;
            ps->xi[o + k] = buf[k].xi;
        }
    }
    free(buf);
}

/* ----------------------------------------------- per-path NEE factors */

typedef struct {
    int nv, nl, dim;
    cplx *e_l0, *e_l0_ms, *e_v0, *e_v0_ms, *af_vl;   /* path level            */
    cplx *e_v_ms, *e_l_ms;                           /* vertex level          */
    double *rv, *rl;                                 /* entrance directions  */
    double *tmp_re, *tmp_im;
    const orc_medium *med;                           /* voxel medium or NULL */
} nee_ws;

static void nee_alloc(nee_ws *ws, int nv, int nl, int dim)
{
    ws->med = NULL;
    ws->nv = nv; ws->nl = nl; ws->dim = dim;
    ws->e_l0 = calloc((size_t)nl, sizeof(cplx));
    ws->e_l0_ms = calloc((size_t)nl, sizeof(cplx));
    ws->e_l_ms = calloc((size_t)nl, sizeof(cplx));
    ws->e_v0 = calloc((size_t)nv, sizeof(cplx));
    ws->e_v0_ms = calloc((size_t)nv, sizeof(cplx));
    ws->e_v_ms = calloc((size_t)nv, sizeof(cplx));
    ws->af_vl = calloc((size_t)nv * nl, sizeof(cplx));
    ws->rv = calloc((size_t)nv * 3, sizeof(double));
    ws->rl = calloc((size_t)nl * 3, sizeof(double));
    int m = nv > nl ? nv : nl;
    ws->tmp_re = calloc((size_t)m, sizeof(double));
    ws->tmp_im = calloc((size_t)m, sizeof(double));
}

static void nee_free(nee_ws *ws)
{
    free(ws->e_l0); free(ws->e_l0_ms); free(ws->e_l_ms); free(ws->e_v0); free(ws->e_v0_ms);
    free(ws->e_v_ms); free(ws->af_vl); free(ws->rv); free(ws->rl); free(ws->tmp_re); free(ws->tmp_im);
}

static void phase_att_c(const orc_problem *pr, const double *x, const double *d, int m, int is_ff,
                        double sign, nee_ws *ws, cplx *out)
{
    if (ws->med)
        orc_eval_phase_att_hetro(pr, ws->med, x, d, m, is_ff, sign, ws->tmp_re, ws->tmp_im);
    else
        orc_eval_phase_att(x, d, m, pr->dim, is_ff, sign, pr->sigt, pr->lambda, pr->box_min, pr->box_max,
                           ws->tmp_re, ws->tmp_im);
    for (int i = 0; i < m; ++i) out[i] = CMPLX(ws->tmp_re[i], ws->tmp_im[i]);
}

/* MCcov.m:127-179: path-level quantities at the first vertex. */
static void nee_path(const orc_problem *pr, const double *x1, const double *w1, cplx w0p, nee_ws *ws)
{
    const int dim = pr->dim, nv = pr->nv, nl = pr->nl;
    const double ff_sign = pr->is_ff_v ? -1.0 : 1.0;
    /* MCcov.m:99-105, 128-135 */
    for (int i = 0; i < nv; ++i) {
        double s = 0, r[3] = {0, 0, 0};
        for (int j = 0; j < dim; ++j) {
            r[j] = pr->is_ff_v ? pr->v[i * dim + j] : pr->v[i * dim + j] - x1[j];
            s += r[j] * r[j];
        }
        double nrm = pr->is_ff_v ? 1.0 : sqrt(s);
        for (int j = 0; j < dim; ++j) ws->rv[i * dim + j] = r[j] / nrm;
    }
    for (int i = 0; i < nl; ++i) {
        double s = 0, r[3] = {0, 0, 0};
        for (int j = 0; j < dim; ++j) {
            r[j] = pr->is_ff_l ? pr->l[i * dim + j] : x1[j] - pr->l[i * dim + j];
            s += r[j] * r[j];
        }
        double nrm = pr->is_ff_l ? 1.0 : sqrt(s);
        for (int j = 0; j < dim; ++j) ws->rl[i * dim + j] = r[j] / nrm;
    }
    /* MCcov.m:91-96, 138-142: single-scattering amplitude a(rl_j . rv_i) */
    for (int j = 0; j < nl; ++j)
        for (int i = 0; i < nv; ++i)
            ws->af_vl[i + (size_t)nv * j] = eval_amp_c(dot3(&ws->rl[j * dim], &ws->rv[i * dim], dim), &pr->amp, dim);
    /* MCcov.m:157-167 */
    phase_att_c(pr, x1, pr->l, nl, pr->is_ff_l, 1.0, ws, ws->e_l0);
    for (int j = 0; j < nl; ++j) {
        cplx af = eval_amp_c(dot3(w1, &ws->rl[j * dim], dim), &pr->amp, dim) / w0p;
        ws->e_l0_ms[j] = ws->e_l0[j] * af;
    }
    /* MCcov.m:171-179 */
    phase_att_c(pr, x1, pr->v, nv, pr->is_ff_v, ff_sign, ws, ws->e_v0);
    if (pr->do_cbs) {
        double mw[3] = {-w1[0], -w1[1], -w1[2]};
        for (int i = 0; i < nv; ++i) {
            cplx af = eval_amp_c(dot3(mw, &ws->rv[i * dim], dim), &pr->amp, dim) / w0p;
            ws->e_v0_ms[i] = ws->e_v0[i] * af;
        }
    }
}

/* MCcov.m:205-219: vertex-level quantities for pL > 1. */
static void nee_vertex(const orc_problem *pr, const double *x, const double *ow, nee_ws *ws)
{
    const int dim = pr->dim, nv = pr->nv, nl = pr->nl;
    const double ff_sign = pr->is_ff_v ? -1.0 : 1.0;
    phase_att_c(pr, x, pr->v, nv, pr->is_ff_v, ff_sign, ws, ws->e_v_ms);
    for (int i = 0; i < nv; ++i)
        ws->e_v_ms[i] *= eval_amp_c(dot3(ow, &ws->rv[i * dim], dim), &pr->amp, dim);
    if (pr->do_cbs) {
        double mow[3] = {-ow[0], -ow[1], -ow[2]};
        phase_att_c(pr, x, pr->l, nl, pr->is_ff_l, 1.0, ws, ws->e_l_ms);
        for (int j = 0; j < nl; ++j)
            ws->e_l_ms[j] *= eval_amp_c(dot3(mow, &ws->rl[j * dim], dim), &pr->amp, dim);
    }
}

/* ------------------------------------------------------- MCcov proper */

typedef struct {
    cplx *m00, *mnn, *mn0, *m0n;   /* nv x nv */
} mcov_ws;

static void outer(const cplx *a, const cplx *b, int n, cplx *m)
{
    /* m = a * b' (MATLAB conjugate transpose), column-major */
    for (int j = 0; j < n; ++j) {
        cplx cb = conj(b[j]);
        for (int i = 0; i < n; ++i) m[i + (size_t)n * j] = a[i] * cb;
    }
}

/* Accumulate one path (head + vertices) into unnormalised Ms, Mm following
 * MCcov.m:181-264 term by term. */
static void mccov_path(const orc_problem *pr, const path_head *h, const vtx *vs, int64_t nvtx,
                       nee_ws *ws, mcov_ws *mw, cplx *Ms, cplx *Mm, cplx *mean1)
{
    const int nv = pr->nv, nl = pr->nl;
    const size_t nv2 = (size_t)nv * nv;
    nee_path(pr, h->x1, h->w1, h->w0p, ws);
    if (mean1 && nvtx > 0) {
        /* MCCovHetro.m:152, 210-211: mean1 += weight/px af_ang_vl .* (e_v0 * u_l0.'),
         * u_l0 = evalphaseatt with sigt = 0 */
        double *ure = malloc(sizeof(double) * (size_t)nl), *uim = malloc(sizeof(double) * (size_t)nl);
        orc_eval_phase_att(h->x1, pr->l, nl, pr->dim, pr->is_ff_l, 1.0, 0.0, pr->lambda, pr->box_min, pr->box_max,
                           ure, uim);
        const double wp = vs[0].weight / h->px;
        for (int j = 0; j < nl; ++j)
            for (int i = 0; i < nv; ++i)
                mean1[i + (size_t)nv * j] +=
                    wp * ws->af_vl[i + (size_t)nv * j] * (ws->e_v0[i] * CMPLX(ure[j], uim[j]));
        free(ure);
        free(uim);
    }
    /* MCcov.m:190 */
    outer(ws->e_v0, ws->e_v0, nv, mw->m00);
    for (int64_t n = 1; n <= nvtx; ++n) {
        const vtx *vv = &vs[n - 1];
        if (n > 1) {
            nee_vertex(pr, vv->x, vv->ow, ws);
            if (pr->do_cbs) {
                outer(ws->e_v_ms, ws->e_v0_ms, nv, mw->mn0);
                outer(ws->e_v0_ms, ws->e_v_ms, nv, mw->m0n);
            }
            outer(ws->e_v_ms, ws->e_v_ms, nv, mw->mnn);
        }
        if (n == 2 && pr->do_cbs) outer(ws->e_v0_ms, ws->e_v0_ms, nv, mw->m00);
        const double wp = vv->weight / h->px;
        if (n == 1) {
            /* MCcov.m:234-241 */
            for (int j1 = 0; j1 < nl; ++j1)
                for (int j2 = 0; j2 < nl; ++j2) {
                    cplx s = conj(ws->e_l0[j2]) * ws->e_l0[j1];
                    cplx *dst = Ms + nv2 * ((size_t)j1 + (size_t)nl * j2);
                    const cplx *a1 = ws->af_vl + (size_t)nv * j1;
                    const cplx *a2 = ws->af_vl + (size_t)nv * j2;
                    for (int v2 = 0; v2 < nv; ++v2) {
                        cplx ca2 = conj(a2[v2]);
                        for (int v1 = 0; v1 < nv; ++v1)
                            dst[v1 + (size_t)nv * v2] += (wp * mw->m00[v1 + (size_t)nv * v2] * s) * (a1[v1] * ca2);
                    }
                }
        } else if (pr->do_cbs) {
            /* MCcov.m:245-254 */
            for (int j1 = 0; j1 < nl; ++j1)
                for (int j2 = 0; j2 < nl; ++j2) {
                    cplx s00 = conj(ws->e_l_ms[j2]) * ws->e_l_ms[j1];
                    cplx snn = conj(ws->e_l0_ms[j2]) * ws->e_l0_ms[j1];
                    cplx sn0 = conj(ws->e_l_ms[j2]) * ws->e_l0_ms[j1];
                    cplx s0n = conj(ws->e_l0_ms[j2]) * ws->e_l_ms[j1];
                    cplx *dst = Mm + nv2 * ((size_t)j1 + (size_t)nl * j2);
                    for (size_t e = 0; e < nv2; ++e)
                        dst[e] += wp * (mw->m00[e] * s00 + mw->mnn[e] * snn + mw->mn0[e] * sn0 + mw->m0n[e] * s0n);
                }
        } else {
            /* MCcov.m:256-262 */
            for (int j1 = 0; j1 < nl; ++j1)
                for (int j2 = 0; j2 < nl; ++j2) {
                    cplx snn = conj(ws->e_l0_ms[j2]) * ws->e_l0_ms[j1];
                    cplx *dst = Mm + nv2 * ((size_t)j1 + (size_t)nl * j2);
                    for (size_t e = 0; e < nv2; ++e) dst[e] += wp * (mw->mnn[e] * snn);
                }
        }
    }
}

static double box_volume(const orc_problem *pr)
{
    double V = 1;
    for (int j = 0; j < pr->dim; ++j) V *= pr->box_max[j] - pr->box_min[j];
    return V;
}

static void mccov_range(const orc_problem *pr, uint64_t p0, int64_t np, cplx *Ms, cplx *Mm)
{
    double unmean[3], meanl[3];
    mean_dir(pr, unmean, meanl);
    nee_ws ws;
    nee_alloc(&ws, pr->nv, pr->nl, pr->dim);
    size_t nv2 = (size_t)pr->nv * pr->nv;
    mcov_ws mw = {calloc(nv2, sizeof(cplx)), calloc(nv2, sizeof(cplx)), calloc(nv2, sizeof(cplx)),
                  calloc(nv2, sizeof(cplx))};
    vtx *buf = NULL;
    int64_t cap = 0;
    path_head h;
    for (int64_t i = 0; i < np; ++i) {
        int64_t n = walk(pr, p0 + (uint64_t)i, unmean, meanl, &h, &buf, &cap);
        mccov_path(pr, &h, buf, n, &ws, &mw, Ms, Mm, NULL);
    }
    free(buf);
    free(mw.m00); free(mw.mnn); free(mw.mn0); free(mw.m0n);
    nee_free(&ws);
}

static void split_out(const cplx *src, size_t n, double scale, double *re, double *im)
{
    for (size_t i = 0; i < n; ++i) { re[i] = creal(src[i]) * scale; im[i] = cimag(src[i]) * scale; }
}

/* Unnormalised sums over paths [p0, p0 + np) into Ms, Mm (zeroed), over
 * nthreads OpenMP path shards each with its own accumulators. */
static void mccov_sums_into(const orc_problem *pr, uint64_t p0, int64_t np, int nthreads, cplx *Ms, cplx *Mm)
{
    size_t ne = (size_t)pr->nv * pr->nv * pr->nl * pr->nl;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel num_threads(nthreads)
        {
            int t = omp_get_thread_num(), T = omp_get_num_threads();
            int64_t a = np * t / T, b = np * (t + 1) / T;
            cplx *ls = calloc(ne, sizeof(cplx)), *lm = calloc(ne, sizeof(cplx));
            mccov_range(pr, p0 + (uint64_t)a, b - a, ls, lm);
#pragma omp critical
            for (size_t e = 0; e < ne; ++e) { Ms[e] += ls[e]; Mm[e] += lm[e]; }
            free(ls); free(lm);
        }
    } else
#endif
        mccov_range(pr, p0, np, Ms, Mm);
    (void)ne;
}

void orc_mccov_sums(const orc_problem *pr, uint64_t p0, int64_t np, int nthreads,
                    double *ms_re, double *ms_im, double *mm_re, double *mm_im)
{
    size_t ne = (size_t)pr->nv * pr->nv * pr->nl * pr->nl;
    cplx *Ms = calloc(ne, sizeof(cplx)), *Mm = calloc(ne, sizeof(cplx));
    mccov_sums_into(pr, p0, np, nthreads, Ms, Mm);
    split_out(Ms, ne, 1.0, ms_re, ms_im);
    split_out(Mm, ne, 1.0, mm_re, mm_im);
    free(Ms); free(Mm);
}

/* MCcov.m:293-299 */
static void mccov_normalise(const orc_problem *pr, int64_t max_itr, cplx *Ms, cplx *Mm, size_t ne,
                            double *ms_re, double *ms_im, double *mm_re, double *mm_im)
{
    double V = box_volume(pr);
    for (size_t e = 0; e < ne; ++e) {
        cplx m = Mm[e];
        if (pr->do_cbs) m = m / 2;
        m = m / (double)max_itr * V * pr->sigt;
        mm_re[e] = creal(m); mm_im[e] = cimag(m);
        cplx s = Ms[e] / (double)max_itr * V * pr->sigt;
        ms_re[e] = creal(s); ms_im[e] = cimag(s);
    }
}

void orc_mccov(const orc_problem *pr, int64_t max_itr,
               double *ms_re, double *ms_im, double *mm_re, double *mm_im)
{
    size_t ne = (size_t)pr->nv * pr->nv * pr->nl * pr->nl;
    cplx *Ms = calloc(ne, sizeof(cplx)), *Mm = calloc(ne, sizeof(cplx));
    mccov_range(pr, 0, max_itr, Ms, Mm);
    mccov_normalise(pr, max_itr, Ms, Mm, ne, ms_re, ms_im, mm_re, mm_im);
    free(Ms); free(Mm);
}

/* orc_mccov with the paths summed over nthreads shards (the sums of the
 * shards are added in a thread-dependent order, so results agree with
 * orc_mccov to rounding). */
void orc_mccov_mt(const orc_problem *pr, int64_t max_itr, int nthreads,
                  double *ms_re, double *ms_im, double *mm_re, double *mm_im)
{
    size_t ne = (size_t)pr->nv * pr->nv * pr->nl * pr->nl;
    cplx *Ms = calloc(ne, sizeof(cplx)), *Mm = calloc(ne, sizeof(cplx));
    mccov_sums_into(pr, 0, max_itr, nthreads, Ms, Mm);
    mccov_normalise(pr, max_itr, Ms, Mm, ne, ms_re, ms_im, mm_re, mm_im);
    free(Ms); free(Mm);
}

static void load_path(const orc_pathset *ps, int64_t i, path_head *h, vtx **buf, int64_t *cap, int64_t *n)
{
    int64_t o = ps->offset[i], e = ps->offset[i + 1];
    *n = e - o;
    if (*n > *cap) { *cap = *n; *buf = realloc(*buf, sizeof(vtx) * (size_t)*cap); }
    for (int j = 0; j < 3; ++j) { h->x1[j] = ps->x1[3 * i + j]; h->w1[j] = ps->w1[3 * i + j]; }
    h->px = ps->px[i];
    h->w0p = CMPLX(ps->w0p_re[i], ps->w0p_im[i]);
    for (int64_t k = 0; k < *n; ++k) {
        for (int j = 0; j < 3; ++j) { (*buf)[k].x[j] = ps->x[3 * (o + k) + j]; (*buf)[k].ow[j] = ps->ow[3 * (o + k) + j]; }
        (*buf)[k].weight = ps->weight[o + k];
        (*buf)[k].xi = ps->xi[o + k];
    }
}

void orc_mccov_replay(const orc_problem *pr, const orc_pathset *ps, int64_t max_itr,
                      double *ms_re, double *ms_im, double *mm_re, double *mm_im)
{
    size_t ne = (size_t)pr->nv * pr->nv * pr->nl * pr->nl;
    cplx *Ms = calloc(ne, sizeof(cplx)), *Mm = calloc(ne, sizeof(cplx));
    nee_ws ws;
    nee_alloc(&ws, pr->nv, pr->nl, pr->dim);
    size_t nv2 = (size_t)pr->nv * pr->nv;
    mcov_ws mw = {calloc(nv2, sizeof(cplx)), calloc(nv2, sizeof(cplx)), calloc(nv2, sizeof(cplx)),
                  calloc(nv2, sizeof(cplx))};
    vtx *buf = NULL;
    int64_t cap = 0, n;
    path_head h;
    for (int64_t i = 0; i < ps->n_paths; ++i) {
        load_path(ps, i, &h, &buf, &cap, &n);
        mccov_path(pr, &h, buf, n, &ws, &mw, Ms, Mm, NULL);
    }
    mccov_normalise(pr, max_itr, Ms, Mm, ne, ms_re, ms_im, mm_re, mm_im);
    free(buf);
    free(mw.m00); free(mw.mnn); free(mw.mn0); free(mw.m0n);
    nee_free(&ws);
    free(Ms); free(Mm);
}

/* ------------------------------------------------------- MCfield proper */

/* MCfield.m:190-222 for one path, unnormalised sum into u (nv x nl). */
static void mcfield_path(const orc_problem *pr, const path_head *h, const vtx *vs, int64_t nvtx,
                         nee_ws *ws, cplx *u)
{
    const int nv = pr->nv, nl = pr->nl;
    nee_path(pr, h->x1, h->w1, h->w0p, ws);
    for (int64_t n = 1; n <= nvtx; ++n) {
        const vtx *vv = &vs[n - 1];
        if (n > 1) nee_vertex(pr, vv->x, vv->ow, ws);
        /* MCfield.m:218-220 */
        double amp = sqrt(vv->weight / h->px);
        cplx ph = cexp(CMPLX(0, 2 * M_PI * vv->xi));
        for (int j = 0; j < nl; ++j)
            for (int i = 0; i < nv; ++i) {
                cplx t;
                if (n == 1) {
                    /* MCfield.m:209: af_ang_vl .* (e_v0 * e_l0.') */
                    t = ws->af_vl[i + (size_t)nv * j] * (ws->e_v0[i] * ws->e_l0[j]);
                } else {
                    /* MCfield.m:211-214 */
                    t = ws->e_v_ms[i] * ws->e_l0_ms[j];
                    if (pr->do_cbs) t = 1 / sqrt(2.0) * (t + ws->e_v0_ms[i] * ws->e_l_ms[j]);
                }
                u[i + (size_t)nv * j] += (amp * t) * ph;
            }
    }
}

static void mcfield_range(const orc_problem *pr, uint64_t p0, int64_t np, cplx *u)
{
    double unmean[3], meanl[3];
    mean_dir(pr, unmean, meanl);
    nee_ws ws;
    nee_alloc(&ws, pr->nv, pr->nl, pr->dim);
    vtx *buf = NULL;
    int64_t cap = 0;
    path_head h;
    for (int64_t i = 0; i < np; ++i) {
        int64_t n = walk(pr, p0 + (uint64_t)i, unmean, meanl, &h, &buf, &cap);
        mcfield_path(pr, &h, buf, n, &ws, u);
    }
    free(buf);
    nee_free(&ws);
}

void orc_mcfield_sums(const orc_problem *pr, uint64_t p0, int64_t np, int nthreads, double *u_re, double *u_im)
{
    size_t ne = (size_t)pr->nv * pr->nl;
    cplx *u = calloc(ne, sizeof(cplx));
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel num_threads(nthreads)
        {
            int t = omp_get_thread_num(), T = omp_get_num_threads();
            int64_t a = np * t / T, b = np * (t + 1) / T;
            cplx *lu = calloc(ne, sizeof(cplx));
            mcfield_range(pr, p0 + (uint64_t)a, b - a, lu);
#pragma omp critical
            for (size_t e = 0; e < ne; ++e) u[e] += lu[e];
            free(lu);
        }
    } else
#endif
        mcfield_range(pr, p0, np, u);
    split_out(u, ne, 1.0, u_re, u_im);
    free(u);
}

void orc_mcfield(const orc_problem *pr, int64_t max_itr, double *u_re, double *u_im)
{
    size_t ne = (size_t)pr->nv * pr->nl;
    cplx *u = calloc(ne, sizeof(cplx));
    mcfield_range(pr, 0, max_itr, u);
    /* MCfield.m:253-254 */
    double s = sqrt(1.0 / (double)max_itr * box_volume(pr) * pr->sigt);
    split_out(u, ne, s, u_re, u_im);
    free(u);
}

/* code/MCfieldOnWave.m:113-267 for one path: the lights are combined with
 * the weights Wl before the outer product, so u is nv x 1.  The single-
 * scattering term is e_v0 .* (af_ang_vl * (e_l0 .* Wl)) (MCfieldOnWave.m:214,
 * written here as the sum over lights for every scattering type; the
 * reference's isotropic branch reduces along the wrong dimension unless
 * Nl = 1), e_l0_ms and e_l_ms are sums over lights (:177, :199). */
static void mcfield_wave_path(const orc_problem *pr, const cplx *wl, const path_head *h, const vtx *vs,
                              int64_t nvtx, nee_ws *ws, cplx *u)
{
    const int nv = pr->nv, nl = pr->nl;
    nee_path(pr, h->x1, h->w1, h->w0p, ws);
    cplx s0 = 0;   /* sum(e_l0 .* af_l .* Wl) */
    for (int j = 0; j < nl; ++j) s0 += ws->e_l0_ms[j] * wl[j];
    for (int64_t n = 1; n <= nvtx; ++n) {
        const vtx *vv = &vs[n - 1];
        cplx sl = 0;
        if (n > 1) {
            nee_vertex(pr, vv->x, vv->ow, ws);
            if (pr->do_cbs)
                for (int j = 0; j < nl; ++j) sl += ws->e_l_ms[j] * wl[j];
        }
        double amp = sqrt(vv->weight / h->px);
        cplx ph = cexp(CMPLX(0, 2 * M_PI * vv->xi));
        for (int i = 0; i < nv; ++i) {
            cplx t;
            if (n == 1) {
                cplx acc = 0;
                for (int j = 0; j < nl; ++j) acc += ws->af_vl[i + (size_t)nv * j] * (ws->e_l0[j] * wl[j]);
                t = ws->e_v0[i] * acc;
            } else {
                t = ws->e_v_ms[i] * s0;
                if (pr->do_cbs) t = 1 / sqrt(2.0) * (t + ws->e_v0_ms[i] * sl);
            }
            u[i] += (amp * t) * ph;
        }
    }
}

void orc_mcfield_wave(const orc_problem *pr, const double *wl_re, const double *wl_im, int64_t max_itr,
                      double *u_re, double *u_im)
{
    double unmean[3], meanl[3];
    mean_dir(pr, unmean, meanl);
    cplx *wl = malloc(sizeof(cplx) * (size_t)pr->nl);
    for (int j = 0; j < pr->nl; ++j) wl[j] = CMPLX(wl_re[j], wl_im ? wl_im[j] : 0.0);
    cplx *u = calloc((size_t)pr->nv, sizeof(cplx));
    nee_ws ws;
    nee_alloc(&ws, pr->nv, pr->nl, pr->dim);
    vtx *buf = NULL;
    int64_t cap = 0;
    path_head h;
    for (int64_t i = 0; i < max_itr; ++i) {
        int64_t n = walk(pr, (uint64_t)i, unmean, meanl, &h, &buf, &cap);
        mcfield_wave_path(pr, wl, &h, buf, n, &ws, u);
    }
    free(buf);
    nee_free(&ws);
    /* MCfieldOnWave.m:265-267 */
    double s = sqrt(1.0 / (double)max_itr * box_volume(pr) * pr->sigt);
    split_out(u, (size_t)pr->nv, s, u_re, u_im);
    free(u);
    free(wl);
}

void orc_mcfield_replay(const orc_problem *pr, const orc_pathset *ps, int64_t max_itr, double *u_re, double *u_im)
{
    size_t ne = (size_t)pr->nv * pr->nl;
    cplx *u = calloc(ne, sizeof(cplx));
    nee_ws ws;
    nee_alloc(&ws, pr->nv, pr->nl, pr->dim);
    vtx *buf = NULL;
    int64_t cap = 0, n;
    path_head h;
    for (int64_t i = 0; i < ps->n_paths; ++i) {
        load_path(ps, i, &h, &buf, &cap, &n);
        mcfield_path(pr, &h, buf, n, &ws, u);
    }
    double s = sqrt(1.0 / (double)max_itr * box_volume(pr) * pr->sigt);
    split_out(u, ne, s, u_re, u_im);
    free(buf);
    nee_free(&ws);
    free(u);
}

/* --------------------------------------------------- tabulated tables */

/* invCDF.m:3-17 */
void orc_inv_cdf(const double *cdf, int n, int dim, double *icdf)
{
    double stp = 1.0 / (n - 1);
    for (int j = 0; j < n; ++j) {
        double g = j * stp;
        int best = 0;
        double bv = fabs(cdf[0] - g);
        for (int i = 1; i < n; ++i) {
            double d = fabs(cdf[i] - g);
            if (d < bv) { bv = d; best = i; }
        }
        double idx1 = best + 1; /* MATLAB 1-based index */
        icdf[j] = (dim == 2) ? idx1 / n * 2 * M_PI : idx1 / n * M_PI;
    }
}

/* tabulatedAmplitudeScatter.m:53-99 */
int orc_tabulated_scatter(const double *theta, const double *f_re, const double *f_im, int n,
                          double *cs, double *sample_pdf, double *sample_cdf, double *sample_icdf,
                          double *eval_pdf, double *eval_amp_re, double *eval_amp_im)
{
    for (int i = 1; i < n; ++i)
        if (!(theta[i] >= theta[i - 1])) return -1;
    int D;
    if (theta[0] == 0 && theta[n - 1] == 2 * M_PI) D = 2;
    else if (theta[0] == 0 && theta[n - 1] == M_PI) D = 3;
    else return -2;
    double s = 0, sp = 0;
    for (int i = 0; i < n; ++i) {
        double a = hypot(f_re[i], f_im[i]);
        double a2 = a * a;
        double w = (D == 3) ? a2 * sin(theta[i]) : a2;
        s += w;
        sample_pdf[i] = w;
        sp += (D == 3) ? a2 * sin(theta[i]) : a2;
    }
    *cs = (D == 3) ? sp / n * 2 * M_PI * M_PI : sp / n * 2 * M_PI;
    double c = 0;
    for (int i = 0; i < n; ++i) {
        sample_pdf[i] /= s;
        c += sample_pdf[i];
        sample_cdf[i] = c;
    }
    orc_inv_cdf(sample_cdf, n, D, sample_icdf);
    for (int i = 0; i < n; ++i) {
        double a = hypot(f_re[i], f_im[i]);
        eval_pdf[i] = a * a / *cs;
        double ang = atan2(f_im[i], f_re[i]);
        double m = sqrt(eval_pdf[i]);
        eval_amp_re[i] = m * cos(ang);
        eval_amp_im[i] = m * sin(ang);
    }
    return D;
}

/* ------------------------------------------------ MCfieldFourierQuad */

static int colon_len(double a, double s, double b)
{
    /* MATLAB a:s:b element count (with MATLAB's relative tolerance) */
    double q = (b - a) / s;
    return (int)floor(q + q * 1e-12 + 1e-10) + 1;
}

static int quad_grid_setup(const double *box_min, const double *box_max, double v_max, double v_stp,
                           double lambda, int parity_fix, orc_quad_grid *g)
{
    /* MCfieldFourierQuad.m:80-121 (MCfieldFourierQuadOnWave.m:84-131) */
    g->nv0 = colon_len(-v_max, v_stp, v_max);
    g->nv = g->nv0;
    g->box_stp = lambda / (v_max * 2 + v_stp);
    double bw0 = box_max[0] - box_min[0];
    g->nx = (int)ceil(bw0 / g->box_stp / 2) * 2 + 1;
    /* MCfieldFourierQuadOnWave.m:100-102: same parity as the view grid */
    if (parity_fix && (g->nv - g->nx) % 2 != 0) g->nx += 1;
    double bwn = g->nx * g->box_stp;
    for (int j = 0; j < 3; ++j) { g->box_min[j] = box_min[j]; g->box_max[j] = box_max[j]; }
    for (int j = 0; j < 2; ++j) {
        g->box_min[j] = box_min[j] - (bwn - bw0) / 2;
        g->box_max[j] = box_max[j] + (bwn - bw0) / 2;
    }
    g->gx0 = g->box_min[0] + g->box_stp / 2;
    g->gy0 = g->box_min[1] + g->box_stp / 2;
    g->nx_pad = (g->nv - g->nx) / 2 > 0 ? (g->nv - g->nx) / 2 : 0;
    int nv_pad = (g->nx - g->nv) / 2 > 0 ? (g->nx - g->nv) / 2 : 0;
    g->nva_pad = nv_pad;
    g->nfft = g->nx + 2 * g->nx_pad;
    /* z_bin = lambda / max(1-|vz|) / 10 over the view grid */
    double worst = 0;
    for (int i = 0; i < g->nv; ++i)
        for (int j = 0; j < g->nv; ++j) {
            double vx = -(-v_max + i * v_stp), vy = -(-v_max + j * v_stp);
            double vz = sqrt(1 - vx * vx - vy * vy);
            double q = 1 - fabs(vz);
            if (q > worst) worst = q;
        }
    g->z_bin = lambda / worst / 5;
    g->z_bin = g->z_bin / 2;
    g->nz = colon_len(g->box_min[2], g->z_bin, g->box_max[2]) + 1;
    return g->nv0;
}

int orc_quad_grid_setup(const double *box_min, const double *box_max, double v_max, double v_stp,
                        double lambda, orc_quad_grid *g)
{
    return quad_grid_setup(box_min, box_max, v_max, v_stp, lambda, 0, g);
}

int orc_quad_grid_setup_wave(const double *box_min, const double *box_max, double v_max, double v_stp,
                             double lambda, orc_quad_grid *g)
{
    return quad_grid_setup(box_min, box_max, v_max, v_stp, lambda, 1, g);
}

typedef struct { double z1; int64_t idx; } zkey;
static int zkey_cmp(const void *a, const void *b)
{
    const zkey *x = a, *y = b;
    if (x->z1 < y->z1) return -1;
    if (x->z1 > y->z1) return 1;
    return (x->idx > y->idx) - (x->idx < y->idx);   /* stable, like MATLAB sort */
}

/* 1-D unnormalised DFT of length n with stride, into out. */
static void dft1(const cplx *in, size_t stride, int n, cplx *out, const cplx *tw)
{
    for (int k = 0; k < n; ++k) {
        cplx s = 0;
        for (int p = 0; p < n; ++p) s += in[(size_t)p * stride] * tw[((size_t)k * p) % (size_t)n];
        out[k] = s;
    }
}

/* MCfieldFourierQuad.m:1-303, or with wl != NULL MCfieldFourierQuadOnWave.m:
 * 1-337 (lights combined per vertex with the weights Wl, one output plane,
 * the grid-parity fix of :100-102 and the vsign conventions of :143, :290,
 * :302). */
static int64_t fourier_quad(double sigt, double albedo, const double *box_min_in, const double *box_max_in,
                            const double *l, int nl, const cplx *wl, double v_max, double v_stp, int64_t max_itr,
                            double lambda, int do_cbs, int smp_flg, const double *lmean0, double vsign,
                            const orc_phase *amp, uint64_t seed, double *u_re, double *u_im)
{
    orc_quad_grid G;
    quad_grid_setup(box_min_in, box_max_in, v_max, v_stp, lambda, wl != NULL, &G);
    const int nout = wl ? 1 : nl;
    const int nv0 = G.nv0, nx = G.nx, nfft = G.nfft;
    /* view grids MCfieldFourierQuad.m:108-113 */
    double *vz0 = malloc(sizeof(double) * (size_t)nv0 * nv0);
    for (int j = 0; j < nv0; ++j)
        for (int i = 0; i < nv0; ++i) {
            double vx = -(-v_max + i * v_stp), vy = -(-v_max + j * v_stp);
            vz0[i + (size_t)nv0 * j] = vsign * sqrt(1 - vx * vx - vy * vy);
        }
    double *gz = malloc(sizeof(double) * (size_t)G.nz);
    for (int i = 0; i < G.nz - 1; ++i) gz[i] = G.box_min[2] + i * G.z_bin;
    gz[G.nz - 1] = G.box_max[2];
    const double lz = l[3 * (nl - 1) + 2];   /* l(end) */
    const double l_diratt = (lz > 0) ? 1 / fabs(lz) : 0;
    const double l_idiratt = (lz < 0) ? 1 / fabs(lz) : 0;
    const double kk = 2 * M_PI / lambda;

    /* path generation MCfieldFourierQuad.m:146-218 */
    orc_problem pr;
    memset(&pr, 0, sizeof pr);
    pr.dim = 3; pr.sigt = sigt; pr.albedo = albedo; pr.lambda = lambda;
    for (int j = 0; j < 3; ++j) { pr.box_min[j] = G.box_min[j]; pr.box_max[j] = G.box_max[j]; }
    pr.nl = nl; pr.l = l; pr.nv = 0; pr.v = NULL;
    pr.is_ff_l = 1; pr.is_ff_v = 1; pr.do_cbs = do_cbs; pr.smp_flg = smp_flg;
    /* the walk's phase function: isotropic in the reference; amp selects
     * HG or a tabulated table instead (directions only: the scatter itself
     * carries no amplitude) */
    if (amp)
        pr.amp = *amp;
    else
        pr.amp.type = 1;
    pr.has_lmean0 = 1;
    double lm[3] = {0, 0, 1};
    if (lmean0) for (int j = 0; j < 3; ++j) lm[j] = lmean0[j];
    for (int j = 0; j < 3; ++j) pr.lmean0[j] = lm[j];
    pr.seed = seed;
    double unmean[3], meanl[3];
    mean_dir(&pr, unmean, meanl);

    int64_t cap = 0, nstore = 0, scap = 1024;
    vtx *buf = NULL;
    double *xL = malloc(sizeof(double) * 3 * (size_t)scap), *x1L = malloc(sizeof(double) * 3 * (size_t)scap);
    cplx *pxL = malloc(sizeof(cplx) * (size_t)scap);
    path_head h;
    for (int64_t p = 0; p < max_itr; ++p) {
        int64_t n = walk(&pr, (uint64_t)p, unmean, meanl, &h, &buf, &cap);
        if (nstore + n > scap) {
            while (nstore + n > scap) scap *= 2;
            xL = realloc(xL, sizeof(double) * 3 * (size_t)scap);
            x1L = realloc(x1L, sizeof(double) * 3 * (size_t)scap);
            pxL = realloc(pxL, sizeof(cplx) * (size_t)scap);
        }
        for (int64_t k = 0; k < n; ++k) {
            for (int j = 0; j < 3; ++j) { xL[3 * nstore + j] = buf[k].x[j]; x1L[3 * nstore + j] = h.x1[j]; }
            cplx c = sqrt(buf[k].weight / h.px) * cexp(CMPLX(0, 2 * M_PI * buf[k].xi));
            if (do_cbs && k > 0) c = 1 / sqrt(2.0) * c;
            pxL[nstore] = c;
            ++nstore;
        }
    }
    free(buf);

    /* twiddles exp(-2 pi i kp/n) */
    cplx *tw = malloc(sizeof(cplx) * (size_t)nfft);
    for (int k = 0; k < nfft; ++k) tw[k] = cexp(CMPLX(0, -2 * M_PI * k / nfft));
    const size_t nf2 = (size_t)nfft * nfft, nvv = (size_t)nv0 * nv0;
    cplx *u = calloc(nvv * (size_t)nout, sizeof(cplx));
    cplx *uxy = malloc(sizeof(cplx) * nf2), *tmp = malloc(sizeof(cplx) * nf2);
    cplx *colo = malloc(sizeof(cplx) * (size_t)nfft);
    zkey *keys = malloc(sizeof(zkey) * (size_t)(nstore ? nstore : 1));
    const int hs = nfft / 2;   /* floor(n/2): the shift of fftshift / ifftshift */

    for (int tt = 0; tt < 2; ++tt) {
        if (tt == 1) {
            if (!do_cbs) continue;
            double *t = xL; xL = x1L; x1L = t;   /* MCfieldFourierQuad.m:226-232 */
        }
        for (int64_t i = 0; i < nstore; ++i) { keys[i].z1 = x1L[3 * i + 2]; keys[i].idx = i; }
        qsort(keys, (size_t)nstore, sizeof(zkey), zkey_cmp);
        int64_t zmax_i = 0;   /* number of consumed vertices */
        for (int j1 = 0; j1 < G.nz - 1; ++j1) {
            int64_t c = zmax_i;
            double z_max = gz[j1 + 1];
            while (c < nstore && keys[c].z1 < z_max) ++c;
            int64_t lo = zmax_i, hi = c;   /* [lo, hi) */
            zmax_i = c;
            if (hi <= lo) continue;
            double tz = (gz[j1] + gz[j1 + 1]) / 2;
            for (int j2 = 0; j2 < nout; ++j2) {
                const double *lj = &l[3 * j2];
                memset(uxy, 0, sizeof(cplx) * nf2);
                for (int64_t q = lo; q < hi; ++q) {
                    int64_t id = keys[q].idx;
                    const double *xa = &xL[3 * id], *x1 = &x1L[3 * id];
                    long ix = lround((x1[0] - G.gx0) / G.box_stp);
                    long iy = lround((x1[1] - G.gy0) / G.box_stp);
                    if (ix < 0) ix = 0;
                    if (ix > nx - 1) ix = nx - 1;
                    if (iy < 0) iy = 0;
                    if (iy > nx - 1) iy = nx - 1;
                    double att = -sigt / 2 * ((xa[2] - G.box_min[2]) * l_diratt + (G.box_max[2] - xa[2]) * l_idiratt);
                    cplx val;
                    if (wl) {
                        /* MCfieldFourierQuadOnWave.m:281-302 */
                        cplx twl = 0;
                        for (int j = 0; j < nl; ++j) twl += wl[j] * cexp(CMPLX(0, vsign * kk * dot3(&l[3 * j], xa, 3)));
                        cplx e1 = cexp(CMPLX(0, vsign * kk * (-x1[2] + tz)));
                        val = pxL[id] * e1 * twl * exp(att);
                    } else {
                        cplx e1 = cexp(CMPLX(0, -kk * (-x1[2] + tz)));
                        cplx e2 = cexp(CMPLX(att, kk * dot3(lj, xa, 3)));
                        val = pxL[id] * e1 * e2;
                    }
                    /* padded placement: padarray by nx_pad on each side */
                    size_t pi = (size_t)(ix + G.nx_pad), pj = (size_t)(iy + G.nx_pad);
                    uxy[pi + (size_t)nfft * pj] += val;
                }
                /* tfu = fliplr(flipud(fftshift(fft2(ifftshift(u_xy))))) */
                /* ifftshift: x'[p] = x[(p + floor(n/2)) mod n] */
                for (int j = 0; j < nfft; ++j)
                    for (int i = 0; i < nfft; ++i)
                        tmp[i + (size_t)nfft * j] = uxy[(size_t)((i + hs) % nfft) + (size_t)nfft * ((j + hs) % nfft)];
                /* DFT along dim 1 (columns), then dim 2 */
                for (int j = 0; j < nfft; ++j) {
                    dft1(&tmp[(size_t)nfft * j], 1, nfft, colo, tw);
                    memcpy(&tmp[(size_t)nfft * j], colo, sizeof(cplx) * (size_t)nfft);
                }
                for (int i = 0; i < nfft; ++i) {
                    dft1(&tmp[i], (size_t)nfft, nfft, colo, tw);
                    for (int j = 0; j < nfft; ++j) tmp[i + (size_t)nfft * j] = colo[j];
                }
                /* fftshift: y[i] = F[(i - floor(n/2)) mod n]; flips: H[a] = y[n-1-a] */
                for (int b = G.nva_pad; b < nfft - G.nva_pad; ++b)
                    for (int a = G.nva_pad; a < nfft - G.nva_pad; ++a) {
                        int ya = nfft - 1 - a, yb = nfft - 1 - b;
                        int fa = (ya - hs + nfft) % nfft, fb = (yb - hs + nfft) % nfft;
                        cplx val = tmp[(size_t)fa + (size_t)nfft * fb];
                        int oa = a - G.nva_pad, ob = b - G.nva_pad;
                        size_t o = (size_t)oa + (size_t)nv0 * ob;
                        /* zexp0 = -+2 pi i/lambda vz0 vsign - sigt/2 (diratt0 - idiratt0)
                         * (MCfieldFourierQuad.m:134; + in MCfieldFourierQuadOnWave.m:143) */
                        double vz = vz0[o];
                        double dir = (vz < 0) ? 1 / fabs(vz) : 0, idir = (vz > 0) ? 1 / fabs(vz) : 0;
                        cplx zexp = CMPLX(-sigt / 2 * (dir - idir), (wl ? 2 : -2) * M_PI / lambda * vz * vsign);
                        u[o + nvv * (size_t)j2] += val * cexp(zexp * tz);
                    }
            }
        }
    }
    /* boxexp0 MCfieldFourierQuad.m:137-138, 291-293 */
    for (size_t o = 0; o < nvv; ++o) {
        double vz = vz0[o];
        double dir = (vz < 0) ? 1 / fabs(vz) : 0, idir = (vz > 0) ? 1 / fabs(vz) : 0;
        cplx be = exp(-sigt / 2 * (-G.box_min[2] * dir + G.box_max[2] * idir)) *
                  cexp(CMPLX(0, 2 * M_PI / lambda * (-G.box_min[2] * (vz < 0) * vz + G.box_max[2] * (vz > 0) * vz)));
        for (int j2 = 0; j2 < nout; ++j2) u[o + nvv * (size_t)j2] *= be;
    }
    /* MCfieldFourierQuad.m:302-303 : maxItr is the stored vertex count */
    /* box_w(1:2) = box_w_n for both lateral axes (MCfieldFourierQuad.m:98) */
    double bwn = G.nx * G.box_stp;
    double V = bwn * bwn * (G.box_max[2] - G.box_min[2]);
    double s = nstore ? sqrt(1.0 / (double)nstore * V * sigt) : 0;
    split_out(u, nvv * (size_t)nout, s, u_re, u_im);
    free(vz0); free(gz); free(xL); free(x1L); free(pxL); free(tw); free(u); free(uxy); free(tmp);
    free(colo); free(keys);
    return nstore;
}

int64_t orc_mcfield_fourier_quad(double sigt, double albedo, const double *box_min, const double *box_max,
                                 const double *l, int nl, double v_max, double v_stp, int64_t max_itr,
                                 double lambda, int do_cbs, int smp_flg, const double *lmean0,
                                 double vsign, const orc_phase *amp, uint64_t seed, double *u_re, double *u_im)
{
    return fourier_quad(sigt, albedo, box_min, box_max, l, nl, NULL, v_max, v_stp, max_itr, lambda, do_cbs,
                        smp_flg, lmean0, vsign, amp, seed, u_re, u_im);
}

int64_t orc_mcfield_fourier_quad_wave(double sigt, double albedo, const double *box_min, const double *box_max,
                                      const double *l, int nl, const double *wl_re, const double *wl_im,
                                      double v_max, double v_stp, int64_t max_itr, double lambda, int do_cbs,
                                      int smp_flg, const double *lmean0, double vsign, const orc_phase *amp,
                                      uint64_t seed, double *u_re, double *u_im)
{
    cplx *wl = malloc(sizeof(cplx) * (size_t)nl);
    for (int j = 0; j < nl; ++j) wl[j] = CMPLX(wl_re[j], wl_im ? wl_im[j] : 0.0);
    int64_t r = fourier_quad(sigt, albedo, box_min, box_max, l, nl, wl, v_max, v_stp, max_itr, lambda, do_cbs,
                             smp_flg, lmean0, vsign, amp, seed, u_re, u_im);
    free(wl);
    return r;
}

/* ================================================== heterogeneous media */
/* code/MCCovHetro.m:1-276 with coreFunctions/{findBin,smpWoodcock,
 * evalphaseattHetro,expSmpXHetro,cubeProj,invCDFvec,smpicdf}.m.
 *
 * Random numbers (same Philox keys as the homogeneous walk, so a uniform
 * one-voxel medium walks MCcov's paths exactly):
 *   vertex 0: g0 = (u_z, u_x), g1 = (u_y, d0), g2 = (d1, d2), g3 = (d3, a)
 *             a picks the voxel (smpFlg 1); smpFlg 2 uses (g0.y, g1.x) for the
 *             plane patch and (g0.x, g3.y) for the first Woodcock trial;
 *             Woodcock trials k >= 1 at vertex 0 use group 16 + k; the lrad
 *             retries of smpFlg 1 use groups 1024 + 2k, 1025 + 2k.
 *   vertex n: g0 = (xi, first Woodcock distance), g1 = (u_roulette, d0),
 *             g2 = (d1, d2), g3 = (d3, first Woodcock acceptance); trials
 *             k >= 1 use group 16 + k. */

static int med_nvox(const orc_medium *m, int dim)
{
    int n = 1;
    for (int j = 0; j < dim; ++j) n *= m->nb[j];
    return n;
}

/* findBin.m:1-21 (0-based index; the cell of a point on the lower face is
 * clamped to the first cell, MATLAB would index 0) */
static int find_bin(const orc_problem *pr, const orc_medium *m, const double *x, double *bmn, double *bmx)
{
    int idx = 0, stride = 1;
    for (int j = 0; j < pr->dim; ++j) {
        double q = ceil((x[j] - pr->box_min[j]) / m->bin[j]);
        int ii = (int)q;
        if (ii < 1) ii = 1;
        if (ii > m->nb[j]) ii = m->nb[j];
        idx += (ii - 1) * stride;
        stride *= m->nb[j];
        if (bmn) {
            bmn[j] = pr->box_min[j] + m->bin[j] * (ii - 1);
            bmx[j] = bmn[j] + m->bin[j];
        }
    }
    return idx;
}

static double med_albedo(const orc_problem *pr, const orc_medium *m, const double *x)
{
    if (!m->albedo) return pr->albedo;
    return m->albedo[find_bin(pr, m, x, NULL, NULL)];
}

static double med_sigmax(const orc_problem *pr, const orc_medium *m)
{
    double s = 0;
    for (int i = 0; i < med_nvox(m, pr->dim); ++i) s = fmax(s, m->sigt[i]);
    return s;
}

/* evalphaseattHetro.m:29-48: optical depth from x along -v (towards the
 * source), cell by cell, capped at the distance nd. */
static double march_depth(const orc_problem *pr, const orc_medium *m, const double *x, const double *v, double nd,
                          int cap)
{
    const int dim = pr->dim;
    double mv[3] = {-v[0], -v[1], -v[2]};
    double nx[3] = {x[0], x[1], x[2]};
    double bdv = 0, wbdv = 0;
    for (int it = 0; it < cap; ++it) {
        double bmn[3], bmx[3];
        int b = find_bin(pr, m, nx, bmn, bmx);
        double dist = orc_cube_dist(nx, bmn, bmx, mv, dim);
        double ep = 0;
        for (int j = 0; j < dim; ++j) ep = fmax(ep, fabs(bmx[j] - bmn[j]));
        ep = 0.00010000000000000002 * ep;   /* 0.1^4 as computed by pow */
        for (int j = 0; j < dim; ++j) nx[j] = nx[j] + (dist + ep) * mv[j];
        bdv = bdv + dist;
        wbdv = wbdv + dist * m->sigt[b];
        if (bdv >= nd) {
            double md = bdv - nd;
            bdv = nd;
            wbdv = wbdv - md * m->sigt[b];
            break;
        }
        if (outside(nx, pr->box_min, pr->box_max, dim)) break;
    }
    return wbdv;
}

static int march_cap(const orc_medium *m, int dim)
{
    return 4 * (m->nb[0] + m->nb[1] + (dim == 3 ? m->nb[2] : 0)) + 16;
}

/* evalphaseattHetro.m:1-51: march from x towards the source through the
 * cells (cubeDist per cell, stepping 1e-4 cell sizes past each face), the
 * optical depth sum d*sigt(cell) capped at the source distance nd. */
void orc_eval_phase_att_hetro(const orc_problem *pr, const orc_medium *m, const double *x, const double *d,
                              int nd_count, int is_ff, double sign, double *out_re, double *out_im)
{
    const int dim = pr->dim;
    const double k = M_PI * 2.0 / pr->lambda;
    const int cap = march_cap(m, dim);
    for (int i = 0; i < nd_count; ++i) {
        double v[3] = {0, 0, 0};
        for (int j = 0; j < dim; ++j) v[j] = sign * d[i * dim + j];
        double pv, rv, nd;
        if (is_ff) {
            pv = dot3(x, v, dim);
            rv = 1.0;
            nd = 1e12;
        } else {
            double dd[3] = {0, 0, 0}, s = 0;
            for (int j = 0; j < dim; ++j) { dd[j] = x[j] - v[j]; s += dd[j] * dd[j]; }
            nd = sqrt(s);
            for (int j = 0; j < dim; ++j) v[j] = dd[j] / nd;
            pv = nd;
            rv = 1.0 / pow(nd + 0.01 * pr->lambda, (dim - 1) / 2.0);
        }
        double wbdv = march_depth(pr, m, x, v, nd, cap);
        double mag = exp(-0.5 * wbdv) * rv;
        double ph = k * pv;
        out_re[i] = mag * cos(ph);
        out_im[i] = mag * sin(ph);
    }
}

/* smpWoodcock.m:1-15 (delta tracking with sigt_max) */
static double woodcock(const orc_problem *pr, const orc_medium *m, double smax, const double *x, const double *w,
                       uint64_t path, uint32_t vertex, double u_dist0, double u_acc0)
{
    double d = 0;
    for (int it = 0; it < 1000000; ++it) {
        double ud = u_dist0, ua = u_acc0;
        if (it > 0) {
            double g[2];
            orc_uniform_pair(pr->seed, path, vertex, 16u + (uint32_t)it, g);
            ud = g[0];
            ua = g[1];
        }
        d = d - log(1 - ud) / smax;
        double nx[3] = {0, 0, 0};
        for (int j = 0; j < pr->dim; ++j) nx[j] = x[j] + d * w[j];
        if (outside(nx, pr->box_min, pr->box_max, pr->dim)) break;
        int b = find_bin(pr, m, nx, NULL, NULL);
        if (ua <= m->sigt[b] / smax) break;
    }
    return d;
}

double orc_woodcock(const orc_problem *pr, const orc_medium *m, const double *x, const double *w, uint64_t path)
{
    double g0[2], g3[2];
    orc_uniform_pair(pr->seed, path, 1, 0, g0);
    orc_uniform_pair(pr->seed, path, 1, 3, g3);
    return woodcock(pr, m, med_sigmax(pr, m), x, w, path, 1, g0[1], g3[1]);
}

/* invCDFvec.m:1-8 with MATLAB's hist(y, N): N equal bins over [min, max],
 * values on an inner edge counted in the upper bin, then the cumulative
 * counts.  icdf has n_out entries. */
int orc_hetro_icdf(const double *sigt, int nvox, int n_out, int32_t *icdf)
{
    double tot = 0;
    for (int i = 0; i < nvox; ++i) tot += sigt[i];
    if (!(tot > 0)) return -1;
    double *cdf = malloc(sizeof(double) * (size_t)nvox);
    double c = 0;
    for (int i = 0; i < nvox; ++i) {
        c += sigt[i] / tot;   /* cumsum(sigt(:)'/sum(sigt(:))) */
        cdf[i] = c;
    }
    double lo = cdf[0], hi = cdf[0];
    for (int i = 0; i < nvox; ++i) { lo = fmin(lo, cdf[i]); hi = fmax(hi, cdf[i]); }
    int32_t *cnt = calloc((size_t)n_out, sizeof(int32_t));
    for (int i = 0; i < nvox; ++i) {
        int b;
        if (hi == lo) {
            b = n_out / 2;   /* hist of a constant puts everything in the middle bin */
        } else {
            double bw = (hi - lo) / n_out;
            b = 0;
            /* largest b with lo + b*bw <= cdf (edges: lo + bw*(1:N-1)) */
            double q = floor((cdf[i] - lo) / bw);
            b = (int)q;
            if (b > n_out - 1) b = n_out - 1;
            if (b < 0) b = 0;
            while (b > 0 && cdf[i] < lo + bw * b) --b;
            while (b < n_out - 1 && cdf[i] >= lo + bw * (b + 1)) ++b;
        }
        cnt[b] += 1;
    }
    int32_t run = 0;
    for (int b = 0; b < n_out; ++b) { run += cnt[b]; icdf[b] = run; }
    free(cnt);
    free(cdf);
    return 0;
}

/* an orthonormal basis of the complement of lmean (the columns of
 * null(lmean') up to an in-plane rotation) */
static void null_basis(const double *lm, int dim, double u[3][3])
{
    if (dim == 2) {
        u[0][0] = -lm[1];
        u[1][0] = lm[0];
        return;
    }
    /* the axis least aligned with lmean */
    int a = 0;
    for (int j = 1; j < 3; ++j)
        if (fabs(lm[j]) < fabs(lm[a])) a = j;
    double e[3] = {0, 0, 0};
    e[a] = 1;
    double c1[3] = {lm[1] * e[2] - lm[2] * e[1], lm[2] * e[0] - lm[0] * e[2], lm[0] * e[1] - lm[1] * e[0]};
    double n1 = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
    for (int j = 0; j < 3; ++j) c1[j] /= n1;
    double c2[3] = {lm[1] * c1[2] - lm[2] * c1[1], lm[2] * c1[0] - lm[0] * c1[2], lm[0] * c1[1] - lm[1] * c1[0]};
    for (int j = 0; j < 3; ++j) {
        u[j][0] = c1[j];
        u[j][1] = c2[j];
    }
}

/* cubeProj.m:1-20 : move x along v to just inside the box; 0 when the
 * point is already past the entry (the reference returns a scalar there) */
static int cube_proj(const double *x, const double *bmin, const double *bmax, const double *v, int dim, double *xd)
{
    const double thr = 1.0000000000000004e-07;
    double d = -INFINITY;
    for (int j = 0; j < dim; ++j) {
        double ad = fmax(fabs(v[j]), thr);
        double bd = (v[j] <= 0) ? (x[j] - bmax[j]) / ad : (-x[j] + bmin[j]) / ad;
        d = fmax(d, bd);
    }
    if (!(d >= 0)) return 0;
    double ep = 0;
    for (int j = 0; j < dim; ++j) ep = fmax(ep, fabs(bmax[j] - bmin[j]));
    ep = 0.00010000000000000002 * ep;
    for (int j = 0; j < dim; ++j) xd[j] = x[j] + (d + ep) * v[j];
    return 1;
}

/* One path of MCCovHetro.m:101-244.  Returns the vertex count, 0 when the
 * first vertex falls outside the box (the iteration is spent, MCCovHetro.m:121-123). */
static int64_t walk_hetro(const orc_problem *pr, const orc_medium *m, const int32_t *icdf, int n_icdf, double V,
                          double smax, uint64_t path, const double *meanl, path_head *h, vtx **buf, int64_t *cap)
{
    const int dim = pr->dim;
    const double kill_thr = 0.2;
    double g0[2], g1[2], g2[2], g3[2];
    orc_uniform_pair(pr->seed, path, 0, 0, g0);
    orc_uniform_pair(pr->seed, path, 0, 1, g1);
    orc_uniform_pair(pr->seed, path, 0, 2, g2);
    orc_uniform_pair(pr->seed, path, 0, 3, g3);
    double x[3] = {0, 0, 0}, px;
    if (pr->smp_flg == 1) {
        /* MCCovHetro.m:103-114 : cell from the quantised icdf (smpicdf.m), uniform in the cell */
        int i = (int)ceil(g3[1] * n_icdf);
        if (i < 1) i = 1;
        /* a table entry of 0 (only from hist of a constant cdf, one cell) is
         * not a valid MATLAB index: take the first cell */
        int b = icdf[i - 1] - 1;
        if (b < 0) b = 0;
        int rem = b;
        double bmn[3] = {0, 0, 0};
        for (int j = 0; j < dim; ++j) {
            int ii = rem % m->nb[j];
            rem /= m->nb[j];
            bmn[j] = pr->box_min[j] + m->bin[j] * ii;
        }
        double r[3] = {g0[1], dim == 3 ? g1[0] : g0[0], g0[0]};
        for (int j = 0; j < dim; ++j) x[j] = r[j] * m->bin[j] + bmn[j];
        px = 1 / V;
        if (m->has_lrad) {
            /* MCCovHetro.m:45-48, 107-113 (one light) */
            double ls[3];
            for (int j = 0; j < dim; ++j) ls[j] = (pr->box_min[j] + pr->box_max[j]) / 2;
            ls[dim - 1] = (pr->l[dim - 1] > 0) ? pr->box_min[dim - 1] : pr->box_max[dim - 1];
            for (int it = 0;; ++it) {
                double lx = dot3(pr->l, x, dim), r2 = 0;
                for (int j = 0; j < dim; ++j) {
                    double q = x[j] - lx * pr->l[j] - ls[j];
                    r2 += q * q;
                }
                if (!(sqrt(r2) > m->lrad) || it > 100000) break;
                double a[2], c[2];
                orc_uniform_pair(pr->seed, path, 0, 1024u + 2u * (uint32_t)it, a);
                orc_uniform_pair(pr->seed, path, 0, 1025u + 2u * (uint32_t)it, c);
                double rr[3] = {a[0], a[1], c[0]};
                for (int j = 0; j < dim; ++j) x[j] = rr[j] * (pr->box_max[j] - pr->box_min[j]) + pr->box_min[j];
                px = 1;
            }
        }
    } else {
        /* expSmpXHetro.m:1-31 */
        double lrad = 0;
        if (m->has_lrad) {
            lrad = m->lrad;
        } else {
            for (int j = 0; j < dim; ++j) lrad += (pr->box_max[j] - pr->box_min[j]) * (pr->box_max[j] - pr->box_min[j]);
            lrad = sqrt(lrad);
        }
        double u[3][3] = {{0}};
        null_basis(meanl, dim, u);
        double c = 0;
        for (int j = 0; j < dim; ++j) c = fmax(c, fmax(fabs(pr->box_max[j]), fabs(pr->box_min[j])));
        c = c * dim;
        double r[2] = {(g0[1] - 0.5) * lrad, (g1[0] - 0.5) * lrad};
        double x0[3] = {0, 0, 0};
        for (int j = 0; j < dim; ++j) {
            double s = 0;
            for (int q = 0; q < dim - 1; ++q) s += u[j][q] * r[q];
            x0[j] = s - meanl[j] * c;
        }
        int ok = cube_proj(x0, pr->box_min, pr->box_max, meanl, dim, x);
        if (!ok || outside(x, pr->box_min, pr->box_max, dim)) return 0;
        double d = woodcock(pr, m, smax, x, meanl, path, 0, g0[0], g3[1]);
        for (int j = 0; j < dim; ++j) x[j] = x[j] + d * meanl[j];
        if (outside(x, pr->box_min, pr->box_max, dim)) return 0;
        /* px = abs(evalphaseattHetro(x, lmean, 1, sigt, 1, ...))^2 / lrad^(dim-1);
         * the modulus of the unit phase is taken as exactly 1 */
        double mag = exp(-0.5 * march_depth(pr, m, x, meanl, 1e12, march_cap(m, dim)));
        px = mag * mag / pow(lrad, dim - 1);
    }
    if (outside(x, pr->box_min, pr->box_max, dim)) return 0;
    double du[4] = {g1[1], g2[0], g2[1], g3[0]};
    double w[3] = {0, 0, 0};
    cplx w0p;
    if (!pr->has_amp0) {
        iso_dir(du, dim, w);
        w0p = 1.0 / sqrt(pow(2.0, dim - 1) * M_PI);
    } else {
        orc_phase p0 = pr->amp0;
        p0.type = pr->amp.type;
        orc_sample_dir(meanl, &p0, dim, du, w);
        w0p = eval_amp_c(dot3(meanl, w, dim), &p0, dim);
    }
    memcpy(h->x1, x, sizeof x);
    h->px = px;
    memcpy(h->w1, w, sizeof w);
    h->w0p = w0p;

    double weight = med_albedo(pr, m, x);   /* MCCovHetro.m:164 */
    double ow[3] = {0, 0, 0};
    int64_t n = 0;
    for (;;) {
        ++n;
        double a[2], b[2], c[2], e[2];
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 0, a);
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 3, e);
        if (n > *cap) {
            *cap = *cap ? *cap * 2 : 64;
            *buf = (vtx *)realloc(*buf, sizeof(vtx) * (size_t)*cap);
        }
        vtx *vv = &(*buf)[n - 1];
        memcpy(vv->x, x, sizeof x);
        memcpy(vv->ow, ow, sizeof ow);
        vv->weight = weight;
        vv->xi = a[0];
        /* MCCovHetro.m:228-233 */
        double d = woodcock(pr, m, smax, x, w, path, (uint32_t)n, a[1], e[1]);
        for (int j = 0; j < dim; ++j) x[j] = x[j] + d * w[j];
        if (outside(x, pr->box_min, pr->box_max, dim)) break;
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 1, b);
        /* MCCovHetro.m:234-241 */
        if (weight < kill_thr) {
            if (b[0] > med_albedo(pr, m, x)) break;
        } else {
            weight = weight * med_albedo(pr, m, x);
        }
        orc_uniform_pair(pr->seed, path, (uint32_t)n, 2, c);
        double u4[4] = {b[1], c[0], c[1], e[0]};
        memcpy(ow, w, sizeof w);
        orc_sample_dir(ow, &pr->amp, dim, u4, w);
    }
    return n;
}

/* The per-medium constants of MCCovHetro.m:64-66 and the path ids
 * [path_offset, path_offset + max_itr). */
int64_t orc_mccov_hetro_sums(const orc_problem *pr, const orc_medium *m, uint64_t p0, int64_t np,
                             double *ms_re, double *ms_im, double *mm_re, double *mm_im, double *mean_re,
                             double *mean_im)
{
    const int nvox = med_nvox(m, pr->dim);
    const int n_icdf = nvox * 100;
    int32_t *icdf = malloc(sizeof(int32_t) * (size_t)n_icdf);
    double tot = 0;
    for (int i = 0; i < nvox; ++i) tot += m->sigt[i];
    double vol = 1;
    for (int j = 0; j < pr->dim; ++j) vol *= m->bin[j];
    const double V = tot * vol;   /* V = sum(sigt(:))*prod(box_bin) */
    if (orc_hetro_icdf(m->sigt, nvox, n_icdf, icdf)) {
        free(icdf);
        return -1;
    }
    const double smax = med_sigmax(pr, m);
    double unmean[3], meanl[3];
    orc_problem q = *pr;
    q.has_lmean0 = 0;   /* MCCovHetro.m:40-41 */
    mean_dir(&q, unmean, meanl);
    const int nv = pr->nv, nl = pr->nl;
    const size_t ne = (size_t)nv * nv * nl * nl, nf = (size_t)nv * nl;
    cplx *Ms = calloc(ne, sizeof(cplx)), *Mm = calloc(ne, sizeof(cplx)), *mean1 = calloc(nf, sizeof(cplx));
    nee_ws ws;
    nee_alloc(&ws, nv, nl, pr->dim);
    ws.med = m;
    mcov_ws mw = {calloc((size_t)nv * nv, sizeof(cplx)), calloc((size_t)nv * nv, sizeof(cplx)),
                  calloc((size_t)nv * nv, sizeof(cplx)), calloc((size_t)nv * nv, sizeof(cplx))};
    vtx *buf = NULL;
    int64_t cap = 0, nvert = 0;
    path_head h;
    for (int64_t i = 0; i < np; ++i) {
        int64_t n = walk_hetro(pr, m, icdf, n_icdf, V, smax, p0 + (uint64_t)i, meanl, &h, &buf, &cap);
        if (n == 0) continue;
        nvert += n;
        mccov_path(pr, &h, buf, n, &ws, &mw, Ms, Mm, mean1);
    }
    split_out(Ms, ne, 1.0, ms_re, ms_im);
    split_out(Mm, ne, 1.0, mm_re, mm_im);
    split_out(mean1, nf, 1.0, mean_re, mean_im);
    free(buf); free(icdf); free(Ms); free(Mm); free(mean1);
    free(mw.m00); free(mw.mnn); free(mw.mn0); free(mw.m0n);
    nee_free(&ws);
    return nvert;
}

/* MCCovHetro.m:248-276 normalisation: Mm/2 with CBS, /maxItr; mean1 /maxItr
 * (/sqrt(cs) with cs = 1, see the header). */
int64_t orc_mccov_hetro(const orc_problem *pr, const orc_medium *m, int64_t max_itr, double *ms_re, double *ms_im,
                        double *mm_re, double *mm_im, double *mean_re, double *mean_im)
{
    int64_t nvert = orc_mccov_hetro_sums(pr, m, 0, max_itr, ms_re, ms_im, mm_re, mm_im, mean_re, mean_im);
    if (nvert < 0) return nvert;
    const size_t ne = (size_t)pr->nv * pr->nv * pr->nl * pr->nl, nf = (size_t)pr->nv * pr->nl;
    const double s = 1.0 / (double)max_itr, sm = (pr->do_cbs ? 0.5 : 1.0) * s;
    for (size_t i = 0; i < ne; ++i) {
        ms_re[i] *= s; ms_im[i] *= s;
        mm_re[i] *= sm; mm_im[i] *= sm;
    }
    for (size_t i = 0; i < nf; ++i) { mean_re[i] *= s; mean_im[i] *= s; }
    return nvert;
}


/* ------------------------------------------------------- MCsampleHetro */

/* code/MCsampleHetro.m:82-216 over paths [p0, p0 + np): MCCovHetro's walks
 * (walk_hetro; lmean0 honoured, MCsampleHetro.m:35-40; no lrad), each path's
 * field terms as MCfield.m:207-222 with E from the cell-by-cell march
 * (MCsampleHetro.m:141-186).  Unnormalised u (nv x nl); returns the vertex
 * count, or -1 for an invalid medium. */
int64_t orc_mcfield_hetro_sums(const orc_problem *pr, const orc_medium *m, uint64_t p0, int64_t np, double *u_re,
                               double *u_im)
{
    if (m->has_lrad) return -1;
    const int nvox = med_nvox(m, pr->dim);
    const int n_icdf = nvox * 100;
    int32_t *icdf = malloc(sizeof(int32_t) * (size_t)n_icdf);
    double tot = 0;
    for (int i = 0; i < nvox; ++i) tot += m->sigt[i];
    double vol = 1;
    for (int j = 0; j < pr->dim; ++j) vol *= m->bin[j];
    const double V = tot * vol;   /* MCsampleHetro.m:60 */
    if (orc_hetro_icdf(m->sigt, nvox, n_icdf, icdf)) {
        free(icdf);
        return -1;
    }
    const double smax = med_sigmax(pr, m);
    double unmean[3], meanl[3];
    mean_dir(pr, unmean, meanl);
    const size_t ne = (size_t)pr->nv * pr->nl;
    cplx *u = calloc(ne, sizeof(cplx));
    nee_ws ws;
    nee_alloc(&ws, pr->nv, pr->nl, pr->dim);
    ws.med = m;
    vtx *buf = NULL;
    int64_t cap = 0, nvert = 0;
    path_head h;
    for (int64_t i = 0; i < np; ++i) {
        int64_t n = walk_hetro(pr, m, icdf, n_icdf, V, smax, p0 + (uint64_t)i, meanl, &h, &buf, &cap);
        if (n == 0) continue;   /* MCsampleHetro.m:100-102 */
        nvert += n;
        mcfield_path(pr, &h, buf, n, &ws, u);
    }
    split_out(u, ne, 1.0, u_re, u_im);
    free(buf); free(icdf); free(u);
    nee_free(&ws);
    return nvert;
}

/* MCsampleHetro.m:222: u*sqrt(1/maxItr) (the volume is inside px). */
int64_t orc_mcfield_hetro(const orc_problem *pr, const orc_medium *m, int64_t max_itr, double *u_re, double *u_im)
{
    int64_t nvert = orc_mcfield_hetro_sums(pr, m, 0, max_itr, u_re, u_im);
    if (nvert < 0) return nvert;
    const size_t ne = (size_t)pr->nv * pr->nl;
    const double s = sqrt(1.0 / (double)max_itr);
    for (size_t i = 0; i < ne; ++i) {
        u_re[i] *= s;
        u_im[i] *= s;
    }
    return nvert;
}
