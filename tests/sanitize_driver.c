/* Drives every entry point of the FP64 oracle on tiny problems, for a build
 * under AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitize.py).
 * It checks nothing numerically: the sanitizers report any out-of-bounds
 * access, use of freed memory, signed overflow or invalid shift, and the
 * driver fails if an output is NaN/Inf. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scmc_oracle.h"

static int finite_all(const char *what, const double *a, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(a[i])) {
            fprintf(stderr, "%s: non-finite output at %zu\n", what, i);
            return 0;
        }
    return 1;
}

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static void base_problem(orc_problem *pr, int dim, const double *l, int nl, const double *v, int nv) {
    memset(pr, 0, sizeof *pr);
    pr->dim = dim;
    pr->sigt = 4.0;
    pr->albedo = 0.95;
    pr->lambda = 1.0;
    for (int k = 0; k < 3; ++k) {
        pr->box_min[k] = -1.0;
        pr->box_max[k] = 1.0;
    }
    pr->nl = nl;
    pr->nv = nv;
    pr->l = l;
    pr->v = v;
    pr->is_ff_l = 1;
    pr->is_ff_v = 1;
    pr->do_cbs = 1;
    pr->smp_flg = 2;
    pr->amp.type = 3;
    pr->amp.g = 0.7;
    pr->seed = 11;
}

int main(void) {
    int ok = 1;
    /* two lights, three views, unit directions */
    const double l3[6] = {0, 0, 1, 0.2, 0, 0.9797958971132712};
    const double v3[9] = {0, 0, 1, 0, 0.3, 0.9539392014169456, 0.1, -0.1, 0.9899494936611665};
    const double l2[4] = {0, 1, 0.6, 0.8};
    const double v2[6] = {0, 1, 0.28, 0.96, -0.6, 0.8};
    const int nl = 2, nv = 3;
    const size_t ncov = (size_t)nv * nv * nl * nl, nfld = (size_t)nv * nl;
    double *a = zalloc(ncov), *b = zalloc(ncov), *c = zalloc(ncov), *d = zalloc(ncov);
    double *e = zalloc(nfld), *f = zalloc(nfld);

    for (int dim = 2; dim <= 3; ++dim) {
        orc_problem pr;
        base_problem(&pr, dim, dim == 3 ? l3 : l2, nl, dim == 3 ? v3 : v2, nv);
        orc_mccov(&pr, 40, a, b, c, d);
        ok &= finite_all("mccov", a, ncov) & finite_all("mccov", c, ncov);
        orc_mccov_sums(&pr, 7, 20, 2, a, b, c, d);
        ok &= finite_all("mccov_sums", a, ncov);
        orc_mcfield(&pr, 40, e, f);
        ok &= finite_all("mcfield", e, nfld);
        orc_mcfield_sums(&pr, 3, 20, 2, e, f);
        ok &= finite_all("mcfield_sums", e, nfld);

        /* trace + replay */
        int64_t np = 16, counts[16];
        int64_t nvx = orc_trace_count(&pr, 0, np, counts);
        orc_pathset ps;
        memset(&ps, 0, sizeof ps);
        ps.n_paths = np;
        ps.n_vertices = nvx;
        ps.offset = (int64_t *)calloc((size_t)np + 1, sizeof(int64_t));
        for (int64_t p = 0; p < np; ++p) ps.offset[p + 1] = ps.offset[p] + counts[p];
        ps.x1 = zalloc(3 * (size_t)np);
        ps.px = zalloc((size_t)np);
        ps.w1 = zalloc(3 * (size_t)np);
        ps.w0p_re = zalloc((size_t)np);
        ps.w0p_im = zalloc((size_t)np);
        ps.x = zalloc(3 * (size_t)nvx);
        ps.ow = zalloc(3 * (size_t)nvx);
        ps.weight = zalloc((size_t)nvx);
        ps.xi = zalloc((size_t)nvx);
        orc_trace_fill(&pr, 0, np, &ps);
        orc_mccov_replay(&pr, &ps, np, a, b, c, d);
        ok &= finite_all("mccov_replay", a, ncov);
        orc_mcfield_replay(&pr, &ps, np, e, f);
        ok &= finite_all("mcfield_replay", e, nfld);
        free(ps.offset); free(ps.x1); free(ps.px); free(ps.w1); free(ps.w0p_re); free(ps.w0p_im);
        free(ps.x); free(ps.ow); free(ps.weight); free(ps.xi);

        /* wave */
        const double wre[2] = {0.6, -0.3}, wim[2] = {0.1, 0.4};
        orc_mcfield_wave(&pr, wre, wim, 40, e, f);
        ok &= finite_all("mcfield_wave", e, (size_t)nv);

        /* near-field lights and views, uniform sampling, no CBS */
        pr.is_ff_l = 0;
        pr.is_ff_v = 0;
        pr.smp_flg = 1;
        pr.do_cbs = 0;
        const double lnear3[6] = {0, 0, -3, 0.5, 0, -3}, vnear3[9] = {0, 0, 3, 0.5, 0.5, 3, -0.5, 0, 3};
        const double lnear2[4] = {0, -3, 0.5, -3}, vnear2[6] = {0, 3, 0.5, 3, -0.5, 3};
        pr.l = dim == 3 ? lnear3 : lnear2;
        pr.v = dim == 3 ? vnear3 : vnear2;
        orc_mccov(&pr, 30, a, b, c, d);
        ok &= finite_all("mccov_near", a, ncov);
        orc_mcfield(&pr, 30, e, f);
        ok &= finite_all("mcfield_near", e, nfld);
    }

    /* tabulated amplitude phase function, 3-D */
    {
        enum { NT = 257 };
        double *th = zalloc(NT), *fre = zalloc(NT), *fim = zalloc(NT);
        double *cs = zalloc(1), *spdf = zalloc(NT), *scdf = zalloc(NT), *sicdf = zalloc(NT);
        double *epdf = zalloc(NT), *are = zalloc(NT), *aim = zalloc(NT);
        for (int i = 0; i < NT; ++i) {
            th[i] = M_PI * i / (NT - 1);
            fre[i] = exp(-2.0 * th[i]) + 0.05;
            fim[i] = 0.1 * sin(th[i]);
        }
        int D = orc_tabulated_scatter(th, fre, fim, NT, cs, spdf, scdf, sicdf, epdf, are, aim);
        if (D != 3) {
            fprintf(stderr, "tabulated_scatter returned %d\n", D);
            ok = 0;
        }
        orc_problem pr;
        base_problem(&pr, 3, l3, nl, v3, nv);
        pr.amp.type = 2;
        pr.amp.n_amp = NT;
        pr.amp.amp_re = are;
        pr.amp.amp_im = aim;
        pr.amp.n_icdf = NT;
        pr.amp.icdf = sicdf;
        orc_mccov(&pr, 30, a, b, c, d);
        ok &= finite_all("mccov_tab", a, ncov);
        orc_mcfield(&pr, 30, e, f);
        ok &= finite_all("mcfield_tab", e, nfld);
        free(th); free(fre); free(fim); free(cs); free(spdf); free(scdf); free(sicdf);
        free(epdf); free(are); free(aim);
    }

    /* Fourier-accelerated field on a small grid (isotropic, 3-D) */
    {
        const double bmin[3] = {-2, -2, 0}, bmax[3] = {2, 2, 0.6};
        const double lq[6] = {0.05, 0, 0.99874921777190895, -0.1, 0.08, 0.99176609339196};
        orc_quad_grid g, gw;
        int nv0 = orc_quad_grid_setup(bmin, bmax, 0.3, 0.05, 0.5, &g);
        int nw0 = orc_quad_grid_setup_wave(bmin, bmax, 0.3, 0.05, 0.5, &gw);
        if (nv0 <= 0 || nw0 <= 0) {
            fprintf(stderr, "quad grid setup returned %d / %d\n", nv0, nw0);
            ok = 0;
        } else {
            size_t nq = (size_t)nv0 * nv0 * nl, nqw = (size_t)nw0 * nw0;
            double *qre = zalloc(nq), *qim = zalloc(nq);
            int64_t nvq = orc_mcfield_fourier_quad(3.0, 0.95, bmin, bmax, lq, nl, 0.3, 0.05, 40, 0.5, 1, 2, NULL,
                                                   -1.0, NULL, 4, qre, qim);
            ok &= nvq > 0 && finite_all("fourier_quad", qre, nq);
            orc_phase hg;
            memset(&hg, 0, sizeof hg);
            hg.type = 3;
            hg.g = 0.8;
            nvq = orc_mcfield_fourier_quad(3.0, 0.95, bmin, bmax, lq, nl, 0.3, 0.05, 40, 0.5, 1, 2, NULL, -1.0, &hg, 4,
                                           qre, qim);
            ok &= nvq > 0 && finite_all("fourier_quad", qre, nq);
            free(qre);
            free(qim);
            qre = zalloc(nqw);
            qim = zalloc(nqw);
            const double wre[2] = {1.0, 0.5};
            int64_t nvw = orc_mcfield_fourier_quad_wave(3.0, 0.95, bmin, bmax, lq, nl, wre, NULL, 0.3, 0.05, 40, 0.5,
                                                        0, 1, NULL, 1.0, NULL, 4, qre, qim);
            ok &= nvw > 0 && finite_all("fourier_quad_wave", qre, nqw);
            free(qre);
            free(qim);
        }
    }

    /* voxel media */
    {
        double sig[32], alb[32];
        for (int i = 0; i < 32; ++i) {
            sig[i] = 1.0 + (i % 5);
            alb[i] = 0.8 + 0.005 * i;
        }
        orc_medium m;
        memset(&m, 0, sizeof m);
        m.nb[0] = 4;
        m.nb[1] = 4;
        m.nb[2] = 2;
        m.bin[0] = 0.5;
        m.bin[1] = 0.5;
        m.bin[2] = 1.0;
        m.sigt = sig;
        m.albedo = alb;
        orc_problem pr;
        base_problem(&pr, 3, l3, nl, v3, nv);
        double *mre = zalloc(nfld), *mim = zalloc(nfld);
        int64_t n1 = orc_mccov_hetro(&pr, &m, 30, a, b, c, d, mre, mim);
        ok &= n1 >= 0 && finite_all("mccov_hetro", a, ncov) & finite_all("mccov_hetro_mean", mre, nfld);
        int64_t n2 = orc_mcfield_hetro(&pr, &m, 30, e, f);
        ok &= n2 >= 0 && finite_all("mcfield_hetro", e, nfld);
        double pos[3] = {0.1, -0.2, 0.3}, dir[3] = {0, 0, 1}, tre[1], tim[1];
        orc_eval_phase_att_hetro(&pr, &m, pos, dir, 1, 1, 1.0, tre, tim);
        ok &= finite_all("phase_att_hetro", tre, 1);
        double t = orc_woodcock(&pr, &m, pos, dir, 3);
        ok &= finite_all("woodcock", &t, 1);
        int32_t icdf[64];
        ok &= orc_hetro_icdf(sig, 32, 64, icdf) >= 0;
        m.has_lrad = 1;
        m.lrad = 0.5;
        m.albedo = NULL;
        n1 = orc_mccov_hetro(&pr, &m, 20, a, b, c, d, mre, mim);
        ok &= n1 >= 0 && finite_all("mccov_hetro_lrad", a, ncov);
        free(mre);
        free(mim);
    }

    free(a); free(b); free(c); free(d); free(e); free(f);
    if (!ok) return 2;
    printf("sanitize driver: all entry points ran\n");
    return 0;
}
