/*
 * scmc_oracle.h -- FP64 CPU restatement of the SCMC Monte Carlo speckle core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (mcspeckle_amd/) never links it.
 *
 * Every function restates a reference MATLAB file; the file:line it follows is
 * given beside each declaration.  Random numbers come from Philox4x32-10
 * keyed by (seed) with counter (vertex, group, path_lo, path_hi), so a path is
 * a pure function of (seed, path id) and any scheduler reproduces it.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   philox             : Random123 known-answer vectors.
 *   cube_dist          : slab closed form z/|dz|, (T-z)/|dz|.
 *   eval_phase_att     : phase == k x.d exactly, |e| == exp(-sigma/2 D).
 *   eval_amp (HG)      : integral of |a|^2 over the sphere/circle == 1.
 *   sample_dir (HG)    : E[cos] == g (3-D); |w| == 1; w.ow == cos.
 *   exp_smp_x          : E[1/px] == 1 (importance weight of a density).
 *   mccov              : brute-force term-by-term restatement of the
 *                        covariance formula on tiny inputs; Hermitian/PSD
 *                        invariants; single-scatter closed form within 3 SE.
 *   mcfield            : E[u u^H] == mccov on the same paths (statistical),
 *                        exact identity of its terms with the covariance
 *                        factors on a replayed path set.
 *   tabulated_scatter  : scatteringAmplitude.mat fixture; cdf monotone,
 *                        icdf index scaling.
 *   fourier_quad       : direct DFT definition vs numpy.fft on the same
 *                        grid; scatter-add vs brute force.
 *   near-field sources : point-source closed form of evalphaseatt; sources
 *                        at distance R reproduce the far-field results.
 *   mcfield_wave       : equals mcfield contracted with Wl on the same paths.
 *   fourier_quad_wave  : numpy.fft restatement at odd and even transform
 *                        sizes; linear in Wl.
 *   mccov_hetro        : a uniform one-cell medium walks MCcov's paths and
 *                        reproduces MCcov; Woodcock distances follow the
 *                        exact transmittance of a layered medium (KS);
 *                        single-scatter closed form of a two-layer slab;
 *                        the ray march equals a direct sum over layers.
 *   mcfield_hetro      : a uniform one-cell medium reproduces MCfield on the
 *                        same paths (far and near field, CBS, lmean0);
 *                        E[|u|^2] of a two-layer slab equals the trace of
 *                        MCCovHetro's Ms + Mm/2 (statistical).
 */
#ifndef SCMC_ORACLE_H
#define SCMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Phase function descriptor (interface/scatter: isotropic, HG, tabulated). */
typedef struct {
    int type;              /* 1 isotropic, 2 tabulated amplitude, 3 HG  */
    double g;              /* HG anisotropy                              */
    int n_amp;             /* tabulated: length of evalAmp               */
    const double *amp_re;  /* evalAmp real part                          */
    const double *amp_im;  /* evalAmp imaginary part                     */
    int n_icdf;            /* tabulated: length of sampleIcdf            */
    const double *icdf;    /* sampleIcdf (angles)                        */
} orc_phase;

/* Voxel medium of MCCovHetro.m:1 (sigt / albedo grids over box_bin cells). */
typedef struct {
    int nb[3];                /* cells per axis, round((box_max-box_min)./box_bin) */
    double bin[3];            /* box_bin                                          */
    const double *sigt;       /* prod(nb) extinction values, column-major         */
    const double *albedo;     /* prod(nb) values, or NULL for the scalar albedo   */
    int has_lrad;             /* lrad given (MCCovHetro.m:45, expSmpXHetro.m:4)   */
    double lrad;
} orc_medium;

/* Problem descriptor: the positional arguments of MCcov.m:1 / MCfield.m:1. */
typedef struct {
    int dim;                       /* 2 or 3                                */
    double sigt, albedo, lambda;
    double box_min[3], box_max[3];
    int nl, nv;
    const double *l;               /* dim x nl column-major                 */
    const double *v;               /* dim x nv column-major                 */
    int is_ff_l, is_ff_v;
    int do_cbs, smp_flg;           /* smp_flg 1 uniform, 2 exponential      */
    orc_phase amp;                 /* sct_type + ampfunc                    */
    int has_amp0;                  /* ampfunc0 given (15th argument)        */
    orc_phase amp0;
    int has_lmean0;                /* MCfield.m:77-81 lmean0 (16th argument) */
    double lmean0[3];
    uint64_t seed;
} orc_problem;

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Two 53-bit uniforms in [0,1) for (seed, path, vertex, group). */
void orc_uniform_pair(uint64_t seed, uint64_t path, uint32_t vertex, uint32_t group, double out[2]);

/* coreFunctions/cubeDist.m:4-11 */
double orc_cube_dist(const double *x, const double *bmin, const double *bmax, const double *d, int dim);
/* coreFunctions/evalphaseatt.m:1-20 ; d is dim x m, out is m complex */
void orc_eval_phase_att(const double *x, const double *d, int m, int dim, int is_ff, double sign,
                        double sigt, double lambda, const double *bmin, const double *bmax,
                        double *out_re, double *out_im);
/* coreFunctions/evalampfunc_general.m:1-19 */
void orc_eval_amp(double cosang, const orc_phase *ph, int dim, double *re, double *im);
/* coreFunctions/evaluateHG.m:25-39 (cosine input) */
double orc_evaluate_hg(double cosang, double g, int dim);
/* coreFunctions/sampleHG.m:1-18 ; u is one uniform */
double orc_sample_hg(double g, int dim, double u);
/* coreFunctions/rotateBytheta.m:1-28 ; u is one uniform (phi or sign) */
void orc_rotate_by_theta(const double *ow, double costheta, int dim, double u, double *w);
/* coreFunctions/smpampfunc_general.m:1-21 ; u[4] uniforms */
void orc_sample_dir(const double *ow, const orc_phase *ph, int dim, const double u[4], double *w);
/* coreFunctions/expSmpX.m:1-12 ; u[3] = (u_z, u_x, u_y) */
void orc_exp_smp_x(const double *bmin, const double *bmax, const double *lmean, double sigt, int dim,
                   const double u[3], double *x, double *px);

/* Path sets: the sampled random walks, vertex-major. */
typedef struct {
    int64_t n_paths, n_vertices;
    int64_t *offset;    /* n_paths + 1 : first vertex of each path              */
    double *x1;         /* 3 per path  : first scattering position               */
    double *px;         /* 1 per path  : first-vertex density ratio              */
    double *w1;         /* 3 per path  : first scattering direction              */
    double *w0p_re;     /* 1 per path  : first-direction amplitude (complex)     */
    double *w0p_im;
    double *x;          /* 3 per vertex: vertex position                         */
    double *ow;         /* 3 per vertex: incoming direction (zero at vertex 1)   */
    double *weight;     /* 1 per vertex: albedo weight W_n                       */
    double *xi;         /* 1 per vertex: MCfield random-phase uniform            */
} orc_pathset;

/* Count vertices of paths [p0, p0+np); counts[np]. Returns total. */
int64_t orc_trace_count(const orc_problem *pr, uint64_t p0, int64_t np, int64_t *counts);
/* Fill a path set whose offset[] was set by the caller from counts. */
void orc_trace_fill(const orc_problem *pr, uint64_t p0, int64_t np, orc_pathset *ps);

/* code/MCcov.m:1-299 : full run over paths [0, max_itr), normalised.
 * Out: Ms, Mm as nv x nv x nl x nl complex, column-major. */
void orc_mccov(const orc_problem *pr, int64_t max_itr,
               double *ms_re, double *ms_im, double *mm_re, double *mm_im);
void orc_mccov_mt(const orc_problem *pr, int64_t max_itr, int nthreads,
                  double *ms_re, double *ms_im, double *mm_re, double *mm_im);
/* Same, over paths [p0, p0+np), UNnormalised sums (for sharding and timing),
 * nthreads > 1 splits the paths over OpenMP threads with private sums. */
void orc_mccov_sums(const orc_problem *pr, uint64_t p0, int64_t np, int nthreads,
                    double *ms_re, double *ms_im, double *mm_re, double *mm_im);
/* MCcov.m:232-264 accumulation replayed from a path set, normalised by
 * max_itr (MCcov.m:293-299). */
void orc_mccov_replay(const orc_problem *pr, const orc_pathset *ps, int64_t max_itr,
                      double *ms_re, double *ms_im, double *mm_re, double *mm_im);

/* code/MCfield.m:1-254 : full run, normalised. Out: u nv x nl complex. */
void orc_mcfield(const orc_problem *pr, int64_t max_itr, double *u_re, double *u_im);
void orc_mcfield_sums(const orc_problem *pr, uint64_t p0, int64_t np, int nthreads,
                      double *u_re, double *u_im);
void orc_mcfield_replay(const orc_problem *pr, const orc_pathset *ps, int64_t max_itr,
                        double *u_re, double *u_im);

/* code/MCfieldOnWave.m:1-267 : the field of the illumination wave
 * sum_l Wl(l) * light l.  Out: u nv x 1 complex. */
void orc_mcfield_wave(const orc_problem *pr, const double *wl_re, const double *wl_im, int64_t max_itr,
                      double *u_re, double *u_im);

/* code/MCfieldFourierQuad.m grid set-up (lines 80-138).  Returns the view
 * count nv0 and the lateral grid size nx; fills the enlarged box. */
typedef struct {
    int nv, nv0, nx, nz, nfft, nx_pad, nva_pad;
    double box_stp, z_bin;
    double box_min[3], box_max[3];
    double gx0, gy0;
} orc_quad_grid;
int orc_quad_grid_setup(const double *box_min, const double *box_max, double v_max, double v_stp,
                        double lambda, orc_quad_grid *g);
/* The same with MCfieldFourierQuadOnWave.m:100-102 (nx takes the parity of
 * the view grid, so the transform size may be even). */
int orc_quad_grid_setup_wave(const double *box_min, const double *box_max, double v_max, double v_stp,
                             double lambda, orc_quad_grid *g);
/* code/MCfieldFourierQuad.m:1-303 (3-D).  l is 3 x nl, lmean0 may be NULL
 * (defaults to [0;0;1]).  amp: the walk's phase function, NULL for the
 * reference's isotropic walk (HG or tabulated: directions are sampled as in
 * MCcov, the scatter is unchanged).  Out: u nv0 x nv0 x nl complex,
 * column-major; returns the number of stored vertices (the normaliser). */
int64_t orc_mcfield_fourier_quad(double sigt, double albedo, const double *box_min, const double *box_max,
                                 const double *l, int nl, double v_max, double v_stp, int64_t max_itr,
                                 double lambda, int do_cbs, int smp_flg, const double *lmean0,
                                 double vsign, const orc_phase *amp, uint64_t seed, double *u_re, double *u_im);

/* code/MCfieldFourierQuadOnWave.m:1-337 : the lights are combined per vertex
 * with the weights Wl (length nl, wl_im may be NULL).  Out: u nv0 x nv0. */
int64_t orc_mcfield_fourier_quad_wave(double sigt, double albedo, const double *box_min, const double *box_max,
                                      const double *l, int nl, const double *wl_re, const double *wl_im,
                                      double v_max, double v_stp, int64_t max_itr, double lambda, int do_cbs,
                                      int smp_flg, const double *lmean0, double vsign, const orc_phase *amp,
                                      uint64_t seed, double *u_re, double *u_im);

/* code/MCCovHetro.m:1-276 (heterogeneous media).  pr supplies everything but
 * sigt/albedo grids (pr->albedo is used when m->albedo is NULL; pr->sigt is
 * ignored).  Outputs Ms, Mm (nv x nv x nl x nl) and mean1 (nv x nl); returns
 * the vertex count, or -1 for an all-zero sigt grid.  mean1's normaliser
 * sqrt(cs) is taken as 1: the reference reads ampfunc.pdf, a field the
 * tabulatedAmplitudeScatter struct does not have, whose evalAmp is already
 * normalised by cs. */
int64_t orc_mccov_hetro(const orc_problem *pr, const orc_medium *m, int64_t max_itr, double *ms_re, double *ms_im,
                        double *mm_re, double *mm_im, double *mean_re, double *mean_im);
/* unnormalised sums over path ids [p0, p0 + np) */
int64_t orc_mccov_hetro_sums(const orc_problem *pr, const orc_medium *m, uint64_t p0, int64_t np,
                             double *ms_re, double *ms_im, double *mm_re, double *mm_im, double *mean_re,
                             double *mean_im);
/* code/MCsampleHetro.m:1-227 (field sample of a voxel medium); m->has_lrad
 * must be 0.  Returns the vertex count or -1. */
int64_t orc_mcfield_hetro(const orc_problem *pr, const orc_medium *m, int64_t max_itr, double *u_re, double *u_im);
int64_t orc_mcfield_hetro_sums(const orc_problem *pr, const orc_medium *m, uint64_t p0, int64_t np, double *u_re,
                               double *u_im);
/* evalphaseattHetro.m:1-51 for n directions / positions (sign * d[:, i]). */
void orc_eval_phase_att_hetro(const orc_problem *pr, const orc_medium *m, const double *x, const double *d, int n,
                              int is_ff, double sign, double *out_re, double *out_im);
/* smpWoodcock.m:1-15 from x along w with the Philox stream of vertex 1 of
 * the given path (the test hook of the sampler). */
double orc_woodcock(const orc_problem *pr, const orc_medium *m, const double *x, const double *w, uint64_t path);
/* invCDFvec(cumsum(sigt(:)'/sum(sigt(:))), n_out) : icdf has n_out entries. */
int orc_hetro_icdf(const double *sigt, int nvox, int n_out, int32_t *icdf);

/* interface/scatter/tabulatedAmplitudeScatter.m:53-99 + coreFunctions/invCDF.m
 * Returns D (2 or 3) or a negative error code. Outputs are length n. */
int orc_tabulated_scatter(const double *theta, const double *f_re, const double *f_im, int n,
                          double *cs, double *sample_pdf, double *sample_cdf, double *sample_icdf,
                          double *eval_pdf, double *eval_amp_re, double *eval_amp_im);
void orc_inv_cdf(const double *cdf, int n, int dim, double *icdf);

#ifdef __cplusplus
}
#endif
#endif
