/*
 * scmc.h -- C-ABI of the MI355X speckle Monte Carlo core (libscmc.so).
 *
 * This is the drop-in boundary for the reference's engines.  In the
 * reference, scmc.m calls the engines by function name with positional
 * arguments (scmc.m:333,339,351-354,359-362 for MCcov; scmc.m:388,394,
 * 404-406,412-414 for MCfield); a MEX file of the same name placed beside
 * MCcov.m / MCfield.m shadows the .m file, and the MEX gateways
 * (mcspeckle_amd/mex/) translate mxArrays into the plain structs below.
 *
 * Conventions
 *  - Plain C types only.  Arrays are column-major like MATLAB.
 *  - Directions/positions are dim x n (dim = 2 or 3); the 1 x n angle form
 *    of MCcov.m:56-70 is converted to [sin; cos] rows by the caller.
 *  - Complex outputs are split (re, im) when the imaginary pointer is
 *    non-NULL (R2017b mxGetPr/mxGetPi) and interleaved (re, im, re, im, ...)
 *    in the real pointer when the imaginary pointer is NULL (R2018a+).
 *  - Every entry point returns 0 on success or a negative SCMC_E* code;
 *    scmc_last_error() returns the message (thread-local).  No entry point
 *    throws or aborts across the boundary.
 *  - Random numbers: Philox4x32-10, key = seed, counter = (vertex, group,
 *    path id); path ids are [path_offset, path_offset + max_itr), so shards
 *    of one run can be computed by different processes or GPUs and summed.
 */
#ifndef SCMC_H
#define SCMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCMC_ABI_VERSION 3

enum scmc_status {
    SCMC_OK = 0,
    SCMC_E_ARG = -1,      /* invalid argument (reference error messages)   */
    SCMC_E_DEVICE = -2,   /* HIP runtime / no device                        */
    SCMC_E_NOMEM = -3,    /* device allocation failed / over HBM budget     */
    SCMC_E_UNSUPPORTED = -4,
    SCMC_E_CANCELLED = -5,
    SCMC_E_BUSY = -6,     /* request not finished (scmc_poll)               */
    SCMC_E_NOTFOUND = -7
};

enum scmc_precision {
    SCMC_FP64 = 0,   /* FP64 everywhere (tracer, NEE, covariance on FP64 MFMA)      */
    SCMC_MIXED = 1   /* FP64 tracer + phase, FP32 NEE, FP16 hi/lo MFMA, FP32 acc      */
};

enum scmc_sct_type { SCMC_ISOTROPIC = 1, SCMC_TABULATED = 2, SCMC_HG = 3 };

/* ampfunc / ampfunc0 (MCcov.m:31-35): g for HG, a table for tabulated
 * (interface/scatter/tabulatedAmplitudeScatter.m:89-99: evalAmp, sampleIcdf). */
typedef struct {
    int32_t type;            /* scmc_sct_type                                  */
    double g;                /* HG anisotropy                                  */
    int32_t n_amp;           /* tabulated: numel(evalAmp)                      */
    const double *amp_re;    /* evalAmp real part                              */
    const double *amp_im;    /* evalAmp imaginary part (may be NULL = real)    */
    int32_t n_icdf;          /* tabulated: numel(sampleIcdf)                   */
    const double *icdf;      /* sampleIcdf                                     */
} scmc_phase;

/* The positional arguments of MCcov.m:1 / MCfield.m:1 plus the RNG key. */
typedef struct {
    int32_t dim;             /* 2 or 3 (numel(box_min))                        */
    double sigt;             /* extinction coefficient                         */
    double albedo;
    double lambda;
    double box_min[3];
    double box_max[3];
    int32_t nl, nv;
    const double *l;         /* dim x nl : light directions or positions       */
    const double *v;         /* dim x nv : view directions or positions        */
    int32_t is_ff_l, is_ff_v;
    int32_t do_cbs;
    int32_t smp_flg;         /* 1 uniform, 2 exponential first scatter         */
    scmc_phase amp;          /* sct_type + ampfunc                             */
    int32_t has_amp0;        /* ampfunc0 given                                 */
    scmc_phase amp0;         /* type field ignored: sct_type is amp.type       */
    int32_t has_lmean0;      /* MCfield's optional lmean0                      */
    double lmean0[3];
    uint64_t seed;
    uint64_t path_offset;    /* first Philox path id                           */
} scmc_problem;

/* Voxel medium of MCCovHetro.m:1 (sigt, albedo grids over box_bin cells). */
typedef struct {
    int32_t nb[3];           /* cells per axis: round((box_max-box_min)./box_bin)  */
    double box_bin[3];       /* cell size                                          */
    const double *sigt;      /* prod(nb) extinction values, column-major (x first) */
    const double *albedo;    /* prod(nb) albedos, or NULL: pr->albedo everywhere   */
    int32_t has_lrad;        /* optional lrad argument (MCCovHetro.m:45)           */
    double lrad;
} scmc_medium;

/* Per-solve statistics returned by the engines (may be passed as NULL). */
typedef struct {
    int64_t paths;
    int64_t vertices;        /* total scattering vertices K                    */
    int64_t chunks;
    double ms_trace;         /* tracer (count + emit)                          */
    double ms_nee;           /* next-event estimation factors                  */
    double ms_accum;         /* covariance update / field accumulation         */
    double ms_total;         /* wall time of the call                          */
    double fp16_scale_log2;  /* mixed: last chunk's power-of-two scale         */
    double ms_herk;          /* covariance update launches, HIP-event time on
                                the solve stream: mixed, one launch per chunk
                                for multiple and single scattering; FP64, the
                                multiple-scattering launches                   */
    int64_t herk_launches;   /* number of those launches                        */
    double herk_flops;       /* their algorithmic flops, 8 N^2 K (K: the rank-1
                                terms those launches add)                      */
    double ms_expand;        /* mixed: FP16 hi/lo row expansion before them     */
    int64_t nonfinite;       /* vertices (covariance) or paths (field) whose
                                factors came out NaN/Inf (e.g. a NaN path weight
                                in a replayed set); the results then hold NaN,
                                as the reference's would                        */
    double ms_queued;        /* serving: submit to admission (the rest of
                                ms_total is the request's own run)              */
} scmc_stats;

/* Replayable path set (vertex-major), the format of the oracle's sampler and
 * of scmc_trace.  offset has n_paths + 1 entries. */
typedef struct {
    int64_t n_paths, n_vertices;
    const int64_t *offset;
    const double *x1, *px, *w1, *w0p_re, *w0p_im;     /* per path  */
    const double *x, *ow, *weight, *xi;               /* per vertex */
} scmc_pathset;

/* ---- engines ---------------------------------------------------------- */

/* [Ms,Mm] = MCcov(...) -- code/MCcov.m:1-299.  Outputs nv x nv x nl x nl.
 * Normalised exactly as MCcov.m:292-299 (CBS halving, V*sigt/maxItr). */
int scmc_mccov(const scmc_problem *pr, int64_t max_itr, int32_t precision,
               double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats);

/* u = MCfield(...) -- code/MCfield.m:1-254.  Output nv x nl, normalised as
 * MCfield.m:252-254. */
int scmc_mcfield(const scmc_problem *pr, int64_t max_itr, int32_t precision,
                 double *u_re, double *u_im, scmc_stats *stats);

/* u = MCfieldFourierQuad(sigt,albedo,box_min,box_max,l,v_max,v_stp,maxItr,
 * lambda,doCBS,smpFlg,lmean0,vsign) -- code/MCfieldFourierQuad.m:1-303.
 * 3-D only.  pr supplies sigt, albedo, box, l (3 x nl), lambda, do_cbs,
 * smp_flg, lmean0 (has_lmean0), seed; pr->amp selects the scattering
 * function of the walk (isotropic reproduces the reference; HG/tabulated are
 * an extension).  Output nv0 x nv0 x nl with nv0 from scmc_quad_grid. */
int scmc_mcfield_fourier_quad(const scmc_problem *pr, double v_max, double v_stp, int64_t max_itr,
                              double vsign, int32_t precision, double *u_re, double *u_im,
                              scmc_stats *stats);
/* Grid sizes of MCfieldFourierQuad.m:80-121 for the given arguments. */
int scmc_quad_grid(const double box_min[3], const double box_max[3], double v_max, double v_stp,
                   double lambda, int32_t *nv0, int32_t *nx, int32_t *nz);

/* u = MCfieldOnWave(sigt,albedo,box_min,box_max,l,v,is_ff_l,is_ff_v,maxItr,
 * lambda,doCBS,smpFlg,sct_type,ampfunc,ampfunc0,lmean0,Wl) --
 * code/MCfieldOnWave.m:1-267: the field of the illumination wave
 * sum_l Wl(l) * light l.  wl_re/wl_im hold nl weights (wl_im may be NULL).
 * Output nv x 1.  The single-scattering term sums over lights for every
 * scattering type (the reference's isotropic branch, MCfieldOnWave.m:214,
 * only does so for Nl = 1). */
int scmc_mcfield_wave(const scmc_problem *pr, const double *wl_re, const double *wl_im, int64_t max_itr,
                      int32_t precision, double *u_re, double *u_im, scmc_stats *stats);

/* u = MCfieldFourierQuadOnWave(sigt,albedo,box_min,box_max,l,v_max,v_stp,
 * maxItr,lambda,doCBS,smpFlg,lmean0,Wl,vsign) --
 * code/MCfieldFourierQuadOnWave.m:1-337.  As scmc_mcfield_fourier_quad with
 * the lights combined per vertex with the weights Wl (:281-302), the grid
 * parity fix of :100-102 and the sign conventions of :119, :143, :302.
 * Output nv0 x nv0 with nv0 from scmc_quad_grid_wave. */
int scmc_mcfield_fourier_quad_wave(const scmc_problem *pr, const double *wl_re, const double *wl_im,
                                   double v_max, double v_stp, int64_t max_itr, double vsign,
                                   int32_t precision, double *u_re, double *u_im, scmc_stats *stats);
/* Grid sizes of MCfieldFourierQuadOnWave.m:84-131. */
int scmc_quad_grid_wave(const double box_min[3], const double box_max[3], double v_max, double v_stp,
                        double lambda, int32_t *nv0, int32_t *nx, int32_t *nz);

/* [Ms,Mm,mean1] = MCCovHetro(sigt,albedo,box_min,box_max,box_bin,l,v,is_ff_l,
 * is_ff_v,maxItr,lambda,doCBS,smpFlg,sct_type,ampfunc[,ampfunc0[,lrad]]) --
 * code/MCCovHetro.m:1-276 (heterogeneous media: Woodcock tracking, cell-by-
 * cell attenuation).  pr supplies the box, sources, flags, phase functions
 * and seed (pr->sigt is not used; pr->albedo when med->albedo is NULL;
 * has_lmean0 is ignored).  Ms, Mm are nv x nv x nl x nl, normalised as
 * MCCovHetro.m:248-265 (Mm/2 with CBS, /maxItr); mean1 is nv x nl
 * (MCCovHetro.m:271-276 with cs = 1).  Any output pointer may be NULL. */
int scmc_mccov_hetro(const scmc_problem *pr, const scmc_medium *med, int64_t max_itr, int32_t precision,
                     double *ms_re, double *ms_im, double *mm_re, double *mm_im,
                     double *mean_re, double *mean_im, scmc_stats *stats);

/* u = MCsampleHetro(sigt,albedo,box_min,box_max,box_bin,l,v,is_ff_l,is_ff_v,
 * maxItr,lambda,doCBS,smpFlg,sct_type,ampfunc[,ampfunc0[,lmean0]]) --
 * code/MCsampleHetro.m:1-227 (the field sample of a voxel medium: the walks
 * of MCCovHetro, E by the cell-by-cell march, u*sqrt(1/maxItr),
 * MCsampleHetro.m:222).  pr as for scmc_mccov_hetro, except that has_lmean0
 * / lmean0 are honoured (MCsampleHetro.m:35-40); med->has_lrad must be 0
 * (MCsampleHetro has no lrad argument).  Output nv x nl.  The debug pairs
 * (dbgpair, corrM) are not provided.  Both precision modes evaluate E in
 * FP64. */
int scmc_mcfield_hetro(const scmc_problem *pr, const scmc_medium *med, int64_t max_itr, int32_t precision,
                       double *u_re, double *u_im, scmc_stats *stats);

/* u = MCfield(...) without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, walks and output as scmc_mcfield (homogeneous media); every
 * device buffer lies in one file-scope device array of 128 MiB per GPU
 * (held until the process ends), so the call issues no hipMalloc/hipFree.
 * Problems whose fixed part (accumulators and per-view work, about
 * 1 KiB x nv x nl) leaves no room for 64 paths are rejected with SCMC_E_ARG
 * before any device work.  Calls in one process run one after another; calls
 * from several threads wait for each other, and the array keeps the last
 * call's contents (each call clears what must start at zero). */
int scmc_mcfield_noalloc(const scmc_problem *pr, int64_t max_itr, int32_t precision, double *u_re, double *u_im,
                         scmc_stats *stats);

/* ---- sessions: accumulate path ranges, read results at any time -------- */
/* A session owns the device accumulators of one MCcov (kind 0) or MCfield
 * (kind 1) problem.  scmc_session_run adds paths [first, first + n) (ids
 * relative to pr->path_offset) and returns when they are accumulated;
 * scmc_session_result normalises by max_itr without consuming the sums, so a
 * run can be checkpointed, resumed or sharded (the parfor repeats of
 * scmc.m:343-367 are disjoint path ranges of one session).  Calls on one
 * session from several threads run one after another; destroying a session
 * while another thread's call on it is in progress releases the handle at
 * once and the session itself when that call returns. */
int scmc_session_create(const scmc_problem *pr, int32_t kind, int32_t precision, int64_t *session);
int scmc_session_run(int64_t session, int64_t first, int64_t n_paths);
int scmc_session_result(int64_t session, int64_t max_itr, double *a_re, double *a_im, double *b_re, double *b_im);
int scmc_session_stats(int64_t session, scmc_stats *stats);
int scmc_session_destroy(int64_t session);

/* ---- replay / export (path-set parity harness) ------------------------ */

/* Counts vertices of paths [path_offset, path_offset + n_paths) and, when
 * the output arrays are non-NULL, writes the path set sampled on the GPU.
 * *n_vertices receives the total.  Arrays sized as scmc_pathset. */
int scmc_trace(const scmc_problem *pr, int64_t n_paths, int64_t *n_vertices, int64_t *offset,
               double *x1, double *px, double *w1, double *w0p_re, double *w0p_im,
               double *x, double *ow, double *weight, double *xi);
/* MCcov contributions (MCcov.m:181-264) of a given path set, normalised by
 * max_itr; the GPU covariance kernels with the tracer bypassed.  stats (may
 * be NULL) receives the path / vertex counts and the non-finite count. */
int scmc_mccov_replay(const scmc_problem *pr, const scmc_pathset *ps, int64_t max_itr, int32_t precision,
                      double *ms_re, double *ms_im, double *mm_re, double *mm_im, scmc_stats *stats);
/* MCfield contributions (MCfield.m:194-222) of a given path set. */
int scmc_mcfield_replay(const scmc_problem *pr, const scmc_pathset *ps, int64_t max_itr, int32_t precision,
                        double *u_re, double *u_im, scmc_stats *stats);

/* ---- serving: continuous batching of covariance and field solves -------- */
/* A host scheduler with max_active slots: each admitted request runs on its
 * slot's own HIP streams, so the work of all admitted requests overlaps on the
 * GPU; requests are admitted in arrival order as slots and HBM budget free
 * up (a later request that fits may pass one that does not). */

typedef struct {
    int32_t device;          /* HIP device ordinal                             */
    int32_t max_active;      /* max requests in flight                          */
    double hbm_budget_gb;    /* admission budget for request buffers            */
    int64_t paths_per_slice; /* paths per request per scheduling round (0=auto) */
    int32_t shared_launches; /* 1: the mixed updates of small requests (N <= 1024)
                                running at the same time share launches;
                                0 (default): each request launches its own     */
    int32_t resident_loop;   /* 1: the mixed updates of every request go to one
                                resident update kernel that takes their work
                                items from a job table in host memory as the
                                requests publish them (no launch per update);
                                0 (default): each request launches its own     */
    int32_t resident_workgroups; /* workgroups of the resident kernel
                                (0: one per CU, leaving the other half of each
                                CU to the kernels that prepare the updates)    */
} scmc_server_config;

int scmc_server_start(const scmc_server_config *cfg);
int scmc_server_stop(void);
/* Submit an MCcov request (mixed or FP64).  The problem's arrays are copied
 * before return.  *request_id receives the handle. */
int scmc_submit_mccov(const scmc_problem *pr, int64_t max_itr, int32_t precision, int64_t *request_id);
/* Submit an MCfield request (MCfield.m:1-254, mixed or FP64); its field u
 * (nv x nl, layout as scmc_mcfield) comes back in the first output pair of
 * scmc_poll / scmc_wait, and the second pair is left untouched. */
int scmc_submit_mcfield(const scmc_problem *pr, int64_t max_itr, int32_t precision, int64_t *request_id);
/* 0 when finished (results copied out, handle released), SCMC_E_BUSY when
 * still running.  Output layout as scmc_mccov (covariance requests) or
 * scmc_mcfield (field requests, first pair); timing in *stats (ms_total =
 * submit-to-finish latency, ms_queued = submit to admission). */
int scmc_poll(int64_t request_id, double *ms_re, double *ms_im, double *mm_re, double *mm_im,
              scmc_stats *stats);
int scmc_wait(int64_t request_id, double *ms_re, double *ms_im, double *mm_re, double *mm_im,
              scmc_stats *stats);
int scmc_cancel(int64_t request_id);
/* Bytes of device memory a request with far-field sources would hold while
 * admitted (a near-field request also holds its single-scatter rows). */
int64_t scmc_request_footprint(int32_t nv, int32_t nl, int32_t precision);

/* ---- misc -------------------------------------------------------------- */

const char *scmc_last_error(void);
const char *scmc_version(void);
/* Number of visible HIP devices (0 on a machine without a GPU). */
int scmc_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
