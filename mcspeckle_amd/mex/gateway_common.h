// gateway_common.h -- argument handling shared by the MEX gateways.
//
// Restates the argument conventions of code/MCcov.m:45-70 (box, 2-D angle
// rows, source counts) and maps the MATLAB values onto the scmc_problem of
// include/scmc.h.  Errors raise through mexErrMsgIdAndTxt with the
// reference's messages.
#pragma once
#include "mex.h"
#include "scmc.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace scmc_mex {

struct Args {
    scmc_problem pr;
    std::vector<double> l, v, amp_re, amp_im, icdf, amp0_re, amp0_im, icdf0;
};

inline double scalar(const mxArray *a, const char *id, const char *what)
{
    if (!a || mxGetNumberOfElements(a) < 1) mexErrMsgIdAndTxt(id, "%s must be a scalar", what);
    return mxGetScalar(a);
}

// A real double array with exactly n elements (lmean0: MCfield.m:77-82 reads
// it as a dim-vector), copied into out; anything else raises.
inline void read_vector(const mxArray *a, int n, double *out, const char *id, const char *what)
{
    if (!mxIsDouble(a) || mxIsComplex(a) || (int)mxGetNumberOfElements(a) != n)
        mexErrMsgIdAndTxt(id, "%s must be a real double vector of %d elements", what, n);
    const double *p = mxGetPr(a);
    for (int j = 0; j < n; ++j) out[j] = p[j];
}

// MCcov.m:48-53
inline int read_box(const mxArray *bmin, const mxArray *bmax, double *lo, double *hi, const char *id)
{
    const mwSize n = mxGetNumberOfElements(bmax);
    if ((n != 2 && n != 3) || mxGetN(bmax) != 1 || mxGetM(bmin) != mxGetM(bmax) || mxGetN(bmin) != mxGetN(bmax))
        mexErrMsgIdAndTxt(id, "Invalid box size");
    const int dim = (int)mxGetM(bmin);
    for (int j = 0; j < dim; ++j) {
        lo[j] = mxGetPr(bmin)[j];
        hi[j] = mxGetPr(bmax)[j];
    }
    return dim;
}

// MCcov.m:56-70: dim x N, or (2-D) 1 x N angles -> [sin; cos]
inline int read_sources(const mxArray *a, int dim, std::vector<double> &out, const char *id, const char *msg)
{
    const mwSize m = mxGetM(a), n = mxGetN(a);
    const double *p = mxGetPr(a);
    if ((int)m == dim) {
        out.assign(p, p + m * n);
    } else if (dim == 2 && m == 1) {
        out.resize(2 * n);
        for (mwSize i = 0; i < n; ++i) {
            out[2 * i] = sin(p[i]);
            out[2 * i + 1] = cos(p[i]);
        }
    } else {
        mexErrMsgIdAndTxt(id, "%s", msg);
    }
    return (int)n;
}

// ampfunc: 0 (isotropic), g (HG) or a struct with evalAmp and sampleIcdf
inline void read_phase(const mxArray *a, int type, scmc_phase &ph, std::vector<double> &re, std::vector<double> &im,
                       std::vector<double> &ic, const char *id)
{
    memset(&ph, 0, sizeof ph);
    ph.type = type;
    if (type == SCMC_HG) ph.g = scalar(a, id, "g");
    if (type != SCMC_TABULATED) return;
    if (!mxIsStruct(a)) mexErrMsgIdAndTxt(id, "tabulated ampfunc must be a struct");
    const mxArray *ea = mxGetField(a, 0, "evalAmp"), *si = mxGetField(a, 0, "sampleIcdf");
    if (!ea || !si) mexErrMsgIdAndTxt(id, "ampfunc must have fields evalAmp and sampleIcdf");
    const mwSize n = mxGetNumberOfElements(ea), m = mxGetNumberOfElements(si);
    re.assign(mxGetPr(ea), mxGetPr(ea) + n);
    im.assign(n, 0.0);
    if (mxIsComplex(ea)) im.assign(mxGetPi(ea), mxGetPi(ea) + n);
    ic.assign(mxGetPr(si), mxGetPr(si) + m);
    ph.n_amp = (int32_t)n;
    ph.amp_re = re.data();
    ph.amp_im = im.data();
    ph.n_icdf = (int32_t)m;
    ph.icdf = ic.data();
}

// Philox key drawn from MATLAB's global stream, so rng(seed) in scmc.m:329
// fixes the result and successive calls differ as they do with rand.
// Wl of the illumination-wave variants: nl complex weights, any shape with
// numel == nl (real or complex).
inline void read_weights(const mxArray *a, int nl, std::vector<double> &re, std::vector<double> &im, const char *id)
{
    if ((int)mxGetNumberOfElements(a) != nl) mexErrMsgIdAndTxt(id, "Wl must have one weight per light");
    const double *pr = mxGetPr(a), *pi = mxGetPi(a);
    re.assign(pr, pr + nl);
    if (pi)
        im.assign(pi, pi + nl);
    else
        im.assign((size_t)nl, 0.0);
}

inline uint64_t seed_from_matlab()
{
    mxArray *out = nullptr;
    mxArray *in[2] = {mxCreateDoubleMatrix(1, 1, mxREAL), mxCreateDoubleMatrix(1, 1, mxREAL)};
    mxGetPr(in[0])[0] = 1;
    mxGetPr(in[1])[0] = 2;
    mexCallMATLAB(1, &out, 2, in, "rand");
    const double *u = mxGetPr(out);
    uint64_t lo = (uint64_t)(u[0] * 4294967296.0), hi = (uint64_t)(u[1] * 4294967296.0);
    mxDestroyArray(in[0]);
    mxDestroyArray(in[1]);
    mxDestroyArray(out);
    return lo | (hi << 32);
}

// Precision of the GPU path: SCMC_PRECISION=mixed selects the FP16 hi/lo
// MFMA covariance; the default is FP64 (the reference's arithmetic).
inline int precision_from_env()
{
    const char *e = getenv("SCMC_PRECISION");
    return (e && strcmp(e, "mixed") == 0) ? SCMC_MIXED : SCMC_FP64;
}

inline void check_rc(int rc, const char *id)
{
    if (rc != SCMC_OK) mexErrMsgIdAndTxt(id, "%s", scmc_last_error());
}

// The voxel medium of MCCovHetro.m / MCsampleHetro.m: box_bin (dim sizes),
// the sigt grid (numel = prod(round((box_max-box_min)./box_bin)), MATLAB
// order) and albedo as a scalar (into p.albedo) or a grid of the same size
// (MCCovHetro.m:27-29, MCsampleHetro.m:25-27).  p.dim and the box are read.
inline void read_medium(const mxArray *sig, const mxArray *alb, const mxArray *bin, scmc_problem &p,
                        scmc_medium &med, const char *id)
{
    if ((int)mxGetNumberOfElements(bin) != p.dim) mexErrMsgIdAndTxt(id, "Invalid box_bin size");
    memset(&med, 0, sizeof med);
    const double *bb = mxGetPr(bin);
    int64_t ncell = 1;
    for (int j = 0; j < 3; ++j) {
        med.nb[j] = 1;
        med.box_bin[j] = 1.0;
    }
    for (int j = 0; j < p.dim; ++j) {
        med.box_bin[j] = bb[j];
        med.nb[j] = (int32_t)std::lround((p.box_max[j] - p.box_min[j]) / bb[j]);
        ncell *= med.nb[j];
    }
    if ((int64_t)mxGetNumberOfElements(sig) != ncell) mexErrMsgIdAndTxt(id, "sigt does not match the cell grid");
    med.sigt = mxGetPr(sig);
    p.sigt = 1.0;
    if (mxGetNumberOfElements(alb) == 1) {
        p.albedo = mxGetPr(alb)[0];
    } else {
        if ((int64_t)mxGetNumberOfElements(alb) != ncell) mexErrMsgIdAndTxt(id, "albedo does not match the cell grid");
        med.albedo = mxGetPr(alb);
        p.albedo = 1.0;
    }
}

} // namespace scmc_mex
