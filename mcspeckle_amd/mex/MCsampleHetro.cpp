// MCsampleHetro.cpp -- MEX gateway: u = MCsampleHetro(sigt,albedo,box_min,
// box_max,box_bin,l,v,is_ff_l,is_ff_v,maxItr,lambda,doCBS,smpFlg,sct_type,
// ampfunc[,ampfunc0[,lmean0]]) (code/MCsampleHetro.m:1-40).  sigt is the cell
// grid, albedo a scalar or a grid of the same size (MCsampleHetro.m:25-27).
// The debug pairs (dbgpair, and the corrM output) are not provided.
#include "gateway_common.h"

using namespace scmc_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *id = "scmc:MCsampleHetro";
    if (nrhs < 15) mexErrMsgIdAndTxt("MATLAB:narginchk:notEnoughInputs", "Not enough input arguments.");
    if (nrhs > 18) mexErrMsgIdAndTxt("MATLAB:narginchk:tooManyInputs", "Too many input arguments.");
    if ((nrhs == 18 && !mxIsEmpty(prhs[17])) || nlhs > 1)
        mexErrMsgIdAndTxt(id, "dbgpair / corrM are not provided by this build");
    Args A;
    memset(&A.pr, 0, sizeof A.pr);
    scmc_problem &p = A.pr;
    p.dim = read_box(prhs[2], prhs[3], p.box_min, p.box_max, id);
    scmc_medium med;
    read_medium(prhs[0], prhs[1], prhs[4], p, med, id);
    // MCsampleHetro.m:19-24: 1 x N rows are angles
    p.nl = read_sources(prhs[5], p.dim, A.l, id, "Invalid light source input");
    p.nv = read_sources(prhs[6], p.dim, A.v, id, "Invalid view source input");
    p.l = A.l.data();
    p.v = A.v.data();
    p.is_ff_l = scalar(prhs[7], id, "is_ff_l") != 0;
    p.is_ff_v = scalar(prhs[8], id, "is_ff_v") != 0;
    const double max_itr = scalar(prhs[9], id, "maxItr");
    p.lambda = scalar(prhs[10], id, "lambda");
    p.do_cbs = scalar(prhs[11], id, "doCBS") != 0;
    p.smp_flg = (int32_t)scalar(prhs[12], id, "smpFlg");
    const int type = (int)scalar(prhs[13], id, "sct_type");
    read_phase(prhs[14], type, p.amp, A.amp_re, A.amp_im, A.icdf, id);
    if (nrhs >= 16 && !mxIsEmpty(prhs[15])) {
        p.has_amp0 = 1;
        read_phase(prhs[15], type, p.amp0, A.amp0_re, A.amp0_im, A.icdf0, id);
    }
    // MCsampleHetro.m:35-40
    if (nrhs >= 17 && !mxIsEmpty(prhs[16])) {
        p.has_lmean0 = 1;
        read_vector(prhs[16], p.dim, p.lmean0, id, "lmean0");
    }
    p.seed = seed_from_matlab();
    mxArray *u = mxCreateDoubleMatrix((mwSize)p.nv, (mwSize)p.nl, mxCOMPLEX);
    const int rc = scmc_mcfield_hetro(&p, &med, (int64_t)max_itr, precision_from_env(), mxGetPr(u), mxGetPi(u), nullptr);
    if (rc != SCMC_OK) {
        mxDestroyArray(u);
        check_rc(rc, id);
    }
    plhs[0] = u;
}
