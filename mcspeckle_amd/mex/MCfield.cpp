// MCfield.cpp -- MEX gateway: u = MCfield(sigt,albedo,box_min,box_max,l,v,
// is_ff_l,is_ff_v,maxItr,lambda,doCBS,smpFlg,sct_type,ampfunc[,ampfunc0
// [,lmean0]]).  Argument checks follow code/MCfield.m:41-82.
#include "gateway_common.h"

using namespace scmc_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    const char *id = "scmc:MCfield";
    if (nrhs < 14) mexErrMsgIdAndTxt("MATLAB:narginchk:notEnoughInputs", "Not enough input arguments.");
    if (nrhs > 16) mexErrMsgIdAndTxt("MATLAB:narginchk:tooManyInputs", "Too many input arguments.");
    Args A;
    memset(&A.pr, 0, sizeof A.pr);
    scmc_problem &p = A.pr;
    p.sigt = scalar(prhs[0], id, "sigt");
    p.albedo = scalar(prhs[1], id, "albedo");
    p.dim = read_box(prhs[2], prhs[3], p.box_min, p.box_max, id);
    p.nl = read_sources(prhs[4], p.dim, A.l, id, "Invalid light source input");
    p.nv = read_sources(prhs[5], p.dim, A.v, id, "Invalid view source input");
    p.l = A.l.data();
    p.v = A.v.data();
    p.is_ff_l = scalar(prhs[6], id, "is_ff_l") != 0;
    p.is_ff_v = scalar(prhs[7], id, "is_ff_v") != 0;
    const double max_itr = scalar(prhs[8], id, "maxItr");
    p.lambda = scalar(prhs[9], id, "lambda");
    p.do_cbs = scalar(prhs[10], id, "doCBS") != 0;
    p.smp_flg = (int32_t)scalar(prhs[11], id, "smpFlg");
    const int type = (int)scalar(prhs[12], id, "sct_type");
    read_phase(prhs[13], type, p.amp, A.amp_re, A.amp_im, A.icdf, id);
    if (nrhs >= 15 && !mxIsEmpty(prhs[14])) {
        p.has_amp0 = 1;
        read_phase(prhs[14], type, p.amp0, A.amp0_re, A.amp0_im, A.icdf0, id);
    }
    if (nrhs == 16 && !mxIsEmpty(prhs[15])) {
        p.has_lmean0 = 1;
        read_vector(prhs[15], p.dim, p.lmean0, id, "lmean0");
    }
    p.seed = seed_from_matlab();
    mxArray *u = mxCreateDoubleMatrix((mwSize)p.nv, (mwSize)p.nl, mxCOMPLEX);
    const int rc = scmc_mcfield(&p, (int64_t)max_itr, precision_from_env(), mxGetPr(u), mxGetPi(u), nullptr);
    if (rc != SCMC_OK) {
        mxDestroyArray(u);
        check_rc(rc, id);
    }
    plhs[0] = u;
}
