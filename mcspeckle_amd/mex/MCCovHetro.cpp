// MCCovHetro.cpp -- MEX gateway: [Ms,Mm,mean1] = MCCovHetro(sigt,albedo,
// box_min,box_max,box_bin,l,v,is_ff_l,is_ff_v,maxItr,lambda,doCBS,smpFlg,
// sct_type,ampfunc[,ampfunc0[,lrad]]) (code/MCCovHetro.m:1-37).  sigt is the
// cell grid (numel = prod(round((box_max-box_min)./box_bin))); albedo is a
// scalar or a grid of the same size (MCCovHetro.m:27-29).
#include "gateway_common.h"

using namespace scmc_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *id = "scmc:MCCovHetro";
    if (nrhs < 15) mexErrMsgIdAndTxt("MATLAB:narginchk:notEnoughInputs", "Not enough input arguments.");
    if (nrhs > 17) mexErrMsgIdAndTxt("MATLAB:narginchk:tooManyInputs", "Too many input arguments.");
    Args A;
    memset(&A.pr, 0, sizeof A.pr);
    scmc_problem &p = A.pr;
    const mxArray *sig = prhs[0], *alb = prhs[1];
    p.dim = read_box(prhs[2], prhs[3], p.box_min, p.box_max, id);
    scmc_medium med;
    read_medium(sig, alb, prhs[4], p, med, id);
    // MCCovHetro.m:19-24: 1 x N rows are angles
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
    if (nrhs == 17 && !mxIsEmpty(prhs[16])) {
        med.has_lrad = 1;
        med.lrad = scalar(prhs[16], id, "lrad");
    }
    p.seed = seed_from_matlab();
    const mwSize dims[4] = {(mwSize)p.nv, (mwSize)p.nv, (mwSize)p.nl, (mwSize)p.nl};
    mxArray *Ms = mxCreateNumericArray(4, dims, mxDOUBLE_CLASS, mxCOMPLEX);
    mxArray *Mm = mxCreateNumericArray(4, dims, mxDOUBLE_CLASS, mxCOMPLEX);
    mxArray *mean1 = mxCreateDoubleMatrix((mwSize)p.nv, (mwSize)p.nl, mxCOMPLEX);
    const int rc = scmc_mccov_hetro(&p, &med, (int64_t)max_itr, precision_from_env(), mxGetPr(Ms), mxGetPi(Ms),
                                    mxGetPr(Mm), mxGetPi(Mm), mxGetPr(mean1), mxGetPi(mean1), nullptr);
    if (rc != SCMC_OK) {
        mxDestroyArray(Ms);
        mxDestroyArray(Mm);
        mxDestroyArray(mean1);
        check_rc(rc, id);
    }
    plhs[0] = Ms;
    if (nlhs > 1)
        plhs[1] = Mm;
    else
        mxDestroyArray(Mm);
    if (nlhs > 2)
        plhs[2] = mean1;
    else
        mxDestroyArray(mean1);
}
