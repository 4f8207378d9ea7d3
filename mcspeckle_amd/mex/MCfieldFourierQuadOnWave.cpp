// MCfieldFourierQuadOnWave.cpp -- MEX gateway: u = MCfieldFourierQuadOnWave(
// sigt,albedo,box_min,box_max,l,v_max,v_stp,maxItr,lambda,doCBS,smpFlg,
// lmean0,Wl[,vsign]) (code/MCfieldFourierQuadOnWave.m:1, 44-74: 3-D only,
// lmean0 [] -> [0;0;1], vsign default -1, one complex weight per light).
#include "gateway_common.h"

using namespace scmc_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    const char *id = "scmc:MCfieldFourierQuadOnWave";
    if (nrhs < 13) mexErrMsgIdAndTxt("MATLAB:narginchk:notEnoughInputs", "Not enough input arguments.");
    if (nrhs > 14) mexErrMsgIdAndTxt("MATLAB:narginchk:tooManyInputs", "Too many input arguments.");
    Args A;
    memset(&A.pr, 0, sizeof A.pr);
    scmc_problem &p = A.pr;
    p.sigt = scalar(prhs[0], id, "sigt");
    p.albedo = scalar(prhs[1], id, "albedo");
    p.dim = read_box(prhs[2], prhs[3], p.box_min, p.box_max, id);
    if (p.dim != 3) mexErrMsgIdAndTxt(id, "MCfieldFourierQuadOnWave is 3-D only");
    p.nl = read_sources(prhs[4], 3, A.l, id, "Invalid light source input");
    p.l = A.l.data();
    const double v_max = scalar(prhs[5], id, "v_max"), v_stp = scalar(prhs[6], id, "v_stp");
    const double max_itr = scalar(prhs[7], id, "maxItr");
    p.lambda = scalar(prhs[8], id, "lambda");
    p.do_cbs = scalar(prhs[9], id, "doCBS") != 0;
    p.smp_flg = (int32_t)scalar(prhs[10], id, "smpFlg");
    p.amp.type = SCMC_ISOTROPIC;
    if (!mxIsEmpty(prhs[11])) {
        p.has_lmean0 = 1;
        read_vector(prhs[11], 3, p.lmean0, id, "lmean0");
    }
    std::vector<double> wr, wi;
    read_weights(prhs[12], p.nl, wr, wi, id);
    const double vsign = (nrhs == 14) ? scalar(prhs[13], id, "vsign") : -1.0;
    p.is_ff_l = 1;
    p.is_ff_v = 1;
    p.seed = seed_from_matlab();
    int32_t nv0 = 0, nx = 0, nz = 0;
    check_rc(scmc_quad_grid_wave(p.box_min, p.box_max, v_max, v_stp, p.lambda, &nv0, &nx, &nz), id);
    mxArray *u = mxCreateDoubleMatrix((mwSize)nv0, (mwSize)nv0, mxCOMPLEX);
    const int rc = scmc_mcfield_fourier_quad_wave(&p, wr.data(), wi.data(), v_max, v_stp, (int64_t)max_itr, vsign,
                                                  precision_from_env(), mxGetPr(u), mxGetPi(u), nullptr);
    if (rc != SCMC_OK) {
        mxDestroyArray(u);
        check_rc(rc, id);
    }
    plhs[0] = u;
}
