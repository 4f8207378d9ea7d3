/*
 * mex.h -- this project's own minimal declaration of the subset of the
 * documented MATLAB C MEX / mx API that the scmc gateways use (separate
 * real/imaginary storage, as in R2017b).  Used to compile and test the
 * gateways without MATLAB (see tests/mex/fake_libmx.cpp); with MATLAB, the
 * `mex` command compiles the same gateway sources against MATLAB's mex.h.
 */
#ifndef SCMC_FAKE_MEX_H
#define SCMC_FAKE_MEX_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef enum { mxUNKNOWN_CLASS = 0, mxSTRUCT_CLASS = 2, mxLOGICAL_CLASS = 3, mxDOUBLE_CLASS = 6 } mxClassID;

mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
mwSize mxGetNumberOfElements(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsLogical(const mxArray *a);
int mxIsComplex(const mxArray *a);
int mxIsStruct(const mxArray *a);
int mxIsEmpty(const mxArray *a);
double *mxGetPr(const mxArray *a);
double *mxGetPi(const mxArray *a);
double mxGetScalar(const mxArray *a);
mxArray *mxGetField(const mxArray *s, mwIndex i, const char *name);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c);
void mxDestroyArray(mxArray *a);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexCallMATLAB(int nlhs, mxArray *plhs[], int nrhs, mxArray *prhs[], const char *name);

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif
#endif
