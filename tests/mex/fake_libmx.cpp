// fake_libmx.cpp -- TEST INFRASTRUCTURE: a host-memory implementation of the
// mx/mex API subset declared in mcspeckle_amd/mex/fake/mex.h, so the MEX
// gateways can be called from pytest without MATLAB.  mexErrMsgIdAndTxt
// throws; fake_call catches it and reports the identifier and message.
#include "mex.h"

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

struct mxArray_tag {
    mxClassID cls = mxDOUBLE_CLASS;
    std::vector<mwSize> dims;
    std::vector<double> re, im;
    bool complex = false;
    std::map<std::string, mxArray *> fields;
};

struct fake_error {
    std::string id, msg;
};

static uint64_t g_rng = 0x1234;
static std::string g_err_id, g_err_msg;

static double next_uniform()
{
    // splitmix64 -> [0,1)
    uint64_t z = (g_rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

extern "C" {

mwSize mxGetM(const mxArray *a) { return a->dims.empty() ? 0 : a->dims[0]; }
mwSize mxGetN(const mxArray *a)
{
    mwSize n = 1;
    for (size_t i = 1; i < a->dims.size(); ++i) n *= a->dims[i];
    return a->dims.size() < 2 ? 0 : n;
}
mwSize mxGetNumberOfElements(const mxArray *a)
{
    if (a->cls == mxSTRUCT_CLASS) return 1;
    mwSize n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->dims.size(); }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims.data(); }
int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsLogical(const mxArray *a) { return a->cls == mxLOGICAL_CLASS; }
int mxIsComplex(const mxArray *a) { return a->complex; }
int mxIsStruct(const mxArray *a) { return a->cls == mxSTRUCT_CLASS; }
int mxIsEmpty(const mxArray *a) { return a->cls != mxSTRUCT_CLASS && mxGetNumberOfElements(a) == 0; }
double *mxGetPr(const mxArray *a) { return const_cast<double *>(a->re.data()); }
double *mxGetPi(const mxArray *a) { return a->complex ? const_cast<double *>(a->im.data()) : nullptr; }
double mxGetScalar(const mxArray *a) { return a->re.empty() ? 0.0 : a->re[0]; }
mxArray *mxGetField(const mxArray *s, mwIndex, const char *name)
{
    auto it = s->fields.find(name);
    return it == s->fields.end() ? nullptr : it->second;
}
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c)
{
    mxArray *a = new mxArray_tag();
    a->cls = cls;
    a->dims.assign(dims, dims + ndim);
    mwSize n = mxGetNumberOfElements(a);
    a->re.assign(n, 0.0);
    a->complex = c == mxCOMPLEX;
    if (a->complex) a->im.assign(n, 0.0);
    return a;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    mwSize d[2] = {m, n};
    return mxCreateNumericArray(2, d, mxDOUBLE_CLASS, c);
}
void mxDestroyArray(mxArray *a)
{
    if (!a) return;
    for (auto &kv : a->fields) mxDestroyArray(kv.second);
    delete a;
}
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw fake_error{id ? id : "", buf};
}
int mexCallMATLAB(int nlhs, mxArray *plhs[], int nrhs, mxArray *prhs[], const char *name)
{
    if (strcmp(name, "rand") != 0 || nlhs != 1) mexErrMsgIdAndTxt("fake:mexCallMATLAB", "unsupported call %s", name);
    mwSize m = nrhs > 0 ? (mwSize)mxGetScalar(prhs[0]) : 1, n = nrhs > 1 ? (mwSize)mxGetScalar(prhs[1]) : m;
    mxArray *o = mxCreateDoubleMatrix(m, n, mxREAL);
    for (auto &x : o->re) x = next_uniform();
    plhs[0] = o;
    return 0;
}

// ---- test-side helpers
mxArray *fake_matrix(mwSize m, mwSize n, const double *re, const double *im)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, im ? mxCOMPLEX : mxREAL);
    if (re) memcpy(a->re.data(), re, sizeof(double) * m * n);
    if (im) memcpy(a->im.data(), im, sizeof(double) * m * n);
    return a;
}
mxArray *fake_struct(void)
{
    mxArray *a = new mxArray_tag();
    a->cls = mxSTRUCT_CLASS;
    a->dims = {1, 1};
    return a;
}
void fake_set_field(mxArray *s, const char *name, mxArray *v) { s->fields[name] = v; }
void fake_rng_reset(uint64_t seed) { g_rng = seed; }
double fake_rand(void) { return next_uniform(); }
const char *fake_err_id(void) { return g_err_id.c_str(); }
const char *fake_err_msg(void) { return g_err_msg.c_str(); }
typedef void (*mexfn)(int, mxArray **, int, const mxArray **);
// 0 on success, 1 when the gateway raised an error (see fake_err_id/msg)
int fake_call(void *fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs)
{
    try {
        ((mexfn)fn)(nlhs, plhs, nrhs, prhs);
    } catch (const fake_error &e) {
        g_err_id = e.id;
        g_err_msg = e.msg;
        return 1;
    }
    return 0;
}
}
