// A minimal stand-in for the MATLAB host that MEX gateways run in: it provides
// the mx*/mex* functions of include/mex_api.h so the gateways can be loaded and
// driven from Python tests.  Test infrastructure only.
//
// Arrays are column-major with up to 8 dimensions; structs are 1x1 with named
// fields.  mexErrMsgTxt throws, and fm_call turns that into a return code and
// message, like MATLAB's error() at the prompt.  mexPrintf output and
// mexEvalString commands are captured.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex_api.h"

struct mxArray_tag {
    mxClassID cls = mxDOUBLE_CLASS;
    std::vector<mwSize> dims;
    std::vector<unsigned char> data;
    std::vector<std::string> names;      // struct field names
    std::vector<mxArray *> values;       // struct field values (owned)
};

namespace {
std::string g_out;
std::vector<std::string> g_eval;
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

size_t elsize(mxClassID c)
{
    switch (c) {
    case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
    case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
    case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
    default: return 1;
    }
}

size_t numel(const mxArray *a)
{
    if (a->cls == mxSTRUCT_CLASS) return 1;
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}

mxArray *make_numeric(mwSize nd, const mwSize *dims, mxClassID c)
{
    mxArray *a = new mxArray_tag();
    a->cls = c;
    a->dims.assign(dims, dims + nd);
    while (a->dims.size() > 2 && a->dims.back() == 1) a->dims.pop_back();
    while (a->dims.size() < 2) a->dims.push_back(1);
    a->data.assign(numel(a) * elsize(c), 0);
    return a;
}
}  // namespace

extern "C" {

// ---- mx API -------------------------------------------------------------
mxArray *mxGetField(const mxArray *pa, mwIndex i, const char *name)
{
    if (!pa || pa->cls != mxSTRUCT_CLASS || i != 0) return nullptr;
    for (size_t k = 0; k < pa->names.size(); k++)
        if (pa->names[k] == name) return pa->values[k];
    return nullptr;
}

double mxGetScalar(const mxArray *pa)
{
    if (!pa || numel(pa) == 0 || pa->cls == mxSTRUCT_CLASS) return 0.0;
    const unsigned char *p = pa->data.data();
    switch (pa->cls) {
    case mxDOUBLE_CLASS: return *(const double *)p;
    case mxSINGLE_CLASS: return *(const float *)p;
    case mxINT32_CLASS: return *(const int *)p;
    case mxLOGICAL_CLASS: case mxUINT8_CLASS: return *(const unsigned char *)p;
    default: return 0.0;
    }
}

void *mxGetData(const mxArray *pa) { return pa ? (void *)pa->data.data() : nullptr; }
double *mxGetPr(const mxArray *pa) { return (double *)mxGetData(pa); }
size_t mxGetNumberOfElements(const mxArray *pa) { return pa ? numel(pa) : 0; }
size_t mxGetM(const mxArray *pa) { return pa ? pa->dims[0] : 0; }

size_t mxGetN(const mxArray *pa)
{
    if (!pa) return 0;
    size_t n = 1;
    for (size_t k = 1; k < pa->dims.size(); k++) n *= pa->dims[k];
    return n;
}

mwSize mxGetNumberOfDimensions(const mxArray *pa) { return pa ? pa->dims.size() : 0; }
const mwSize *mxGetDimensions(const mxArray *pa) { return pa ? pa->dims.data() : nullptr; }
mxClassID mxGetClassID(const mxArray *pa) { return pa ? pa->cls : mxUNKNOWN_CLASS; }
int mxIsEmpty(const mxArray *pa) { return !pa || (pa->cls != mxSTRUCT_CLASS && numel(pa) == 0); }

mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity)
{
    mwSize d[2] = {m, n};
    return make_numeric(2, d, c);
}

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity f)
{
    return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, f);
}

mxArray *mxCreateNumericArray(mwSize nd, const mwSize *dims, mxClassID c, mxComplexity)
{
    return make_numeric(nd, dims, c);
}

mxArray *mxCreateDoubleScalar(double v)
{
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    *(double *)a->data.data() = v;
    return a;
}

mxArray *mxCreateStructMatrix(mwSize, mwSize, int nfields, const char **names)
{
    mxArray *a = new mxArray_tag();
    a->cls = mxSTRUCT_CLASS;
    a->dims = {1, 1};
    for (int k = 0; k < nfields; k++) {
        a->names.push_back(names[k]);
        a->values.push_back(nullptr);
    }
    return a;
}

void fm_free(mxArray *a);

void mxSetField(mxArray *pa, mwIndex, const char *name, mxArray *value)
{
    for (size_t k = 0; k < pa->names.size(); k++)
        if (pa->names[k] == name) {
            if (pa->values[k] && pa->values[k] != value) fm_free(pa->values[k]);
            pa->values[k] = value;
            return;
        }
    pa->names.push_back(name);
    pa->values.push_back(value);
}

// ---- mex API ------------------------------------------------------------
int mexPrintf(const char *fmt, ...)
{
    char buf[4096];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_out += buf;
    return n;
}

void mexErrMsgTxt(const char *msg) { throw MexError(msg ? msg : ""); }

int mexEvalString(const char *cmd)
{
    g_eval.push_back(cmd ? cmd : "");
    return 0;
}

int mexAtExit(void (*)(void)) { return 0; }

// ---- test-side helpers ----------------------------------------------------
mxArray *fm_numeric(int cls, int nd, const long long *dims, const void *src)
{
    std::vector<mwSize> d(dims, dims + nd);
    mxArray *a = make_numeric(nd, d.data(), (mxClassID)cls);
    if (src) std::memcpy(a->data.data(), src, a->data.size());
    return a;
}

mxArray *fm_struct(void) { return mxCreateStructMatrix(1, 1, 0, nullptr); }

void fm_struct_set(mxArray *s, const char *name, mxArray *v) { mxSetField(s, 0, name, v); }

void fm_free(mxArray *a)
{
    if (!a) return;
    for (mxArray *v : a->values) fm_free(v);
    delete a;
}

size_t fm_nbytes(const mxArray *a) { return a ? a->data.size() : 0; }

int fm_call(void (*fn)(int, mxArray **, int, const mxArray **), int nlhs, mxArray **plhs,
            int nrhs, const mxArray **prhs, char *err, int errlen)
{
    if (err && errlen) err[0] = 0;
    try {
        fn(nlhs, plhs, nrhs, prhs);
    } catch (const MexError &e) {
        if (err && errlen) snprintf(err, errlen, "%s", e.what());
        return 1;
    } catch (const std::exception &e) {
        if (err && errlen) snprintf(err, errlen, "C++ exception: %s", e.what());
        return 2;
    }
    return 0;
}

const char *fm_output(void) { return g_out.c_str(); }
void fm_clear(void)
{
    g_out.clear();
    g_eval.clear();
}
int fm_eval_count(void) { return (int)g_eval.size(); }
const char *fm_eval_at(int i) { return (i >= 0 && i < (int)g_eval.size()) ? g_eval[i].c_str() : ""; }

}  // extern "C"
