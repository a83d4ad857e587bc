/*
 * mex_api.h - the subset of the MATLAB MEX / mx C API that the gateways in
 * asd_pocs_mex_cuda_demo_amd/csrc/mex use: names and signatures only.
 *
 * Built against this header, a gateway leaves every mx- and mex- symbol
 * undefined; the host process provides them at load time (MATLAB itself, or
 * the test host in tests/fakemex).  A build inside MATLAB compiles the
 * gateways against MATLAB's own mex.h instead; the calls are the same.
 */
#ifndef ASD_MEX_API_H
#define ASD_MEX_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef ptrdiff_t mwSignedIndex;

typedef enum {
    mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS,
    mxVOID_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS,
    mxINT16_CLASS, mxUINT16_CLASS, mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS,
    mxUINT64_CLASS, mxFUNCTION_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

/* arrays */
mxArray *mxGetField(const mxArray *pa, mwIndex i, const char *fieldname);
double mxGetScalar(const mxArray *pa);
void *mxGetData(const mxArray *pa);
double *mxGetPr(const mxArray *pa);
size_t mxGetNumberOfElements(const mxArray *pa);
size_t mxGetM(const mxArray *pa);
size_t mxGetN(const mxArray *pa);
mwSize mxGetNumberOfDimensions(const mxArray *pa);
const mwSize *mxGetDimensions(const mxArray *pa);
mxClassID mxGetClassID(const mxArray *pa);
int mxIsEmpty(const mxArray *pa);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID classid,
                              mxComplexity flag);
mxArray *mxCreateDoubleScalar(double value);
mxArray *mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char **fieldnames);
void mxSetField(mxArray *pa, mwIndex i, const char *fieldname, mxArray *value);

/* interpreter services */
int mexPrintf(const char *fmt, ...);
void mexErrMsgTxt(const char *msg); /* does not return */
int mexEvalString(const char *cmd);
int mexAtExit(void (*exit_fcn)(void));

/* the gateway entry point every MEX file exports */
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif

#endif /* ASD_MEX_API_H */
