// [y, sr] = Ax_cone_gpu(X, para): MEX gateway of the forward projector.
// Replaces src/AxAtx/Ax_main_test.cpp:19-165 (same inputs, outputs, checks and
// messages); the compute is Ax_cone_mf_gpu_new in libasdpocs.so.
#include "mex_common.hpp"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs > 2 || nrhs <= 1) {
        mexPrintf("Usage: [y,sr] = Ax_cone_mf_gpu(X,parameters)\n");
        mexPrintf("X: image of size nx*ny*nz\n");
        mexPrintf("parameters: nx ny nz SO OD vxlsize detsize phi na nb\n");
        mexPrintf("output  y: projection data\n");
        mexPrintf("output sr: row sum of system matrix\n");
        mexErrMsgTxt("Error usage!");
        return;
    }
    float *X = (float *)mxGetData(prhs[0]);
    asdmex::Geometry g = asdmex::read_geometry(prhs[1]);
    size_t numel = mxGetNumberOfElements(prhs[0]);
    if ((long long)numel != (long long)g.nx * g.ny * g.nz) {
        mexPrintf("Size of the first input data (X) mismatch nx*ny*nz! Please check!\n");
        mexErrMsgTxt("Error size of input argument X!");
        return;
    }
    if (mxGetClassID(prhs[0]) != mxSINGLE_CLASS)
        mexErrMsgTxt("Error input argument X: single precision expected!");
    size_t M = (size_t)g.na * g.nb;
    plhs[0] = mxCreateNumericMatrix(M, 1, mxSINGLE_CLASS, mxREAL);
    plhs[1] = mxCreateNumericMatrix(M, 1, mxSINGLE_CLASS, mxREAL);
    float *y = (float *)mxGetData(plhs[0]);
    float *sr = (float *)mxGetData(plhs[1]);
    Ax_cone_mf_gpu_new(X, y, sr, g.cos_phi, g.sin_phi, g.y_det.data(), g.z_det.data(), g.SO,
                       g.OD, g.scale, g.dz, g.nx, g.ny, g.nz, g.na, g.nb);
    asdmex::check_status(asd_last_status());
}
