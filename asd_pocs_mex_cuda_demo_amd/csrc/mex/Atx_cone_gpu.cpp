// [x, sc] = Atx_cone_gpu(Y, para): MEX gateway of the backprojector.
// Replaces src/AxAtx/Atx_main_test.cpp:20-177; the compute is
// Atx_cone_mf_gpu_new in libasdpocs.so (dy_det = dz_det = detsize/vxlsize).
#include "mex_common.hpp"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs > 2 || nrhs <= 1) {
        mexPrintf("Usage: [x,sc] = Atx_cone_mf_gpu(Y,parameters)\n");
        mexPrintf("Y: a frame of projection of size na*nb\n");
        mexPrintf("parameters: nx ny nz SO OD vxlsize detsize phi na nb\n");
        mexPrintf("output  x: back-projection data (size nx*ny*nz)\n");
        mexPrintf("output sc: col sum of system matrix\n");
        mexErrMsgTxt("Error usage!");
        return;
    }
    float *Y = (float *)mxGetData(prhs[0]);
    asdmex::Geometry g = asdmex::read_geometry(prhs[1]);
    size_t numel = mxGetNumberOfElements(prhs[0]);
    if ((long long)numel != (long long)g.na * g.nb) {
        mexPrintf("Size of the first input data (Y) mismatch na*nb! Please check!\n");
        mexErrMsgTxt("Error size of input argument Y!");
        return;
    }
    if (mxGetClassID(prhs[0]) != mxSINGLE_CLASS)
        mexErrMsgTxt("Error input argument Y: single precision expected!");
    size_t N = (size_t)g.nx * g.ny * g.nz;
    plhs[0] = mxCreateNumericMatrix(N, 1, mxSINGLE_CLASS, mxREAL);
    plhs[1] = mxCreateNumericMatrix(N, 1, mxSINGLE_CLASS, mxREAL);
    float *X = (float *)mxGetData(plhs[0]);
    float *sc = (float *)mxGetData(plhs[1]);
    Atx_cone_mf_gpu_new(X, Y, sc, g.cos_phi, g.sin_phi, g.y_det.data(), g.z_det.data(), g.SO,
                        g.OD, g.scale, g.det, g.det, g.dz, g.nx, g.ny, g.nz, g.na, g.nb);
    asdmex::check_status(asd_last_status());
}
