// I2 = imgTVGradMin3D(I, alpha, nTVIter, pos[, I0]): MEX gateway of the 3-D TV
// steepest descent.  Replaces src/tvmin/imgTVGradMin3D.cpp:18-189; the compute
// is asd_tv_min_3d (FP64, on the GPU).  A 2-D input is treated as Nz = 1
// (the reference reads a third dimension that is not there).
#include "mex_common.hpp"

static void show_usage()
{
    mexPrintf("---imgTVGradMin3D - minimize image TV by gradient minimization for 3D data---\n");
    mexPrintf("usage 1: \t I2 = imgTVGradMin3D(I, alpha, nTVIter, pos, I0)\n");
    mexPrintf("usage 2: \t I2 = imgTVGradMin3D(I, alpha, nTVIter, pos)\n");
}

static void dims3(const mxArray *a, mwSize d[3])
{
    mwSize nd = mxGetNumberOfDimensions(a);
    const mwSize *dim = mxGetDimensions(a);
    d[0] = nd > 0 ? dim[0] : 1;
    d[1] = nd > 1 ? dim[1] : 1;
    d[2] = nd > 2 ? dim[2] : 1;
    for (mwSize k = 3; k < nd; k++) d[2] *= dim[k];
}

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 4 && nrhs != 5) {
        show_usage();
        return;
    }
    if (mxGetClassID(prhs[0]) != mxDOUBLE_CLASS) {
        mexPrintf("Invalid data format. Please make sure the input image data is in double format.\n");
        return;
    }
    mwSize d[3];
    dims3(prhs[0], d);
    const double *img = mxGetPr(prhs[0]);
    double alpha = mxGetScalar(prhs[1]);
    int nTVIter = (int)mxGetScalar(prhs[2]);
    int pos = (int)mxGetScalar(prhs[3]);
    const double *img_old = nullptr;
    if (nrhs == 5) {
        mwSize d0[3];
        dims3(prhs[4], d0);
        if (d[0] != d0[0] || d[1] != d0[1] || d[2] != d0[2]) {
            mexPrintf("img0 dimensions do not match with img\n");
            return;
        }
        img_old = mxGetPr(prhs[4]);
    }
    plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), mxGetDimensions(prhs[0]),
                                   mxDOUBLE_CLASS, mxREAL);
    if (d[0] > 0 && d[1] > 0 && d[2] > 0)
        asdmex::check_status(asd_tv_min_3d(img, (int)d[0], (int)d[1], (int)d[2], alpha, nTVIter,
                                           pos, img_old, mxGetPr(plhs[0])));
}
