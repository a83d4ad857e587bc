// I2 = imgTVGradMin(I, alpha, nTVIter, pos[, I0]): MEX gateway of the 2-D TV
// steepest descent.  Replaces src/tvmin/imgTVGradMin.cpp:16-140 with the same
// usage text, checks (print and return, leaving the output unassigned) and
// the " +TV-PC " notice; the compute is asd_tv_min_2d (FP64, on the GPU).
#include "mex_common.hpp"

static void show_usage()
{
    mexPrintf("---imgTVGradMin - minimize image TV by gradient minimization ---\n");
    mexPrintf("usage 1: \t I2 = imgTVGradMin(I, alpha, nTVIter, pos, I0)\n");
    mexPrintf("usage 2: \t I2 = imgTVGradMin(I, alpha, nTVIter, pos)\n");
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
    const double *img = mxGetPr(prhs[0]);
    double alpha = mxGetScalar(prhs[1]);
    int nTVIter = (int)mxGetScalar(prhs[2]);
    int pos = (int)mxGetScalar(prhs[3]);
    int rowN = (int)mxGetM(prhs[0]);
    int colN = (int)mxGetN(prhs[0]);
    const double *img_old = nullptr;
    if (nrhs == 5) {
        if ((int)mxGetM(prhs[4]) != rowN || (int)mxGetN(prhs[4]) != colN) {
            mexPrintf("img0 dimensions do not match with img\n");
            return;
        }
        img_old = mxGetPr(prhs[4]);
    }
    plhs[0] = mxCreateDoubleMatrix(rowN, colN, mxREAL);
    if (rowN > 0 && colN > 0)
        asdmex::check_status(asd_tv_min_2d(img, rowN, colN, alpha, nTVIter, pos, img_old,
                                           mxGetPr(plhs[0])));
    if (pos == 1 && nTVIter > 0) {
        mexPrintf(" +TV-PC ");
        mexEvalString("drawnow;");
    }
}
