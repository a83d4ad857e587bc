// [m_res, out] = ASD_POCS_gpu(P, para): the whole ASD_POCS_cuda.m loop as one
// MEX call, with the image and sinogram resident on the GPU between views.
// Same `para` fields and defaults as utilities/ASD_POCS_cuda.m:20-96 (plus
// para.phi_vec in degrees), same `out` fields as :173-180.
#include <cstring>
#include <vector>

#include "mex_common.hpp"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 2) {
        mexPrintf("Usage: [m_res, out] = ASD_POCS_gpu(P, para)\n");
        mexErrMsgTxt("Error usage!");
        return;
    }
    const mxArray *P = prhs[0], *pa = prhs[1];
    if (mxGetClassID(P) != mxDOUBLE_CLASS) mexErrMsgTxt("P must be double");
    asd_para p;
    asd_para_defaults(&p);
    p.nx = (int)asdmex::need_field(pa, "nx");
    p.ny = (int)asdmex::need_field(pa, "ny");
    p.nz = (int)asdmex::need_field(pa, "nz");
    p.SO = asdmex::need_field(pa, "SO");
    p.OD = asdmex::need_field(pa, "OD");
    p.vxlsize = asdmex::need_field(pa, "vxlsize");
    p.detsize = asdmex::need_field(pa, "detsize");
    p.na = (int)asdmex::need_field(pa, "na");
    p.nb = (int)asdmex::need_field(pa, "nb");
    p.maxnumIter = (int)asdmex::opt_field(pa, "maxnumIter", p.maxnumIter);
    p.numOfgradients = (int)asdmex::opt_field(pa, "numOfgradients", p.numOfgradients);
    p.alpha = asdmex::opt_field(pa, "alpha", p.alpha);
    p.alpha_red = asdmex::opt_field(pa, "alpha_red", p.alpha_red);
    p.beta = asdmex::opt_field(pa, "beta", p.beta);
    p.beta_red = asdmex::opt_field(pa, "beta_red", p.beta_red);
    p.r_max = asdmex::opt_field(pa, "r_max", p.r_max);
    p.epsilon = asdmex::opt_field(pa, "epsilon", p.epsilon);
    p.stopRelChg = asdmex::opt_field(pa, "stopRelChg", p.stopRelChg);
    // ASD_POCS_cuda.m:24 tests para.initBySART == true, i.e. == 1
    p.initBySART = asdmex::opt_field(pa, "initBySART", 0.0) == 1.0;
    p.lambda = asdmex::opt_field(pa, "lambda", 1.0);
    p.precision = (int)asdmex::opt_field(pa, "precision", 0.0);
    if (p.beta >= 2.0 || p.beta <= 0.0)
        mexPrintf("Warning: Bad value for para.lambda! For convergence, lambda should be in (0.0 2.0)!\n");
    const mxArray *phi = mxGetField(pa, 0, "phi_vec");
    if (!phi) mexErrMsgTxt("Error input argument parameter.phi_vec!");
    int nv = (int)mxGetN(P);
    if ((size_t)nv != mxGetNumberOfElements(phi)) mexErrMsgTxt("Size of Projection and phi_vec mismatch!");
    if ((long long)mxGetM(P) != (long long)p.na * p.nb) mexErrMsgTxt("Error size of input argument P!");
    size_t N = (size_t)p.nx * p.ny * p.nz;
    const double *m0 = nullptr, *I0 = nullptr;
    const mxArray *fm0 = mxGetField(pa, 0, "m0");
    if (fm0 && !mxIsEmpty(fm0)) {
        if (mxGetNumberOfElements(fm0) != N) mexErrMsgTxt("Error size of para.m0!");
        m0 = mxGetPr(fm0);
    }
    const mxArray *fI0 = mxGetField(pa, 0, "I0");
    if (fI0 && !mxIsEmpty(fI0)) {
        if (mxGetNumberOfElements(fI0) != N) mexErrMsgTxt("Error size of para.I0!");
        I0 = mxGetPr(fI0);
    }
    mwSize dims[3] = {(mwSize)p.nx, (mwSize)p.ny, (mwSize)p.nz};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    std::vector<double> rmse(p.maxnumIter > 0 ? p.maxnumIter : 1);
    asd_out out;
    asdmex::check_status(asd_pocs(mxGetPr(P), mxGetPr(phi), nv, &p, m0, I0, mxGetPr(plhs[0]),
                                  rmse.data(), &out));
    if (nlhs > 1) {
        const char *names[] = {"IterationNum", "dataTolerance", "beta", "alpha", "reconImg", "RMSE"};
        int nf = I0 ? 6 : 5;
        mxArray *o = mxCreateStructMatrix(1, 1, nf, names);
        mxSetField(o, 0, "IterationNum", mxCreateDoubleScalar(out.IterationNum));
        mxSetField(o, 0, "dataTolerance", mxCreateDoubleScalar(out.dataTolerance));
        mxSetField(o, 0, "beta", mxCreateDoubleScalar(out.beta));
        mxSetField(o, 0, "alpha", mxCreateDoubleScalar(out.alpha));
        mxArray *img = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
        std::memcpy(mxGetPr(img), mxGetPr(plhs[0]), N * sizeof(double));
        mxSetField(o, 0, "reconImg", img);
        if (I0) {
            mxArray *r = mxCreateDoubleMatrix(out.IterationNum, 1, mxREAL);
            for (int i = 0; i < out.IterationNum; i++) mxGetPr(r)[i] = rmse[i];
            mxSetField(o, 0, "RMSE", r);
        }
        plhs[1] = o;
    }
}
