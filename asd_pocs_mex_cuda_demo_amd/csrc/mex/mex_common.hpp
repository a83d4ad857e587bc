// Shared helpers of the MEX gateways: `para` field access with the
// reference's messages and the voxel-unit geometry of the gateways
// (src/AxAtx/Ax_main_test.cpp:48-161, Atx_main_test.cpp:59-172).
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "asdpocs.h"
#include "mex_api.h"

namespace asdmex {

// mxGetScalar of para.<name>, or the reference's two-line error.
inline double need_field(const mxArray *para, const char *name)
{
    const mxArray *f = mxGetField(para, 0, name);
    if (f == nullptr) {
        mexPrintf("parameters.%s is not assigned! Please check!\n", name);
        std::string msg = std::string("Error input argument parameter.") + name + "!";
        mexErrMsgTxt(msg.c_str());
    }
    return mxGetScalar(f);
}

// para.<name> when present and not empty, else the default.
inline double opt_field(const mxArray *para, const char *name, double dflt)
{
    const mxArray *f = mxGetField(para, 0, name);
    if (f == nullptr || mxIsEmpty(f)) return dflt;
    return mxGetScalar(f);
}

struct Geometry {
    int nx, ny, nz, na, nb;
    float cos_phi, sin_phi, SO, OD, scale, det, dz;
    std::vector<float> y_det, z_det;
};

// Reads the ten geometry fields in the reference's order and converts them
// like the gateways: float casts, cos/sin of the float angle, lengths in
// voxel units, detector tables (i - na/2.0 + 0.5) * det.
inline Geometry read_geometry(const mxArray *para)
{
    Geometry g;
    float SO = (float)need_field(para, "SO");
    float OD = (float)need_field(para, "OD");
    g.scale = (float)need_field(para, "vxlsize");
    g.nx = (int)need_field(para, "nx");
    g.ny = (int)need_field(para, "ny");
    g.nz = (int)need_field(para, "nz");
    float phi = (float)need_field(para, "phi");
    float det = (float)need_field(para, "detsize");
    g.na = (int)need_field(para, "na");
    g.nb = (int)need_field(para, "nb");
    g.cos_phi = (float)std::cos((double)phi);
    g.sin_phi = (float)std::sin((double)phi);
    g.dz = 1.0f;
    g.SO = SO / g.scale;
    g.OD = OD / g.scale;
    g.det = det / g.scale;
    if (g.na > 0 && g.nb > 0) {
        g.y_det.resize(g.na);
        g.z_det.resize(g.nb);
        for (int i = 0; i < g.na; i++)
            g.y_det[i] = (float)(((double)(float)i - (double)(float)g.na / 2.0 + 0.5) * (double)g.det);
        for (int j = 0; j < g.nb; j++)
            g.z_det[j] = (float)(((double)(float)j - (double)(float)g.nb / 2.0 + 0.5) * (double)g.det);
    }
    return g;
}

// A failed library call becomes a MATLAB error (the reference launchers
// report nothing).
inline void check_status(int rc)
{
    if (rc != ASD_OK) {
        std::string msg = std::string("asdpocs: ") + asd_last_error();
        mexErrMsgTxt(msg.c_str());
    }
}

}  // namespace asdmex
