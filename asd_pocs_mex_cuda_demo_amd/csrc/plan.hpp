// The device-resident ASD-POCS plan shared by the driver (driver.hip) and
// the on-device input simulation (simulate.hip).
#pragma once

#include <vector>

#include "asdpocs.h"
#include "projector_kernels.hpp"
#include "tv.hpp"

namespace asd {

// Stage timing (events on the plan stream) for the dominant kernels.
enum : int { ST_AX = 0, ST_ATX = 1, ST_TVGRAD = 2, ST_N = 3 };  // ST_TVGRAD: a whole TV descent

constexpr int HIST_CAP = 1 << 16;  // RMSE history entries kept on the device

struct DevCtl {
    double alpha;    // current TV step factor
    double beta;     // current SART relaxation of the ASD loop
    double relax;    // relaxation of a stand-alone sweep (SART_cuda_inline's lambda)
    double dp, dd2, dg;
    int k;           // outer iterations completed
    int stopped;     // the stop rule fired
};

struct CtlRules {
    double alpha_red, beta_red, r_max, epsilon, stopRelChg;
};

}  // namespace asd

using asd::ATX_NVARIANT;
using asd::CtlRules;
using asd::DevCtl;
using asd::DeviceBuf;
using asd::ST_N;
using asd::TvScratch;
using asd::ViewGeom;

struct asd_plan {
    asd_para para;
    int nv = 0;
    size_t N = 0, M = 0;
    std::vector<ViewGeom> views;
    hipStream_t stream = nullptr;
    // Images m, mpre, mk1, f0, I0 and the TV buffers v, u2, v2 hold FP32
    // values, or FP64 ones in the reference-arithmetic plan (precision 4),
    // which also keeps ms = single(m), the image its forward projector reads.
    DeviceBuf ydet, zdet, P, m, mpre, mk1, f0, I0, cor, res2, rt, v, u2, v2, ctlbuf, hist, ms;
    std::vector<double> u16_lo, u16_step;  // per-view scale of the 16-bit fixed-point sinogram
    TvScratch red;
    bool have_P = false, have_I0 = false;
    bool I0_sim = false;            // the RMSE reference is a simulated phantom
    asd::AxShape ax;                // forward projector launch shape
    bool ax_forced = false;         // ax set by asd_plan_set_ax_shape
    int *h_stop = nullptr;          // pinned copies of the stop flag (4 checkpoints)
    hipEvent_t stop_ev[4] = {};
    int atx_variant = -1;           // backprojector variant picked by the autotuner
    double atx_us[ATX_NVARIANT];    // its measurements (-1: not tried)
    DevCtl h;  // host copy of the control block after the last sync
    // hipGraph of one outer iteration
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // optional stage timing (eager path only)
    bool timing = false;
    bool capturing = false;
    struct Ev { hipEvent_t a, b; int stage; };
    std::vector<Ev> ev;
    double stage_ms[ST_N] = {0, 0, 0};
    long stage_cnt[ST_N] = {0, 0, 0};

    DevCtl *ctl() { return ctlbuf.as<DevCtl>(); }
    // the reference-arithmetic plan: FP64 sinogram, image state and TV
    bool ref64() const { return para.precision == 4; }
    size_t img_bytes() const { return ref64() ? 8 : 4; }  // per image element
    CtlRules rules() const
    {
        return CtlRules{para.alpha_red, para.beta_red, para.r_max, para.epsilon, para.stopRelChg};
    }
};

namespace asd {
// Drops the captured iteration graph (kernel arguments changed).
void plan_drop_graph(asd_plan *p);
}  // namespace asd
