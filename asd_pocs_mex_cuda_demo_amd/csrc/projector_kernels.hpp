// Projector kernel templates shared by the plain operators and the fused
// SART stages.  The epilogue functor decides what a finished ray sum or voxel
// sum becomes (plain y/sr, or the SART correction / image update).
#pragma once

#include "geom_device.hpp"

namespace asd {

// Forward projector: LPR lanes march one ray together (lane j takes slabs
// j, j+LPR, ...), then a butterfly over the LPR lanes sums their partials.
// With LPR = 1 a thread walks its ray in slab order like the reference kernel
// (cu:375, :576); with LPR = 64 a whole wave serves one ray, which keeps a 2-D
// view (a few hundred rays) spread over the chip, and lanes walking along x
// read consecutive voxels.
template <int LPR, bool KAHAN, class Epi>
__global__ __launch_bounds__(256) void ax_kernel(ViewGeom g, const float *__restrict__ X,
                                                 const float *__restrict__ y_det,
                                                 const float *__restrict__ z_det, Epi epi)
{
    constexpr int RPB = 256 / LPR;  // rays per block
    if (epi.skip()) return;
    const long nrays = (long)g.na * g.nb;
    const int lane = threadIdx.x % LPR;
    const long id = (long)blockIdx.x * RPB + threadIdx.x / LPR;
    const bool active = id < nrays;
    Acc<KAHAN> sa, sl;
    if (active) {
        int ia = (int)(id % g.na), ib = (int)(id / g.na);
        AxRay r = ax_ray_setup(g, y_det[ia], z_det[ib]);
        int nu = r.xdriven ? g.nx : g.ny;
        for (int iu = lane; iu < nu; iu += LPR) ax_slab<KAHAN>(r, X, iu, g.nx, g.ny, sa, sl);
    }
    float acc, accl;
    if constexpr (LPR > 1 && KAHAN) {
        // compensated lane partials are combined in FP64
        double a = (double)sa.s - (double)sa.c, b = (double)sl.s - (double)sl.c;
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, LPR);
            b += __shfl_xor(b, off, LPR);
        }
        acc = (float)a;
        accl = (float)b;
    } else {
        acc = sa.value();
        accl = sl.value();
        if constexpr (LPR > 1) {
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) {
                acc += __shfl_xor(acc, off, LPR);
                accl += __shfl_xor(accl, off, LPR);
            }
        }
    }
    if (active && lane == 0) epi(id, acc, accl);
}

template <bool KAHAN = false, class Epi>
int dispatch_ax(int lpr, const ViewGeom &g, const float *x, const float *y_det,
                const float *z_det, const Epi &epi, hipStream_t s)
{
    long nrays = (long)g.na * g.nb;
    auto blocks = [&](int l) { return (unsigned)((nrays * l + 255) / 256); };
    switch (lpr) {
    case 1: ax_kernel<1, KAHAN, Epi><<<blocks(1), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    case 2: ax_kernel<2, KAHAN, Epi><<<blocks(2), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    case 4: ax_kernel<4, KAHAN, Epi><<<blocks(4), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    case 8: ax_kernel<8, KAHAN, Epi><<<blocks(8), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    case 16: ax_kernel<16, KAHAN, Epi><<<blocks(16), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    case 32: ax_kernel<32, KAHAN, Epi><<<blocks(32), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    default: ax_kernel<64, KAHAN, Epi><<<blocks(64), 256, 0, s>>>(g, x, y_det, z_det, epi); break;
    }
    ASD_LAUNCH_CHECK("ax_kernel");
    return ASD_OK;
}

// Backprojector: one thread per voxel (x fastest) walks its detector
// footprint, ib outer and ia inner like the reference (cu:255-269), and sums
// l*y and l in that order.  Blocks are remapped so each XCD owns a contiguous
// slab of voxels and the rays they share stay in its L2.
template <class Epi, bool KAHAN = false>
__global__ __launch_bounds__(256) void atx_kernel(ViewGeom g, const float *__restrict__ Y,
                                                  const RayRec *__restrict__ rt, Epi epi)
{
#pragma clang fp contract(off)
    if (epi.skip()) return;
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    const long v = (long)b * 256 + threadIdx.x;
    const long N = (long)g.nx * g.ny * g.nz;
    if (v >= N) return;
    const int ix = (int)(v % g.nx);
    const long r = v / g.nx;
    const int iy = (int)(r % g.ny), iz = (int)(r / g.ny);
    Footprint f = footprint(g, ix, iy, iz);
    const float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
    const float px = sx - f.xc, py = sy - f.yc, pz = 0.0f - f.zc;
    const float hz = g.dz / 2.0f;
    Acc<KAHAN> acc, accl;
    for (int ib = f.b_lo; ib <= f.b_hi; ib++) {
        const long row = (long)ib * g.na;
        for (int ia = f.a_lo; ia <= f.a_hi; ia++) {
            RayRec R = rt[row + ia];
            float l = find_l_3d_rec(R, px, py, pz, hz, g.dz);
            acc.add(l * Y[row + ia]);
            accl.add(l);
        }
    }
    epi(v, acc.value(), accl.value());
}

}  // namespace asd
