/*
 * CPU oracle for the ASD-POCS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libasdpocs.so) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of the reference's hot path:
 *   - find_l / find_l_3d          src/AxAtx/Ax_cone_gpu.cu:16-117, 120-209
 *   - forward projector (Ax)      src/AxAtx/Ax_cone_gpu.cu:339-776 (+ launcher :778-805)
 *   - backprojector (Atx)         src/AxAtx/Ax_cone_gpu.cu:214-271 (+ set2zero/scalex/launcher :273-331)
 *   - MEX geometry set-up         src/AxAtx/Ax_main_test.cpp:138-161, Atx_main_test.cpp:149-175
 *   - 2-D TV steepest descent     src/tvmin/imgTVGradMin.cpp:57-119   (FP64)
 *   - 3-D TV steepest descent     src/tvmin/imgTVGradMin3D.cpp:76-170 (FP64)
 *   - SART view update            utilities/ASD_POCS_cuda.m:110-134, SART_cuda_inline.m:56-87 (FP64)
 *   - ASD outer step              utilities/ASD_POCS_cuda.m:103-171 (FP64)
 *
 * Projectors run in FP32 like the reference kernels; TV and the driver run in
 * FP64 like the reference's C++ and MATLAB.  The CUDA thread grid becomes an
 * OpenMP loop over rays (Ax) or voxels (Atx); every output element has one
 * writer, so the result does not depend on the thread count.
 *
 * Parity pins (what checks this restatement; the tests live in tests/):
 *   find_l, find_l_3d   FP64 Liang-Barsky clip of the line against the box, brute force
 *                       on random rays, plus the a==0 / b==0 2-D fallback.
 *   Ax 2-D              closed form: y = mu*vxl*|ray ∩ box| for a box indicator image,
 *                       sr = vxl*|ray ∩ image| for every image; brute-force exact
 *                       system matrix on small images.
 *   Ax 3-D              brute-force exact per-slab clip of the "monotone" segment
 *                       (y and z extents sorted independently, the reference's split rule).
 *   Atx                 adjoint identity <Ax x, y> = <x, Atx y> in 2-D, and a brute-force
 *                       all-bins exact backprojection in 3-D.
 *   TV 2-D / 3-D        v equals the finite-difference gradient of the zero-padded
 *                       isotropic TV functional (times 2 in 3-D); ||I_new - I|| = alpha*d.
 *   driver              invariants (positivity, data error decreasing on noiseless data,
 *                       alpha rule) and agreement with an independent numpy driver.
 *
 * The reference cannot be compiled here: Ax_cone_gpu.cu needs the CUDA runtime
 * and the TV sources need MATLAB's mex.h, neither of which is in this image.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n; }

int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#define NTH orc_get_threads()

static inline float fabs_f(float a) { return a > 0 ? a : -a; }

/* ------------------------------------------------------------------------ */
/* Line ∩ box lengths (reference cu:16-209)                                  */
/* ------------------------------------------------------------------------ */

/* Length of the infinite 2-D line through p1=(x1_0,y1_0), p2=(x2_0,y2_0)
 * inside the dx*dy rectangle centred at (x,y).  Axis-parallel lines give the
 * full side when the line lies within the closed slab (cu:26-40).  Otherwise
 * the faces normal to the dominant axis are tried first, then the others, and
 * the first two hits define the chord (cu:45-113). */
float orc_find_l(float x1_0, float y1_0, float x2_0, float y2_0,
                 float dx, float dy, float x, float y)
{
    float a = x2_0 - x1_0, b = y2_0 - y1_0;
    float hx = dx / 2.0f, hy = dy / 2.0f;
    if (a == 0) return (fabs_f(x1_0 - x) <= hx) ? dy : 0.0f;
    if (b == 0) return (fabs_f(y1_0 - y) <= hy) ? dx : 0.0f;

    float px = x1_0 - x, py = y1_0 - y;
    float hitx[2], hity[2];
    int n = 0;
    /* x-faces: y on the line at x = -hx and x = +hx (relative to the centre) */
    float sxy = b / a, cy = sxy * (-px) + py, hy_off = sxy * hx;
    /* y-faces: x on the line at y = -hy and y = +hy */
    float syx = a / b, cx = syx * (-py) + px, hx_off = syx * hy;
    if (fabs_f(a) > fabs_f(b)) {
        if (fabs_f(cy - hy_off) <= hy) { hitx[n] = -hx; hity[n] = cy - hy_off; n++; }
        if (fabs_f(cy + hy_off) <= hy) { hitx[n] = hx; hity[n] = cy + hy_off; n++; }
        if (n < 2 && fabs_f(cx - hx_off) <= hx) { hity[n] = -hy; hitx[n] = cx - hx_off; n++; }
        if (n < 2 && fabs_f(cx + hx_off) <= hx) { hity[n] = hy; hitx[n] = cx + hx_off; n++; }
    } else {
        if (fabs_f(cx - hx_off) <= hx) { hity[n] = -hy; hitx[n] = cx - hx_off; n++; }
        if (fabs_f(cx + hx_off) <= hx) { hity[n] = hy; hitx[n] = cx + hx_off; n++; }
        if (n < 2 && fabs_f(cy - hy_off) <= hy) { hitx[n] = -hx; hity[n] = cy - hy_off; n++; }
        if (n < 2 && fabs_f(cy + hy_off) <= hy) { hitx[n] = hx; hity[n] = cy + hy_off; n++; }
    }
    if (n != 2) return 0.0f;
    float ex = hitx[1] - hitx[0], ey = hity[1] - hity[0];
    return (float)sqrt(ex * ex + ey * ey);
}

/* Face hit test of one face pair for find_l_3d: the line at u = -h and u = +h
 * (u the face-normal coordinate) has transverse coordinates c0 -/+ o0 and
 * c1 -/+ o1; a hit needs both inside the closed half-widths w0, w1. */
typedef struct { float p[2][3]; int n; } hits3;

static inline void face_pair(hits3 *H, int axis, float h, float c0, float o0,
                             float c1, float o1, float w0, float w1, int ax0, int ax1)
{
    /* "-" face then "+" face; stop at two hits (cu:144-147 and siblings) */
    for (int s = -1; s <= 1 && H->n < 2; s += 2) {
        float t0 = c0 + (float)s * o0, t1 = c1 + (float)s * o1;
        if (fabs_f(t0) <= w0 && fabs_f(t1) <= w1) {
            H->p[H->n][axis] = (float)s * h;
            H->p[H->n][ax0] = t0;
            H->p[H->n][ax1] = t1;
            H->n++;
        }
    }
}

/* Length of the 3-D line p1->p2 inside the dx*dy*dz box centred at (x,y,z)
 * (cu:120-209).  When a==0 or b==0 the reference falls back to the 2-D chord
 * in the other vertical plane and ignores the dropped coordinate (cu:131-135).
 * Faces are tried in dominant in-plane order (x,y,z or y,x,z); the first two
 * hits define the chord. */
float orc_find_l_3d(float x1_0, float y1_0, float z1_0, float x2_0, float y2_0, float z2_0,
                    float dx, float dy, float dz, float x, float y, float z)
{
    float a = x2_0 - x1_0, b = y2_0 - y1_0, c = z2_0 - z1_0;
    float hx = dx / 2.0f, hy = dy / 2.0f, hz = dz / 2.0f;
    if (a == 0) return orc_find_l(y1_0, z1_0, y2_0, z2_0, dy, dz, y, z);
    if (b == 0) return orc_find_l(x1_0, z1_0, x2_0, z2_0, dx, dz, x, z);

    float px = x1_0 - x, py = y1_0 - y, pz = z1_0 - z;
    hits3 H; H.n = 0;
    /* the slope quotients exactly as the reference forms them */
    float s_ba = b / a, s_ca = c / a;
    float s_ab = a / b, s_cb = c / b;
    float s_ac = a / c, s_bc = b / c;
    /* x faces: (y,z) = (s_ba*(-px)+py, s_ca*(-px)+pz) -/+ (s_ba*hx, s_ca*hx) */
    float xy_c = s_ba * (-px) + py, xy_o = s_ba * hx, xz_c = s_ca * (-px) + pz, xz_o = s_ca * hx;
    /* y faces: (x,z) */
    float yx_c = s_ab * (-py) + px, yx_o = s_ab * hy, yz_c = s_cb * (-py) + pz, yz_o = s_cb * hy;
    /* z faces: (x,y) */
    float zx_c = s_ac * (-pz) + px, zx_o = s_ac * hz, zy_c = s_bc * (-pz) + py, zy_o = s_bc * hz;

    if (fabs_f(a) > fabs_f(b)) {
        face_pair(&H, 0, hx, xy_c, xy_o, xz_c, xz_o, hy, hz, 1, 2);
        if (H.n < 2) face_pair(&H, 1, hy, yx_c, yx_o, yz_c, yz_o, hx, hz, 0, 2);
    } else {
        face_pair(&H, 1, hy, yx_c, yx_o, yz_c, yz_o, hx, hz, 0, 2);
        if (H.n < 2) face_pair(&H, 0, hx, xy_c, xy_o, xz_c, xz_o, hy, hz, 1, 2);
    }
    if (H.n < 2) face_pair(&H, 2, hz, zx_c, zx_o, zy_c, zy_o, hx, hy, 0, 1);
    if (H.n != 2) return 0.0f;
    float ex = H.p[1][0] - H.p[0][0], ey = H.p[1][1] - H.p[0][1], ez = H.p[1][2] - H.p[0][2];
    return (float)sqrt(ex * ex + ey * ey + ez * ez);
}

/* ------------------------------------------------------------------------ */
/* Gateway geometry (Ax_main_test.cpp:138-161, Atx_main_test.cpp:149-172)     */
/* ------------------------------------------------------------------------ */

/* phi, SO, OD, vxlsize, detsize as MATLAB hands them (double), rounded to
 * float exactly like the gateway's (float)mxGetScalar casts.  Outputs are in
 * voxel units; y_det has na entries, z_det nb entries. */
void orc_gateway_geometry(double phi, double SO, double OD, double vxlsize, double detsize,
                          int na, int nb, float *cos_phi, float *sin_phi, float *SOv,
                          float *ODv, float *detv, float *scale, float *y_det, float *z_det)
{
    float phif = (float)phi, sof = (float)SO, odf = (float)OD, sc = (float)vxlsize;
    float det = (float)detsize;
    *cos_phi = (float)cos((double)phif);
    *sin_phi = (float)sin((double)phif);
    *SOv = sof / sc;
    *ODv = odf / sc;
    det = det / sc;
    *detv = det;
    *scale = sc;
    for (int i = 0; i < na; i++)
        y_det[i] = (float)(((double)(float)i - (double)(float)na / 2.0 + 0.5) * (double)det);
    for (int j = 0; j < nb; j++)
        z_det[j] = (float)(((double)(float)j - (double)(float)nb / 2.0 + 0.5) * (double)det);
}

/* ------------------------------------------------------------------------ */
/* Forward projector (cu:339-776)                                            */
/* ------------------------------------------------------------------------ */

/* Contribution of one unit slab of a ray.  Along the slab the transverse
 * extent [t1,t2] (sorted) and the axial extent [z1,z2] (sorted, sorted
 * independently of t: the reference's ordering, cu:377-397) are split at the
 * cell crossings r_t = (ct2-t1)/(t2-t1) and r_z = (cz2-z1)/(z2-z1) (1 when no
 * crossing).  The three pieces [0,min], [min,max], [max,1] fall in cells
 * (ct1,cz1), (ct2,cz1) or (ct1,cz2), and (ct2,cz2); each in-range piece adds
 * w*l*x to the ray sum and w*l to the row sum.  This rule reproduces all
 * sixteen labelled boundary cases of the reference (// 11 ... // 44). */
static inline void slab_pieces(float t1, float t2, float zz1, float zz2, int nt, int zlo,
                               int zhi, float *w, int *ct, int *cz)
{
    int ct1 = (int)floorf(t1), ct2 = (int)floorf(t2);
    int cz1 = (int)floorf(zz1), cz2 = (int)floorf(zz2);
    float rt = (ct1 == ct2) ? 1.0f : ((float)ct2 - t1) / (t2 - t1);
    float rz = (cz1 == cz2) ? 1.0f : ((float)cz2 - zz1) / (zz2 - zz1);
    float lo = rt > rz ? rz : rt, hi = rt > rz ? rt : rz;
    w[0] = lo; ct[0] = ct1; cz[0] = cz1;
    w[1] = hi - lo;
    if (rt > rz) { ct[1] = ct1; cz[1] = cz2; } else { ct[1] = ct2; cz[1] = cz1; }
    w[2] = 1.0f - hi; ct[2] = ct2; cz[2] = cz2;
    for (int k = 0; k < 3; k++)
        if (ct[k] < 0 || ct[k] >= nt || cz[k] < zlo || cz[k] >= zhi) ct[k] = -1;
    /* The reference adds only the pieces its case analysis names; a piece of
     * zero weight (no crossing) adds +0 and leaves the sums unchanged. */
}

/* The sums of ray (ia, ib) over the pieces in slices [zlo, zhi) (the whole
 * volume: 0, nz); X holds those slices only (slice zlo first). */
static void ax_one_ray(const float *X, float *yout, float *srout, int ia, int ib,
                       float cos_phi, float sin_phi, const float *y_det, const float *z_det,
                       float SO, float OD, float scale, float dz, int nx, int ny, int nz, int na,
                       int zlo, int zhi)
{
    int nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
    float x1 = cos_phi * (-SO), y1 = sin_phi * (-SO), z1 = 0.0f;
    float x2 = cos_phi * OD - sin_phi * y_det[ia];
    float y2 = sin_phi * OD + cos_phi * y_det[ia];
    float z2 = z_det[ib];
    float acc = 0.0f, accl = 0.0f;
    float tmp = (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2);
    float q = tmp + (z1 - z2) * (z1 - z2);
    int xdriven = fabs_f(x1 - x2) > fabs_f(y1 - y2);
    /* drive along u (x or y); t is the other in-plane axis */
    float u1 = xdriven ? x1 : y1, t1s = xdriven ? y1 : x1;
    float du = xdriven ? (x2 - x1) : (y2 - y1), dt = xdriven ? (y2 - y1) : (x2 - x1);
    int nu = xdriven ? nx : ny, nt = xdriven ? ny : nx;
    int nu2 = xdriven ? nx2 : ny2, nt2 = xdriven ? ny2 : nx2;
    float s1 = dt / du, s2 = (z2 - z1) / du;
    for (int iu = 0; iu < nu; iu++) {
        float ua = (float)(iu - nu2), ub = ua + 1;
        float ta, tb, za, zb;
        if (s1 >= 0) { ta = t1s + s1 * (ua - u1) + nt2; tb = t1s + s1 * (ub - u1) + nt2; }
        else         { ta = t1s + s1 * (ub - u1) + nt2; tb = t1s + s1 * (ua - u1) + nt2; }
        if (s2 >= 0) { za = (z1 + s2 * (ua - u1)) / dz + nz2; zb = (z1 + s2 * (ub - u1)) / dz + nz2; }
        else         { za = (z1 + s2 * (ub - u1)) / dz + nz2; zb = (z1 + s2 * (ua - u1)) / dz + nz2; }
        float w[3]; int ct[3], cz[3];
        slab_pieces(ta, tb, za, zb, nt, zlo, zhi, w, ct, cz);
        if (ct[0] < 0 && ct[1] < 0 && ct[2] < 0) continue;
        float d = tb - ta;
        float l = (float)sqrt((d * d + 1) * q / tmp);
        for (int k = 0; k < 3; k++) {
            if (ct[k] < 0) continue;
            int ix = xdriven ? iu : ct[k], iy = xdriven ? ct[k] : iu;
            size_t idx = (size_t)(cz[k] - zlo) * ny * nx + (size_t)iy * nx + ix;
            float wl = w[k] * l;
            acc += wl * X[idx];
            accl += wl;
        }
    }
    *yout = acc * scale;
    *srout = accl * scale;
}

/* One view of the forward projector; same arguments as the reference
 * launcher Ax_cone_mf_gpu_new (cu:778), all pointers host-side. */
void orc_ax(const float *X, float *y, float *sr, float cos_phi, float sin_phi,
            const float *y_det, const float *z_det, float SO, float OD, float scale, float dz,
            int nx, int ny, int nz, int na, int nb)
{
    long nrays = (long)na * nb;
#pragma omp parallel for schedule(dynamic, 16) num_threads(NTH)
    for (long id = 0; id < nrays; id++) {
        int ia = (int)(id % na), ib = (int)(id / na);
        ax_one_ray(X, &y[id], &sr[id], ia, ib, cos_phi, sin_phi, y_det, z_det, SO, OD, scale,
                   dz, nx, ny, nz, na, 0, nz);
    }
}

/* ------------------------------------------------------------------------ */
/* Backprojector (cu:214-331)                                                */
/* ------------------------------------------------------------------------ */

/* Sums of voxel v, written to *xo and *sco. */
static void atx_one_voxel(float *xo, const float *y, float *sco, long v, float cos_phi,
                          float sin_phi, const float *y_det, const float *z_det, float SO,
                          float OD, float scale, float dy_det, float dz_det, float dz, int nx,
                          int ny, int nz, int na, int nb)
{
    int ix = (int)(v % nx);
    long iyz = v / nx;
    int iy = (int)(iyz % ny), iz = (int)(iyz / ny);
    float SD = SO + OD;
    int na2 = na / 2, nb2 = nb / 2, nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
    float d = (float)sqrt((1 + dz * dz) / 2);
    float zc = (float)(iz + 0.5 - nz2) * dz;
    float yc = (float)(iy + 0.5 - ny2);
    float xc = (float)(ix + 0.5 - nx2);
    float xr = cos_phi * xc + sin_phi * yc;
    float yr = -sin_phi * xc + cos_phi * yc;
    float t = SD / ((xr + SO) * dy_det);
    int a_hi = (int)floorf((yr + 1) * t + na2);
    int a_lo = (int)floorf((yr - 1) * t + na2);
    t = SD / ((xr + SO) * dz_det);
    int b_hi = (int)floorf((zc + d) * t + nb2);
    int b_lo = (int)floorf((zc - d) * t + nb2);
    if (a_lo < 0) a_lo = 0;
    if (b_lo < 0) b_lo = 0;
    if (a_hi > na - 1) a_hi = na - 1;
    if (b_hi > nb - 1) b_hi = nb - 1;
    float sx = cos_phi * (-SO), sy = sin_phi * (-SO);
    float acc = 0.0f, accl = 0.0f;
    for (int ib = b_lo; ib <= b_hi; ib++) {
        for (int ia = a_lo; ia <= a_hi; ia++) {
            float ex = cos_phi * OD - sin_phi * y_det[ia];
            float ey = sin_phi * OD + cos_phi * y_det[ia];
            float l = orc_find_l_3d(sx, sy, 0.0f, ex, ey, z_det[ib], 1.0f, 1.0f, dz, xc, yc, zc);
            acc += l * y[(size_t)ib * na + ia];
            accl += l;
        }
    }
    *xo = acc * scale;
    *sco = accl * scale;
}

/* One view of the backprojector; same arguments as the reference launcher
 * Atx_cone_mf_gpu_new (cu:299), all pointers host-side. */
void orc_atx(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
             const float *y_det, const float *z_det, float SO, float OD, float scale,
             float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na, int nb)
{
    long N = (long)nx * ny * nz;
#pragma omp parallel for schedule(static) num_threads(NTH)
    for (long v = 0; v < N; v++)
        atx_one_voxel(&X[v], y, &sc[v], v, cos_phi, sin_phi, y_det, z_det, SO, OD, scale,
                      dy_det, dz_det, dz, nx, ny, nz, na, nb);
}

/* ------------------------------------------------------------------------ */
/* Subsets of one view, for full-size views whose whole evaluation would take */
/* minutes on the CPU: the same per-ray and per-voxel arithmetic as above.    */
/* ------------------------------------------------------------------------ */

/* y[k], sr[k] = the forward projector's sums of ray ids[k] (id = ib*na + ia),
 * exactly as orc_ax forms them. */
void orc_ax_rays(const float *X, float *y, float *sr, const long *ids, long nids, float cos_phi,
                 float sin_phi, const float *y_det, const float *z_det, float SO, float OD,
                 float scale, float dz, int nx, int ny, int nz, int na, int nb)
{
#pragma omp parallel for schedule(dynamic, 16) num_threads(NTH)
    for (long k = 0; k < nids; k++) {
        long id = ids[k];
        ax_one_ray(X, &y[k], &sr[k], (int)(id % na), (int)(id / na), cos_phi, sin_phi, y_det,
                   z_det, SO, OD, scale, dz, nx, ny, nz, na, 0, nz);
    }
}

/* The forward projector's partial sums over the z-slab [z0, z0 + nzs) of an
 * nx x ny x nz volume (the slab's pieces of every ray, with the whole
 * volume's geometry), of the rays ids[k]; X holds the slab's slices only. */
void orc_ax_slab_rays(const float *X, float *y, float *sr, const long *ids, long nids,
                      float cos_phi, float sin_phi, const float *y_det, const float *z_det,
                      float SO, float OD, float scale, float dz, int nx, int ny, int nz, int na,
                      int nb, int z0, int nzs)
{
#pragma omp parallel for schedule(dynamic, 16) num_threads(NTH)
    for (long k = 0; k < nids; k++) {
        long id = ids[k];
        ax_one_ray(X, &y[k], &sr[k], (int)(id % na), (int)(id / na), cos_phi, sin_phi, y_det,
                   z_det, SO, OD, scale, dz, nx, ny, nz, na, z0, z0 + nzs);
    }
}

/* x[k], sc[k] = the backprojector's sums of voxel ids[k] (idx = (iz*ny +
 * iy)*nx + ix), exactly as orc_atx forms them. */
void orc_atx_voxels(float *x, const float *y, float *sc, const long *ids, long nids,
                    float cos_phi, float sin_phi, const float *y_det, const float *z_det,
                    float SO, float OD, float scale, float dy_det, float dz_det, float dz,
                    int nx, int ny, int nz, int na, int nb)
{
#pragma omp parallel for schedule(static) num_threads(NTH)
    for (long k = 0; k < nids; k++)
        atx_one_voxel(&x[k], y, &sc[k], ids[k], cos_phi, sin_phi, y_det, z_det, SO, OD, scale,
                      dy_det, dz_det, dz, nx, ny, nz, na, nb);
}

/* Bins [a_lo, a_hi] x [b_lo, b_hi] of voxel v's footprint (cu:230-257), as
 * atx_one_voxel bounds them. */
static void voxel_footprint(long v, float cos_phi, float sin_phi, float SO, float OD,
                            float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na,
                            int nb, int *a_lo, int *a_hi, int *b_lo, int *b_hi)
{
    int ix = (int)(v % nx);
    long iyz = v / nx;
    int iy = (int)(iyz % ny), iz = (int)(iyz / ny);
    float SD = SO + OD;
    int na2 = na / 2, nb2 = nb / 2, nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
    float d = (float)sqrt((1 + dz * dz) / 2);
    float zc = (float)(iz + 0.5 - nz2) * dz;
    float yc = (float)(iy + 0.5 - ny2);
    float xc = (float)(ix + 0.5 - nx2);
    float xr = cos_phi * xc + sin_phi * yc;
    float yr = -sin_phi * xc + cos_phi * yc;
    float t = SD / ((xr + SO) * dy_det);
    *a_hi = (int)floorf((yr + 1) * t + na2);
    *a_lo = (int)floorf((yr - 1) * t + na2);
    t = SD / ((xr + SO) * dz_det);
    *b_hi = (int)floorf((zc + d) * t + nb2);
    *b_lo = (int)floorf((zc - d) * t + nb2);
    if (*a_lo < 0) *a_lo = 0;
    if (*b_lo < 0) *b_lo = 0;
    if (*a_hi > na - 1) *a_hi = na - 1;
    if (*b_hi > nb - 1) *b_hi = nb - 1;
}

/* One SART view update (ASD_POCS_cuda.m:116-131 with relaxation `relax`) of
 * the FP64 image m, evaluated only at voxels ids[k]: m_new[k] = max(0,
 * m[v] + relax * bp[v] / sumC[v]) with bp the backprojection of
 * cor = (P_v - proj) ./ sumR (sumR == 0 -> 1), proj and sumR from
 * single(m).  Only the corrections of bins in those voxels' footprints are
 * formed (each from its ray's sums, as orc_ax forms them), so the result
 * equals the full update at those voxels.  phi, SO, OD, vxlsize, detsize as
 * the gateway receives them. */
void orc_sart_view_voxels(const double *m, const double *Pv, double phi, double SO_mm,
                          double OD_mm, double vxlsize, double detsize, int nx, int ny, int nz,
                          int na, int nb, double relax, const long *ids, long nids,
                          double *m_new)
{
    size_t N = (size_t)nx * ny * nz, M = (size_t)na * nb;
    float cp, sp, SO, OD, det, scale;
    float *ydet = (float *)malloc(na * sizeof(float)), *zdet = (float *)malloc(nb * sizeof(float));
    orc_gateway_geometry(phi, SO_mm, OD_mm, vxlsize, detsize, na, nb, &cp, &sp, &SO, &OD, &det,
                         &scale, ydet, zdet);
    float *mf = (float *)malloc(N * sizeof(float));
    float *cor = (float *)calloc(M, sizeof(float));
    unsigned char *need = (unsigned char *)calloc(M, 1);
#pragma omp parallel for schedule(static) num_threads(NTH)
    for (size_t i = 0; i < N; i++) mf[i] = (float)m[i];
    /* the bins the listed voxels read (each flag is only ever set to 1) */
#pragma omp parallel for schedule(static) num_threads(NTH)
    for (long k = 0; k < nids; k++) {
        int a0, a1, b0, b1;
        voxel_footprint(ids[k], cp, sp, SO, OD, det, det, 1.0f, nx, ny, nz, na, nb, &a0, &a1,
                        &b0, &b1);
        for (int ib = b0; ib <= b1; ib++)
            for (int ia = a0; ia <= a1; ia++) need[(size_t)ib * na + ia] = 1;
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(NTH)
    for (long id = 0; id < (long)M; id++) {
        if (!need[id]) continue;
        float proj, sumR;
        ax_one_ray(mf, &proj, &sumR, (int)(id % na), (int)(id / na), cp, sp, ydet, zdet, SO, OD,
                   scale, 1.0f, nx, ny, nz, na, 0, nz);
        double r = (double)sumR;
        if (r == 0) r = 1;
        double c0 = Pv[id] - (double)proj;
        cor[id] = (float)(c0 / r);
    }
#pragma omp parallel for schedule(static) num_threads(NTH)
    for (long k = 0; k < nids; k++) {
        float bp, sumC;
        long v = ids[k];
        atx_one_voxel(&bp, cor, &sumC, v, cp, sp, ydet, zdet, SO, OD, scale, det, det, 1.0f,
                      nx, ny, nz, na, nb);
        double c = (double)sumC;
        if (c == 0) c = 1;
        double nm = m[v] + relax * ((double)bp / c);
        m_new[k] = nm < 0 ? 0 : nm;
    }
    free(ydet); free(zdet); free(mf); free(cor); free(need);
}

/* ------------------------------------------------------------------------ */
/* TV steepest descent (imgTVGradMin.cpp:57-119, imgTVGradMin3D.cpp:76-170)  */
/* ------------------------------------------------------------------------ */

static inline double sq(double a) { return a * a; }

/* Summation order of the FP64 sums of squares behind every norm of the TV
 * descent and the driver (d, ||v||, dp, dg, dd2, RMSE).  The reference forms
 * them with serial loops (imgTVGradMin.cpp:58-59, :96; imgTVGradMin3D.cpp:150-153)
 * and MATLAB's norm / sum (ASD_POCS_cuda.m:122, 141, 145, 166), whose order is
 * not specified.  0 (default): serial, index order.  1: pairwise (index-order
 * blocks of 8, halved recursively).  Test-only: the two orders differ only
 * in rounding, so comparing them measures how far the algorithm itself
 * amplifies an FP64 rounding difference. */
static int g_sum_order = 0;

void orc_set_sum_order(int order) { g_sum_order = order; }

static double pairwise_sumsq(const double *a, const double *b, size_t n)
{
    if (n <= 8) {
        double s = 0;
        for (size_t i = 0; i < n; i++) s += sq(a[i] - (b ? b[i] : 0.0));
        return s;
    }
    size_t h = n / 2;
    return pairwise_sumsq(a, b, h) + pairwise_sumsq(a + h, b ? b + h : NULL, n - h);
}

/* sum over i of (a[i] - b[i])^2 (b NULL: zeros) in the selected order */
static double sumsq_diff(const double *a, const double *b, size_t n)
{
    if (g_sum_order == 1) return pairwise_sumsq(a, b, n);
    double s = 0;
    for (size_t i = 0; i < n; i++) {
        double e = a[i] - (b ? b[i] : 0.0);
        s += e * e;
    }
    return s;
}

/* TV gradient of a 2-D image, rows r (fastest) x cols c, zero outside. */
void orc_tv2d_grad(const double *u, int R, int C, double *v)
{
    const double eps = 1e-8;
#define U2(r, c) (((r) < 0 || (r) >= R || (c) < 0 || (c) >= C) ? 0.0 : u[(size_t)(c) * R + (r)])
    for (int c = 0; c < C; c++)
        for (int r = 0; r < R; r++) {
            double p = U2(r, c);
            double rm = U2(r - 1, c), cm = U2(r, c - 1);
            double rp = U2(r + 1, c), rp_cm = U2(r + 1, c - 1);
            double cp = U2(r, c + 1), rm_cp = U2(r - 1, c + 1);
            double g1 = ((p - rm) + (p - cm)) / sqrt(eps + sq(p - rm) + sq(p - cm));
            double g2 = -(rp - p) / sqrt(eps + sq(rp - p) + sq(rp - rp_cm));
            double g3 = -(cp - p) / sqrt(eps + sq(cp - p) + sq(cp - rm_cp));
            v[(size_t)c * R + r] = g1 + g2 + g3;
        }
#undef U2
}

/* TV gradient of a 3-D volume (x fastest, then y, then z), zero outside. */
void orc_tv3d_grad(const double *u, int Nx, int Ny, int Nz, double *v)
{
    const double eps = 1e-8;
#define U3(x, y, z) (((x) < 0 || (x) >= Nx || (y) < 0 || (y) >= Ny || (z) < 0 || (z) >= Nz) ? 0.0 \
                     : u[((size_t)(z) * Ny + (y)) * Nx + (x)])
#pragma omp parallel for schedule(static) num_threads(NTH)
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                double p = U3(x, y, z);
                double xm = U3(x - 1, y, z), ym = U3(x, y - 1, z), zm = U3(x, y, z - 1);
                double zp = U3(x, y, z + 1), xm_zp = U3(x - 1, y, z + 1), ym_zp = U3(x, y - 1, z + 1);
                double yp = U3(x, y + 1, z), xm_yp = U3(x - 1, y + 1, z), yp_zm = U3(x, y + 1, z - 1);
                double xp = U3(x + 1, y, z), xp_ym = U3(x + 1, y - 1, z), xp_zm = U3(x + 1, y, z - 1);
                double g1 = (6 * p - 2 * (xm + ym + zm)) /
                            sqrt(eps + sq(p - xm) + sq(p - ym) + sq(p - zm));
                double g2 = (2 * p - 2 * zp) / sqrt(eps + sq(zp - xm_zp) + sq(zp - ym_zp) + sq(zp - p));
                double g3 = (2 * p - 2 * yp) / sqrt(eps + sq(yp - xm_yp) + sq(yp - p) + sq(yp - yp_zm));
                double g4 = (2 * p - 2 * xp) / sqrt(eps + sq(xp - p) + sq(xp - xp_ym) + sq(xp - xp_zm));
                v[((size_t)z * Ny + y) * Nx + x] = g1 + g2 + g3 + g4;
            }
#undef U3
}

/* nTVIter normalised gradient steps of length alpha*||I - I0|| (I0 = zeros
 * when img_old is NULL), then positivity when pos==1 and nTVIter>0.
 * dims: (R, C, 1) selects the 2-D stencil, Nz >= 1 with three_d selects 3-D.
 * out may alias img.  Returns 0. */
static int tv_common(const double *img, int Nx, int Ny, int Nz, int three_d, double alpha,
                     int nTVIter, int pos, const double *img_old, double *out)
{
    size_t n = (size_t)Nx * Ny * Nz;
    double *v = (double *)malloc(n * sizeof(double));
    if (out != img) memcpy(out, img, n * sizeof(double));
    double d = sqrt(sumsq_diff(out, img_old, n));
    for (int it = 0; it < nTVIter; it++) {
        if (three_d) orc_tv3d_grad(out, Nx, Ny, Nz, v);
        else orc_tv2d_grad(out, Nx, Ny, v);
        double nrm = sqrt(sumsq_diff(v, NULL, n));
        for (size_t i = 0; i < n; i++) out[i] -= alpha * d * v[i] / nrm;
    }
    if (pos == 1 && nTVIter > 0)
        for (size_t i = 0; i < n; i++)
            if (out[i] < 0) out[i] = 0;
    free(v);
    return 0;
}

int orc_tv2d(const double *img, int rowN, int colN, double alpha, int nTVIter, int pos,
             const double *img_old, double *out)
{
    return tv_common(img, rowN, colN, 1, 0, alpha, nTVIter, pos, img_old, out);
}

int orc_tv3d(const double *img, int Nx, int Ny, int Nz, double alpha, int nTVIter, int pos,
             const double *img_old, double *out)
{
    return tv_common(img, Nx, Ny, Nz, 1, alpha, nTVIter, pos, img_old, out);
}

/* ------------------------------------------------------------------------ */
/* SART sweep and ASD-POCS driver (ASD_POCS_cuda.m, SART_cuda_inline.m)      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz, na, nb;
    double SO, OD, vxlsize, detsize;
    int maxnumIter, numOfgradients;
    double alpha, alpha_red, beta, beta_red, r_max, epsilon, stopRelChg;
} orc_para;

typedef struct {
    int IterationNum;
    double dataTolerance, beta, alpha;
} orc_out;

#define ORC_TRACE_COLS 5

/* One SART sweep over all views in order (ASD_POCS_cuda.m:110-134):
 * m is FP64 and cast to single for the operators, like MATLAB's single().
 * Returns the accumulated data error dd2. */
static double sart_sweep(double *m, const double *P, const double *phi_deg, int nv,
                         const orc_para *pa, double beta)
{
    int nx = pa->nx, ny = pa->ny, nz = pa->nz, na = pa->na, nb = pa->nb;
    size_t N = (size_t)nx * ny * nz, M = (size_t)na * nb;
    float *mf = (float *)malloc(N * sizeof(float));
    float *proj = (float *)malloc(M * sizeof(float)), *sumR = (float *)malloc(M * sizeof(float));
    float *corf = (float *)malloc(M * sizeof(float));
    double *res = (double *)malloc(M * sizeof(double));
    float *bp = (float *)malloc(N * sizeof(float)), *sumC = (float *)malloc(N * sizeof(float));
    float *ydet = (float *)malloc(na * sizeof(float)), *zdet = (float *)malloc(nb * sizeof(float));
    double dd2 = 0;
    for (int v = 0; v < nv; v++) {
        double theta = phi_deg[v] / 180 * M_PI;
        double phi = -1 * theta;
        float cp, sp, SO, OD, det, scale;
        orc_gateway_geometry(phi, pa->SO, pa->OD, pa->vxlsize, pa->detsize, na, nb, &cp, &sp,
                             &SO, &OD, &det, &scale, ydet, zdet);
        for (size_t i = 0; i < N; i++) mf[i] = (float)m[i];
        orc_ax(mf, proj, sumR, cp, sp, ydet, zdet, SO, OD, scale, 1.0f, nx, ny, nz, na, nb);
        for (size_t i = 0; i < M; i++) {
            double r = (double)sumR[i];
            if (r == 0) r = 1;
            double c0 = P[(size_t)v * M + i] - (double)proj[i];
            res[i] = c0;
            corf[i] = (float)(c0 / r);
        }
        dd2 += sumsq_diff(res, NULL, M);
        orc_atx(bp, corf, sumC, cp, sp, ydet, zdet, SO, OD, scale, det, det, 1.0f, nx, ny, nz,
                na, nb);
        for (size_t i = 0; i < N; i++) {
            double c = (double)sumC[i];
            if (c == 0) c = 1;
            double nm = m[i] + beta * ((double)bp[i] / c);
            m[i] = nm < 0 ? 0 : nm;
        }
    }
    free(mf); free(proj); free(sumR); free(corf); free(res); free(bp); free(sumC); free(ydet);
    free(zdet);
    return dd2;
}

static double nrm2diff(const double *a, const double *b, size_t n)
{
    return sqrt(sumsq_diff(a, b, n));
}

/* ASD_POCS_cuda(P, para) with m0 given (zeros when m0 is NULL).
 * P: (na*nb) x nv, view-major (each view contiguous), FP64.
 * m_res: N doubles, the post-SART image of the last iteration (m:138).
 * rmse: maxnumIter doubles (filled up to IterationNum) when I0 != NULL.
 * max_views: process only the first max_views views of each sweep when > 0
 * (used to time a bounded sample). */
static int asd_pocs_impl(const double *P, const double *phi_deg, int nv, const orc_para *pa,
                         const double *m0, const double *I0, double *m_res, double *rmse,
                         orc_out *out, double *trace);

int orc_asd_pocs(const double *P, const double *phi_deg, int nv, const orc_para *pa,
                 const double *m0, const double *I0, double *m_res, double *rmse, orc_out *out)
{
    return asd_pocs_impl(P, phi_deg, nv, pa, m0, I0, m_res, rmse, out, NULL);
}

/* The same loop, also recording per outer iteration k (row k-1 of `trace`,
 * ORC_TRACE_COLS doubles): dp = ||m - m_pre|| after the sweep, dg = ||m_tv -
 * f0|| after the TV descent (0 when the stop rule ended the iteration
 * first), dd2, and alpha and beta after the iteration's rules.  Test-only:
 * it lets the pins observe the step-adaptation inputs the loop never
 * returns. */
int orc_asd_pocs_trace(const double *P, const double *phi_deg, int nv, const orc_para *pa,
                       const double *m0, const double *I0, double *m_res, double *rmse,
                       orc_out *out, double *trace)
{
    return asd_pocs_impl(P, phi_deg, nv, pa, m0, I0, m_res, rmse, out, trace);
}

static int asd_pocs_impl(const double *P, const double *phi_deg, int nv, const orc_para *pa,
                         const double *m0, const double *I0, double *m_res, double *rmse,
                         orc_out *out, double *trace)
{
    size_t N = (size_t)pa->nx * pa->ny * pa->nz;
    double *m = (double *)calloc(N, sizeof(double));
    double *m_pre = (double *)malloc(N * sizeof(double));
    double *m_k1 = (double *)malloc(N * sizeof(double));
    double *f0 = (double *)malloc(N * sizeof(double));
    if (m0) memcpy(m, m0, N * sizeof(double));
    double alpha = pa->alpha, beta = pa->beta;
    double dd2 = 0;
    int k;
    for (k = 1; k <= pa->maxnumIter; k++) {
        memcpy(m_pre, m, N * sizeof(double));
        dd2 = sart_sweep(m, P, phi_deg, nv, pa, beta);
        double dd = sqrt(dd2);
        memcpy(m_res, m, N * sizeof(double));
        if (I0 && rmse) {
            rmse[k - 1] = sqrt(sumsq_diff(m, I0, N) / (double)N);
        }
        double dp = nrm2diff(m, m_pre, N);
        double *tr = trace ? trace + (size_t)(k - 1) * ORC_TRACE_COLS : NULL;
        if (tr) { tr[0] = dp; tr[1] = 0; tr[2] = dd2; tr[3] = alpha; tr[4] = beta; }
        if (dp / (double)N <= pa->stopRelChg) break;
        if (k == 1) memcpy(m_k1, m_pre, N * sizeof(double));
        memcpy(f0, m, N * sizeof(double));
        if (pa->nz >= 2)
            orc_tv3d(m, pa->nx, pa->ny, pa->nz, alpha, pa->numOfgradients, 1, m_k1, m);
        else
            orc_tv2d(m, pa->nx, pa->ny, alpha, pa->numOfgradients, 1, m_k1, m);
        double dg = nrm2diff(m, f0, N);
        if (dg > pa->r_max * dp && dd > pa->epsilon) alpha = alpha * pa->alpha_red;
        beta = beta * pa->beta_red;
        if (tr) { tr[1] = dg; tr[3] = alpha; tr[4] = beta; }
    }
    if (k > pa->maxnumIter) k = pa->maxnumIter;
    out->IterationNum = k;
    out->dataTolerance = dd2;
    out->beta = beta;
    out->alpha = alpha;
    free(m); free(m_pre); free(m_k1); free(f0);
    return 0;
}

/* SART_cuda_inline(P, para): exactly one sweep with relaxation lambda
 * (SART_cuda_inline.m:56-87).  m (N doubles) holds m0 on entry (zeros when
 * the caller has none) and the result on exit.  Returns dataTolerance. */
double orc_sart(const double *P, const double *phi_deg, int nv, const orc_para *pa,
                double lambda, double *m)
{
    return sart_sweep(m, P, phi_deg, nv, pa, lambda);
}
