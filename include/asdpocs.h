/*
 * asdpocs.h - C ABI of the MI355X ASD-POCS hot path (libasdpocs.so).
 *
 * Every entry point takes plain pointers and sizes.  Geometry arguments of the
 * projector launchers are in voxel units, exactly as the reference's MEX
 * gateways hand them to its CUDA launchers.  Volumes are stored x fastest
 * (MATLAB column-major nx*ny*nz), projections ia fastest (na*nb), and a
 * sinogram is (na*nb) x nv with each view contiguous.
 *
 * Functions returning int return ASD_OK (0) or an ASD_ERR_* code; the text of
 * the last error on the calling thread is available from asd_last_error().
 */
#ifndef ASDPOCS_H
#define ASDPOCS_H

#ifdef __cplusplus
extern "C" {
#endif

#define ASD_OK 0
#define ASD_ERR_ARG 1    /* invalid argument (sizes, NULL pointers, geometry) */
#define ASD_ERR_HIP 2    /* a HIP runtime call or kernel launch failed */
#define ASD_ERR_ALLOC 3  /* device or host allocation failed */
#define ASD_ERR_STATE 4  /* plan used before it was given projections */

/* ---------------------------------------------------------------------- */
/* Drop-in projector launchers (host pointers, synchronous)                */
/* ---------------------------------------------------------------------- */

/* One view of the ray-driven forward projector: y = vxl * sum_vox l * X,
 * sr = vxl * sum_vox l, for na*nb rays (id = ib*na + ia).
 * Replaces Ax_cone_mf_gpu_new, src/AxAtx/Ax_cone_gpu.cu:778-805 (same
 * signature, same argument meaning, host pointers in and out).  The reference
 * returns no status; asd_last_status() reports this call's. */
void Ax_cone_mf_gpu_new(float *X, float *y, float *sr, float cos_phi, float sin_phi,
                        float *y_det, float *z_det, float SO, float OD, float scale, float dz,
                        int nx, int ny, int nz, int na, int nb);

/* One view of the voxel-driven backprojector: X = vxl * sum_bins l * y,
 * sc = vxl * sum_bins l, over each voxel's detector footprint.
 * Replaces Atx_cone_mf_gpu_new, src/AxAtx/Ax_cone_gpu.cu:299-331 (same
 * signature, host pointers). */
void Atx_cone_mf_gpu_new(float *X, float *y, float *sc, float cos_phi, float sin_phi,
                         float *y_det, float *z_det, float SO, float OD, float scale,
                         float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na,
                         int nb);

/* Status and message of the last call made on this thread. */
int asd_last_status(void);
const char *asd_last_error(void);

/* ---------------------------------------------------------------------- */
/* Device-resident projector forms                                         */
/* ---------------------------------------------------------------------- */

/* Same operators on device pointers (HBM-resident X, y, y_det, z_det),
 * enqueued on `stream` (a hipStream_t, NULL = the null stream); they return
 * after the launch.  Same role as the kernels launched at
 * src/AxAtx/Ax_cone_gpu.cu:796 and :318-326, without the per-call
 * malloc/copy/free of the reference launchers. */
int asd_ax_device(const float *X, float *y, float *sr, float cos_phi, float sin_phi,
                  const float *y_det, const float *z_det, float SO, float OD, float scale,
                  float dz, int nx, int ny, int nz, int na, int nb, void *stream);
int asd_atx_device(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
                   const float *y_det, const float *z_det, float SO, float OD, float scale,
                   float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na, int nb,
                   void *stream);

/* The operators on one z-slab of an nx x ny x nz volume, the unit of a
 * volume sharded over GPUs by slices (no counterpart in the reference, whose
 * kernels take the whole volume, src/AxAtx/Ax_cone_gpu.cu:315-316, :796).
 * The geometry is the whole volume's; the slab is slices [z0, z0 + nzs),
 * and X (nx * ny * nzs floats, x fastest) holds only those slices.
 * asd_ax_slab_device: y and sr are the slab's partial ray sums (the pieces
 * of each ray inside the slab); the partial sums of all slabs add up to
 * asd_ax_device's y and sr (to the order of the additions).
 * asd_atx_slab_device: X and sc receive the slab's voxels, each equal bit
 * for bit to asd_atx_device's value for that voxel. */
int asd_ax_slab_device(const float *X, float *y, float *sr, float cos_phi, float sin_phi,
                       const float *y_det, const float *z_det, float SO, float OD, float scale,
                       float dz, int nx, int ny, int nz, int na, int nb, int z0, int nzs,
                       void *stream);
int asd_atx_slab_device(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
                        const float *y_det, const float *z_det, float SO, float OD, float scale,
                        float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na,
                        int nb, int z0, int nzs, void *stream);

/* asd_atx_device without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (work under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments
 * and results as asd_atx_device; its ray-slope table lives in a fixed
 * device array of 2^20 records (32 MiB, held by each device the library
 * runs on until the process ends), so na * nb must not exceed 1048576
 * (ASD_ERR_ARG otherwise).  Calls on one device share that array: calls
 * that overlap on one device (on different streams) race; calls on one
 * stream are ordered and safe. */
int asd_atx_device_noalloc(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
                           const float *y_det, const float *z_det, float SO, float OD,
                           float scale, float dy_det, float dz_det, float dz, int nx, int ny,
                           int nz, int na, int nb, void *stream);

/* ---------------------------------------------------------------------- */
/* TV steepest descent                                                     */
/* ---------------------------------------------------------------------- */

/* I2 = imgTVGradMin(I, alpha, nTVIter, pos[, I0]) on host FP64 arrays of
 * rowN x colN (column-major).  img_old may be NULL (zeros).  out may alias img.
 * Replaces the compute of src/tvmin/imgTVGradMin.cpp:36-123. */
int asd_tv_min_2d(const double *img, int rowN, int colN, double alpha, int nTVIter, int pos,
                  const double *img_old, double *out);

/* I2 = imgTVGradMin3D(I, alpha, nTVIter, pos[, I0]) on host FP64 arrays of
 * Nx x Ny x Nz (x fastest).  Replaces src/tvmin/imgTVGradMin3D.cpp:44-173. */
int asd_tv_min_3d(const double *img, int Nx, int Ny, int Nz, double alpha, int nTVIter,
                  int pos, const double *img_old, double *out);

/* In-place FP64 TV descent of a device array (three_d = 0: nx x ny 2-D stencil). */
int asd_tv_min_device(double *img, const double *img_old, int nx, int ny, int nz, int three_d,
                      double alpha, int nTVIter, int pos, void *stream);

/* ---------------------------------------------------------------------- */
/* ASD-POCS / SART drivers                                                 */
/* ---------------------------------------------------------------------- */

/* Mirror of the `para` struct read by ASD_POCS_cuda.m (utilities/ASD_POCS_cuda.m:20-96)
 * and the MEX gateways.  asd_para_defaults() fills the reference defaults. */
typedef struct asd_para {
    int nx, ny, nz;         /* volume size (voxels) */
    double SO, OD;          /* source-isocentre and isocentre-detector distances (mm) */
    double vxlsize;         /* voxel size (mm) */
    double detsize;         /* detector bin size (mm) */
    int na, nb;             /* detector bins (horizontal, vertical) */
    int maxnumIter;         /* default 5 */
    int numOfgradients;     /* TV sub-iterations ng, default 20 */
    double alpha;           /* TV step factor, default 0.02 */
    double alpha_red;       /* default 0.95 */
    double beta;            /* SART relaxation, default 1 */
    double beta_red;        /* default 0.995 */
    double r_max;           /* default 0.95 */
    double epsilon;         /* default eps = 2^-52 */
    double stopRelChg;      /* default eps */
    int initBySART;         /* 1: m0 = one SART sweep (SART_cuda_inline.m) */
    double lambda;          /* SART relaxation of the warm start, default 1 */
    int precision;          /* 0: plain FP32 (default);
                             * 1: 16-bit fixed-point sinogram per view, compensated sums;
                             * 2: FP16 sinogram, compensated sums;
                             * 3: as 1, and the backprojector sets up its rays in FP64
                             *    per voxel tile (lengths within ~1e-6 voxel of exact);
                             * 4: the reference's arithmetic: sinogram, image, m_pre,
                             *    m_k1, f0 and TV in FP64 as ASD_POCS_cuda.m holds
                             *    them, single(m) fed to a forward projector that sums
                             *    each ray in slab order, the update in FP64;
                             * 5: FP16 sinogram, plain FP32 sums */
} asd_para;

typedef struct asd_out {
    int IterationNum;       /* out.IterationNum */
    double dataTolerance;   /* out.dataTolerance (dd2 of the last sweep) */
    double beta;            /* out.beta */
    double alpha;           /* out.alpha */
} asd_out;

void asd_para_defaults(asd_para *para);

/* [m_res, out] = ASD_POCS_cuda(P, para).  P: host FP64 (na*nb) x nv, view
 * contiguous; phi_vec: nv angles in degrees.  m0 / I0 may be NULL.  m_res:
 * N doubles (the post-SART image of the last iteration).  rmse: maxnumIter
 * doubles or NULL.  Replaces utilities/ASD_POCS_cuda.m:1-182 with the image,
 * sinogram and all per-view work resident in HBM. */
int asd_pocs(const double *P, const double *phi_vec, int nv, const asd_para *para,
             const double *m0, const double *I0, double *m_res, double *rmse, asd_out *out);

/* [m, out] = SART_cuda_inline(P, para): one sweep with relaxation
 * para->lambda.  m: N doubles, m0 on entry (zeros if the caller has none),
 * the result on exit.  Replaces utilities/SART_cuda_inline.m:56-87. */
int asd_sart(const double *P, const double *phi_vec, int nv, const asd_para *para, double *m,
             double *dataTolerance);

/* Device-resident plan: geometry tables, sinogram and image stay in HBM
 * across calls (the unit the iteration benchmark times). */
typedef struct asd_plan asd_plan;
int asd_plan_create(const asd_para *para, const double *phi_vec, int nv, asd_plan **plan);
int asd_plan_set_projections(asd_plan *plan, const double *P_host);
int asd_plan_reset(asd_plan *plan, const double *m0_host, const double *I0_host);
/* Run up to n_iter outer iterations (continuing the counters of earlier calls). */
int asd_plan_iterate(asd_plan *plan, int n_iter);
int asd_plan_get_image(asd_plan *plan, double *m_res_host);
int asd_plan_get_out(asd_plan *plan, asd_out *out, double *rmse_host);
/* One SART sweep over all views with relaxation `lambda` on the plan's
 * current image (SART_cuda_inline.m:56-87); *dataTolerance = its data error. */
int asd_plan_sweep(asd_plan *plan, double lambda, double *dataTolerance);
/* Current image (after TV of the last iteration, or after a sweep). */
int asd_plan_get_current(asd_plan *plan, double *m_host);
/* Synchronise the plan's stream (for timing). */
int asd_plan_sync(asd_plan *plan);
/* Fills the plan with simulated data on the device.  No counterpart in the
 * reference, whose demos build their inputs in MATLAB (demo_test.m:35-52).
 *  - The ellipsoid phantom (n_ell rows of A, a, b, c, x0, y0, z0, phi,
 *    theta, psi in degrees, the layout of utilities/phantom3d.m:183-200;
 *    coordinates in [-1, 1] across the grid; values summed in row order,
 *    times scale) becomes the plan's RMSE reference; a later asd_plan_reset
 *    with I0 == NULL keeps it.  3-D volumes follow phantom3d.m:76-108 (x
 *    along the second index, the Euler rotation applied before the
 *    offsets); a one-slice image rotates each ellipse by phi about its
 *    centre, x along the first index.
 *  - Every view is forward projected.  With I0_photons > 0 each projection
 *    value p becomes -ln(max(k, 1) / I0_photons), k ~ Poisson(I0_photons *
 *    exp(-p)), from a Philox4x32-10 stream per value keyed by seed.
 *  - The result is encoded into the plan's sinogram storage as
 *    asd_plan_set_projections would encode it. */
int asd_plan_simulate(asd_plan *plan, const double *ellipsoids, int n_ell, double scale,
                      double I0_photons, unsigned long long seed);
/* The plan's sinogram decoded to FP64, views contiguous (nv x na*nb). */
int asd_plan_get_projections(asd_plan *plan, double *P_host);
/* The plan's RMSE reference image (FP32 values as FP64). */
int asd_plan_get_reference(asd_plan *plan, double *I0_host);
/* Per-kernel event timing on the plan stream: stage 0 = forward projector,
 * 1 = backprojector + update, 2 = TV gradient.  Enabling resets the totals. */
int asd_plan_set_timing(asd_plan *plan, int enable);
int asd_plan_get_timing(asd_plan *plan, int stage, double *total_ms, long *count);
/* 1 (default): replay one outer iteration as a captured hipGraph;
 * 0: enqueue its kernels one by one (always the case while timing is on). */
int asd_plan_set_graph(asd_plan *plan, int enable);
/* Backprojector variant the plan's autotuner picked on its first sweep
 * (0 direct, 1 16x16x1 tiles, 2 8x8x4, 3 16x4x4, 4 voxel-bin pairs, 5 direct
 * with two-pair load batches, 6 direct one pair at a time, 7 to 9 two-pair
 * walks of 16x16x1 / 8x8x4 / 8x4x8 tiles regrouped by footprint size; -1
 * before the first sweep) and the measured time of each (atx_us[10],
 * microseconds per view, the mean over the three views timed: the first and
 * those an eighth and a quarter of the way round; -1 = not tried).  All
 * variants give bit-identical results. */
int asd_plan_tuning(asd_plan *plan, int *atx_variant, double *atx_us);
/* Forces the backprojector variant (0..9, as above; the tiles with four
 * slices need nz > 1) or, with -1, asks for a new measurement on the next
 * sweep. */
int asd_plan_set_atx_variant(asd_plan *plan, int atx_variant);
/* Forces the forward projector's launch shape: lanes per ray (1..256, a
 * power of two) with 1 row per bundle, or 8 lanes with bundles of 2 or 3
 * detector rows; (0, 0) restores the plan's default (for precision 4, the
 * slab-order projector).  The shape changes the order in which each ray's
 * slab products are summed (the last bits of the sums), not which products
 * there are. */
int asd_plan_set_ax_shape(asd_plan *plan, int lanes_per_ray, int rows_per_bundle);
/* Launch shape of the forward projector (lanes per ray, or per bundle of
 * rows_per_bundle detector rows that share a column; 0 lanes: the
 * slab-order projector of precision 4) and the plan's hipStream_t.  Any
 * pointer may be NULL. */
int asd_plan_info(asd_plan *plan, int *lanes_per_ray, int *rows_per_bundle, void **stream);
void asd_plan_destroy(asd_plan *plan);

/* ---------------------------------------------------------------------- */
/* Batched reconstructions                                                 */
/* ---------------------------------------------------------------------- */

/* Many independent ASD-POCS reconstructions (each with its own para, size
 * and view count) advanced together: one kernel launch per stage for all
 * cases, one hipGraph per outer iteration of the batch.  All cases share
 * para.precision.  No counterpart in the reference, which runs one
 * reconstruction per MATLAB call (utilities/ASD_POCS_cuda.m). */
typedef struct asd_batch asd_batch;
int asd_batch_create(int ncases, const asd_para *paras, const double *const *phi_vecs,
                     const int *nvs, asd_batch **batch);
/* Sinogram of case c ((na*nb) x nv, view contiguous), its start image m0
 * (NULL: zeros) and reference image I0 for the RMSE (NULL: none). */
int asd_batch_set_case(asd_batch *batch, int c, const double *P, const double *m0,
                       const double *I0);
int asd_batch_iterate(asd_batch *batch, int n_iter);
int asd_batch_get_image(asd_batch *batch, int c, double *m_res);
int asd_batch_get_out(asd_batch *batch, int c, asd_out *out, double *rmse);
int asd_batch_set_graph(asd_batch *batch, int enable);
/* The backprojector stage's footprint pairs per load batch, picked by the
 * batch's autotuner on its first asd_batch_iterate (0 before), and the
 * tuning times of 1, 2 and 4 pairs in µs (-1 before); every choice gives the
 * same bits.  Either pointer may be NULL. */
int asd_batch_tuning(asd_batch *batch, int *chunk, double *us3);
/* Forces the backprojector stage's pairs per load batch (1, 2 or 4; 0: let
 * the autotuner pick on the next asd_batch_iterate). */
int asd_batch_set_atx_chunk(asd_batch *batch, int chunk);
int asd_batch_sync(asd_batch *batch);
void asd_batch_destroy(asd_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* ASDPOCS_H */
