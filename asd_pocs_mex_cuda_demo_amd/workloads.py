"""The five reconstruction workloads the benchmark and the per-workload GPU
tests run, and their synthetic inputs.

  fan2d_256   2-D fan beam, 256^2 at 1.0 mm, 90 views over 360 deg, 512 bins
              at 1.0 mm, SID 500 / SDD 1000 mm, noiseless projections, 20
              iterations (the reference's own CPU-runnable case)
  fan2d_512   2-D fan beam, 512^2 at 0.5 mm, 120 views, 768 bins at 0.8 mm,
              SID 541 / SDD 949 mm, Poisson noise I0 = 1e5, 50 iterations
  cone3d_256  3-D cone beam, 256^3 at 1.0 mm, 180 views, 384 x 384 bins at
              1.2 mm, SID 500 / SDD 1000 mm, Poisson noise I0 = 5e4, 30
              iterations
  batch64     64 independent 2-D fan-beam cases (batch_cases)
  cone3d_512  3-D cone beam, 512^3 at 0.5 mm, 720 views, 1024 x 768 bins at
              0.6 mm, SID 600 / SDD 1100 mm, Shepp-Logan plus 200 seeded
              spheres, FP16-stored projections with plain FP32 sums
              (para.precision 5), 20 iterations; simulated on the device
              (phantom, projections)

Every workload runs ASD-POCS with alpha = 0.2, 20 TV sub-iterations and the
driver's other defaults (beta = 1 decaying by 0.995, alpha_red = 0.95,
r_max = 0.95), over views at (1..nv) * 360 / nv degrees (demo_test.m:39-40).
"""
import numpy as np

from . import ops, phantoms

CONFIGS = {
    "fan2d_256": dict(n=256, nz=1, vxl=1.0, nv=90, na=512, nb=1, det=1.0, SO=500.0, SDD=1000.0,
                      I0=0.0, iters=20),
    "fan2d_512": dict(n=512, nz=1, vxl=0.5, nv=120, na=768, nb=1, det=0.8, SO=541.0, SDD=949.0,
                      I0=1e5, iters=50),
    "cone3d_256": dict(n=256, nz=256, vxl=1.0, nv=180, na=384, nb=384, det=1.2, SO=500.0,
                       SDD=1000.0, I0=5e4, iters=30),
    "cone3d_512": dict(n=512, nz=512, vxl=0.5, nv=720, na=1024, nb=768, det=0.6, SO=600.0,
                       SDD=1100.0, I0=0.0, iters=20, precision=5, sim="gpu", spheres=200),
}
SEED = 20241203
BATCH_SEED = 20241204


def para(name, **over):
    """The para struct of a workload (ASD_POCS_cuda.m field names) with its
    view angles in para['phi_vec']; keyword arguments override fields."""
    w = CONFIGS[name]
    p = dict(nx=w["n"], ny=w["n"], nz=w["nz"], SO=w["SO"], OD=w["SDD"] - w["SO"],
             vxlsize=w["vxl"], detsize=w["det"], na=w["na"], nb=w["nb"], alpha=0.2,
             numOfgradients=20, maxnumIter=w["iters"], precision=w.get("precision", 0))
    p["phi_vec"] = np.arange(1, w["nv"] + 1) * 360.0 / w["nv"]
    p.update(over)
    return p


def phantom(name):
    """The modified Shepp-Logan phantom of a workload (x fastest), FP64."""
    w = CONFIGS[name]
    if w["nz"] == 1:
        return phantoms.shepp_logan_2d(w["n"])
    return phantoms.shepp_logan_3d(w["n"], nz=w["nz"])


def sinogram(p, img, I0, seed=SEED):
    """Projections ((na*nb) x nv) of img with the drop-in forward projector,
    then Poisson noise at I0 photons per bin and the log transform (none
    when I0 is 0)."""
    P = np.zeros((p["na"] * p["nb"], len(p["phi_vec"])))
    x = np.asarray(img, np.float32)
    for i, th in enumerate(p["phi_vec"]):
        P[:, i] = ops.Ax_cone_gpu(x, dict(p, phi=-th / 180 * np.pi))[0]
    return phantoms.poisson_log(P, I0, seed) if I0 else P


def problem(name, seed=SEED, **over):
    """(para, phantom, sinogram) of a workload built on the host."""
    p = para(name, **over)
    img = phantom(name)
    return p, img, sinogram(p, img, CONFIGS[name]["I0"], seed)


def sphere_table(k, seed, width_mm):
    """k seeded spheres (radius 1-4 mm, attenuation 0.01-0.03 /mm, centres in
    the inner 60% of the phantom) as phantom rows, values relative to the
    phantom's scale of 0.02 /mm, coordinates in units of half the width."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(1.0, 4.0, k) / (width_mm / 2)
    mu = rng.uniform(0.01, 0.03, k) / 0.02
    c = rng.uniform(-0.6, 0.6, (k, 3)) * np.array([0.69, 0.92, 0.81])
    return np.column_stack([mu, r, r, r, c, np.zeros((k, 3))])


def phantom_table(name, seed=SEED):
    """The ellipsoid table of a workload's phantom (with its spheres)."""
    w = CONFIGS[name]
    return np.vstack([phantoms.ellipses(),
                      sphere_table(w.get("spheres", 0), seed, w["n"] * w["vxl"])])


def sim_plan(name, precision=None, seed=SEED, phi_vec=None, **over):
    """A plan of the workload whose phantom and projections are simulated on
    the device (asd_plan_simulate); returns (para, plan)."""
    w = CONFIGS[name]
    p = para(name, **over)
    if precision is not None:
        p["precision"] = precision
    if phi_vec is not None:
        p["phi_vec"] = np.asarray(phi_vec, np.float64)
    plan = ops.Plan(p, p["phi_vec"])
    plan.simulate(phantom_table(name, seed), scale=0.02, I0=w["I0"], seed=seed)
    return p, plan


def batch_cases(ncases=64, seed=BATCH_SEED):
    """64 independent 2-D fan-beam cases: N from {128,192,256,384,512},
    views uniform in [60, 360] over 360 deg, 1.5N bins, perturbed modified
    Shepp-Logan phantoms (intensity +-10%, axes +-5%), Poisson I0 = 1e5.
    Geometry (not fixed by the workload): SID 541 mm, SDD 949 mm, pixel
    256/N mm (a 256 mm field), bin 1.6 x pixel.  Returns [(para, img, P)]."""
    rng = np.random.default_rng(seed)
    cases = []
    for c in range(ncases):
        n = int(rng.choice([128, 192, 256, 384, 512]))
        nv = int(rng.integers(60, 361))
        vxl = 256.0 / n
        p = dict(nx=n, ny=n, nz=1, SO=541.0, OD=408.0, vxlsize=vxl, detsize=1.6 * vxl,
                 na=int(1.5 * n), nb=1, alpha=0.2, numOfgradients=20, maxnumIter=20)
        p["phi_vec"] = np.arange(1, nv + 1) * 360.0 / nv
        img = phantoms.shepp_logan_2d(n, table=phantoms.ellipses(rng))
        cases.append((p, img, sinogram(p, img, 1e5, seed + c)))
    return cases


def quantize_u16(P):
    """The 16-bit fixed-point codec of para.precision 1 and 3 applied to a
    sinogram ((na*nb) x nv), decoded back to FP64: per view
    lo + step * round((P - lo) / step), step = (hi - lo) / 65535."""
    P = np.asarray(P, np.float64)
    out = np.empty_like(P)
    for v in range(P.shape[1]):
        lo, hi = P[:, v].min(), P[:, v].max()
        step = (hi - lo) / 65535.0 if hi > lo else 0.0
        q = np.clip(np.floor((P[:, v] - lo) / step + 0.5), 0, 65535) if step > 0 else 0.0
        out[:, v] = lo + step * q
    return out
