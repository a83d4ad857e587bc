"""ASD-POCS iteration benchmark on one MI355X.

Headline workload (default, the largest configuration): cone3d_512, 3-D
cone-beam ASD-POCS on a 512^3 volume (0.5 mm voxels, Shepp-Logan plus 200
seeded spheres), 720 views over 360 deg, 1024 x 768 bins at 0.6 mm, SID
600 mm / SDD 1100 mm, FP16-stored projections with FP32 sums,
simulated on the device; 20 iterations of ASD-POCS with 20 TV sub-iterations
(alpha = 0.2).  A step is one outer iteration (a SART sweep over the 720
views with positivity, the TV descent and the step adaptation) with the
sinogram and image resident in HBM.  The other workloads of
asd_pocs_mex_cuda_demo_amd/workloads.py are selected with --workload.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--workload NAME]
                    [--precision P] [--no-cpu] [--no-extras]

Prints one JSON line: value = iterations/s of the headline workload, the
roofline of its dominant kernel (HIP events on the plan's stream, per
launch), the CPU oracle timed on a bounded sample of the same workload, and
(unless --no-extras) the rates of the other single-plan workloads, the
64-case batch against 64 sequential plans, and the accuracy of the 512^2
fan-beam reconstruction against the CPU oracle over its 50 iterations.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ASD-POCS iters/sec and projector GB/s (% HBM roofline), batched-mixed vs plain at RMSE"
HBM_PEAK_GBS = 8000.0
WORKLOAD_CHOICES = ["fan2d_256", "fan2d_512", "cone3d_256", "cone3d_512", "batch64"]

CODEC = {0: "SinoF32", 1: "SinoU16", 2: "SinoF16", 3: "SinoU16", 4: "SinoF64", 5: "SinoF16"}
ATX_VARIANTS = ["direct", "tile16x16x1", "tile8x8x4", "tile16x4x4", "pairs", "direct2",
                "direct1", "sorted16", "sorted8x8x4", "sorted8x4x8"]


def atx_kernel_name(variant, kahan, precision=0, three_d=True, general=False):
    """The backprojector kernel a plan launches for its views.  general: the
    views' rays are all general rays (cone views with fewer than 2048 bins per
    detector axis; a view with an a == 0 or b == 0 ray, which none of the
    workloads has, keeps atx_sorted_kernel), where the sorted variants run
    atx_sorted_general_kernel."""
    k = "true" if kahan else "false"
    if precision == 3:  # the tiled backprojector with FP64 per-tile ray setup
        t = "8, 8, 4" if three_d else "16, 16, 1"
        return f"atx_tile_kernel<{t}, true, EpiSartAtx>"
    if precision == 4:  # FP64 image and its FP32 shadow
        return atx_kernel_name(variant, kahan, 0, three_d, general).replace(
            "EpiSartAtx,", "EpiSartAtx64,").replace("EpiSartAtx>", "EpiSartAtx64>")
    if not three_d and variant in ("direct", "direct2", "direct1"):
        # one-slice views: the walk compiled for fan rays
        c = {"direct": 4, "direct2": 2, "direct1": 1}[variant]
        return f"atx_kernel<EpiSartAtx, {k}, {c}, true>"
    if general and variant in ("sorted16", "sorted8x8x4", "sorted8x4x8"):
        t = {"sorted16": "16, 16, 2", "sorted8x8x4": "8, 8, 8", "sorted8x4x8": "8, 4, 16"}[variant]
        return f"atx_sorted_general_kernel<{t}, EpiSartAtx, {k}>"
    return {"direct": f"atx_kernel<EpiSartAtx, {k}, 4>",
            "tile16x16x1": f"atx_tile_plain_kernel<16, 16, 1, {k}, EpiSartAtx>",
            "tile8x8x4": f"atx_tile_plain_kernel<8, 8, 4, {k}, EpiSartAtx>",
            "tile16x4x4": f"atx_tile_plain_kernel<16, 4, 4, {k}, EpiSartAtx>",
            "pairs": f"atx_pairs_kernel<EpiSartAtx, {k}>",
            "direct2": f"atx_kernel<EpiSartAtx, {k}, 2>",
            "direct1": f"atx_kernel<EpiSartAtx, {k}, 1>",
            "sorted16": f"atx_sorted_kernel<16, 16, 1, EpiSartAtx, {k}, 2>",
            "sorted8x8x4": f"atx_sorted_kernel<8, 8, 4, EpiSartAtx, {k}, 2>",
            "sorted8x4x8": f"atx_sorted_kernel<8, 4, 8, EpiSartAtx, {k}, 2>"}.get(variant)


def ax_kernel_name(lpr, rows, precision):
    k = "true" if precision in (1, 2, 3) else "false"
    epi = f"EpiSartAx<{CODEC[precision]}>"
    if precision == 4:  # each ray summed in slab order, one wave per ray
        return f"ax_ordered_kernel<false, {epi} >"
    if rows > 1:
        return f"ax_rows_kernel<{lpr}, {rows}, {1 if rows == 3 else 2}, {k}, false, {epi} >"
    return f"ax_kernel<{lpr}, {k}, false, {epi} >"


def profiled(workload, kernel, what):
    """A per-launch figure of `kernel` from a committed rocprofv3 summary of
    this workload: what = "traffic" (HBM bytes per launch, FETCH_SIZE +
    WRITE_SIZE scaled by a plain copy) from profiles/<workload>_hbm_traffic.md,
    or "valu" (the VALU issue fraction and L2 hit rate) from
    profiles/<workload>_valu_l2.md; None when there is no row for it."""
    name = {"traffic": "hbm_traffic", "valu": "valu_l2"}[what]
    path = os.path.join(ROOT, "profiles", f"{workload}_{name}.md")
    if not kernel or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| `") and f"`{kernel}`" in line:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if what == "traffic":
                return int(cells[-1])
            return {"valu_issue_frac": float(cells[-2]), "l2_hit_rate": float(cells[-1])}
    return None


def stage_bytes(N, M, ng, sino=4):
    """Algorithmic (compulsory) HBM bytes per timed unit of each stage:
    forward projector launch reads m (4N) and P_v (`sino` bytes per value)
    and writes cor (4M); backprojector+update launch reads cor (4M) and
    reads/writes m (8N); a TV descent runs ng x (gradient: read u, write v =
    8N; step: read u, v, write u = 12N)."""
    return {0: 4 * N + (sino + 4) * M, 1: 8 * N + 4 * M, 2: ng * 20 * N}


def build_plan(name, precision):
    """(para, plan, img, P) of a workload; img and P are None when the plan
    simulated them on the device (its reference image and sinogram stay
    there)."""
    from asd_pocs_mex_cuda_demo_amd import ops
    from asd_pocs_mex_cuda_demo_amd import workloads as W
    if W.CONFIGS[name].get("sim") == "gpu":
        p, plan = W.sim_plan(name, precision)
        return p, plan, None, None
    p, img, P = W.problem(name, precision=precision)
    plan = ops.Plan(p, p["phi_vec"])
    plan.set_projections(P)
    return p, plan, img, P


def time_iterations(plan, img, steps, warmup):
    """`warmup` untimed iterations, then `steps` timed ones (the captured
    iteration graph replayed) from a fresh start; returns (seconds,
    iterations done, rmse history)."""
    plan.reset(None, img)
    plan.iterate(warmup)
    plan.reset(None, img)
    plan.sync()
    t0 = time.perf_counter()
    plan.iterate(steps)
    plan.sync()
    dt = time.perf_counter() - t0
    out, rmse = plan.out()
    return dt, out["IterationNum"], rmse


def cpu_threads():
    """Threads of the CPU baseline: the host share of this GPU where the
    environment states it (OMP_NUM_THREADS), else every CPU this process may
    run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    aff = len(os.sched_getaffinity(0))
    return min(int(env), aff) if env.isdigit() and int(env) > 0 else aff


def host_info():
    info = {"cpu_model": platform.processor() or None, "online_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["cpu_model"] = v
            elif k in ("Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.lower().replace("(s)", "s").replace(" ", "_")] = int(v)
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    return info


def cpu_sample_2d(oracle, p, P, iters, threads):
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    oracle.asd_pocs(P, dict(p, maxnumIter=iters))
    return time.perf_counter() - t0


def cpu_sample_3d(oracle, p, P, nviews, tv_steps, stride, threads):
    """Seconds of one outer iteration of the oracle on a 3-D workload,
    extrapolated from `nviews` SART view updates and `tv_steps` TV3D steps
    (the reference's FP32 projector arithmetic and FP64 TV/driver).  With
    stride > 1 every stride-th ray and voxel of the views and every
    stride-th slice of the TV step are evaluated and the time scaled by
    stride (a uniform subset of the same work)."""
    oracle.set_threads(threads)
    nv = len(p["phi_vec"])
    N = p["nx"] * p["ny"] * p["nz"]
    M = p["na"] * p["nb"]
    m = np.zeros(N, np.float32)
    rays = np.arange(0, M, stride, dtype=np.int64)
    vox = np.arange(0, N, stride, dtype=np.int64)
    t0 = time.perf_counter()
    for i in range(nviews):
        q = dict(p, phi=-p["phi_vec"][i] / 180 * np.pi)
        if stride == 1:
            proj, sumR = oracle.ax_para(m, q)
            sel = slice(None)
        else:
            proj, sumR = oracle.ax_rays_para(m, q, rays)
            sel = rays
        sumR = sumR.astype(np.float64)
        sumR[sumR == 0] = 1
        cor = np.zeros(M, np.float32)  # the sampled rays' corrections (all of them at stride 1)
        cor[sel] = (P[sel, i] - proj) / sumR
        if stride == 1:
            bp, sumC = oracle.atx_para(cor, q)
        else:
            bp, sumC = oracle.atx_voxels_para(cor, q, vox)
        sumC = sumC.astype(np.float64)
        sumC[sumC == 0] = 1
        if stride == 1:
            m = np.maximum(m + bp / sumC, 0).astype(np.float32)
    t_views = (time.perf_counter() - t0) * stride
    nz = max(p["nz"] // stride, 2)
    vol = np.full((p["nx"], p["ny"], nz), 1e-3) + np.random.default_rng(1).random(
        (p["nx"], p["ny"], nz)) * 1e-4
    t1 = time.perf_counter()
    oracle.tv3d(vol, 0.2, tv_steps, 1, np.zeros_like(vol))
    t_tv = (time.perf_counter() - t1) * (p["nz"] / nz) * 20 / tv_steps
    return t_views / nviews * nv + t_tv, t_views, t_tv


def cpu_baseline(name, p, P, threads):
    """The CPU oracle on this host: a bounded sample of the same workload on
    `threads` threads, and the same sample's single-thread time (on a
    uniform 1/16 subset where a full one would take minutes)."""
    from oracle import oracle
    nz = p["nz"]
    rec = {"unit": "iters/sec", "cores": threads, "kind": "port"}
    if nz == 1:
        iters = 5
        dt = cpu_sample_2d(oracle, p, P, iters, threads)
        dt1 = cpu_sample_2d(oracle, p, P, 1, 1)
        rec.update(value=iters / dt,
                   sample=f"{iters} full ASD-POCS iterations of the same workload (FP32 "
                          f"projectors in the reference's order, FP64 TV and driver; OpenMP over "
                          f"rays and voxels), {dt:.2f} s wall",
                   one_thread={"value": 1 / dt1, "sample": f"1 full iteration, {dt1:.2f} s"})
    else:
        big = p["nx"] * p["ny"] * nz > 64 ** 3 * 64
        stride = 16 if big else 1
        t_it, tv_, tt_ = cpu_sample_3d(oracle, p, P, 1, 1, stride, threads)
        rec.update(value=1 / t_it,
                   sample=f"1 SART view update ({'every 16th ray and voxel, x16: ' if big else ''}"
                          f"{tv_:.1f} s) and 1 TV3D step ({tt_:.1f} s scaled to 20), extrapolated "
                          f"to one iteration of {len(p['phi_vec'])} views ({t_it:.0f} s)")
        t1, tv1, tt1 = cpu_sample_3d(oracle, p, P, 1, 1, 16, 1)
        rec["one_thread"] = {"value": 1 / t1,
                             "sample": f"every 16th ray and voxel of 1 view update (x16: "
                                       f"{tv1:.1f} s) and 1/16 of a TV3D step (x16, x20: "
                                       f"{tt1:.1f} s), extrapolated ({t1:.0f} s per iteration)"}
    oracle.set_threads(threads)
    return rec


def headline(args):
    from asd_pocs_mex_cuda_demo_amd import workloads as W
    name = args.workload
    w = W.CONFIGS[name]
    precision = args.precision if args.precision is not None else w.get("precision", 0)
    p, plan, img, P = build_plan(name, precision)
    if args.atx_variant:
        plan.set_atx_variant(ATX_VARIANTS.index(args.atx_variant))
    if args.ax_shape:
        plan.set_ax_shape(*[int(v) for v in args.ax_shape.split(",")])
    N = p["nx"] * p["ny"] * p["nz"]
    M = p["na"] * p["nb"]
    dt, done, rmse = time_iterations(plan, img, args.steps, args.warmup)
    # the same iterations enqueued kernel by kernel, each projector launch
    # stamped with HIP events by its dispatch on the plan's stream
    prof = max(1, min(args.steps, args.prof_steps))
    plan.reset(None, img)
    plan.set_timing(True)
    plan.sync()
    t1 = time.perf_counter()
    plan.iterate(prof)
    plan.sync()
    dt_ev = time.perf_counter() - t1
    stage = {s: plan.timing(s) for s in (0, 1, 2)}
    plan.set_timing(False)
    nbytes = stage_bytes(N, M, p["numOfgradients"], {0: 4, 4: 8}.get(precision, 2))
    tuning = plan.tuning()
    lpr, rows, _ = plan.info()
    names = {0: ax_kernel_name(lpr, rows, precision),
             1: atx_kernel_name(tuning["atx_variant"], precision in (1, 2, 3), precision,
                                p["nz"] > 1, p["nz"] > 1 and max(p["na"], p["nb"]) < 2048),
             2: "TV descent (tv_grad_kernel x ng + tv_step_kernel)"}
    per_stage = {}
    for s, (ms, cnt) in stage.items():
        if cnt:
            avg = ms / cnt
            per_stage[names[s] or f"stage {s}"] = {
                "avg_us": avg * 1e3, "launches": cnt, "total_ms": ms,
                "algorithmic_bytes": nbytes[s], "GBps": nbytes[s] / (avg * 1e-3) / 1e9}
    dom = max(stage, key=lambda s: stage[s][0])
    ms_d, cnt_d = stage[dom]
    achieved = nbytes[dom] / (ms_d / cnt_d * 1e-3) / 1e9 if cnt_d else 0.0
    kernel = names[dom]
    roof = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled(name, kernel, "traffic"),
            "algorithmic_bytes_per_launch": nbytes[dom], "avg_launch_us": ms_d / cnt_d * 1e3}
    valu = profiled(name, kernel, "valu")
    if valu:
        roof.update(valu)
    result = {
        "metric": METRIC, "value": done / dt, "unit": "iters/sec", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / done * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {0: "f32", 1: "f32 (16-bit fixed-point sinogram, compensated sums)",
                  2: "f32 (fp16 sinogram, compensated sums)",
                  3: "f32 (16-bit fixed-point sinogram, compensated sums, f64 tile ray setup)",
                  4: "f64 state, f32 projectors (the reference's arithmetic)",
                  5: "f32 (fp16 sinogram)"}[precision],
        "data": ("synthetic, simulated on the device: modified Shepp-Logan phantom"
                 + (f" + {w['spheres']} seeded spheres" if w.get("spheres") else "")
                 + (f", Poisson noise I0={w['I0']:g} (seeded), log transform" if w["I0"]
                    else ", noiseless projections")) if P is None else
                ("synthetic: modified Shepp-Logan phantom"
                 + (f", Poisson noise I0={w['I0']:g} (seeded), log transform" if w["I0"]
                    else ", noiseless projections")),
        "config": {"workload": name, "nx": p["nx"], "ny": p["ny"], "nz": p["nz"],
                   "views": len(p["phi_vec"]), "na": p["na"], "nb": p["nb"],
                   "iterations": args.steps, "ng": p["numOfgradients"], "precision": precision},
        "roofline": roof,
        "stages": per_stage,
        "ms_per_step_eager_with_events": dt_ev / prof * 1e3,
        "rmse_final": float(rmse[-1]) if len(rmse) else None,
        "ax_lanes_per_ray": lpr, "ax_rows_per_bundle": rows, "atx_tuning": tuning,
    }
    if not args.no_cpu:
        if P is None:
            P = plan.projections()
        cb = cpu_baseline(name, p, P, cpu_threads())
        result["cpu_baseline"] = cb
        result["speedup_vs_cpu_baseline"] = result["value"] / cb["value"]
        result["host"] = host_info()
    plan.close()
    return result


def workload_rate(name, cpu, threads):
    """Iterations/s of a single-plan workload over its own iteration count
    (one warm-up iteration), with the CPU oracle's rate beside it."""
    from asd_pocs_mex_cuda_demo_amd import workloads as W
    p, plan, img, P = build_plan(name, W.CONFIGS[name].get("precision", 0))
    iters = W.CONFIGS[name]["iters"]
    dt, done, rmse = time_iterations(plan, img, iters, 1)
    r = {"iters_per_sec": done / dt, "ms_per_step": dt / done * 1e3, "iterations": done,
         "rmse_final": float(rmse[-1]) if len(rmse) else None}
    plan.close()
    if cpu:
        from oracle import oracle
        if p["nz"] == 1:
            t = cpu_sample_2d(oracle, p, P, 2, threads) / 2
            r["cpu_sample"] = "2 full iterations"
        else:
            t, _, _ = cpu_sample_3d(oracle, p, P, 2, 2, 1, threads)
            r["cpu_sample"] = "2 SART view updates + 2 TV3D steps, extrapolated"
        r["cpu_iters_per_sec"] = 1.0 / t
        r["speedup_vs_cpu"] = r["iters_per_sec"] * t
    return r


def batch_rates(steps=20):
    """The 64-case batch: batched-mixed (one launch per stage for all cases,
    16-bit sinogram, compensated sums) and batched-plain against the same
    cases as 64 sequential plain plans; case-iterations per second."""
    from asd_pocs_mex_cuda_demo_amd import ops
    from asd_pocs_mex_cuda_demo_amd import workloads as W
    cases = W.batch_cases()
    K = len(cases)
    t_plain = 0.0
    for p, img, P in cases:
        pl = ops.Plan(dict(p, maxnumIter=steps), p["phi_vec"])
        pl.set_projections(P)
        dt, _, _ = time_iterations(pl, img, steps, 1)
        t_plain += dt
        pl.close()
    res = {"cases": K, "iterations": steps,
           "sizes": sorted(set(int(p["nx"]) for p, _, _ in cases)),
           "views_min": int(min(len(p["phi_vec"]) for p, _, _ in cases)),
           "views_max": int(max(len(p["phi_vec"]) for p, _, _ in cases)),
           "plain_sequential_case_iters_per_sec": K * steps / t_plain}
    for prec, key in ((0, "batched_plain"), (1, "batched_mixed")):
        B = ops.Batch([dict(p, maxnumIter=steps, precision=prec) for p, _, _ in cases])
        for c, (p, img, P) in enumerate(cases):
            B.set_case(c, P, None, img)
        B.iterate(1)
        for c, (p, img, P) in enumerate(cases):
            B.set_case(c, P, None, img)
        B.sync()
        t0 = time.perf_counter()
        B.iterate(steps)
        B.sync()
        res[f"{key}_case_iters_per_sec"] = K * steps / (time.perf_counter() - t0)
        res[f"{key}_atx_tuning"] = {"pairs_per_load_batch": B.tuning()["atx_chunk"],
                                    "us": {str(k): v for k, v in B.tuning()["atx_us"].items()}}
        B.close()
    res["speedup_batched_mixed_vs_sequential_plain"] = (
        res["batched_mixed_case_iters_per_sec"] / res["plain_sequential_case_iters_per_sec"])
    res["speedup_batched_mixed_vs_batched_plain"] = (
        res["batched_mixed_case_iters_per_sec"] / res["batched_plain_case_iters_per_sec"])
    return res


def accuracy_fan2d_512(threads):
    """The 512^2 fan-beam reconstruction (50 iterations) against the CPU
    oracle: the reference-arithmetic plan (precision 4) and the plain FP32
    plan, beside how far the oracle moves itself when only the order of its
    FP64 sums changes, or its data by 1e-9 relative."""
    from asd_pocs_mex_cuda_demo_amd import ops
    from asd_pocs_mex_cuda_demo_amd import workloads as W
    from oracle import oracle
    oracle.set_threads(threads)
    p, img, P = W.problem("fan2d_512")
    q = dict(p, I0=img, precision=0)
    mo, oo = oracle.asd_pocs(P, q)

    def gap(o):
        return np.abs(np.asarray(o["RMSE"]) - oo["RMSE"]) / oo["RMSE"]

    def rel(m):
        return float(np.linalg.norm(np.ravel(m, order="F") - mo) / np.linalg.norm(mo))
    oracle.set_sum_order(1)
    try:
        _, o_pw = oracle.asd_pocs(P, q)
    finally:
        oracle.set_sum_order(0)
    Pp = P * (1 + 1e-9 * np.random.default_rng(1).standard_normal(P.shape))
    m9, o9 = oracle.asd_pocs(Pp, q)
    m4, o4 = ops.ASD_POCS_cuda(P, dict(q, precision=4))
    m0, o0 = ops.ASD_POCS_cuda(P, q)
    g4, g0, gpw, g9 = gap(o4), gap(o0), gap(o_pw), gap(o9)
    return {
        "workload": "fan2d_512", "iterations": len(oo["RMSE"]),
        "oracle": "FP32 projectors in the reference's expression and summation order, FP64 "
                  "TV and driver with serial sums",
        "reference_arithmetic_plan": {
            "image_bit_identical": bool(np.array_equal(np.ravel(m4, order="F"), mo)),
            "rmse_rel_diff_max": float(g4.max()), "alpha_equal": o4["alpha"] == oo["alpha"]},
        "plain_fp32_plan": {"rmse_rel_diff_first_iteration": float(g0[0]),
                            "rmse_rel_diff_max": float(g0.max()),
                            "rmse_rel_diff_final": float(g0[-1]), "image_rel_l2": rel(m0)},
        "oracle_vs_itself": {
            "pairwise_fp64_sums": {"rmse_rel_diff_max": float(gpw.max()),
                                   "rmse_rel_diff_first_three": [float(x) for x in gpw[:3]]},
            "data_perturbed_1e-9": {"rmse_rel_diff_max": float(g9.max()),
                                    "image_rel_l2": rel(m9)}},
    }


def log(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cone3d_512", choices=WORKLOAD_CHOICES)
    ap.add_argument("--precision", type=int, default=None,
                    help="sinogram storage / arithmetic variant (default: the workload's)")
    ap.add_argument("--prof-steps", type=int, default=2,
                    help="iterations of the per-kernel event timing pass")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline workload (for profiling runs)")
    ap.add_argument("--ax-shape", default=None,
                    help="force the forward projector's shape as LANES,ROWS (for example 8,3)")
    ap.add_argument("--atx-variant", default=None, choices=ATX_VARIANTS,
                    help="force the backprojector variant instead of the autotuner's pick")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device to device three times (a plain copy "
                         "kernel of known size, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is described, not built).", file=sys.stderr)
        sys.exit(2)
    from asd_pocs_mex_cuda_demo_amd import _native
    _native.lib()  # fail loudly if the HIP library is missing
    if args.copy_calibration:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "batch64":
        r = batch_rates(args.steps)
        result = {"metric": METRIC, "value": r["batched_mixed_case_iters_per_sec"],
                  "unit": "case-iters/sec", "n_gpus": 1, "steps": args.steps, "warmup": 1,
                  "ms_per_step": r["cases"] / r["batched_mixed_case_iters_per_sec"] * 1e3,
                  "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "dtype": "f32 (16-bit fixed-point sinogram, compensated sums)",
                  "data": "synthetic: 64 perturbed modified Shepp-Logan phantoms, Poisson "
                          "I0=1e5 (seeded)",
                  "config": {"workload": "batch64", **{k: r[k] for k in
                                                       ("cases", "iterations", "sizes",
                                                        "views_min", "views_max")}},
                  "batch": r}
        print(json.dumps(result), flush=True)
        return
    log(f"{args.workload}: {args.warmup} warm-up and {args.steps} timed iterations")
    result = headline(args)
    log(f"{args.workload}: {result['value']:.4g} {result['unit']}")
    if not args.no_extras:
        threads = cpu_threads()
        extras = {}
        for name in ("fan2d_256", "fan2d_512", "cone3d_256"):
            if name != args.workload:
                extras[name] = workload_rate(name, not args.no_cpu, threads)
                log(f"{name}: {extras[name]['iters_per_sec']:.4g} iterations/s")
        result["workloads"] = extras
        result["batch64"] = batch_rates()
        log("batch64 done")
        if not args.no_cpu:
            result["accuracy_vs_oracle"] = accuracy_fan2d_512(threads)
            log("accuracy against the oracle done")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
