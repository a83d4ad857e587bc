"""GPU parity: every operator through the C ABI against the CPU oracle.

Tolerances: the projectors evaluate the reference's FP32 expressions in the
same order with contraction off, so the backprojector and the 3-D thread-per-ray
forward projector match the oracle to rounding of the summation order only
(relative L2 <= 1e-5 is the bar; observed is far below).  TV runs in FP64
(relative L2 <= 1e-12).  The ASD-POCS driver keeps the image in FP32 where the
oracle's driver keeps FP64, so reconstructions agree to <= 1e-4 relative.
"""
import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import ops, phantoms

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def geom(nx, ny, nz, na, nb, SO, OD, vxl, det, phi):
    return dict(nx=nx, ny=ny, nz=nz, na=na, nb=nb, SO=SO, OD=OD, vxlsize=vxl, detsize=det,
                phi=phi)


CASES = [
    # 2-D fan beam (nz = nb = 1), the demo-style geometry and the 512^2 one
    ("fan64", geom(64, 64, 1, 128, 1, 405.941, 500.31, 0.5, 0.148 * 2, -0.3)),
    ("fan512", geom(512, 512, 1, 768, 1, 541.0, 408.0, 0.5, 0.8, -1.9)),
    ("fan_odd_na", geom(48, 40, 1, 97, 1, 500.0, 500.0, 1.0, 1.0, -2.6)),
    # 2-D views whose rays span 128 lanes (two rays per block, two waves
    # each) and 256 lanes (four waves per ray, like the 512^2 view above)
    ("fan300", geom(300, 280, 1, 400, 1, 500.0, 450.0, 0.7, 1.0, -1.1)),
    ("fan1040", geom(1040, 200, 1, 64, 1, 800.0, 500.0, 0.5, 1.0, -0.2)),
    # 3-D cone beam
    ("cone32", geom(32, 32, 16, 48, 24, 500.0, 500.0, 1.0, 1.2, -0.7)),
    ("cone_odd", geom(24, 20, 9, 37, 15, 300.0, 400.0, 1.0, 1.5, -4.0)),
]


@pytest.mark.parametrize("name,para", CASES, ids=[c[0] for c in CASES])
def test_ax_matches_oracle(orc, name, para):
    rng = np.random.default_rng(1)
    N = para["nx"] * para["ny"] * para["nz"]
    X = rng.random(N).astype(np.float32)
    y, sr = ops.Ax_cone_gpu(X, para)
    yo, sro = orc.ax_para(X, para)
    assert rel_l2(y, yo) <= 1e-5, rel_l2(y, yo)
    assert rel_l2(sr, sro) <= 1e-5, rel_l2(sr, sro)
    assert np.abs(y - yo).max() <= 1e-5 * np.abs(yo).max() + 1e-30


@pytest.mark.parametrize("phi", [-0.3, -1.2, -2.9, -4.4])
@pytest.mark.parametrize("nz,nb", [(3, 800), (1, 1)])
def test_forward_projector_slab_arithmetic_is_bit_identical(orc, nz, nb, phi):
    """A 3-voxel-wide volume gets one lane per ray, so each ray is summed in
    slab order like the oracle: every slab's length factor (the per-ray
    division by tmp, finished from its reciprocal), its square root and its
    cell-crossing fractions (the division without scaling) must give the
    oracle's IEEE results bit for bit, over many rays and crossings."""
    para = geom(3, 3, nz, 1000, nb, 9.0, 7.0, 1.0, 0.01, phi)
    rng = np.random.default_rng(11)
    X = rng.random(3 * 3 * nz).astype(np.float32)
    y, sr = ops.Ax_cone_gpu(X, para)
    yo, sro = orc.ax_para(X, para)
    assert np.array_equal(sr, sro), np.abs(sr - sro).max()
    assert np.array_equal(y, yo), np.abs(y - yo).max()


def test_row_bundles_match_single_rays_and_oracle(orc):
    """A 512 x 512 detector gets the row-bundle forward projector (8 lanes on
    a bundle of three detector rows of one column, each slab's transverse
    half shared).  Each ray's slab products and their lane order are the
    single-ray kernel's, so the result equals the device functions' host run
    at 8 lanes per ray bit for bit, and the oracle to summation rounding; a
    SART sweep through a plan of that shape matches the oracle's sweep."""
    from tests.test_device_geometry_cpu import emul, P as ptr
    nx, ny, nz, na, nb = 32, 32, 24, 512, 512
    base = geom(nx, ny, nz, na, nb, 500.0, 500.0, 1.0, 0.15, 0.0)
    pl = ops.Plan(base, [0.0, 90.0])
    assert pl.info()[:2] == (8, 3)
    pl.close()
    # a view of at least 2^19 rays takes bundles of two rows (the same sums)
    pl = ops.Plan(dict(base, na=1024, nb=768), [0.0])
    assert pl.info()[:2] == (8, 2)
    pl.close()
    rng = np.random.default_rng(5)
    X = rng.random(nx * ny * nz).astype(np.float32)
    for phi in (-0.4, -2.0):
        para = dict(base, phi=phi)
        y, sr = ops.Ax_cone_gpu(X, para)
        g = orc.gateway_geometry(phi, 500.0, 500.0, 1.0, 0.15, na, nb)
        ye = np.zeros(na * nb, np.float32)
        sre = np.zeros_like(ye)
        emul().emul_ax(ptr(X), ptr(ye), ptr(sre), g["cos_phi"], g["sin_phi"], ptr(g["y_det"]),
                       ptr(g["z_det"]), g["SO"], g["OD"], g["scale"], np.float32(1), nx, ny, nz,
                       na, nb, 8, 0, 1)
        assert np.array_equal(y, ye) and np.array_equal(sr, sre), phi
        yo, sro = orc.ax_para(X, para)
        assert rel_l2(y, yo) <= 1e-5 and rel_l2(sr, sro) <= 1e-5, phi
    phi_vec = np.array([10.0, 100.0, 190.0, 280.0])
    img = phantoms.shepp_logan_3d(nx, nz=nz).astype(np.float32)
    P = np.stack([ops.Ax_cone_gpu(img, dict(base, phi=-t / 180 * np.pi))[0] for t in phi_vec],
                 axis=1).astype(np.float64)
    para = dict(base, phi_vec=phi_vec)
    m, out = ops.SART_cuda_inline(P, para)
    mo, outo = orc.sart(P, para)
    assert rel_l2(m, mo) <= 1e-5, rel_l2(m, mo)
    assert abs(out["dataTolerance"] - outo["dataTolerance"]) <= 1e-5 * outo["dataTolerance"]


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_forward_shapes_give_identical_iterations(precision):
    """At 8 lanes, single rays and bundles of 2 or 3 detector rows sum every
    ray's slab products in the same order, so two ASD-POCS iterations are
    bit-identical whichever shape the plan is forced to (plain sums, and the
    compensated sums of precision 1 and 2 with their 16-bit sinograms).  Other lane counts change only the
    last bits; shapes outside the compiled set are rejected."""
    para, img, P = _small_problem(n=32, nv=6, na=48, nz=16, nb=40)
    para.update(maxnumIter=2, alpha=0.2, numOfgradients=5, precision=precision)
    res = []
    for shape in [(8, 1), (8, 2), (8, 3), (16, 1)]:
        pl = ops.Plan(para, para["phi_vec"])
        pl.set_projections(P)
        pl.reset(None, img)
        pl.set_atx_variant(0)
        pl.set_ax_shape(*shape)
        assert pl.info()[:2] == shape
        pl.iterate(2)
        pl.sync()
        res.append(pl.image())
        if shape == (16, 1):
            for bad in [(8, 4), (16, 3), (12, 1), (512, 1), (0, 1)]:
                with pytest.raises(Exception):
                    pl.set_ax_shape(*bad)
            pl.set_ax_shape(0, 0)
            assert pl.info()[:2] == (16, 1)
        pl.close()
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
    assert rel_l2(res[3], res[0]) <= 1e-4, rel_l2(res[3], res[0])


def test_wide_addressing_matches_oracle(orc):
    """Volumes of 2^30 voxels or more, or rows of 2^22 voxels or more, are read
    through 64-bit element addresses instead of buffer byte offsets.  Rows of
    4194306 voxels force that form; rays along y walk 16 short slabs near
    the middle of the rows, so every sum has few terms and a piece read from
    a wrong voxel shows."""
    nx, ny = 4194306, 16
    para = geom(nx, ny, 1, 64, 1, 1000.0, 500.0, 1.0, 1.0, -np.pi / 2)
    X = np.random.default_rng(6).random(nx * ny).astype(np.float32)
    y, sr = ops.Ax_cone_gpu(X, para)
    yo, sro = orc.ax_para(X, para)
    assert np.all(sro > 0)
    assert np.abs(y - yo).max() <= 1e-5 * np.abs(yo).max(), np.abs(y - yo).max()
    assert np.abs(sr - sro).max() <= 1e-5 * np.abs(sro).max(), np.abs(sr - sro).max()


def test_wide_addressing_row_bundles(orc):
    """The row-bundle forward projector in the wide form: rows of 4194306
    voxels and a 512 x 256 detector (8 lanes, bundles of three rows).  Its
    sums equal the device functions' host run at 8 lanes bit for bit, and
    the oracle to summation rounding."""
    from tests.test_device_geometry_cpu import emul, P as ptr
    nx, ny, nz, na, nb = 4194306, 16, 2, 512, 256
    para = geom(nx, ny, nz, na, nb, 1000.0, 500.0, 1.0, 0.01, -np.pi / 2)
    X = np.random.default_rng(8).random(nx * ny * nz).astype(np.float32)
    y, sr = ops.Ax_cone_gpu(X, para)
    g = orc.gateway_geometry(-np.pi / 2, 1000.0, 500.0, 1.0, 0.01, na, nb)
    ye = np.zeros(na * nb, np.float32)
    sre = np.zeros_like(ye)
    emul().emul_ax(ptr(X), ptr(ye), ptr(sre), g["cos_phi"], g["sin_phi"], ptr(g["y_det"]),
                   ptr(g["z_det"]), g["SO"], g["OD"], g["scale"], np.float32(1), nx, ny, nz, na,
                   nb, 8, 0, 1)
    assert np.array_equal(y, ye) and np.array_equal(sr, sre)
    yo, sro = orc.ax_para(X, para)
    assert np.count_nonzero(sro) > na * nb // 2
    assert rel_l2(y, yo) <= 1e-5 and rel_l2(sr, sro) <= 1e-5


@pytest.mark.parametrize("name,para", CASES, ids=[c[0] for c in CASES])
def test_atx_matches_oracle(orc, name, para):
    rng = np.random.default_rng(2)
    M = para["na"] * para["nb"]
    Y = rng.standard_normal(M).astype(np.float32)
    x, sc = ops.Atx_cone_gpu(Y, para)
    xo, sco = orc.atx_para(Y, para)
    # same expressions in the same order: bit-identical
    assert np.array_equal(sc, sco), np.abs(sc - sco).max()
    assert np.array_equal(x, xo), np.abs(x - xo).max()


def test_tv2d_matches_oracle(orc):
    rng = np.random.default_rng(3)
    I = rng.random((37, 29))
    I0 = rng.random((37, 29)) * 0.1
    for nit, pos in [(0, 1), (1, 0), (5, 1)]:
        out = ops.imgTVGradMin(I, 0.2, nit, pos, I0)
        ref = orc.tv2d(I, 0.2, nit, pos, I0)
        assert rel_l2(out, ref) <= 1e-12
    out = ops.imgTVGradMin(I, 0.05, 3, 1)
    assert rel_l2(out, orc.tv2d(I, 0.05, 3, 1)) <= 1e-12


def test_tv3d_matches_oracle(orc):
    rng = np.random.default_rng(4)
    I = rng.random((17, 13, 11))
    out = ops.imgTVGradMin3D(I, 0.01, 20, 1)
    ref = orc.tv3d(I, 0.01, 20, 1)
    assert rel_l2(out, ref) <= 1e-12
    I0 = rng.random((17, 13, 11))
    out = ops.imgTVGradMin3D(I, 0.3, 4, 0, I0)
    assert rel_l2(out, orc.tv3d(I, 0.3, 4, 0, I0)) <= 1e-12


def test_tv_drop_ins_are_bit_identical(orc):
    """The drop-in descents sum d^2 and every ||v||^2 in index order and step
    as u - ((alpha*d)*v)/||v|| (imgTVGradMin.cpp:58-59, :96, :104-107), so
    they reproduce the oracle exactly, odd sizes and a stopped descent
    (nTVIter 0) included."""
    rng = np.random.default_rng(11)
    for shape, fn, ref in [((37, 29), ops.imgTVGradMin, orc.tv2d),
                           ((17, 13, 11), ops.imgTVGradMin3D, orc.tv3d)]:
        I = rng.random(shape)
        I0 = rng.random(shape) * 0.1
        for alpha, nit, pos in [(0.2, 20, 1), (0.05, 3, 0), (0.1, 0, 1)]:
            out = fn(I, alpha, nit, pos, I0)
            exp = ref(I, alpha, nit, pos, I0)
            assert np.array_equal(out, exp), (shape, alpha, nit, np.abs(out - exp).max())


def _small_problem(n=64, nv=24, na=96, nz=1, nb=1):
    para = dict(nx=n, ny=n, nz=nz, SO=500.0, OD=400.0, vxlsize=1.0, detsize=1.2, na=na, nb=nb)
    para["phi_vec"] = np.arange(1, nv + 1) * 360.0 / nv
    img = phantoms.shepp_logan_2d(n) if nz == 1 else phantoms.shepp_logan_3d(n, nz=nz)
    P = np.zeros((na * nb, nv))
    for i, th in enumerate(para["phi_vec"]):
        q = dict(para, phi=-th / 180 * np.pi)
        P[:, i] = ops.Ax_cone_gpu(img.astype(np.float32), q)[0]
    return para, img, P


def test_asd_pocs_matches_oracle_driver(orc):
    para, img, P = _small_problem()
    para.update(maxnumIter=4, alpha=0.2, numOfgradients=10, I0=img)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert out["IterationNum"] == outo["IterationNum"] == 4
    # the reconstruction bar: RMSE within 1e-4 relative; the FP32 image (the
    # oracle keeps MATLAB's FP64 image) moves single pixels by ~1e-4
    assert np.allclose(out["RMSE"], outo["RMSE"], rtol=1e-4, atol=0)
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-3, rel_l2(m.ravel(order="F"), mo)
    assert abs(out["alpha"] - outo["alpha"]) <= 1e-12 * abs(outo["alpha"])
    assert abs(out["beta"] - outo["beta"]) <= 1e-12
    assert abs(out["dataTolerance"] - outo["dataTolerance"]) <= 1e-3 * outo["dataTolerance"]
    assert m.min() >= 0


def test_sart_matches_oracle(orc):
    para, img, P = _small_problem(n=48, nv=18, na=80)
    m, out = ops.SART_cuda_inline(P, para)
    mo, outo = orc.sart(P, para)
    assert rel_l2(m, mo) <= 1e-5, rel_l2(m, mo)
    assert abs(out["dataTolerance"] - outo["dataTolerance"]) <= 1e-5 * outo["dataTolerance"]


def test_asd_pocs_3d_matches_oracle(orc):
    para, img, P = _small_problem(n=24, nv=12, na=40, nz=8, nb=20)
    para.update(maxnumIter=2, alpha=0.2, numOfgradients=5)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-4, rel_l2(m.ravel(order="F"), mo)


def test_plan_is_deterministic():
    para, img, P = _small_problem(n=48, nv=12, na=72)
    para.update(maxnumIter=3)
    a, _ = ops.ASD_POCS_cuda(P, para)
    b, _ = ops.ASD_POCS_cuda(P, para)
    assert np.array_equal(a, b)


def test_graph_replays_match_eager_iterations():
    """One outer iteration captured as a hipGraph and replayed, with eager
    kernels and host synchronisation between replays, gives bit-identical
    images, alpha, beta and RMSE to the same iterations enqueued eagerly."""
    para, img, P = _small_problem(n=48, nv=12, na=72)
    para.update(maxnumIter=6, alpha=0.2, numOfgradients=8)
    eager = ops.Plan(para, para["phi_vec"])
    graph = ops.Plan(para, para["phi_vec"])
    eager.set_graph(False)
    graph.set_graph(True)
    for pl in (eager, graph):
        pl.set_projections(P)
        pl.reset(None, img)
    for it in range(5):
        eager.iterate(1)
        graph.iterate(1)
        # eager work on the device between replays
        ops.Ax_cone_gpu(img.astype(np.float32), dict(para, phi=-0.3))
        a, b = eager.image(), graph.image()
        assert np.array_equal(a, b), it
        assert np.array_equal(eager.current(), graph.current()), it
        oa, ra = eager.out()
        ob, rb = graph.out()
        assert oa == ob and np.array_equal(ra, rb)
    # several replays in one call
    eager.iterate(3)
    graph.iterate(3)
    assert np.array_equal(eager.image(), graph.image())
    eager.close()
    graph.close()


@pytest.mark.parametrize("three_d", [False, True], ids=["fan", "cone"])
def test_backprojector_variants_are_bit_identical(three_d):
    """The direct and LDS-tiled backprojectors (the variants the plan's
    autotuner chooses between) give bit-identical iterations, and the
    autotuner reports a measured time for every variant it tried."""
    if three_d:
        para, img, P = _small_problem(n=24, nv=8, na=40, nz=8, nb=20)
        variants = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9]
        tuned = variants
    else:
        para, img, P = _small_problem(n=48, nv=12, na=72)
        variants = [0, 1, 4, 5, 6, 7]
        tuned = [0, 1, 4, 5, 6]  # the sorted tiles are tried on 3-D views only
    para.update(maxnumIter=3, alpha=0.2, numOfgradients=5)
    ref = None
    for v in [-1] + variants:
        pl = ops.Plan(para, para["phi_vec"])
        pl.set_projections(P)
        pl.reset(None, img)
        if v >= 0:
            pl.set_atx_variant(v)
        pl.iterate(3)
        t = pl.tuning()
        img_v, (o, r) = pl.image(), pl.out()
        pl.close()
        if v < 0:
            assert t["atx_variant"] is not None and len(t["atx_us"]) == len(tuned)
            assert all(us > 0 for us in t["atx_us"].values())
        if ref is None:
            ref = (img_v, o, r)
        else:
            assert np.array_equal(img_v, ref[0]), v
            assert o == ref[1] and np.array_equal(r, ref[2]), v


def test_stop_rule_ends_the_loop():
    """stopRelChg large: the first iteration already stops; m_res is the
    post-SART image and IterationNum = 1 (ASD_POCS_cuda.m:146-149)."""
    para, img, P = _small_problem(n=32, nv=8, na=48)
    para.update(maxnumIter=5, stopRelChg=1e9)
    m, out = ops.ASD_POCS_cuda(P, para)
    sweep, _ = ops.SART_cuda_inline(P, dict(para, **{"lambda": 1.0}))
    assert out["IterationNum"] == 1 and out["alpha"] == 0.02 and out["beta"] == 1.0
    np.testing.assert_allclose(m.ravel(order="F"), sweep, rtol=0, atol=0)


def _decoded(P, prec):
    """The sinogram a plan of precision `prec` holds, decoded to FP64."""
    from asd_pocs_mex_cuda_demo_amd import workloads as W
    if prec in (1, 3):
        return W.quantize_u16(P)
    if prec in (2, 5):
        return P.astype(np.float32).astype(np.float16).astype(np.float64)
    return P


def test_precision_variants_against_oracle(orc):
    """Every storage / arithmetic variant against the oracle run on the data
    the variant holds (16-bit fixed point or FP16 decoded), 5 iterations of
    ng = 5 on noisy 64^2 data: the first iteration (no TV step yet) within
    1e-4, every iteration within max(1e-4, 2 * floor), where floor is how far
    the oracle moves itself when only the summation order of its FP64 norms
    changes (the TV descent amplifies any rounding difference; see the
    per-workload tests).  r_max is raised so the alpha rule, a discrete
    decision, does not fire.  Precision 3 changes the backprojector's lengths
    on purpose and is held to 1e-3 after the first iteration."""
    para, img, P = _small_problem(n=64, nv=24, na=96)
    Pn = phantoms.poisson_log(P, 1e5, seed=7)
    para.update(maxnumIter=5, alpha=0.2, numOfgradients=5, I0=img, r_max=1e9)
    _, oo = orc.asd_pocs(Pn, para)
    orc.set_sum_order(1)
    try:
        _, op = orc.asd_pocs(Pn, para)
    finally:
        orc.set_sum_order(0)
    floor = float(np.max(np.abs(op["RMSE"] - oo["RMSE"]) / oo["RMSE"]))
    bar = max(1e-4, 2 * floor)
    for prec in (0, 1, 2, 3, 5):
        _, ov = orc.asd_pocs(_decoded(Pn, prec), para)
        _, o = ops.ASD_POCS_cuda(Pn, dict(para, precision=prec))
        g = np.abs(o["RMSE"] - ov["RMSE"]) / ov["RMSE"]
        assert o["alpha"] == 0.2 and ov["alpha"] == 0.2
        assert g[0] <= 1e-4, (prec, g)
        # precision 3 departs from the reference's lengths by design (FP64 ray
        # setup: ~1e-6 voxel of exact, where the reference's FP32 lengths carry
        # ~1e-4 of cancellation); the TV descent amplifies that systematic
        # difference (5.3e-4 after 5 iterations here), so it is held to 1e-3
        assert g.max() <= (1e-3 if prec == 3 else bar), (prec, g, bar)


def test_precision_variants_against_exact_geometry(orc):
    """One SART view update of each variant against the same update with the
    exact FP64 system matrix (oracle/exact64: every ray-pixel length clipped
    in FP64, the rays' end points as the operators form them in FP32) on a
    512^2-like fan: the reference's FP32 lengths (precision 0, 4, and the
    16-bit / FP16 storage variants) carry ~1e-4 voxel of cancellation, the
    FP64 tile setup of precision 3 removes most of it."""
    from oracle import exact64 as E
    n, na = 64, 96
    para = dict(nx=n, ny=n, nz=1, SO=541.0, OD=408.0, vxlsize=0.5, detsize=0.8, na=na, nb=1)
    th = 37.0
    q = dict(para, phi=-th / 180 * np.pi)
    g = orc.gateway_geometry(q["phi"], para["SO"], para["OD"], 0.5, 0.8, na, 1)
    A = E.system_matrix_2d(g["cos_phi"], g["sin_phi"], float(g["SO"]), float(g["OD"]),
                           g["y_det"], n, n, f32_points=True)
    img = phantoms.shepp_logan_2d(n).ravel(order="F")
    m0 = (0.9 * img).astype(np.float32).astype(np.float64)  # held exactly by every plan
    Pv = 0.5 * (A @ img)  # exact line integrals (vxl = 0.5 mm)
    errs = {}
    for prec in (0, 1, 2, 3, 4, 5):
        Pd = _decoded(Pv[:, None], prec)[:, 0]
        proj, sr = 0.5 * (A @ m0), 0.5 * A.sum(axis=1)
        cor = (Pd - proj) / np.where(sr == 0, 1, sr)
        bp, sc = 0.5 * (A.T @ cor), 0.5 * A.sum(axis=0)
        upd = bp / np.where(sc == 0, 1, sc)
        pl = ops.Plan(dict(para, precision=prec), [th])
        pl.set_projections(Pv[:, None])
        pl.reset(m0, None)
        pl.sweep(1.0)
        got = pl.current() - m0
        pl.close()
        errs[prec] = rel_l2(got, upd)
    print("one-view update, rel L2 vs exact geometry:", errs)
    assert all(e <= 1e-3 for e in errs.values()), errs
    assert errs[3] < errs[0], errs


def test_batch_matches_single_plans():
    """Cases of different sizes and view counts advanced together give the
    per-case results of separate plans (to FP64 reduction-order rounding)."""
    rng = np.random.default_rng(11)
    paras, Ps, imgs = [], [], []
    for n, nv in [(32, 10), (48, 7), (40, 13), (32, 4)]:
        para = dict(nx=n, ny=n, nz=1, SO=500.0, OD=400.0, vxlsize=1.0, detsize=1.3, na=int(1.5 * n),
                    nb=1, alpha=0.2, numOfgradients=int(rng.integers(3, 8)), maxnumIter=3)
        para["phi_vec"] = np.sort(rng.uniform(0, 360, nv))
        img = phantoms.shepp_logan_2d(n)
        P = np.zeros((para["na"], nv))
        for i, th in enumerate(para["phi_vec"]):
            P[:, i] = ops.Ax_cone_gpu(img.astype(np.float32), dict(para, phi=-th / 180 * np.pi))[0]
        paras.append(para)
        Ps.append(P)
        imgs.append(img)
    B = ops.Batch(paras)
    for c in range(len(paras)):
        B.set_case(c, Ps[c], None, imgs[c])
    B.iterate(3)
    for c, para in enumerate(paras):
        m, out = ops.ASD_POCS_cuda(Ps[c], dict(para, I0=imgs[c]))
        ob, rb = B.out(c)
        assert ob["IterationNum"] == 3
        assert rel_l2(B.image(c), m.ravel(order="F")) <= 1e-5, c
        np.testing.assert_allclose(rb, out["RMSE"], rtol=1e-5)
        assert ob["beta"] == pytest.approx(out["beta"], rel=1e-12)
    tune = B.tuning()
    assert tune["atx_chunk"] in (1, 2, 4) and set(tune["atx_us"]) == {1, 2, 4}, tune
    ref = [B.image(c) for c in range(len(paras))]
    B.close()
    # every pairs-per-load-batch choice of the backprojector stage gives the
    # same bits
    for chunk in (1, 2, 4):
        Bc = ops.Batch(paras)
        Bc.set_atx_chunk(chunk)
        for c in range(len(paras)):
            Bc.set_case(c, Ps[c], None, imgs[c])
        Bc.iterate(3)
        assert Bc.tuning()["atx_chunk"] == chunk
        for c in range(len(paras)):
            assert np.array_equal(Bc.image(c), ref[c]), (chunk, c)
        Bc.close()


def test_stage_timing_counts_every_launch_and_changes_nothing():
    """Per-stage timing (kernels launched with dispatch-stamped events) records
    one forward and one backprojector launch per view and one TV descent per
    iteration, each with a positive duration, and leaves the iterations
    bit-identical to an untimed plan."""
    para, img, P = _small_problem(n=48, nv=12, na=72)
    para.update(maxnumIter=2, alpha=0.2, numOfgradients=5)
    res = []
    for timed in (False, True):
        pl = ops.Plan(para, para["phi_vec"])
        pl.set_projections(P)
        pl.reset(None, img)
        pl.set_atx_variant(5)  # same kernels in both runs (no autotuning launches)
        if timed:
            pl.set_timing(True)
        pl.iterate(2)
        pl.sync()
        if timed:
            for stage, expect in ((0, 2 * 12), (1, 2 * 12), (2, 2)):
                ms, cnt = pl.timing(stage)
                assert cnt == expect, (stage, cnt)
                assert ms > 0.0, (stage, ms)
        res.append((pl.image(), pl.out()))
        pl.close()
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1][0] == res[1][1][0] and np.array_equal(res[0][1][1], res[1][1][1])
