"""The projector device functions (geom_device.hpp), compiled for the host,
against the oracle: the same code the GPU kernels run, checked without a GPU.
Backprojector and the thread-per-ray forward projector must be bit-identical
to the oracle (same FP32 expressions, same order, no contraction)."""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def emul():
    global _lib
    if _lib is None:
        from tests.cpu_emul import build
        path = os.path.join(HERE, "cpu_emul", "libemul.so")
        if build.stale(path):
            build.build()
        _lib = C.CDLL(path)
        f, i, vp = C.c_float, C.c_int, C.c_void_p
        _lib.emul_ax.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, i, i, i, i, i, i, i, i]
        _lib.emul_atx.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, f, f, i, i, i, i, i]
    return _lib


def P(a):
    return a.ctypes.data_as(C.c_void_p)


CASES = [
    ("fan", 40, 36, 1, 64, 1, 500.0, 400.0, 1.3),
    ("fan_odd", 33, 30, 1, 47, 1, 300.0, 300.0, 1.0),
    ("cone", 20, 18, 10, 32, 16, 300.0, 300.0, 1.5),
    ("cone_steep", 12, 10, 6, 20, 8, 25.0, 25.0, 1.5),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("phi", [-0.3, -1.9, -3.7, -np.pi / 2])
def test_device_projectors_match_oracle(orc, case, phi):
    _, nx, ny, nz, na, nb, SO, OD, det = case
    g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
    rng = np.random.default_rng(3)
    X = rng.random(nx * ny * nz).astype(np.float32)
    Y = rng.standard_normal(na * nb).astype(np.float32)
    args = (g["cos_phi"], g["sin_phi"], P(g["y_det"]), P(g["z_det"]), g["SO"], g["OD"], g["scale"])
    yo, sro = orc.ax(X, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
                     g["scale"], np.float32(1), nx, ny, nz, na, nb)
    for lpr, kahan in [(1, 0), (8, 0), (64, 0), (128, 0), (256, 0), (64, 1), (256, 1)]:
        y = np.zeros(na * nb, np.float32)
        sr = np.zeros(na * nb, np.float32)
        emul().emul_ax(P(X), P(y), P(sr), *args, np.float32(1), nx, ny, nz, na, nb, lpr, kahan, 1)
        if lpr == 1 and not kahan:
            assert np.array_equal(y, yo) and np.array_equal(sr, sro)
        else:
            assert np.abs(y - yo).max() <= 2e-6 * np.abs(yo).max()
            assert np.abs(sr - sro).max() <= 2e-6 * np.abs(sro).max()
        # bundles of detector rows sharing each slab's transverse half sum
        # the same products in the same order as single rays, and the wide
        # addressing (element indices) reads the same voxels as byte offsets
        for rows, wide in ((2, 0), (3, 0), (1, 1), (2, 1)):
            y2 = np.zeros_like(y)
            sr2 = np.zeros_like(sr)
            emul().emul_set_wide(wide)
            emul().emul_ax(P(X), P(y2), P(sr2), *args, np.float32(1), nx, ny, nz, na, nb, lpr,
                           kahan, rows)
            emul().emul_set_wide(0)
            assert np.array_equal(y2, y) and np.array_equal(sr2, sr), (lpr, kahan, rows, wide)
    x = np.zeros(nx * ny * nz, np.float32)
    sc = np.zeros_like(x)
    emul().emul_atx(P(x), P(Y), P(sc), *args, g["det"], g["det"], np.float32(1), nx, ny, nz, na, nb)
    xo, sco = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
                      g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz, na, nb)
    assert np.array_equal(x, xo) and np.array_equal(sc, sco)


def test_degenerate_rays_match_oracle(orc):
    """phi = 0 with odd na: the central bin has y_det = 0, so b == 0 and the
    backprojector takes the reference's 2-D fallback (cu:134-135)."""
    nx, ny, nz, na, nb = 16, 16, 4, 31, 5
    g = orc.gateway_geometry(0.0, 200.0, 200.0, 1.0, 1.0, na, nb)
    assert g["sin_phi"] == 0.0 and g["y_det"][na // 2] == 0.0
    Y = np.random.default_rng(4).random(na * nb).astype(np.float32)
    args = (g["cos_phi"], g["sin_phi"], P(g["y_det"]), P(g["z_det"]), g["SO"], g["OD"], g["scale"])
    x = np.zeros(nx * ny * nz, np.float32)
    sc = np.zeros_like(x)
    emul().emul_atx(P(x), P(Y), P(sc), *args, g["det"], g["det"], np.float32(1), nx, ny, nz, na, nb)
    xo, sco = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
                      g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz, na, nb)
    assert np.array_equal(x, xo) and np.array_equal(sc, sco)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("phi", [-0.3, -1.9, -3.7])
def test_precision_backprojector_is_exact(orc, case, phi):
    """The precision variant's backprojector (tile-local FP64 ray setup) is
    within ~1e-6 of the exact FP64 backprojection of the same rays (end points
    rounded to FP32 as the operators form them), where the reference's FP32
    arithmetic (the oracle) is off by ~1e-4."""
    from oracle import exact64 as E
    _, nx, ny, nz, na, nb, SO, OD, det = case
    g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
    lib = emul()
    lib.emul_atx_prec.argtypes = [C.c_void_p] * 3 + [C.c_float] * 2 + [C.c_void_p] * 2 + \
        [C.c_float] * 6 + [C.c_int] * 8
    Y = np.random.default_rng(5).random(na * nb).astype(np.float32)
    x = np.zeros(nx * ny * nz, np.float32)
    sc = np.zeros_like(x)
    tiles = (16, 16, 1) if nz == 1 else (8, 8, 4)
    lib.emul_atx_prec(P(x), P(Y), P(sc), g["cos_phi"], g["sin_phi"], P(g["y_det"]), P(g["z_det"]),
                      g["SO"], g["OD"], g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz,
                      na, nb, *tiles)
    xo, _ = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
                    g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz, na, nb)
    if nz == 1:
        A = E.system_matrix_2d(g["cos_phi"], g["sin_phi"], float(g["SO"]), float(g["OD"]),
                               g["y_det"], nx, ny, f32_points=True)
    else:
        A = E.system_matrix(g["cos_phi"], g["sin_phi"], float(g["SO"]), float(g["OD"]),
                            g["y_det"], g["z_det"], nx, ny, nz, f32_points=True)
    xe = A.T @ Y.astype(np.float64)
    if case[0] == "fan_odd":
        # the reference drops the chord of the corner ray at the isocentre;
        # compare away from those four pixels (the oracle pin covers them)
        keep = np.ones(nx * ny, bool)
        for v in range(nx * ny):
            ix, iy = v % nx, v // nx
            if ix in (nx // 2 - 1, nx // 2) and iy in (ny // 2 - 1, ny // 2):
                keep[v] = False
        x, xo, xe = x[keep], xo[keep], xe[keep]
    err_prec = np.abs(x - xe).max() / np.abs(xe).max()
    err_ref = np.abs(xo - xe).max() / np.abs(xe).max()
    assert err_prec <= 5e-6, (err_prec, err_ref)
    assert np.linalg.norm(x - xe) <= np.linalg.norm(xo - xe) + 1e-7 * np.linalg.norm(xe)


def test_backprojector_bit_exact_random_geometries(orc):
    """Random fan and cone geometries (angles, distances, pitches, odd and even
    sizes): the device backprojector, zero-length skip included, stays
    bit-identical to the oracle."""
    rng = np.random.default_rng(21)
    for _ in range(24):
        nx, ny = int(rng.integers(6, 28)), int(rng.integers(6, 28))
        nz = 1 if rng.random() < 0.4 else int(rng.integers(2, 9))
        na = int(rng.integers(5, 60))
        nb = 1 if nz == 1 else int(rng.integers(2, 20))
        SO = float(rng.uniform(20, 800))
        OD = float(rng.uniform(20, 800))
        det = float(rng.uniform(0.3, 3.0))
        phi = float(rng.uniform(-7, 7))
        g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
        Y = rng.standard_normal(na * nb).astype(np.float32)
        x = np.zeros(nx * ny * nz, np.float32)
        sc = np.zeros_like(x)
        emul().emul_atx(P(x), P(Y), P(sc), g["cos_phi"], g["sin_phi"], P(g["y_det"]),
                        P(g["z_det"]), g["SO"], g["OD"], g["scale"], g["det"], g["det"],
                        np.float32(1), nx, ny, nz, na, nb)
        xo, sco = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"],
                          g["OD"], g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz,
                          na, nb)
        assert np.array_equal(x, xo) and np.array_equal(sc, sco), (nx, ny, nz, na, nb, SO, OD, det, phi)


def test_flat_and_unit_pitch_shortcuts_change_no_bit():
    """The forward projector's flat-ray and dz == 1 shortcuts and the
    backprojector's flat-ray chord give bit-identical results to the general
    expressions on random fan geometries (cardinal angles and odd bin counts
    included, where the a == 0 / b == 0 fallbacks appear), and on cone
    geometries (where only the dz == 1 shortcut applies)."""
    rng = np.random.default_rng(33)
    lib = emul()
    lib.emul_set_general.argtypes = [C.c_int]
    phis = [0.0, -np.pi / 2, -np.pi, float(rng.uniform(-7, 7))] + \
        [float(v) for v in rng.uniform(-7, 7, 12)]
    for k, phi in enumerate(phis):
        cone = k % 4 == 3
        nx, ny = int(rng.integers(6, 40)), int(rng.integers(6, 40))
        nz, nb = (int(rng.integers(2, 7)), int(rng.integers(2, 12))) if cone else (1, 1)
        na = int(rng.integers(5, 70))
        SO, OD = float(rng.uniform(20, 800)), float(rng.uniform(20, 800))
        det = float(rng.uniform(0.3, 3.0))
        from oracle import oracle as orc
        g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
        X = rng.random(nx * ny * nz).astype(np.float32)
        Y = rng.standard_normal(na * nb).astype(np.float32)
        out = []
        for general in (0, 1):
            lib.emul_set_general(general)
            y = np.zeros(na * nb, np.float32)
            sr = np.zeros_like(y)
            lib.emul_ax(P(X), P(y), P(sr), g["cos_phi"], g["sin_phi"], P(g["y_det"]),
                        P(g["z_det"]), g["SO"], g["OD"], g["scale"], np.float32(1), nx, ny, nz,
                        na, nb, 1, 0, 1 + general)
            x = np.zeros(nx * ny * nz, np.float32)
            sc = np.zeros_like(x)
            lib.emul_atx(P(x), P(Y), P(sc), g["cos_phi"], g["sin_phi"], P(g["y_det"]),
                         P(g["z_det"]), g["SO"], g["OD"], g["scale"], g["det"], g["det"],
                         np.float32(1), nx, ny, nz, na, nb)
            out.append((y, sr, x, sc))
        lib.emul_set_general(0)
        for a, b in zip(*out):
            assert np.array_equal(a, b), (phi, nx, ny, nz, na, nb)


def test_in_kernel_detector_grid_matches_gateway_tables(orc):
    """The forward projector computes the gateway's detector coordinates in
    place (det_coord); they equal the tables the gateways build
    (Ax_main_test.cpp:156-160), bit for bit, for even and odd bin counts."""
    lib = emul()
    lib.emul_det_coord.argtypes = [C.c_int, C.c_int, C.c_float]
    lib.emul_det_coord.restype = C.c_float
    rng = np.random.default_rng(8)
    for _ in range(20):
        na, nb = int(rng.integers(1, 1600)), int(rng.integers(1, 800))
        vxl, det = float(rng.uniform(0.1, 2)), float(rng.uniform(0.1, 3))
        g = orc.gateway_geometry(-0.4, 500.0, 400.0, vxl, det, na, nb)
        ya = np.array([lib.emul_det_coord(i, na, g["det"]) for i in range(na)], np.float32)
        zb = np.array([lib.emul_det_coord(j, nb, g["det"]) for j in range(nb)], np.float32)
        assert np.array_equal(ya, g["y_det"]) and np.array_equal(zb, g["z_det"])


def test_footprint_cut_is_exact_and_halves_the_work(orc):
    """The backprojector's footprint cut (the voxel's shadow through the source
    plus a 0.05-voxel margin) drops only bins whose length is exactly 0: on
    random fan and cone geometries, near-cardinal angles and odd bin counts
    included, the result is bit-identical to the reference's full footprint
    (and to the oracle), while the (voxel, bin) pairs visited fall by at
    least a third (3-D: about half)."""
    lib = emul()
    lib.emul_set_general.argtypes = [C.c_int]
    lib.emul_footprint_pairs.argtypes = [C.c_float, C.c_float, C.c_void_p] + [C.c_float] * 4 + \
        [C.c_int] * 6
    lib.emul_footprint_pairs.restype = C.c_long
    rng = np.random.default_rng(44)
    cut_pairs = full_pairs = 0
    phis = [float(v) for v in rng.uniform(-7, 7, 14)] + \
        [k * np.pi / 2 + d for k in range(-2, 3) for d in (0.0, 1e-7, -3e-7)]
    for k, phi in enumerate(phis):
        cone = k % 2 == 1
        nx, ny = int(rng.integers(6, 36)), int(rng.integers(6, 36))
        nz, nb = (int(rng.integers(2, 10)), int(rng.integers(2, 30))) if cone else (1, 1)
        na = int(rng.integers(5, 90))
        SO, OD = float(rng.uniform(40, 900)), float(rng.uniform(20, 900))
        det = float(rng.uniform(0.3, 3.0))
        g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
        Y = rng.standard_normal(na * nb).astype(np.float32)
        out = []
        for general in (1, 0):
            lib.emul_set_general(general)
            x = np.zeros(nx * ny * nz, np.float32)
            sc = np.zeros_like(x)
            lib.emul_atx(P(x), P(Y), P(sc), g["cos_phi"], g["sin_phi"], P(g["y_det"]),
                         P(g["z_det"]), g["SO"], g["OD"], g["scale"], g["det"], g["det"],
                         np.float32(1), nx, ny, nz, na, nb)
            out.append((x, sc))
        lib.emul_set_general(0)
        xo, sco = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"],
                          g["OD"], g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz,
                          na, nb)
        for a, b in zip(*out):
            assert np.array_equal(a, b), (phi, nx, ny, nz, na, nb, SO, OD, det)
        assert np.array_equal(out[1][0], xo) and np.array_equal(out[1][1], sco)
        args = (g["cos_phi"], g["sin_phi"], P(g["y_det"]), g["SO"], g["OD"], g["det"], g["det"],
                nx, ny, nz, na, nb)
        cut_pairs += lib.emul_footprint_pairs(*args, 1)
        full_pairs += lib.emul_footprint_pairs(*args, 0)
    assert cut_pairs <= 0.67 * full_pairs, (cut_pairs, full_pairs)


def test_fastdiv_is_exact():
    """The multiply-high division the backprojectors use to split a voxel
    index equals integer division for every divisor up to 2^12, for powers
    of two and their neighbours up to 2^24, and for random divisors, on
    dividends at the 32-bit extremes, around multiples of the divisor and at
    random."""
    lib = emul()
    lib.emul_fastdiv.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_long]
    rng = np.random.default_rng(11)
    ds = list(range(1, 4097))
    ds += [d for k in range(12, 25) for d in ((1 << k) - 1, 1 << k, (1 << k) + 1) if d < (1 << 24)]
    ds += [int(d) for d in rng.integers(1, 1 << 24, 300)]
    for d in ds:
        n = np.concatenate([
            np.array([0, 1, d - 1, d, d + 1, 2**32 - 1, 2**32 - 2, 2**31, 2**31 - 1], np.uint64),
            (np.arange(1, 200, dtype=np.uint64) * d + np.uint64(d - 1)) % np.uint64(2**32),
            rng.integers(0, 2**32, 2000, dtype=np.uint64)]).astype(np.uint32)
        q = np.zeros_like(n)
        lib.emul_fastdiv(d, P(n), P(q), n.size)
        assert np.array_equal(q, n // np.uint32(d)), d


def _pairs(rng, n, near_edges):
    """n (ray, voxel) pairs: random rays of random views (angles, distances,
    detector points) and unit voxels near a point of the ray; with
    near_edges the ray passes within 0 to 3e-4 voxel of an edge or a corner
    of its voxel, where FP32 face tests pass on several faces at once."""
    phi = rng.uniform(-7, 7, n)
    c = np.cos(phi).astype(np.float32)
    s = np.sin(phi).astype(np.float32)
    SO = rng.uniform(30, 1500, n).astype(np.float32)
    OD = rng.uniform(30, 1500, n).astype(np.float32)
    yd = rng.uniform(-400, 400, n).astype(np.float32)
    zd = rng.uniform(-300, 300, n).astype(np.float32)
    Sx, Sy = (c * -SO).astype(np.float64), (s * -SO).astype(np.float64)
    Dx = (c * OD - s * yd).astype(np.float64)
    Dy = (s * OD + c * yd).astype(np.float64)
    t = rng.uniform(0.2, 0.8, n)
    P = np.stack([Sx + t * (Dx - Sx), Sy + t * (Dy - Sy), t * zd], 1)
    if near_edges:
        off = rng.choice([-0.5, 0.5], (n, 3))
        free = rng.integers(0, 4, n)  # the axis along the edge (3: a corner)
        for k in range(3):
            off[free == k, k] = rng.uniform(-0.5, 0.5, (free == k).sum())
        off += rng.choice([0, 1e-7, -1e-7, 1e-6, -1e-6, 1e-5, -1e-5, 1e-4, -1e-4, 3e-4], (n, 3))
    else:
        off = rng.uniform(-0.9, 0.9, (n, 3))
    C = (P + off).astype(np.float32)
    return c, s, SO, OD, yd, zd, C[:, 0].copy(), C[:, 1].copy(), C[:, 2].copy()


@pytest.mark.parametrize("near_edges", [False, True])
def test_two_sided_chord_matches_reference_rule(orc, near_edges):
    """find_l_3d_rec (the backprojector's chord: entry and exit candidates
    selected when one passes on each side, the reference's ordered rule
    otherwise) equals the oracle's find_l_3d bit for bit on 40000 pairs;
    near voxel edges and corners several faces pass their FP32 tests and the
    ordered rule decides (including the chords the reference drops)."""
    lib = emul()
    lib.emul_find_l_3d_pairs.argtypes = [C.c_long] + [C.c_void_p] * 9 + [C.c_int, C.c_void_p]
    n = 40000
    args = _pairs(np.random.default_rng(31 + near_edges), n, near_edges)
    out = np.zeros(n, np.float32)
    lib.emul_find_l_3d_pairs(n, *[a.ctypes.data for a in args], 8, out.ctypes.data)
    c, s, SO, OD, yd, zd, xc, yc, zc = args
    L = orc.lib()
    ref = np.array([L.orc_find_l_3d(float(c[k] * -SO[k]), float(s[k] * -SO[k]), 0.0,
                                    float(np.float32(c[k] * OD[k]) - np.float32(s[k] * yd[k])),
                                    float(np.float32(s[k] * OD[k]) + np.float32(c[k] * yd[k])),
                                    float(zd[k]), 1.0, 1.0, 1.0, float(xc[k]), float(yc[k]),
                                    float(zc[k])) for k in range(n)], np.float32)
    assert (out > 0).sum() > n // 3
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_general_views_take_the_mode_free_walk(orc):
    """Cone views whose rays are all general with exact face offsets
    (view_general) take the walk without a mode dispatch: a view with an odd
    row count (a row with c == 0) is one, a fan view (flat rays) and a view
    with a b == 0 ray are not.  The backprojection of the general views
    equals the oracle bit for bit."""
    lib = emul()
    f, i, vp = C.c_float, C.c_int, C.c_void_p
    lib.emul_view_general.argtypes = [f, f, vp, vp, f, f, f, f, f, i, i, i, i, i]
    for nx, ny, nz, na, nb, SO, OD, det, phi, want in [
            (20, 18, 10, 32, 16, 300.0, 300.0, 1.5, -0.3, 1),
            (20, 18, 9, 31, 15, 200.0, 300.0, 1.1, -2.2, 1),   # odd nb: z_det == 0 row
            (16, 16, 6, 24, 7, 25.0, 25.0, 1.5, -1.2, 1),
            (40, 36, 1, 64, 1, 500.0, 400.0, 1.3, -0.3, 0),   # fan: flat rays
            (16, 16, 4, 31, 5, 200.0, 200.0, 1.0, 0.0, 0)]:   # b == 0 at the central bin
        g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
        got = lib.emul_view_general(g["cos_phi"], g["sin_phi"], P(g["y_det"]), P(g["z_det"]),
                                    g["SO"], g["OD"], g["det"], g["det"], np.float32(1), nx, ny,
                                    nz, na, nb)
        assert got == want, (nx, ny, nz, na, nb, phi)
        Y = np.random.default_rng(nx + na).standard_normal(na * nb).astype(np.float32)
        x = np.zeros(nx * ny * nz, np.float32)
        sc = np.zeros_like(x)
        emul().emul_atx(P(x), P(Y), P(sc), g["cos_phi"], g["sin_phi"], P(g["y_det"]),
                        P(g["z_det"]), g["SO"], g["OD"], g["scale"], g["det"], g["det"],
                        np.float32(1), nx, ny, nz, na, nb)
        xo, sco = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"],
                          g["OD"], g["scale"], g["det"], g["det"], np.float32(1), nx, ny, nz,
                          na, nb)
        assert np.array_equal(x, xo) and np.array_equal(sc, sco)


def test_forward_slab_range_keeps_every_slab_with_a_piece(orc):
    """ax_bundle_slabs (the slabs a row bundle of the forward projector
    marches) leaves out no slab with a piece in the volume, on random cone
    and fan geometries and on the 512^3 workload's geometry (a subset of its
    detector columns), where it drops about a quarter of the slabs."""
    lib = emul()
    f, i, vp = C.c_float, C.c_int, C.c_void_p
    lib.emul_ax_slab_range_misses.restype = C.c_long
    lib.emul_ax_slab_range_misses.argtypes = [f, f, vp, vp, f, f, f, i, i, i, i, i, i, vp, vp]
    rng = np.random.default_rng(41)
    cases = []
    for _ in range(30):
        nz = 1 if rng.random() < 0.3 else int(rng.integers(2, 40))
        cases.append((int(rng.integers(4, 60)), int(rng.integers(4, 60)), nz,
                      int(rng.integers(3, 80)), 1 if nz == 1 else int(rng.integers(2, 40)),
                      float(rng.uniform(20, 900)), float(rng.uniform(20, 900)),
                      float(rng.uniform(0.3, 3.0)), float(rng.uniform(-7, 7)), int(rng.integers(1, 4))))
    for phi in (-0.5 / 180 * np.pi, -37.3 / 180 * np.pi, -np.pi / 2):
        cases.append((512, 512, 512, 1024, 768, 600.0 / 0.5, 500.0 / 0.5, 0.6 / 0.5, phi, 3))
    dropped = []
    for nx, ny, nz, na, nb, SO, OD, det, phi, K in cases:
        g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
        ydet = g["y_det"]
        if na == 1024:  # every 16th detector column of the full-size view
            ydet = np.ascontiguousarray(ydet[::16])
        kept, total = C.c_long(), C.c_long()
        miss = lib.emul_ax_slab_range_misses(g["cos_phi"], g["sin_phi"], P(ydet), P(g["z_det"]),
                                             g["SO"], g["OD"], np.float32(1), nx, ny, nz,
                                             ydet.size, nb, K, C.byref(kept), C.byref(total))
        assert miss == 0, (nx, ny, nz, na, nb, SO, OD, det, phi, K)
        if na == 1024:
            dropped.append(1 - kept.value / total.value)
    assert min(dropped) > 0.15, dropped


def test_forward_tile_order_is_a_bijection():
    """The row-bundle forward projector's longest-first tile order
    (ax_rows_tile) maps the launch's blocks onto its tiles one to one, so
    every bundle is projected once; it reorders the 1024 x 768 detector of the
    512^3 workload and leaves smaller views, and shapes whose bands do not
    divide evenly, in xcd_block's order."""
    lib = emul()
    lib.emul_ax_rows_tile_check.restype = C.c_int
    lib.emul_ax_rows_tile_check.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
    moved = C.c_long()
    cases = [(1024, 768, 2, True), (1024, 768, 3, True), (2048, 1536, 2, True),
             (384, 384, 3, False), (384, 384, 2, False), (1000, 768, 2, False),
             (1024, 770, 2, False), (1024, 1, 1, False), (768, 1, 1, False), (4096, 128, 2, True)]
    for na, nb, rows, reordered in cases:
        assert lib.emul_ax_rows_tile_check(na, nb, rows, C.byref(moved)) == 1, (na, nb, rows)
        assert (moved.value > 0) == reordered, (na, nb, rows, moved.value)


def test_backprojector_tile_order_is_a_bijection():
    """The sorted 3-D backprojector's z-interleaved tile order
    (atx_tile_z_interleave) maps a launch's blocks onto its tiles one to one,
    and where it applies gives every XCD (workgroup b runs on XCD b % 8)
    tiles from min(8, layers / 8) eighths of the volume's height instead of
    one; layer counts that do not split into whole blocks keep the plain
    order."""
    lib = emul()
    lib.emul_atx_tile_z_check.restype = C.c_int
    lib.emul_atx_tile_z_check.argtypes = [C.c_uint, C.c_uint, C.c_void_p]
    spread = (C.c_int * 8)()
    # (tiles per layer, layers, interleaved): 512^3 with 16x16x2, 8x8x8 and
    # 8x4x16 tiles, 256^3 with 16x16x2 and 8x8x8, then shapes left alone
    for per, nl, inter in [(32 * 32, 256, True), (64 * 64, 64, True), (64 * 128, 32, True),
                           (16 * 16, 128, True), (32 * 32, 32, True), (8, 24, True),
                           (6, 8, False), (5, 100, False), (3, 7, False)]:
        assert lib.emul_atx_tile_z_check(per, nl, spread) == 1, (per, nl)
        if inter:
            assert all(v == min(8, nl // 8) for v in spread), (per, nl, list(spread))


def test_footprint_cut_keeps_every_nonzero_pair_at_512_cubed(orc):
    """At the 512^3 workload's geometry (a 1200-voxel source distance, where
    find_l_3d's FP32 rounding is largest), the cut footprint holds every
    (voxel, bin) pair whose reference length is nonzero, on 20000 voxels of
    each of eight views, and keeps at most 10% more pairs than those."""
    lib = emul()
    f, i, vp = C.c_float, C.c_int, C.c_void_p
    lib.emul_footprint_cut_check.restype = C.c_long
    lib.emul_footprint_cut_check.argtypes = [f, f, vp, vp, f, f, f, f, i, i, i, i, i, i,
                                             C.c_uint, vp, vp]
    kept_all = nonzero_all = 0
    for k, deg in enumerate([0.5, 17.3, 37.3, 45.0, 90.5, 133.0, 211.7, 300.1]):
        g = orc.gateway_geometry(-deg / 180 * np.pi, 600.0, 500.0, 0.5, 0.6, 1024, 768)
        kept, nonzero = C.c_long(), C.c_long()
        miss = lib.emul_footprint_cut_check(g["cos_phi"], g["sin_phi"], P(g["y_det"]),
                                            P(g["z_det"]), g["SO"], g["OD"], g["det"], g["det"],
                                            512, 512, 512, 1024, 768, 20000, 11 + k,
                                            C.byref(kept), C.byref(nonzero))
        assert miss == 0, deg
        kept_all += kept.value
        nonzero_all += nonzero.value
    assert nonzero_all > 8 * 20000 * 2 and kept_all <= 1.10 * nonzero_all, (kept_all, nonzero_all)


def test_face_tests_credit_only_near_misses(orc):
    """The premise of the footprint cut's 1e-2 voxel margin: rays aimed at
    voxel edges and corners from a 1200-voxel source distance (detector
    points moved by up to 3e-3), on the reference's voxel grid, get a nonzero
    length from find_l_3d only where the exact FP64 ray misses the voxel by
    less than 1e-3 voxel (observed: about 6e-5)."""
    lib = emul()
    lib.emul_find_l_3d_pairs.argtypes = [C.c_long] + [C.c_void_p] * 9 + [C.c_int, C.c_void_p]
    rng = np.random.default_rng(7)
    n = 200000
    phi = rng.uniform(-7, 7, n)
    c = np.cos(phi).astype(np.float32)
    s = np.sin(phi).astype(np.float32)
    SO = np.full(n, 1200, np.float32)
    OD = np.full(n, 1000, np.float32)
    Cc = np.floor(rng.uniform(-256, 256, (n, 3))) + 0.5
    off = rng.choice([-0.5, 0.5], (n, 3))
    free = rng.integers(0, 4, n)
    for k in range(3):
        off[free == k, k] = rng.uniform(-0.5, 0.5, (free == k).sum())
    S = np.stack([c.astype(np.float64) * -1200, s.astype(np.float64) * -1200, np.zeros(n)], 1)
    v = Cc + off - S
    cd, sd = c.astype(np.float64), s.astype(np.float64)
    t = 2200.0 / (cd * v[:, 0] + sd * v[:, 1])
    yd = ((-sd * v[:, 0] + cd * v[:, 1]) * t + rng.uniform(-3e-3, 3e-3, n)).astype(np.float32)
    zd = (v[:, 2] * t + rng.uniform(-3e-3, 3e-3, n)).astype(np.float32)
    Cf = Cc.astype(np.float32)
    out = np.zeros(n, np.float32)
    args = [np.ascontiguousarray(a) for a in (c, s, SO, OD, yd, zd, Cf[:, 0], Cf[:, 1], Cf[:, 2])]
    lib.emul_find_l_3d_pairs(n, *[a.ctypes.data for a in args], 512, out.ctypes.data)
    # exact L-inf enlargement of the voxel the ray needs to meet it (bisection)
    D = np.stack([(c * OD - s * yd).astype(np.float64), (s * OD + c * yd).astype(np.float64),
                  zd.astype(np.float64)], 1)
    d = D - S
    lo, hi = np.full(n, -0.5), np.full(n, 0.1)
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(50):
            mid = (lo + hi) / 2
            t0 = (Cc - 0.5 - mid[:, None] - S) / d
            t1 = (Cc + 0.5 + mid[:, None] - S) / d
            meet = np.minimum(t0, t1).max(1) <= np.maximum(t0, t1).min(1)
            hi = np.where(meet, mid, hi)
            lo = np.where(meet, lo, mid)
    hit = out > 0
    assert hit.sum() > n // 2 and (hi[hit] > 0).sum() > 10  # some near misses are credited
    assert hi[hit].max() < 1e-3, hi[hit].max()
