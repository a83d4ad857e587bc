"""Command-line front-end (python -m asd_pocs_mex_cuda_demo_amd): argument and
file handling on the CPU, and on the GPU the same results as the Python calls
it wraps."""
import json

import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import __main__ as cli


def test_para_file_fields_and_array_references(tmp_path):
    m0 = np.arange(12.0).reshape(3, 4)
    np.save(tmp_path / "m0.npy", m0)
    (tmp_path / "para.json").write_text(json.dumps(
        {"nx": 3, "ny": 4, "nz": 1, "phi_vec": [1, 2, 3], "m0": str(tmp_path / "m0.npy")}))
    para = cli.load_para(str(tmp_path / "para.json"))
    assert para["nx"] == 3 and para["phi_vec"].dtype == np.float64
    assert np.array_equal(para["m0"], m0)


def test_bad_inputs_are_rejected_before_the_gpu(tmp_path):
    np.save(tmp_path / "p.npy", np.zeros(5))
    (tmp_path / "para.json").write_text("[1, 2]")
    with pytest.raises(SystemExit, match="2-D"):
        cli.main(["asd-pocs", "--projections", str(tmp_path / "p.npy"), "--para",
                  str(tmp_path / "para.json"), "--out", str(tmp_path / "r.npz")])
    np.save(tmp_path / "p.npy", np.zeros((5, 2)))
    with pytest.raises(SystemExit, match="JSON object"):
        cli.main(["sart", "--projections", str(tmp_path / "p.npy"), "--para",
                  str(tmp_path / "para.json"), "--out", str(tmp_path / "r.npz")])
    with pytest.raises(SystemExit):
        cli.main(["reconstruct", "--projections", "x", "--para", "y", "--out", "z"])


@pytest.mark.gpu
def test_cli_matches_the_python_calls(tmp_path):
    from asd_pocs_mex_cuda_demo_amd import ops, phantoms
    n, nv, na = 32, 12, 48
    para = dict(nx=n, ny=n, nz=1, SO=500.0, OD=400.0, vxlsize=1.0, detsize=1.2, na=na, nb=1,
                maxnumIter=3, alpha=0.2, numOfgradients=5)
    phi = np.arange(1, nv + 1) * 360.0 / nv
    img = phantoms.shepp_logan_2d(n).astype(np.float32)
    P = np.stack([ops.Ax_cone_gpu(img, dict(para, phi=-t / 180 * np.pi))[0] for t in phi],
                 axis=1).astype(np.float64)
    np.save(tmp_path / "P.npy", P)
    np.save(tmp_path / "I0.npy", img.astype(np.float64))
    (tmp_path / "para.json").write_text(json.dumps(
        dict(para, phi_vec=phi.tolist(), I0=str(tmp_path / "I0.npy"))))
    assert cli.main(["asd-pocs", "--projections", str(tmp_path / "P.npy"), "--para",
                     str(tmp_path / "para.json"), "--out", str(tmp_path / "r.npz")]) == 0
    r = np.load(tmp_path / "r.npz")
    m, out = ops.ASD_POCS_cuda(P, dict(para, phi_vec=phi, I0=img.astype(np.float64)))
    assert np.array_equal(r["m"], m)
    assert int(r["IterationNum"]) == out["IterationNum"] == 3
    assert np.array_equal(r["RMSE"], out["RMSE"])
    assert cli.main(["sart", "--projections", str(tmp_path / "P.npy"), "--para",
                     str(tmp_path / "para.json"), "--out", str(tmp_path / "s.npz")]) == 0
    s = np.load(tmp_path / "s.npz")
    ms, outs = ops.SART_cuda_inline(P, dict(para, phi_vec=phi))
    assert np.array_equal(s["m"], ms)
    assert float(s["dataTolerance"]) == outs["dataTolerance"]


@pytest.mark.gpu
def test_cli_matches_the_oracle(tmp_path, orc):
    """The command line with the reference arithmetic (precision 4), a warm
    start read from a file and non-default rules equals the oracle's
    ASD_POCS_cuda and SART_cuda_inline bit for bit."""
    from asd_pocs_mex_cuda_demo_amd import phantoms
    n, nv, na = 32, 10, 48
    para = dict(nx=n, ny=n, nz=1, SO=500.0, OD=400.0, vxlsize=1.0, detsize=1.2, na=na, nb=1,
                maxnumIter=3, alpha=0.25, alpha_red=0.9, numOfgradients=5, precision=4)
    phi = np.arange(1, nv + 1) * 360.0 / nv
    img = phantoms.shepp_logan_2d(n)
    P = np.stack([orc.ax_para(img, dict(para, phi=-t / 180 * np.pi))[0] for t in phi],
                 axis=1).astype(np.float64)
    m0 = np.full((n, n), 0.002)
    np.save(tmp_path / "P.npy", P)
    np.save(tmp_path / "m0.npy", m0)
    np.save(tmp_path / "I0.npy", img)
    (tmp_path / "para.json").write_text(json.dumps(
        dict(para, phi_vec=phi.tolist(), I0=str(tmp_path / "I0.npy"),
             m0=str(tmp_path / "m0.npy"))))
    assert cli.main(["asd-pocs", "--projections", str(tmp_path / "P.npy"), "--para",
                     str(tmp_path / "para.json"), "--out", str(tmp_path / "r.npz")]) == 0
    r = np.load(tmp_path / "r.npz")
    q = dict(para, phi_vec=phi, I0=img, m0=m0)
    mo, oo = orc.asd_pocs(P, q)
    assert np.array_equal(r["m"].ravel(order="F"), mo)
    assert np.array_equal(r["RMSE"], oo["RMSE"]) and float(r["alpha"]) == oo["alpha"]
    assert cli.main(["sart", "--projections", str(tmp_path / "P.npy"), "--para",
                     str(tmp_path / "para.json"), "--out", str(tmp_path / "s.npz")]) == 0
    s = np.load(tmp_path / "s.npz")
    ms, _ = orc.sart(P, q)
    assert np.array_equal(s["m"].ravel(order="F"), ms)
