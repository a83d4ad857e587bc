"""Writes oracle_known_answers.npz: small inputs and the CPU oracle's outputs
for them (forward and backprojection of a fan and a cone view, the 2-D and
3-D TV descents, and two ASD-POCS iterations), produced once the oracle's
functions were pinned by tests/test_oracle_pins.py.  tests/test_known_answers.py
checks that the oracle still reproduces them bit for bit, so a later edit of
the oracle that changes any result is caught.

    python tests/golden/make_known_answers.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def cases():
    """The inputs (deterministic) as a dict of arrays and scalars."""
    rng = np.random.default_rng(20241204)
    c = {}
    c["fan_X"] = rng.random(16 * 16).astype(np.float32)
    c["fan_Y"] = rng.standard_normal(24).astype(np.float32)
    c["cone_X"] = rng.random(8 * 8 * 4).astype(np.float32)
    c["cone_Y"] = rng.standard_normal(12 * 6).astype(np.float32)
    c["tv2_I"] = rng.random((6, 5))
    c["tv2_I0"] = rng.random((6, 5)) * 0.1
    c["tv3_I"] = rng.random((4, 4, 3))
    xx, yy = np.meshgrid(np.arange(16) - 7.5, np.arange(16) - 7.5, indexing="ij")
    c["asd_img"] = 0.02 * ((xx / 6.0) ** 2 + (yy / 4.5) ** 2 <= 1.0)
    return c


FAN = dict(nx=16, ny=16, nz=1, na=24, nb=1, SO=60.0, OD=40.0, vxlsize=1.0, detsize=1.3)
CONE = dict(nx=8, ny=8, nz=4, na=12, nb=6, SO=40.0, OD=30.0, vxlsize=1.0, detsize=1.4)
ANGLES = (-0.4, -2.2)


def answers(orc):
    c = cases()
    out = {}
    for name, g, X, Y in (("fan", FAN, c["fan_X"], c["fan_Y"]),
                          ("cone", CONE, c["cone_X"], c["cone_Y"])):
        for k, phi in enumerate(ANGLES):
            q = dict(g, phi=phi)
            out[f"{name}{k}_y"], out[f"{name}{k}_sr"] = orc.ax_para(X, q)
            out[f"{name}{k}_x"], out[f"{name}{k}_sc"] = orc.atx_para(Y, q)
    out["tv2"] = orc.tv2d(c["tv2_I"], 0.2, 4, 1, c["tv2_I0"])
    out["tv3"] = orc.tv3d(c["tv3_I"], 0.05, 3, 1)
    para = dict(FAN, phi_vec=np.arange(1, 9) * 45.0)
    P = np.zeros((FAN["na"], 8))
    for i, th in enumerate(para["phi_vec"]):
        P[:, i] = orc.ax_para(c["asd_img"], dict(FAN, phi=-th / 180 * np.pi))[0]
    m, o = orc.asd_pocs(P, dict(para, maxnumIter=2, alpha=0.2, numOfgradients=3,
                                I0=c["asd_img"]))
    out["asd_m"] = m
    out["asd_rmse"] = o["RMSE"]
    out["asd_alpha_beta_tol"] = np.array([o["alpha"], o["beta"], o["dataTolerance"]])
    return out


def main():
    from oracle import oracle as orc
    np.savez_compressed(os.path.join(HERE, "oracle_known_answers.npz"), **cases(), **answers(orc))
    print("wrote", os.path.join(HERE, "oracle_known_answers.npz"))


if __name__ == "__main__":
    main()
