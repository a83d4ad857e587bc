"""Writes the real-data fixtures from the reference's test data (run here,
where the reference checkout is readable; the fixtures travel with the tests).

  swine_2d_90.npz   prjSwine (testdata/swine_2d.mat: 1537 bins x 360 views at
                    1 degree, fan beam, sod 405.941, sdd 906.251, bin 0.148 mm)
                    at every 4th view (4, 8, ..., 360 degrees), as float32,
                    with the geometry scalars
  head190.npz       img_slice190 (testdata/P10240000_head_slice190.mat, a
                    512 x 512 clinical slice, uint16 storage)

Both files are read with the package's own MAT-file reader; where scipy is
importable the arrays are checked against scipy.io.loadmat first.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from asd_pocs_mex_cuda_demo_amd.matio import read_mat  # noqa: E402

SRC = "/root/reference/testdata"


def load(name):
    path = os.path.join(SRC, name)
    a = read_mat(path)
    try:
        import scipy.io as sio
        b = sio.loadmat(path, mat_dtype=True)
        for k, v in a.items():
            assert v.dtype == b[k].dtype and np.array_equal(v, b[k]), k
    except ImportError:
        pass
    return a


def main():
    s = load("swine_2d.mat")
    P = s["prjSwine"]
    cols = np.arange(4, 361, 4)  # MATLAB columns 4, 8, ..., 360 (degrees)
    np.savez_compressed(os.path.join(HERE, "swine_2d_90.npz"),
                        P=P[:, cols - 1].astype(np.float32), phi_deg=cols.astype(np.float64),
                        det=float(s["detPixelSize"][0, 0]), sod=float(s["sod"][0, 0]),
                        sdd=float(s["sdd"][0, 0]), na=int(s["na"][0, 0]))
    h = load("P10240000_head_slice190.mat")
    img = h["img_slice190"]
    assert np.array_equal(img, np.round(img)) and img.min() >= 0 and img.max() < 65536
    np.savez_compressed(os.path.join(HERE, "head190.npz"), img=img.astype(np.uint16))
    for f in ("swine_2d_90.npz", "head190.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
