"""Command-line front-end of the device-resident reconstruction driver.

    python -m asd_pocs_mex_cuda_demo_amd asd-pocs --projections P.npy --para para.json --out r.npz
    python -m asd_pocs_mex_cuda_demo_amd sart --projections data.mat:prj --para para.json --out r.npz

It mirrors the MATLAB calls the drop-in serves: `[m, out] = ASD_POCS_cuda(P, para)`
(utilities/ASD_POCS_cuda.m:1-180) and `[m, out] = SART_cuda_inline(P, para)`
(utilities/SART_cuda_inline.m:1-107), with the same `para` field names.

- `--projections`: the sinogram P, (na*nb) x nv with one column per view, from a
  `.npy` file or a MAT-v5 variable (`file.mat:name`).
- `--para`: a JSON object of `para` fields (nx, ny, nz, SO, OD, vxlsize, detsize,
  na, nb, phi_vec, maxnumIter, alpha, ...).  Array-valued fields the driver
  reads from arrays (`m0`, `I0`) may be given as `"file.npy"` or
  `"file.mat:name"` strings.
- `--out`: an `.npz` file with `m` (nx x ny x nz, the returned image) and one
  entry per field of `out`.

The driver runs on the GPU through the C ABI; there is no CPU fallback.
"""
import argparse
import json
import sys

import numpy as np


def load_array(spec):
    """A `.npy` file, or variable `name` of a MAT-v5 file given as `file.mat:name`."""
    if ".mat:" in spec:
        path, name = spec.rsplit(":", 1)
        from . import matio
        data = matio.read_mat(path)
        if name not in data:
            raise SystemExit(f"{path} has no variable {name!r} (it has {sorted(data)})")
        return np.asarray(data[name])
    return np.load(spec, allow_pickle=False)


def load_para(path):
    with open(path) as f:
        para = json.load(f)
    if not isinstance(para, dict):
        raise SystemExit(f"{path}: para must be a JSON object")
    for key in ("m0", "I0"):
        if isinstance(para.get(key), str):
            para[key] = load_array(para[key])
    if "phi_vec" in para:
        para["phi_vec"] = np.asarray(para["phi_vec"], np.float64)
    return para


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m asd_pocs_mex_cuda_demo_amd",
                                 description=__doc__.split("\n\n")[0])
    ap.add_argument("command", choices=["asd-pocs", "sart"],
                    help="asd-pocs: ASD_POCS_cuda (ASD-POCS iterations); "
                         "sart: SART_cuda_inline (one SART sweep)")
    ap.add_argument("--projections", required=True,
                    help="sinogram P: file.npy or file.mat:variable")
    ap.add_argument("--para", required=True, help="para fields as a JSON object")
    ap.add_argument("--out", required=True, help="output .npz (m and the out fields)")
    args = ap.parse_args(argv)
    P = np.asarray(load_array(args.projections), np.float64)
    if P.ndim != 2:
        raise SystemExit(f"projections must be a 2-D (na*nb) x nv array, got shape {P.shape}")
    para = load_para(args.para)
    from . import ops
    if args.command == "asd-pocs":
        m, out = ops.ASD_POCS_cuda(P, para)
    else:
        m, out = ops.SART_cuda_inline(P, para)
    fields = {k: np.asarray(v) for k, v in out.items() if k != "reconImg"}
    np.savez(args.out, m=np.asarray(m), **fields)
    summary = {k: (float(v) if np.ndim(v) == 0 else list(np.shape(v))) for k, v in fields.items()}
    print(json.dumps({"command": args.command, "out": args.out, "m_shape": list(np.shape(m)),
                      **summary}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
