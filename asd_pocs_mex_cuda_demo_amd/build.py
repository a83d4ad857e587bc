"""Build recipe of the native library and the MEX gateways (gfx950).

    python -m asd_pocs_mex_cuda_demo_amd.build

produces, in this package directory:
  libasdpocs.so            HIP kernels + C ABI (include/asdpocs.h)
  mex/Ax_cone_gpu.mexa64, mex/Atx_cone_gpu.mexa64, mex/imgTVGradMin.mexa64,
  mex/imgTVGradMin3D.mexa64, mex/ASD_POCS_gpu.mexa64
                           MEX gateways compiled against include/mex_api.h;
                           they resolve the mx*/mex* symbols from the host
                           process (MATLAB, or the test host in tests/fakemex).
"""
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libasdpocs.so")
MEXDIR = os.path.join(PKG, "mex")

HIP_SOURCES = ["common.hip", "projectors.hip", "tv.hip", "driver.hip", "batch.hip",
               "simulate.hip"]
MEX_SOURCES = ["Ax_cone_gpu.cpp", "Atx_cone_gpu.cpp", "imgTVGradMin.cpp", "imgTVGradMin3D.cpp",
               "ASD_POCS_gpu.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # Projector geometry reproduces the reference's FP32 expression order:
    # no contraction, IEEE division and square root, denormals kept.
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    # No packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INC}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build(verbose=False, jobs=None, gateways_only=False):
    """Compile libasdpocs.so and the MEX gateways (only the gateways, against
    the library already built, with gateways_only)."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(MEXDIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    if not gateways_only:
        build_library(jobs)
    build_gateways(jobs)
    if verbose:
        print("built", LIB, "and the MEX gateways")
    return LIB


def build_library(jobs):
    def obj(src):
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        r = _run([HIPCC, *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c",
                  os.path.join(CSRC, src), "-o", o])
        # No kernel may use scratch (private) memory: on gfx950 it goes through
        # L2 to HBM, costing a store and a load per lane per use.
        name = None
        for line in (r.stdout + r.stderr).splitlines():
            m = re.search(r"Function Name: (\S+)", line)
            if m:
                name = m.group(1)
            m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
            if m and int(m.group(1)) > 0:
                raise RuntimeError(f"{src}: kernel {name} uses {m.group(1)} bytes/lane of scratch")
        return o

    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(obj, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)


def build_gateways(jobs):

    # MEX gateways: host C++ against our own mex_api.h, linked to libasdpocs.so
    # with an rpath to this directory; mx*/mex* stay undefined until load.
    # ASDPOCS_SANITIZE=1: the gateways (host C++) with AddressSanitizer and
    # UndefinedBehaviorSanitizer
    san = (["-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=undefined"]
           if os.environ.get("ASDPOCS_SANITIZE") == "1" else [])

    def mex(src):
        name = src.replace(".cpp", ".mexa64")
        out = os.path.join(MEXDIR, name)
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", *san,
              f"-I{INC}", os.path.join(CSRC, "mex", src), "-o", out + ".tmp",
              f"-L{PKG}", "-lasdpocs", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(out + ".tmp", out)
        return out

    present = [s for s in MEX_SOURCES if os.path.exists(os.path.join(CSRC, "mex", s))]
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(mex, present))


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
