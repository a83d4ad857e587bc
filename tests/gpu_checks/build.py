"""Builds tests/gpu_checks/libgpuchecks.so: GPU checks of arithmetic
identities the kernels rely on (test infrastructure), with the library's
own HIP flags."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def build():
    out = os.path.join(HERE, "libgpuchecks.so")
    pkg = os.path.join(ROOT, "asd_pocs_mex_cuda_demo_amd")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-fPIC", "-shared",
                    "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                    "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(pkg, 'csrc')}", os.path.join(HERE, "checks.hip"),
                    "-o", out + ".tmp"], check=True, capture_output=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
