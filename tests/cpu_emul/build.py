"""Builds tests/cpu_emul/libemul.so: the projector device functions compiled
for the host (test infrastructure)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def stale(out):
    """Whether `out` is missing or older than a source it is built from."""
    if not os.path.exists(out):
        return True
    csrc = os.path.join(ROOT, "asd_pocs_mex_cuda_demo_amd", "csrc")
    srcs = [os.path.join(HERE, "emul.hip"), os.path.join(ROOT, "include", "asdpocs.h")]
    srcs += [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".hpp")]
    t = os.path.getmtime(out)
    return any(os.path.getmtime(f) > t for f in srcs)


def build():
    out = os.path.join(HERE, "libemul.so")
    pkg = os.path.join(ROOT, "asd_pocs_mex_cuda_demo_amd")
    # ASDPOCS_SANITIZE=1: the host code (the emulated device functions) with
    # AddressSanitizer and UndefinedBehaviorSanitizer; device code untouched
    san = ["-g", "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
           "-fno-sanitize-recover=undefined"] if os.environ.get("ASDPOCS_SANITIZE") == "1" else []
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-fPIC", "-shared",
                    "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                    *san, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(pkg, 'csrc')}",
                    os.path.join(HERE, "emul.hip"), "-o", out + ".tmp"],
                   check=True, capture_output=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
