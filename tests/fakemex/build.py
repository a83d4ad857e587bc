"""Builds tests/fakemex/libfakemex.so (the MEX test host)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def build():
    out = os.path.join(HERE, "libfakemex.so")
    # ASDPOCS_SANITIZE=1: AddressSanitizer and UndefinedBehaviorSanitizer
    san = (["-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=undefined"]
           if os.environ.get("ASDPOCS_SANITIZE") == "1" else [])
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", *san,
                    f"-I{os.path.join(ROOT, 'include')}", os.path.join(HERE, "fakemex.cpp"),
                    "-o", out + ".tmp"], check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
