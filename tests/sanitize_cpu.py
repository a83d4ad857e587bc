"""Runs the CPU test suite with the CPU-side native code built under
AddressSanitizer and UndefinedBehaviorSanitizer: the oracle
(oracle/liboracle.so), the MEX test host (tests/fakemex/libfakemex.so), the
MEX gateways (asd_pocs_mex_cuda_demo_amd/mex/*.mexa64) and the host build of
the projector device functions (tests/cpu_emul/libemul.so; only its host
code is instrumented).  Python itself is not instrumented, so the runtimes
are preloaded; leak detection is off (the interpreter's own allocations
would drown it).  Any sanitizer report aborts the run.  The normal builds
are restored afterwards.

    python tests/sanitize_cpu.py [extra pytest arguments]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build_all(sanitize):
    env = dict(os.environ, ASDPOCS_SANITIZE="1" if sanitize else "0")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from oracle import oracle; oracle.build()\n"
            "from asd_pocs_mex_cuda_demo_amd import build as b; b.build(gateways_only=True)\n"
            "import runpy\n"
            "runpy.run_path(%r, run_name='__main__')\n"
            "runpy.run_path(%r, run_name='__main__')\n"
            % (ROOT, os.path.join(HERE, "fakemex", "build.py"),
               os.path.join(HERE, "cpu_emul", "build.py")))
    subprocess.run([sys.executable, "-c", code], check=True, env=env)


def runtime(name):
    return subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True,
                          check=True).stdout.strip()


def main(argv):
    build_all(True)
    try:
        env = dict(os.environ,
                   LD_PRELOAD=":".join([runtime("libasan.so"), runtime("libubsan.so")]),
                   ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q",
                            "-m", "not gpu", "-s", "-p", "no:cacheprovider", *argv], env=env,
                           cwd=ROOT)
    finally:
        build_all(False)
    return r.returncode


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
