"""GPU checks of the arithmetic identities the kernels rely on.

The projectors and the FP32 TV take square roots through sqrt_finite_ge_tiny
and sqrt_ge_tiny (asd_internal.hpp): the hardware's reciprocal square root
and one residual step.  Their results must be the correctly rounded sqrtf
bit for bit, as the reference's sqrtf (Ax_cone_gpu.cu:208, :402) and the
oracle give.  That depends on the hardware's rsq, so it is checked on the
GPU, exhaustively over the inputs the kernels pass them.
"""
import ctypes as C
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "gpu_checks", "libgpuchecks.so")


def _check(lo, hi):
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: build it with tests/gpu_checks/build.py")
    L = C.CDLL(LIB)
    L.check_sqrt_forms.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong),
                                   C.POINTER(C.c_uint32)]
    n = (C.c_ulonglong * 3)()
    f = (C.c_uint32 * 2)()
    assert L.check_sqrt_forms(lo, hi, n, f) == 0
    return list(n), list(f)


@pytest.mark.gpu
def test_square_roots_are_correctly_rounded_on_every_input():
    """Every float in [2^-96, FLT_MAX] (1.88e9 inputs): both forms equal
    sqrtf bit for bit, while the bare hardware sqrt differs on a quarter of
    them (the check sees a wrong root)."""
    n, f = _check(0x0F800000, 0x7F800000)
    assert n[0] == 0 and n[1] == 0, (n, [hex(v) for v in f])
    assert n[2] > 10**8, n


@pytest.mark.gpu
def test_square_root_special_values():
    """+0, -0 and +inf: the form with the fixup returns them as their own
    roots, as sqrtf does."""
    for bits in (0x00000000, 0x80000000, 0x7F800000):
        n, _ = _check(bits, bits + 1)
        assert n[1] == 0, hex(bits)
