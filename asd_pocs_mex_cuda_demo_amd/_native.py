"""Loader of libasdpocs.so (the HIP library behind include/asdpocs.h).

The library must have been built (``python -m asd_pocs_mex_cuda_demo_amd.build``);
there is no fallback path: a missing or stale library raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libasdpocs.so")

ASD_OK = 0
ASD_ERR_ARG = 1
ASD_ERR_HIP = 2
ASD_ERR_ALLOC = 3
ASD_ERR_STATE = 4

# Every symbol declared in include/asdpocs.h.
EXPORTS = [
    "Ax_cone_mf_gpu_new", "Atx_cone_mf_gpu_new", "asd_last_status", "asd_last_error",
    "asd_ax_device", "asd_atx_device", "asd_ax_slab_device", "asd_atx_slab_device",
    "asd_atx_device_noalloc", "asd_tv_min_2d", "asd_tv_min_3d", "asd_tv_min_device",
    "asd_para_defaults", "asd_pocs", "asd_sart", "asd_plan_create", "asd_plan_set_projections",
    "asd_plan_reset", "asd_plan_iterate", "asd_plan_get_image", "asd_plan_get_out",
    "asd_plan_sweep", "asd_plan_get_current", "asd_plan_sync", "asd_plan_set_timing",
    "asd_plan_get_timing", "asd_plan_set_graph", "asd_plan_tuning", "asd_plan_set_atx_variant",
    "asd_plan_set_ax_shape",
    "asd_plan_info", "asd_plan_destroy", "asd_plan_simulate", "asd_plan_get_projections",
    "asd_plan_get_reference",
    "asd_batch_create", "asd_batch_set_case", "asd_batch_iterate", "asd_batch_get_image",
    "asd_batch_get_out", "asd_batch_set_graph", "asd_batch_tuning", "asd_batch_set_atx_chunk",
    "asd_batch_sync",
    "asd_batch_destroy",
]


class AsdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"asdpocs error {code}: {msg}")
        self.code = code


class AsdPara(C.Structure):
    """struct asd_para (include/asdpocs.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("SO", C.c_double), ("OD", C.c_double), ("vxlsize", C.c_double),
        ("detsize", C.c_double), ("na", C.c_int), ("nb", C.c_int),
        ("maxnumIter", C.c_int), ("numOfgradients", C.c_int),
        ("alpha", C.c_double), ("alpha_red", C.c_double), ("beta", C.c_double),
        ("beta_red", C.c_double), ("r_max", C.c_double), ("epsilon", C.c_double),
        ("stopRelChg", C.c_double), ("initBySART", C.c_int), ("lambda_", C.c_double),
        ("precision", C.c_int),
    ]


class AsdOut(C.Structure):
    _fields_ = [("IterationNum", C.c_int), ("dataTolerance", C.c_double),
                ("beta", C.c_double), ("alpha", C.c_double)]


_lib = None


def lib():
    """Load libasdpocs.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m asd_pocs_mex_cuda_demo_amd.build`")
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    f, i, vp, d = C.c_float, C.c_int, C.c_void_p, C.c_double
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    L.Ax_cone_mf_gpu_new.restype = None
    L.Ax_cone_mf_gpu_new.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, i, i, i, i, i]
    L.Atx_cone_mf_gpu_new.restype = None
    L.Atx_cone_mf_gpu_new.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, f, f, i, i, i, i, i]
    L.asd_last_status.restype = i
    L.asd_last_status.argtypes = []
    L.asd_last_error.restype = C.c_char_p
    L.asd_last_error.argtypes = []
    L.asd_ax_device.restype = i
    L.asd_ax_device.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, i, i, i, i, i, vp]
    L.asd_atx_device.restype = i
    L.asd_atx_device.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, f, f, i, i, i, i, i, vp]
    L.asd_ax_slab_device.restype = i
    L.asd_ax_slab_device.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, i, i, i, i, i, i, i,
                                     vp]
    L.asd_atx_slab_device.restype = i
    L.asd_atx_slab_device.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, f, f, i, i, i, i, i,
                                      i, i, vp]
    L.asd_atx_device_noalloc.restype = i
    L.asd_atx_device_noalloc.argtypes = [vp, vp, vp, f, f, vp, vp, f, f, f, f, f, f, i, i, i, i,
                                         i, vp]
    L.asd_tv_min_2d.restype = i
    L.asd_tv_min_2d.argtypes = [vp, i, i, d, i, i, vp, vp]
    L.asd_tv_min_3d.restype = i
    L.asd_tv_min_3d.argtypes = [vp, i, i, i, d, i, i, vp, vp]
    L.asd_tv_min_device.restype = i
    L.asd_tv_min_device.argtypes = [vp, vp, i, i, i, i, d, i, i, vp]
    L.asd_para_defaults.restype = None
    L.asd_para_defaults.argtypes = [C.POINTER(AsdPara)]
    L.asd_pocs.restype = i
    L.asd_pocs.argtypes = [vp, vp, i, C.POINTER(AsdPara), vp, vp, vp, vp, C.POINTER(AsdOut)]
    L.asd_sart.restype = i
    L.asd_sart.argtypes = [vp, vp, i, C.POINTER(AsdPara), vp, dp]
    L.asd_plan_create.restype = i
    L.asd_plan_create.argtypes = [C.POINTER(AsdPara), vp, i, C.POINTER(vp)]
    for name, args in [("asd_plan_set_projections", [vp, vp]), ("asd_plan_reset", [vp, vp, vp]),
                       ("asd_plan_iterate", [vp, i]), ("asd_plan_get_image", [vp, vp]),
                       ("asd_plan_get_out", [vp, C.POINTER(AsdOut), vp]),
                       ("asd_plan_sweep", [vp, d, dp]), ("asd_plan_get_current", [vp, vp]),
                       ("asd_plan_sync", [vp]), ("asd_plan_set_timing", [vp, i]),
                       ("asd_plan_set_graph", [vp, i]),
                       ("asd_plan_tuning", [vp, C.POINTER(C.c_int), vp]),
                       ("asd_plan_set_atx_variant", [vp, i]),
                       ("asd_plan_set_ax_shape", [vp, i, i]),
                       ("asd_plan_simulate", [vp, vp, i, d, d, C.c_ulonglong]),
                       ("asd_plan_get_projections", [vp, vp]),
                       ("asd_plan_get_reference", [vp, vp]),
                       ("asd_plan_get_timing", [vp, i, dp, C.POINTER(C.c_long)]),
                       ("asd_plan_info", [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(vp)])]:
        fn = getattr(L, name)
        fn.restype = i
        fn.argtypes = args
    L.asd_plan_destroy.restype = None
    L.asd_plan_destroy.argtypes = [vp]
    L.asd_batch_create.restype = i
    L.asd_batch_create.argtypes = [i, vp, vp, vp, C.POINTER(vp)]
    for name, args in [("asd_batch_set_case", [vp, i, vp, vp, vp]), ("asd_batch_iterate", [vp, i]),
                       ("asd_batch_get_image", [vp, i, vp]),
                       ("asd_batch_get_out", [vp, i, C.POINTER(AsdOut), vp]),
                       ("asd_batch_set_graph", [vp, i]), ("asd_batch_sync", [vp]),
                       ("asd_batch_tuning", [vp, C.POINTER(i), vp]),
                       ("asd_batch_set_atx_chunk", [vp, i])]:
        fn = getattr(L, name)
        fn.restype = i
        fn.argtypes = args
    L.asd_batch_destroy.restype = None
    L.asd_batch_destroy.argtypes = [vp]
    _ = (fp,)
    _lib = L
    return L


def check(rc):
    """Raise AsdError for a non-zero status, with the library's message."""
    if rc != ASD_OK:
        raise AsdError(rc, lib().asd_last_error().decode(errors="replace"))
    return rc


def check_last():
    """Status of the last void drop-in call (Ax/Atx_cone_mf_gpu_new)."""
    return check(lib().asd_last_status())
