"""Python driver of the MEX test host (libfakemex.so): builds mxArrays from
numpy values, calls a gateway's mexFunction, and converts the outputs back."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
MEXDIR = os.path.join(ROOT, "asd_pocs_mex_cuda_demo_amd", "mex")

mxDOUBLE, mxSINGLE, mxSTRUCT, mxINT32 = 6, 7, 2, 12
_fm = None
_mex = {}


def fm():
    global _fm
    if _fm is None:
        path = os.path.join(HERE, "libfakemex.so")
        if not os.path.exists(path):
            from . import build as b  # noqa
        _fm = C.CDLL(path, mode=C.RTLD_GLOBAL)
        vp = C.c_void_p
        _fm.fm_numeric.restype = vp
        _fm.fm_numeric.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_longlong), vp]
        _fm.fm_struct.restype = vp
        _fm.fm_struct_set.argtypes = [vp, C.c_char_p, vp]
        _fm.fm_free.argtypes = [vp]
        _fm.fm_call.restype = C.c_int
        _fm.fm_call.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_char_p,
                                C.c_int]
        _fm.fm_output.restype = C.c_char_p
        _fm.fm_eval_at.restype = C.c_char_p
        _fm.fm_eval_at.argtypes = [C.c_int]
        _fm.mxGetClassID.restype = C.c_int
        _fm.mxGetClassID.argtypes = [vp]
        _fm.mxGetNumberOfDimensions.restype = C.c_size_t
        _fm.mxGetNumberOfDimensions.argtypes = [vp]
        _fm.mxGetDimensions.restype = C.POINTER(C.c_size_t)
        _fm.mxGetDimensions.argtypes = [vp]
        _fm.mxGetData.restype = vp
        _fm.mxGetData.argtypes = [vp]
        _fm.mxGetField.restype = vp
        _fm.mxGetField.argtypes = [vp, C.c_size_t, C.c_char_p]
    return _fm


def load(name):
    fm()
    if name not in _mex:
        _mex[name] = C.CDLL(os.path.join(MEXDIR, name + ".mexa64"))
    return _mex[name]


def to_mx(v):
    """numpy array / scalar / dict -> mxArray* (column-major like MATLAB)."""
    F = fm()
    if isinstance(v, dict):
        s = F.fm_struct()
        for k, x in v.items():
            F.fm_struct_set(s, k.encode(), to_mx(x))
        return s
    a = np.asarray(v)
    if a.dtype == np.float32:
        cls = mxSINGLE
    elif a.dtype == np.int32:
        cls = mxINT32
    else:
        a = a.astype(np.float64)
        cls = mxDOUBLE
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    dims = (C.c_longlong * a.ndim)(*a.shape)
    buf = np.ascontiguousarray(a.ravel(order="F"))
    return F.fm_numeric(cls, a.ndim, dims, buf.ctypes.data_as(C.c_void_p))


def from_mx(p):
    F = fm()
    cls = F.mxGetClassID(p)
    if cls == mxSTRUCT:
        return p  # callers read fields with field()
    nd = F.mxGetNumberOfDimensions(p)
    dims = tuple(F.mxGetDimensions(p)[i] for i in range(nd))
    dt = {mxDOUBLE: np.float64, mxSINGLE: np.float32, mxINT32: np.int32}[cls]
    n = int(np.prod(dims))
    if n == 0:
        return np.zeros(dims, dt)
    raw = np.ctypeslib.as_array(C.cast(F.mxGetData(p), C.POINTER(C.c_char)), (n * np.dtype(dt).itemsize,))
    return np.frombuffer(raw.tobytes(), dt).reshape(dims, order="F")


def field(p, name):
    f = fm().mxGetField(p, 0, name.encode())
    return None if not f else from_mx(f)


class MexCallError(RuntimeError):
    pass


def call(name, nlhs, *args):
    """Call gateway `name` with numpy/dict arguments.  Returns (outputs,
    printed text); raises MexCallError with the mexErrMsgTxt message."""
    F = fm()
    mex = load(name)
    F.fm_clear()
    prhs = (C.c_void_p * max(len(args), 1))(*[to_mx(a) for a in args])
    plhs = (C.c_void_p * max(nlhs, 1))()
    err = C.create_string_buffer(1024)
    fn = C.cast(mex.mexFunction, C.c_void_p)
    rc = F.fm_call(fn, nlhs, plhs, len(args), prhs, err, 1024)
    text = F.fm_output().decode()
    for i in range(len(args)):
        F.fm_free(prhs[i])
    if rc != 0:
        for i in range(nlhs):
            if plhs[i]:
                F.fm_free(plhs[i])
        raise MexCallError(err.value.decode())
    outs = []
    for i in range(nlhs):
        outs.append(None if not plhs[i] else plhs[i])
    return outs, text


def outputs(outs):
    """Convert numeric outputs to numpy (structs stay handles)."""
    return [None if o is None else from_mx(o) for o in outs]
