"""Calls the project's MEX gateways (bb_pd_amd/mex/*Mex.c, built into bb_pd_amd/_build/mex/) the way
MATLAB would, through the project's mex.h stand-in: numpy arrays become mxArrays that wrap their
buffers (column-major, so in-place writes land in the numpy arrays), dicts become 1x1 structs, lists
cell arrays, str char arrays; a mexErrMsgTxt comes back as MexError with its text."""
from __future__ import annotations

import ctypes as C

import numpy as np

from bb_pd_amd import _lib
from bb_pd_amd import build as B


class MexError(RuntimeError):
    pass


_shim = None
_gw = {}


def shim():
    global _shim
    if _shim is None:
        _lib.lib()  # torch's HIP runtime and libbbpd_hip first: the gateways bind to them
        L = C.CDLL(str(B.MEX_OUT / "libmexshim.so"), mode=C.RTLD_GLOBAL)
        V, S = C.c_void_p, C.c_size_t
        for name, res, args in (
                ("mexshim_double", V, [S, S, V]), ("mexshim_string", V, [C.c_char_p]),
                ("mexshim_struct", V, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(V)]),
                ("mexshim_cell", V, [C.c_int, C.POINTER(V)]),
                ("mexshim_call", C.c_int, [V, C.c_int, C.POINTER(V), C.c_int, C.POINTER(V), C.c_char_p, C.c_int]),
                ("mxGetPr", C.POINTER(C.c_double), [V]), ("mxGetData", V, [V]), ("mxGetM", S, [V]),
                ("mxGetN", S, [V]), ("mxGetClassID", C.c_int, [V]), ("mxDestroyArray", None, [V]),
                ("mxCreateNumericMatrix", V, [S, S, C.c_int, C.c_int])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _shim = L
    return _shim


def gateway(name: str):
    if name not in _gw:
        shim()
        lib = C.CDLL(str(B.MEX_OUT / f"{name}.so"))
        _gw[name] = (lib, C.cast(lib.mexFunction, C.c_void_p))
    return _gw[name][1]


class Handle:
    """A uint64 scalar mxArray (bbpdStepMex's handle), kept across calls."""

    def __init__(self, arr):
        self.arr = arr


def _wrap(v, keep):
    L = shim()
    if isinstance(v, Handle):
        return v.arr, False
    if isinstance(v, str):
        return L.mexshim_string(v.encode()), True
    if isinstance(v, dict):
        names = (C.c_char_p * len(v))(*[k.encode() for k in v])
        vals = (C.c_void_p * len(v))(*[_wrap(x, keep)[0] for x in v.values()])
        return L.mexshim_struct(len(v), names, vals), True
    if isinstance(v, list):
        vals = (C.c_void_p * len(v))(*[_wrap(x, keep)[0] for x in v])
        return L.mexshim_cell(len(v), vals), True
    a = v if isinstance(v, np.ndarray) else np.array([float(v)])
    if a.dtype != np.float64:
        raise TypeError("MEX arrays are double")
    if a.ndim == 2 and not a.flags.f_contiguous:
        raise ValueError("2-D arrays must be column-major")
    if a.ndim == 1 and not a.flags.c_contiguous:
        raise ValueError("vectors must be contiguous")
    keep.append(a)
    m, n = (a.shape[0], a.shape[1]) if a.ndim == 2 else (a.shape[0], 1)
    return L.mexshim_double(m, n, a.ctypes.data), True


def _unwrap(p):
    L = shim()
    cls = L.mxGetClassID(p)
    m, n = L.mxGetM(p), L.mxGetN(p)
    if cls == 13:  # uint64: a handle
        return Handle(p)
    a = np.ctypeslib.as_array(L.mxGetPr(p), shape=(n, m)).T.copy(order="F") if m * n else np.zeros((m, n))
    L.mxDestroyArray(p)
    return a[:, 0].copy() if n == 1 else a


def call(name: str, *args, nlhs: int = 0):
    """Run gateway `name` on `args`; returns its nlhs outputs (arrays, a Handle for uint64)."""
    L = shim()
    keep: list = []
    wrapped = [_wrap(a, keep) for a in args]
    prhs = (C.c_void_p * max(len(args), 1))(*[w[0] for w in wrapped])
    plhs = (C.c_void_p * max(nlhs, 1))()
    err = C.create_string_buffer(1024)
    rc = L.mexshim_call(gateway(name), nlhs, plhs, len(args), prhs, err, 1024)
    for (p, own) in wrapped:
        if own:
            L.mxDestroyArray(p)
    if rc:
        raise MexError(err.value.decode())
    outs = [_unwrap(plhs[k]) for k in range(nlhs) if plhs[k]]
    return outs[0] if nlhs == 1 else outs
