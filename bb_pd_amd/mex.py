"""Python mirror of the reference's MEX entry points, on device arrays.

Each function has the name and argument list of a reference MEX file (solvers/functions/mex/*.c) or
MATLAB hot-path function, modifies its array arguments in place like the MEX does, and raises the
reference's error text on a wrong argument count.  Arrays are torch tensors on a HIP device holding
FP64 data in MATLAB's column-major layout: an m x n MATLAB array is a tensor of shape (m, n) with
strides (1, m) (see `to_device`), an m-vector a 1-D tensor.  Every call goes through libbbpd_hip.so;
there is no other code path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def to_device(a: np.ndarray, device="cuda"):
    """Copy a numpy array (any order) to a column-major FP64 device tensor of the same shape."""
    torch = _torch()
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        return torch.from_numpy(np.ascontiguousarray(a)).to(device)
    m, n = a.shape
    t = torch.from_numpy(np.ascontiguousarray(a.T)).to(device)  # (n, m) row-major == (m, n) column-major
    return t.T


def to_host(t) -> np.ndarray:
    """Device tensor back to numpy in the same (column-major) layout."""
    if t.dim() == 1:
        return t.detach().cpu().numpy()
    return np.asfortranarray(t.T.detach().cpu().numpy().T)


def _ptr(t):
    if t is None:
        return None
    if t.dtype != _torch().float64:
        raise TypeError("MEX arrays are double precision")
    if not t.is_cuda:
        raise TypeError("arrays must live on the HIP device")
    if t.dim() == 2 and t.shape[1] > 1 and t.stride() != (1, t.shape[0]):
        raise ValueError("2-D arrays must be column-major (use to_device)")
    if t.dim() == 1 and t.stride(0) != 1 and t.numel() > 1:
        raise ValueError("vectors must be contiguous")
    return C.c_void_p(t.data_ptr())


def _rows(t) -> int:
    return int(t.shape[0])


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _nargs(args, n, word):
    if len(args) != n:
        raise _lib.BBPDError(f"{word} input arguments required")


def calculatedeformedlengthMex(*args):
    """[] = calculatedeformedlengthMex(deformedCoordinates, BONDLIST, UNDEFORMEDLENGTH, deformedLength,
    deformedX, deformedY, deformedZ, stretch) -- calculatedeformedlengthMex.c:50-79"""
    _nargs(args, 8, "Eight")
    x, BL, L0, l, dX, dY, dZ, s = args
    _lib.check(_lib.lib().bbpd_calculatedeformedlength(_ptr(x), _ptr(BL), _ptr(L0), _rows(BL), _rows(x), _ptr(l),
                                                       _ptr(dX), _ptr(dY), _ptr(dZ), _ptr(s), _stream()),
               "calculatedeformedlengthMex")


def calculatebsfexponentialMex(*args):
    """[] = calculatebsfexponentialMex(stretch, s0, sc, k, alpha, bsf, bondSofteningFactor, BONDTYPE,
    flagBondSoftening) -- calculatebsfexponentialMex.c:69-96"""
    _nargs(args, 9, "Nine")
    s, s0, sc, k, alpha, bsf, H, BT, flag = args
    _lib.check(_lib.lib().bbpd_calculatebsfexponential(_ptr(s), _ptr(s0), _ptr(sc), float(k), float(alpha), _ptr(bsf),
                                                       _ptr(H), _ptr(BT), _ptr(flag), _rows(s), _stream()),
               "calculatebsfexponentialMex")


def calculatebondforcesMex(*args):
    """[] = calculatebondforcesMex(bForceX, bForceY, bForceZ, fail, deformedX, deformedY, deformedZ,
    deformedLength, stretch, stretchPlastic, BONDSTIFFNESS, cellVolume, VOLUMECORRECTIONFACTORS,
    bondSofteningFactor) -- calculatebondforcesMex.c:48-80"""
    _nargs(args, 14, "Fourteen")
    fX, fY, fZ, fail, dX, dY, dZ, l, s, sp, c, V, vcf, H = args
    _lib.check(_lib.lib().bbpd_calculatebondforces(_ptr(fX), _ptr(fY), _ptr(fZ), _ptr(fail), _ptr(dX), _ptr(dY),
                                                   _ptr(dZ), _ptr(l), _ptr(s), _ptr(sp), _ptr(c), float(V), _ptr(vcf),
                                                   _ptr(H), _rows(fX), _stream()), "calculatebondforcesMex")


def calculatenodalforcesopenmpMex(*args):
    """[] = calculatenodalforcesopenmpMex(BONDLIST, nodalForce, bForceX, bForceY, bForceZ, nodalForceX,
    nodalForceY, nodalForceZ) -- calculatenodalforcesopenmpMex.c:88-116.  The scratch arrays are
    nNodes x T; T = 8, the thread count that file forces (:39), unless they are narrower."""
    _nargs(args, 8, "Eight")
    BL, F, fX, fY, fZ, sX, sY, sZ = args
    T = int(sX.shape[1]) if sX.dim() == 2 else 1
    T = min(T, 8)
    _lib.check(_lib.lib().bbpd_calculatenodalforces(_rows(BL), _rows(F), _ptr(BL), _ptr(F), _ptr(fX), _ptr(fY),
                                                    _ptr(fZ), _ptr(sX), _ptr(sY), _ptr(sZ), T, _stream()),
               "calculatenodalforcesopenmpMex")


def calculatenodalforcesMex(*args):
    """[] = calculatenodalforcesMex(BONDLIST, nodalForce, bForceX, bForceY, bForceZ)
    -- calculatenodalforcesMex.c:49-73 (serial order)"""
    _nargs(args, 5, "Five")
    BL, F, fX, fY, fZ = args
    _lib.check(_lib.lib().bbpd_calculatenodalforces(_rows(BL), _rows(F), _ptr(BL), _ptr(F), _ptr(fX), _ptr(fY),
                                                    _ptr(fZ), None, None, None, 1, _stream()),
               "calculatenodalforcesMex")


def timeintegrationeulercromerMex(*args):
    """[] = timeintegrationeulercromerMex(nodalAcceleration, nodalForce, DAMPING, nodalVelocity, DENSITY,
    constraintFLAG, DT, nodalDisplacement, deformedCoordinates, undeformedCoordinates)
    -- timeintegrationeulercromerMex.c:56-85"""
    _nargs(args, 10, "Ten")
    a, F, D, v, rho, cons, DT, u, x, X = args
    nod = int(a.shape[1]) if a.dim() == 2 else 1
    _lib.check(_lib.lib().bbpd_timeintegrationeulercromer(_ptr(a), _ptr(F), float(D), _ptr(v), _ptr(rho), _ptr(cons),
                                                          float(DT), _ptr(u), _ptr(x), _ptr(X), _rows(a), nod,
                                                          _stream()), "timeintegrationeulercromerMex")


def calculatebondfailureMex(*args):
    """[] = calculatebondfailureMex(fail, failureFunctionality, BONDTYPE, stretch,
    criticalStretchConcrete, criticalStretchSteel) -- calculatebondfailureMex.c:47-71"""
    _nargs(args, 6, "Six")
    fail, ff, BT, s, csc, css = args
    _lib.check(_lib.lib().bbpd_calculatebondfailure(_ptr(fail), float(ff), _ptr(BT), _ptr(s), float(csc), float(css),
                                                    _rows(fail), _stream()), "calculatebondfailureMex")


def _scalar_or_array(v, like):
    torch = _torch()
    if isinstance(v, (int, float)):
        return torch.full((1,), float(v), dtype=torch.float64, device=like.device), 0
    return v, (0 if v.numel() == 1 else 1)


def calculatebondsofteningfactor(stretch, linearElasticLimit, criticalStretchConcrete, flagBondSoftening,
                                 bondSofteningFactor, BONDTYPE):
    """[bondSofteningFactor, flagBondSoftening] = calculatebondsofteningfactor(...) --
    calculatebondsofteningfactor.m:1-48; updates the two arrays in place and returns them."""
    s0, s0s = _scalar_or_array(linearElasticLimit, stretch)
    sc, scs = _scalar_or_array(criticalStretchConcrete, stretch)
    _lib.check(_lib.lib().bbpd_calculatebondsofteningfactor(_ptr(stretch), _ptr(s0), s0s, _ptr(sc), scs,
                                                            _ptr(flagBondSoftening), _ptr(bondSofteningFactor),
                                                            _ptr(BONDTYPE), _rows(stretch), _stream()),
               "calculatebondsofteningfactor")
    return bondSofteningFactor, flagBondSoftening


def calculatebsftrilinear(stretch, s0, s1, sc, bondSofteningFactor, BONDTYPE, flagBondSoftening):
    """[bondSofteningFactor, flagBondSoftening] = calculatebsftrilinear(...) -- calculatebsftrilinear.m:1-112"""
    _lib.check(_lib.lib().bbpd_calculatebsftrilinear(_ptr(stretch), float(s0), float(s1), float(sc),
                                                     _ptr(bondSofteningFactor), _ptr(BONDTYPE),
                                                     _ptr(flagBondSoftening), _rows(stretch), _stream()),
               "calculatebsftrilinear")
    return bondSofteningFactor, flagBondSoftening


def calculatedampingcoefficient(nodalForce, massVector, nodalForcePrevious, DT, nodalVelocityPreviousHalf,
                                nodalDisplacement) -> float:
    """[cn] = calculatedampingcoefficient(...) -- calculatedampingcoefficient.m:1-112 (NOD = 3)"""
    cn = C.c_double()
    _lib.check(_lib.lib().bbpd_calculatedampingcoefficient(_ptr(nodalForce), _ptr(massVector),
                                                           _ptr(nodalForcePrevious), float(DT),
                                                           _ptr(nodalVelocityPreviousHalf), _ptr(nodalDisplacement),
                                                           _rows(nodalForce), C.byref(cn), _stream()),
               "calculatedampingcoefficient")
    return cn.value


def timeintegrationADR(iTimeStep, DT, nodalVelocityForwardHalf, nodalForce, massVector, cn,
                       nodalVelocityPreviousHalf, nodalDisplacement, CONSTRAINTFLAG):
    """[nodalDisplacement, nodalVelocityForwardHalf] = timeintegrationADR(...) -- timeintegrationADR.m:1-80"""
    _lib.check(_lib.lib().bbpd_timeintegrationADR(int(iTimeStep), float(DT), _ptr(nodalVelocityForwardHalf),
                                                  _ptr(nodalForce), _ptr(massVector), float(cn),
                                                  _ptr(nodalVelocityPreviousHalf), _ptr(nodalDisplacement),
                                                  _ptr(CONSTRAINTFLAG), _rows(nodalForce), _stream()),
               "timeintegrationADR")
    return nodalDisplacement, nodalVelocityForwardHalf


def calculatecontactforce(penetrator: dict, displacementIncrement, undeformedCoordinates, deformedCoordinates,
                          nodalDisplacement, nodalVelocity, DT, cellVolume, DENSITY):
    """[nodalDisplacement, nodalVelocity, deformedCoordinates, penetratorfz] = calculatecontactforce(...)
    -- calculatecontactforce.m:1-121.  penetrator: {'family': device tensor of 1-based ids,
    'centre': (cx, cz), 'radius': R}."""
    fz = C.c_double()
    fam = penetrator["family"]
    cx, cz = penetrator["centre"]
    _lib.check(_lib.lib().bbpd_calculatecontactforce(_ptr(fam), _rows(fam), float(cx), float(cz),
                                                     float(penetrator["radius"]), float(displacementIncrement),
                                                     _ptr(undeformedCoordinates), _ptr(deformedCoordinates),
                                                     _ptr(nodalDisplacement), _ptr(nodalVelocity), float(DT),
                                                     float(cellVolume), _ptr(DENSITY), _rows(undeformedCoordinates),
                                                     C.byref(fz), _stream()), "calculatecontactforce")
    return nodalDisplacement, nodalVelocity, deformedCoordinates, fz.value
