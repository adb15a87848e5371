"""ctypes binding of libbbpd_hip.so (the C ABI declared in include/bbpd.h).

The library is loaded from bb_pd_amd/_build/.  There is no fallback: if the library is missing the
import of anything that computes raises, so no result can come from another code path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libbbpd_hip.so"

BBPD_OK = 0
LAW_LINEAR, LAW_BILINEAR, LAW_TRILINEAR, LAW_EXPONENTIAL = 0, 1, 2, 3
EULER_CROMER, ADR = 0, 1
FP64, MIXED = 0, 1
MAX_BODIES = 8
AUTOTUNE_VARIANTS = 6  # BBPD_AUTOTUNE_VARIANTS
HISTORY_COLUMNS = 8  # BBPD_HISTORY_COLUMNS
NOALLOC_MAX_NODES = 16777216  # BBPD_NOALLOC_MAX_NODES

_dp = C.POINTER(C.c_double)


class BeamSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("DX", C.c_double), ("horizon_factor", C.c_double),
        ("notch", C.c_int), ("notch_x", C.c_double), ("notch_depth", C.c_double),
        ("E", C.c_double),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("nNodes", C.c_int), ("nBonds", C.c_int),
        ("undeformedCoordinates", _dp), ("BONDLIST", _dp), ("UNDEFORMEDLENGTH", _dp),
        ("VOLUMECORRECTIONFACTORS", _dp), ("BONDSTIFFNESS", _dp), ("BONDTYPE", _dp), ("fail", _dp),
        ("DENSITY", _dp), ("CONSTRAINTFLAG", _dp), ("massVector", _dp),
        ("cellVolume", C.c_double), ("horizon", C.c_double), ("RADIJ", C.c_double),
        ("law", C.c_int), ("s0", _dp), ("sc", _dp),
        ("s0_scalar", C.c_double), ("s1", C.c_double), ("sc_scalar", C.c_double), ("k", C.c_double),
        ("alpha", C.c_double), ("sc_steel", C.c_double), ("failureFunctionality", C.c_double),
        ("integrator", C.c_int), ("DT", C.c_double), ("DAMPING", C.c_double),
        ("nBodies", C.c_int), ("family", _dp * MAX_BODIES), ("nFamily", C.c_int * MAX_BODIES),
        ("centreX", C.c_double * MAX_BODIES), ("centreZ", C.c_double * MAX_BODIES),
        ("radius", C.c_double * MAX_BODIES), ("moves", C.c_int * MAX_BODIES),
        ("schedule", C.c_int), ("nTimeSteps", C.c_int), ("appliedDisplacement", C.c_double),
        ("increment", C.c_double), ("itersPerIncrement", C.c_int),
        ("referenceNode", C.c_int), ("cmodNode1", C.c_int), ("cmodNode2", C.c_int),
        ("eqTol", C.c_double), ("itersPerIncrementMin", C.c_int),
    ]


class Observables(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("reaction", C.c_double), ("reaction_static", C.c_double),
        ("deflection", C.c_double), ("cmod", C.c_double), ("cn", C.c_double),
        ("n_softening", C.c_int64), ("n_broken", C.c_int64), ("n_failed", C.c_int64),
    ]


class BatchSpec(C.Structure):
    _fields_ = [
        ("maxRegions", C.c_int), ("maxSlots", C.c_int64), ("historyRows", C.c_int), ("DX", C.c_double),
        ("shape", C.POINTER(Problem)),
    ]


class RegionStatus(C.Structure):
    _fields_ = [
        ("active", C.c_int), ("finished", C.c_int), ("step", C.c_int64), ("peak", C.c_double),
        ("reaction_static", C.c_double), ("nNodes", C.c_int64), ("nBonds", C.c_int64),
    ]


ERR_FULL = -5
ERR_NOMEM = -3

HISTORY_KEYS = ("reaction", "reaction_static", "deflection", "cmod", "disp", "ratio", "cn", "inc_end")


def history_dict(rows):
    """Per-step history rows (BBPD_HISTORY_COLUMNS doubles each) as named columns."""
    d = {k: rows[:, i].copy() for i, k in enumerate(HISTORY_KEYS)}
    d["inc_end"] = d["inc_end"] != 0
    return d

# name -> (restype, argtypes); every symbol include/bbpd.h declares
_V, _I, _D, _L, _P = C.c_void_p, C.c_int, C.c_double, C.c_int64, C.c_void_p
SIGNATURES = {
    "bbpd_last_error": (C.c_char_p, []),
    "bbpd_version": (_I, []),
    "bbpd_device_count": (_I, []),
    "bbpd_calculatedeformedlength": (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _V]),
    "bbpd_calculatebsfexponential": (_I, [_P, _P, _P, _D, _D, _P, _P, _P, _P, _I, _V]),
    "bbpd_calculatebondforces": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _I, _V]),
    "bbpd_calculatenodalforces": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _V]),
    "bbpd_timeintegrationeulercromer": (_I, [_P, _P, _D, _P, _P, _P, _D, _P, _P, _P, _I, _I, _V]),
    "bbpd_calculatebondfailure": (_I, [_P, _D, _P, _P, _D, _D, _I, _V]),
    "bbpd_calculatebondsofteningfactor": (_I, [_P, _P, _I, _P, _I, _P, _P, _P, _I, _V]),
    "bbpd_calculatebsftrilinear": (_I, [_P, _D, _D, _D, _P, _P, _P, _I, _V]),
    "bbpd_calculatedampingcoefficient": (_I, [_P, _P, _P, _D, _P, _P, _I, _dp, _V]),
    "bbpd_calculatedampingcoefficient_noalloc": (_I, [_P, _P, _P, _D, _P, _P, _I, _dp, _V]),
    "bbpd_timeintegrationADR": (_I, [_I, _D, _P, _P, _P, _D, _P, _P, _P, _I, _V]),
    "bbpd_calculatecontactforce": (_I, [_P, _I, _D, _D, _D, _D, _P, _P, _P, _P, _D, _D, _P, _I, _dp, _V]),
    "bbpd_beam_count": (_I, [C.POINTER(BeamSpec), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bbpd_beam_build": (_I, [C.POINTER(BeamSpec), C.c_int64, _P, _P, _P, _P, _P, _P]),
    "bbpd_solver_create": (_I, [C.POINTER(Problem), _I, _I, C.POINTER(_V)]),
    "bbpd_solver_destroy": (_I, [_V]),
    "bbpd_solver_step": (_I, [_V, _I]),
    "bbpd_solver_run_graph": (_I, [_V, _I, _I]),
    "bbpd_solver_sync": (_I, [_V]),
    "bbpd_solver_stream": (_V, [_V]),
    "bbpd_solver_observables": (_I, [_V, C.POINTER(Observables)]),
    "bbpd_solver_get_nodal": (_I, [_V, _P, _P, _P, _P]),
    "bbpd_solver_get_bonds": (_I, [_V, _P, _P, _P]),
    "bbpd_solver_checkpoint": (_I, [_V, C.POINTER(_L), _P, _P, _P, _P, _P, _P, _P, _P]),
    "bbpd_solver_restore": (_I, [_V, _L, _P, _P, _P, _P, _P, _P, _P, _P]),
    "bbpd_solver_history": (_I, [_V, _P, _I, C.POINTER(_I)]),
    "bbpd_solver_get_load_state": (_I, [_V, C.POINTER(_L), C.POINTER(_L)]),
    "bbpd_solver_set_load_state": (_I, [_V, _L, _L]),
    "bbpd_solver_info": (_I, [_V, C.POINTER(_L), C.POINTER(_L), C.POINTER(_L), C.POINTER(_L)]),
    "bbpd_solver_layout": (C.c_char_p, [_V]),
    "bbpd_solver_bond_bytes": (_I, [_V, C.POINTER(_L), C.POINTER(_L), C.POINTER(_L)]),
    "bbpd_solver_autotune": (_I, [_V, _I, _P, C.POINTER(_I)]),
    "bbpd_solver_damage": (_I, [_V, _P, C.POINTER(_D)]),
    "bbpd_batch_damage": (_I, [_V, _I, _P, C.POINTER(_D)]),
    "bbpd_solver_strain_energy": (_I, [_V, _P, C.POINTER(_D)]),
    "bbpd_batch_strain_energy": (_I, [_V, _I, _P, C.POINTER(_D)]),
    "bbpd_batch_create": (_I, [_P, _I, _I, C.POINTER(_V)]),
    "bbpd_batch_admit": (_I, [_V, _P, _D, _D, C.POINTER(_I)]),
    "bbpd_batch_prepare": (_I, [_V, _P, C.POINTER(_V)]),
    "bbpd_batch_admit_prepared": (_I, [_V, _V, _D, _D, C.POINTER(_I)]),
    "bbpd_admission_free": (None, [_V]),
    "bbpd_batch_retire": (_I, [_V, _I]),
    "bbpd_batch_status": (_I, [_V, _P, _I]),
    "bbpd_batch_observables": (_I, [_V, _I, _P]),
    "bbpd_batch_history": (_I, [_V, _I, _P, _I, C.POINTER(_I)]),
    "bbpd_batch_get_nodal": (_I, [_V, _I, _P, _P, _P, _P]),
    "bbpd_batch_get_bonds": (_I, [_V, _I, _P, _P, _P]),
    "bbpd_solver_set_timing": (_I, [_V, _I]),
    "bbpd_solver_kernel_time": (_I, [_V, _dp, C.POINTER(_L)]),
}

_lib = None


class BBPDError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libbbpd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise BBPDError(f"{LIB_PATH} is missing: build it with bb_pd_amd.build.build_library()")
        # One HIP runtime per process: torch ships its own libamdhip64 / libhsa-runtime64 and its
        # libraries ask for them by unversioned name, so they would not reuse an already-loaded
        # /opt/rocm copy and a second runtime would find no device.  Loading torch first registers
        # its runtime under the versioned sonames, which libbbpd_hip then binds to.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != BBPD_OK:
        msg = lib().bbpd_last_error()
        raise BBPDError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
