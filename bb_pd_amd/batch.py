"""Host wrapper of batches (bbpd_batch_* in include/bbpd.h): many beam problems stepped together in one
launch of each kernel, admitted and retired while the others run.

The reference has no batch driver; its only multi-specimen workflow runs specimens one after another
(solvers/scripts/randomfield/dynamicsolverdisplacementcontrolledRF.m:20-114).  Every problem of a
batch shares the lattice spacing, the material law's form and the integrator; geometry, notch,
stiffness, per-bond parameters, contact bodies and load schedule are its own.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .beam import BeamCase, to_problem


class BatchFull(_lib.BBPDError):
    """No region or slot range is free for the problem now."""


class BatchTooSmall(_lib.BBPDError):
    """The problem needs more node slots than the whole batch holds: it can never be admitted."""


class Admission:
    """A problem checked and laid out for a batch on the host (bbpd_batch_prepare): admitting it then
    only finds a region and slots and uploads.  Holds copies of what it needs: the case's arrays may be
    dropped once it exists."""

    def __init__(self, batch: "Batch", handle, nNodes: int, nBonds: int, name: str):
        self._batch, self._h = batch, handle
        self.nNodes, self.nBonds, self.name = nNodes, nBonds, name

    def free(self):
        if getattr(self, "_h", None):
            self._batch._L.bbpd_admission_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    def __init__(self, shape: BeamCase, precision: int = _lib.MIXED, max_regions: int = 64,
                 max_slots: int = 4 << 20, history_rows: int = 65536, device: int = 0):
        L = _lib.lib()
        if L.bbpd_device_count() <= device:
            raise _lib.BBPDError("no HIP device: batches run on the GPU only")
        self._L = L
        self.precision = precision
        self.max_regions = max_regions
        self._keep = []
        self._shape = to_problem(shape, self._keep)
        spec = _lib.BatchSpec(max_regions, max_slots, history_rows, shape.DX, C.pointer(self._shape))
        h = C.c_void_p()
        _lib.check(L.bbpd_batch_create(C.byref(spec), precision, device, C.byref(h)), "bbpd_batch_create")
        self._h = h
        self.cases: dict[int, BeamCase] = {}
        self.sizes: dict[int, tuple] = {}  # region -> (nNodes, nBonds)

    def close(self):
        if getattr(self, "_h", None):
            self._L.bbpd_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- membership
    def _admit_rc(self, rc, what):
        if rc == _lib.ERR_FULL:
            raise BatchFull(self._L.bbpd_last_error().decode())
        if rc == _lib.ERR_NOMEM:
            raise BatchTooSmall(self._L.bbpd_last_error().decode())
        _lib.check(rc, what)

    def admit(self, case, stop_fraction: float = 0.0, stop_displacement: float = 0.0) -> int:
        """Admit a case (bbpd_batch_admit), or an Admission prepared for this batch
        (bbpd_batch_admit_prepared; freed once admitted)."""
        r = C.c_int(-1)
        if isinstance(case, Admission):
            rc = self._L.bbpd_batch_admit_prepared(self._h, case._h, float(stop_fraction), float(stop_displacement),
                                                   C.byref(r))
            self._admit_rc(rc, "bbpd_batch_admit_prepared")
            self.sizes[r.value] = (case.nNodes, case.nBonds)
            case.free()
            return r.value
        keep: list = []
        prob = to_problem(case, keep)
        rc = self._L.bbpd_batch_admit(self._h, C.byref(prob), float(stop_fraction), float(stop_displacement),
                                      C.byref(r))
        self._admit_rc(rc, "bbpd_batch_admit")
        self.cases[r.value] = case
        self.sizes[r.value] = (case.nNodes, case.nBonds)
        return r.value

    def prepare(self, case: BeamCase) -> Admission:
        """Check and lay out a case for this batch on the host only (bbpd_batch_prepare): may run on
        another thread while the batch steps.  Raises BatchTooSmall if it can never fit."""
        keep: list = []
        prob = to_problem(case, keep)
        h = C.c_void_p()
        self._admit_rc(self._L.bbpd_batch_prepare(self._h, C.byref(prob), C.byref(h)), "bbpd_batch_prepare")
        return Admission(self, h, case.nNodes, case.nBonds, case.name)

    def retire(self, region: int):
        _lib.check(self._L.bbpd_batch_retire(self._h, region), "bbpd_batch_retire")
        self.cases.pop(region, None)
        self.sizes.pop(region, None)

    # ---- running (every active region advances one step per launch)
    def step(self, n: int = 1):
        _lib.check(self._L.bbpd_solver_step(self._h, n), "bbpd_solver_step")

    def run_graph(self, n: int, reps: int = 1):
        _lib.check(self._L.bbpd_solver_run_graph(self._h, n, reps), "bbpd_solver_run_graph")

    def sync(self):
        _lib.check(self._L.bbpd_solver_sync(self._h), "bbpd_solver_sync")

    def set_timing(self, on: bool):
        _lib.check(self._L.bbpd_solver_set_timing(self._h, 1 if on else 0), "bbpd_solver_set_timing")

    def kernel_time(self):
        ms, n = C.c_double(), C.c_int64()
        _lib.check(self._L.bbpd_solver_kernel_time(self._h, C.byref(ms), C.byref(n)), "bbpd_solver_kernel_time")
        return ms.value, n.value

    # ---- results
    def status(self) -> list:
        st = (_lib.RegionStatus * self.max_regions)()
        _lib.check(self._L.bbpd_batch_status(self._h, st, self.max_regions), "bbpd_batch_status")
        return [{f: getattr(x, f) for f, _ in x._fields_} for x in st]

    def observables(self, region: int) -> dict:
        o = _lib.Observables()
        _lib.check(self._L.bbpd_batch_observables(self._h, region, C.byref(o)), "bbpd_batch_observables")
        return {f: getattr(o, f) for f, _ in o._fields_}

    def history(self, region: int, cap: int = 1 << 20) -> dict:
        rows = np.zeros((cap, _lib.HISTORY_COLUMNS))
        n = C.c_int()
        _lib.check(self._L.bbpd_batch_history(self._h, region, rows.ctypes.data_as(C.c_void_p), cap, C.byref(n)),
                   "bbpd_batch_history")
        r = rows[:n.value]
        return _lib.history_dict(r)

    def nodal(self, region: int):
        nN = self.sizes[region][0]
        arrs = [np.zeros((nN, 3), order="F") for _ in range(4)]
        p = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        _lib.check(self._L.bbpd_batch_get_nodal(self._h, region, *p), "bbpd_batch_get_nodal")
        return tuple(arrs)

    def bonds(self, region: int):
        nB = self.sizes[region][1]
        arrs = [np.zeros(nB) for _ in range(3)]
        p = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        _lib.check(self._L.bbpd_batch_get_bonds(self._h, region, *p), "bbpd_batch_get_bonds")
        return tuple(arrs)

    def autotune(self, reps: int = 5):
        """Time the bond-kernel variants on the current state (ADR only), keep the fastest; returns
        (index kept, average ms per variant)."""
        ms = (C.c_double * _lib.AUTOTUNE_VARIANTS)()
        ch = C.c_int()
        _lib.check(self._L.bbpd_solver_autotune(self._h, reps, ms, C.byref(ch)), "bbpd_solver_autotune")
        return ch.value, list(ms)

    def damage(self, region: int, energy: bool = True):
        """(damage field of the region's problem, total damage energy or None)."""
        d = np.zeros(self.sizes[region][0])
        e = C.c_double()
        _lib.check(self._L.bbpd_batch_damage(self._h, region, d.ctypes.data_as(C.c_void_p),
                                             C.byref(e) if energy else None), "bbpd_batch_damage")
        return d, (e.value if energy else None)

    def strain_energy(self, region: int):
        """(strain energy density per node of the region's problem, total strain energy)."""
        d = np.zeros(self.sizes[region][0])
        e = C.c_double()
        _lib.check(self._L.bbpd_batch_strain_energy(self._h, region, d.ctypes.data_as(C.c_void_p), C.byref(e)),
                   "bbpd_batch_strain_energy")
        return d, e.value

    def info(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        _lib.check(self._L.bbpd_solver_info(self._h, *[C.byref(x) for x in v]), "bbpd_solver_info")
        return dict(device_bytes=v[0].value, bond_kernel_bytes=v[1].value, n_entries=v[2].value,
                    n_padded_entries=v[3].value)
