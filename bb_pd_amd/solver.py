"""Host wrapper of the fused device-resident solver (bbpd_solver_* in include/bbpd.h).

A `Solver` is built from a BeamCase (reference-layout input arrays) and runs whole explicit steps on
one GPU.  The step sequence restates solvers/scripts/dynamic/dynamicsolverdisplacementcontrolled.m:71-121
(Euler-Cromer) or its ADR variant (calculatedampingcoefficient.m + timeintegrationADR.m).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .beam import BeamCase, to_problem


class Solver:
    def __init__(self, case: BeamCase, precision: int = _lib.FP64, device: int = 0):
        L = _lib.lib()
        if L.bbpd_device_count() <= device:
            raise _lib.BBPDError("no HIP device: the fused solver runs on the GPU only")
        self.case = case
        self.precision = precision
        keep: list = []
        prob = to_problem(case, keep)
        h = C.c_void_p()
        _lib.check(L.bbpd_solver_create(C.byref(prob), precision, device, C.byref(h)), "bbpd_solver_create")
        self._h = h
        self._L = L

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

    # ---- running
    def step(self, n: int = 1):
        _lib.check(self._L.bbpd_solver_step(self._h, n), "bbpd_solver_step")

    def run_graph(self, n: int, reps: int = 1):
        _lib.check(self._L.bbpd_solver_run_graph(self._h, n, reps), "bbpd_solver_run_graph")

    def sync(self):
        _lib.check(self._L.bbpd_solver_sync(self._h), "bbpd_solver_sync")

    @property
    def stream(self) -> int:
        return self._L.bbpd_solver_stream(self._h) or 0

    # ---- results
    def observables(self) -> dict:
        o = _lib.Observables()
        _lib.check(self._L.bbpd_solver_observables(self._h, C.byref(o)), "bbpd_solver_observables")
        return {f: getattr(o, f) for f, _ in o._fields_}

    def nodal(self):
        """(u, v, F, x) in the reference layout: (nNodes, 3) Fortran-order float64."""
        nN = self.case.nNodes
        arrs = [np.zeros((nN, 3), order="F") for _ in range(4)]
        p = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        _lib.check(self._L.bbpd_solver_get_nodal(self._h, *p), "bbpd_solver_get_nodal")
        return tuple(arrs)

    def bonds(self):
        """(bondSofteningFactor, flagBondSoftening, fail) per bond in BONDLIST order."""
        nB = self.case.nBonds
        arrs = [np.zeros(nB) for _ in range(3)]
        p = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        _lib.check(self._L.bbpd_solver_get_bonds(self._h, *p), "bbpd_solver_get_bonds")
        return tuple(arrs)

    def load_state(self):
        """(current load increment, iterations done at it)."""
        inc, it = C.c_int64(), C.c_int64()
        _lib.check(self._L.bbpd_solver_get_load_state(self._h, C.byref(inc), C.byref(it)),
                   "bbpd_solver_get_load_state")
        return inc.value, it.value

    def set_load_state(self, inc: int, iters: int):
        _lib.check(self._L.bbpd_solver_set_load_state(self._h, int(inc), int(iters)), "bbpd_solver_set_load_state")

    def checkpoint(self) -> dict:
        """Device state in the reference layout (see bbpd_solver_checkpoint); under the equilibrium
        staircase also `load_state` (increment, iterations done at it)."""
        nN, nB = self.case.nNodes, self.case.nBonds
        st = {k: np.zeros((nN, 3), order="F") for k in ("u", "v", "x", "F", "Fp")}
        st.update({k: np.zeros(nB) for k in ("H", "flag", "fail")})
        step = C.c_int64()
        p = [st[k].ctypes.data_as(C.c_void_p) for k in ("u", "v", "x", "F", "Fp", "H", "flag", "fail")]
        _lib.check(self._L.bbpd_solver_checkpoint(self._h, C.byref(step), *p), "bbpd_solver_checkpoint")
        st["step"] = step.value
        if self.case.schedule == 2:
            st["load_state"] = self.load_state()
        return st

    def restore(self, step: int, load_state=None, **arrays):
        """Load state (keys u, v, x, F, Fp, H, flag, fail; missing keys stay unchanged); under the
        equilibrium staircase `load_state` = (increment, iterations done at it)."""
        keep = []

        def ptr(k):
            a = arrays.get(k)
            if a is None:
                return None
            a = np.asfortranarray(a, dtype=np.float64) if np.ndim(a) == 2 else np.ascontiguousarray(a, np.float64)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        p = [ptr(k) for k in ("u", "v", "x", "F", "Fp", "H", "flag", "fail")]
        _lib.check(self._L.bbpd_solver_restore(self._h, int(step), *p), "bbpd_solver_restore")
        if load_state is not None:
            self.set_load_state(*load_state)

    def history(self, cap: int = 1 << 20) -> dict:
        """Per-step record since the previous call: arrays `reaction` (contact reaction of body 0),
        `reaction_static` (sum of F_z * cellVolume over body 0's family), `deflection`, `cmod`,
        `disp` (load-point displacement), `ratio` (equilibrium ratio), `cn` (ADR damping coefficient)
        and `inc_end` (True where the step ended a load increment)."""
        rows = np.zeros((cap, _lib.HISTORY_COLUMNS))
        n = C.c_int()
        _lib.check(self._L.bbpd_solver_history(self._h, rows.ctypes.data_as(C.c_void_p), cap, C.byref(n)),
                   "bbpd_solver_history")
        r = rows[:n.value]
        return _lib.history_dict(r)

    def damage(self, energy: bool = True):
        """(damage field per node in the reference order, total damage energy or None)."""
        d = np.zeros(self.case.nNodes)
        e = C.c_double()
        _lib.check(self._L.bbpd_solver_damage(self._h, d.ctypes.data_as(C.c_void_p), C.byref(e) if energy else None),
                   "bbpd_solver_damage")
        return d, (e.value if energy else None)

    def strain_energy(self):
        """(strain energy density per node in the reference order, total strain energy)
        (calculatestrainenergy.m:10-27)."""
        d = np.zeros(self.case.nNodes)
        e = C.c_double()
        _lib.check(self._L.bbpd_solver_strain_energy(self._h, d.ctypes.data_as(C.c_void_p), C.byref(e)),
                   "bbpd_solver_strain_energy")
        return d, e.value

    def layout(self) -> str:
        """"stencil" or "records (<reason>)": the bond-table layout chosen at set-up."""
        return self._L.bbpd_solver_layout(self._h).decode()

    def autotune(self, reps: int = 5):
        """Time the bond-kernel variants on the current state (ADR only), keep the fastest; returns
        (index kept, average ms per variant)."""
        ms = (C.c_double * _lib.AUTOTUNE_VARIANTS)()
        ch = C.c_int()
        _lib.check(self._L.bbpd_solver_autotune(self._h, reps, ms, C.byref(ch)), "bbpd_solver_autotune")
        return ch.value, list(ms)

    def info(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        _lib.check(self._L.bbpd_solver_info(self._h, *[C.byref(x) for x in v]), "bbpd_solver_info")
        return dict(device_bytes=v[0].value, bond_kernel_bytes=v[1].value, n_entries=v[2].value,
                    n_padded_entries=v[3].value)

    def bond_bytes(self) -> dict:
        """Algorithmic bytes of one bond-kernel launch at the current state (bbpd_solver_bond_bytes)."""
        v = [C.c_int64() for _ in range(3)]
        _lib.check(self._L.bbpd_solver_bond_bytes(self._h, *[C.byref(x) for x in v]), "bbpd_solver_bond_bytes")
        return dict(bytes=v[0].value, offsets_set=v[1].value, offsets_total=v[2].value)

    def set_timing(self, on: bool):
        _lib.check(self._L.bbpd_solver_set_timing(self._h, 1 if on else 0), "bbpd_solver_set_timing")

    def kernel_time(self):
        ms = C.c_double()
        n = C.c_int64()
        _lib.check(self._L.bbpd_solver_kernel_time(self._h, C.byref(ms), C.byref(n)), "bbpd_solver_kernel_time")
        return ms.value, n.value
