"""Oracle for the bond-based peridynamics step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use this module, as the
checker or as the timed CPU baseline.  The product (bb_pd_amd) never imports it.

Contents
  * ctypes wrapper of oracle/bbpd_oracle.c, a C restatement of the reference's MEX computational
    routines and MATLAB hot-path functions (each C routine cites its reference file:line);
  * a numpy restatement of the input builders (brute-force |xi| <= delta neighbour search,
    determineintersection.m in floating point, calculatevolumecorrectionfactors.m, buildbonddata.m).

The reference C cannot be compiled here: every MEX source includes MATLAB's mex.h, which this image
does not have, so oracle/_ref is not built.  Each routine is pinned instead by checks that a
plausible slip would fail (tests/test_oracle.py):

  routine                        pinned by
  deformed length / stretch      exact hand values on axis-aligned and 3-4-5 bonds; rigid
                                 translation gives s = 0 exactly; uniform dilation gives s = e
  bond force                     hand-computed product for one bond; Newton's third law; NaN -> 0
  nodal reduction (serial, T)    brute-force sums with math.fsum agree; sum of nodal forces = 0;
                                 T-thread order equals serial order when T = 1; explicit 2-bond
                                 order case where the two orders give different bits
  Euler-Cromer                   closed-form single-DOF free fall and harmonic step
  bilinear law                   bsf = 0 at s0, 1 at sc, force (1-bsf)s linear between;
                                 compression-after-softening quirk (bsf -> 1)
  trilinear law                  force (1-bsf)s equals s0 at s0, beta s0 at s1, 0 at sc
                                 (scripts/trilinear.m:37-61 closed forms)
  exponential law                bsf = 0 at s0, 1 at sc; monotone; the fracture energy implied by
                                 the sc formula equals GF by quadrature of the softening curve
  brittle latch                  latch is irreversible and honours BONDTYPE thresholds
  ADR damping                    Rayleigh quotient of a one-DOF spring: cn = 2 sqrt(k/m); clamp
                                 cn > 2 -> 1.9; zero-velocity DOFs skipped
  ADR update                     closed form of the first step and a two-step hand recurrence
  contact                        projection of a node at known depth onto the circle; v_y = 0;
                                 reaction equals the momentum change; centre move not cumulative
  smoothstep                     0 at i = 1, a at i = N, zero slope at both ends, symmetric midpoint
  neighbour list / notch         brute-force O(N^2) search; exact and floating-point notch tests
  damage energy, damage field    closed forms at s0 and sc; all-broken node has damage 1
  strain energy                  hand value on one bond; uniform dilation e of a bulk lattice node
                                 gives the continuum density 3 E e^2 (c = 12E/(pi delta^4))
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "bbpd_oracle.c"
OUT = HERE / "_build"
LIB = OUT / "liboracle.so"


def build() -> Path:
    """Compile the C restatement (gcc, OpenMP, no FMA contraction)."""
    OUT.mkdir(exist_ok=True)
    tmp = OUT / "liboracle.so.tmp"
    cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", str(SRC), "-o",
           str(tmp), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None
_d = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_d) if a is not None else None


def _f(a):
    """float64 array in the reference layout (column-major) and contiguous."""
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray(a) if a.ndim == 2 else np.ascontiguousarray(a)


# ------------------------------------------------------------------------------------------------
# Routine wrappers (arguments in reference order; arrays modified in place like the MEX)
# ------------------------------------------------------------------------------------------------
def calculatedeformedlength(x, BL, L0, l, dX, dY, dZ, s):
    lib().or_calculatedeformedlength(_p(x), _p(BL), _p(L0), C.c_int(BL.shape[0]), C.c_int(x.shape[0]), _p(l),
                                     _p(dX), _p(dY), _p(dZ), _p(s))


def calculatebsfexponential(s, s0, sc, k, alpha, bsf, H, BT, flag):
    lib().or_calculatebsfexponential(_p(s), _p(s0), _p(sc), C.c_double(k), C.c_double(alpha), _p(bsf), _p(H),
                                     _p(BT), _p(flag), C.c_int(s.shape[0]))


def calculatebondsofteningfactor(s, s0, sc, flag, H, BT):
    s0a, sca = np.atleast_1d(_f(s0)), np.atleast_1d(_f(sc))
    lib().or_calculatebondsofteningfactor(_p(s), _p(s0a), C.c_int(1 if s0a.size > 1 else 0), _p(sca),
                                          C.c_int(1 if sca.size > 1 else 0), _p(flag), _p(H), _p(BT),
                                          C.c_int(s.shape[0]))


def calculatebsftrilinear(s, s0, s1, sc, H, BT, flag):
    lib().or_calculatebsftrilinear(_p(s), C.c_double(s0), C.c_double(s1), C.c_double(sc), _p(H), _p(BT),
                                   _p(flag), C.c_int(s.shape[0]))


def calculatebondfailure(fail, ff, BT, s, csc, css):
    lib().or_calculatebondfailure(_p(fail), C.c_double(ff), _p(BT), _p(s), C.c_double(csc), C.c_double(css),
                                  C.c_int(fail.shape[0]))


def calculatebondforces(fX, fY, fZ, fail, dX, dY, dZ, l, s, sp, c, V, vcf, H):
    lib().or_calculatebondforces(_p(fX), _p(fY), _p(fZ), _p(fail), _p(dX), _p(dY), _p(dZ), _p(l), _p(s), _p(sp),
                                 _p(c), C.c_double(V), _p(vcf), _p(H), C.c_int(fX.shape[0]))


def calculatenodalforces_serial(BL, F, fX, fY, fZ):
    lib().or_calculatenodalforces_serial(C.c_int(BL.shape[0]), C.c_int(F.shape[0]), _p(BL), _p(F), _p(fX), _p(fY),
                                         _p(fZ))


def calculatenodalforces_openmp(BL, F, fX, fY, fZ, sX, sY, sZ, T):
    lib().or_calculatenodalforces_openmp(C.c_int(BL.shape[0]), C.c_int(F.shape[0]), _p(BL), _p(F), _p(fX), _p(fY),
                                         _p(fZ), _p(sX), _p(sY), _p(sZ), C.c_int(T))


def timeintegrationeulercromer(a, F, D, v, rho, cons, DT, u, x, X):
    lib().or_timeintegrationeulercromer(_p(a), _p(F), C.c_double(D), _p(v), _p(rho), _p(cons), C.c_double(DT),
                                        _p(u), _p(x), _p(X), C.c_int(F.shape[0]), C.c_int(F.shape[1]))


def calculatedampingcoefficient(F, m, Fp, DT, vp, u) -> float:
    f = lib().or_calculatedampingcoefficient
    f.restype = C.c_double
    return f(_p(F), _p(m), _p(Fp), C.c_double(DT), _p(vp), _p(u), C.c_int(F.shape[0]))


def calculatedampingcoefficient_masked(F, m, Fp, DT, vp, u, skip) -> float:
    """The quotient over the nodes with skip == 0 (the stepped solver's ADR with contact)."""
    f = lib().or_calculatedampingcoefficient_masked
    f.restype = C.c_double
    sk = np.ascontiguousarray(skip, dtype=np.uint8)
    return f(_p(F), _p(m), _p(Fp), C.c_double(DT), _p(vp), _p(u), C.c_int(F.shape[0]),
             sk.ctypes.data_as(C.c_void_p))


def tree_sum(v) -> float:
    """The device's halving-tree sum of len(v) values (a power of two)."""
    a = np.array(v, dtype=np.float64)
    f = lib().or_tree_sum
    f.restype = C.c_double
    return f(_p(a), C.c_int(a.size))


def slot_sum(by_slot) -> float:
    """The device's sum of one value per node slot (blocks of 256 by the tree, then the block sums)."""
    a = np.ascontiguousarray(by_slot, dtype=np.float64)
    f = lib().or_slot_sum
    f.restype = C.c_double
    return f(_p(a), C.c_long(a.size))


def dampingsums_device(F, m, Fp, DT, vp, u, skip, cons, slot_of, nSlots):
    """(cn1, cn2, sum of F^2 over free DOFs) in the device's order."""
    out = np.zeros(3)
    sk = np.ascontiguousarray(skip, dtype=np.uint8)
    so = np.ascontiguousarray(slot_of, dtype=np.int32)
    lib().or_dampingsums_device(_p(F), _p(m), _p(Fp), C.c_double(DT), _p(vp), _p(u), C.c_int(F.shape[0]),
                                sk.ctypes.data_as(C.c_void_p), _p(_f(cons)) if cons is not None else None,
                                so.ctypes.data_as(C.c_void_p), C.c_long(nSlots), _p(out))
    return tuple(float(x) for x in out)


def damping_from_sums(cn1, cn2) -> float:
    f = lib().or_damping_from_sums
    f.restype = C.c_double
    return f(C.c_double(cn1), C.c_double(cn2))


def static_reaction(F, family, V) -> float:
    """Static reaction (calculatereactionforceFast.m:3-5) in the device's 1024-lane order."""
    f = lib().or_static_reaction
    f.restype = C.c_double
    fam = _f(family)
    return f(_p(F), _p(fam), C.c_int(fam.size), C.c_double(V), C.c_int(F.shape[0]))


def staircase(i, ipi, increment) -> float:
    f = lib().or_staircase
    f.restype = C.c_double
    return f(C.c_long(i), C.c_long(ipi), C.c_double(increment))


def equilibrium_ratio(res2, V, reaction) -> float:
    f = lib().or_equilibrium_ratio
    f.restype = C.c_double
    return f(C.c_double(res2), C.c_double(V), C.c_double(reaction))


def timeintegrationADR(it, DT, vf, F, m, cn, vp, u, cons):
    lib().or_timeintegrationADR(C.c_int(it), C.c_double(DT), _p(vf), _p(F), _p(m), C.c_double(cn), _p(vp), _p(u),
                                _p(cons), C.c_int(F.shape[0]))


def buildmassmatrix(DX, c0, DT, nFAM):
    m = np.zeros(nFAM.shape[0])
    lib().or_buildmassmatrix(C.c_int(nFAM.shape[0]), C.c_double(DX), C.c_double(c0), C.c_double(DT), _p(nFAM), _p(m))
    return m


def calculatecontactforce(family, cx, cz, R, disp, X, x, u, v, DT, V, rho) -> float:
    f = lib().or_calculatecontactforce
    f.restype = C.c_double
    return f(_p(family), C.c_int(family.shape[0]), C.c_double(cx), C.c_double(cz), C.c_double(R),
             C.c_double(disp), _p(X), _p(x), _p(u), _p(v), C.c_double(DT), C.c_double(V), _p(rho),
             C.c_int(X.shape[0]))


def smoothstepdata(cur, start, fin, a0, a1) -> float:
    f = lib().or_smoothstepdata
    f.restype = C.c_double
    return f(C.c_double(cur), C.c_double(start), C.c_double(fin), C.c_double(a0), C.c_double(a1))


def calculatedamage(BL, unbroken, nFAM):
    d = np.zeros(nFAM.shape[0])
    lib().or_calculatedamage(_p(BL), _p(unbroken), _p(nFAM), C.c_int(BL.shape[0]), C.c_int(nFAM.shape[0]), _p(d))
    return d


def calculatestrainenergy(nNodes, fail, c, s, L0, vcf, BL, V):
    """(total strain energy, per-node density) of calculatestrainenergy.m:10-27."""
    d = np.zeros(nNodes)
    f = lib().or_calculatestrainenergy
    f.restype = C.c_double
    e = f(C.c_int(nNodes), _p(_f(fail)), _p(_f(c)), None, _p(_f(s)), _p(_f(L0)), _p(_f(vcf)), _p(_f(BL)),
          C.c_int(BL.shape[0]), C.c_double(V), _p(d))
    return e, d


def calculatedamageenergy(c, s0, sc, L0, s, k, alpha, V, hist):
    lib().or_calculatedamageenergy(_p(c), _p(s0), _p(sc), _p(L0), _p(s), C.c_double(k), C.c_double(alpha),
                                   C.c_double(V), _p(hist), C.c_int(s.shape[0]))


# ------------------------------------------------------------------------------------------------
# Whole-step driver in the reference order (or_step in bbpd_oracle.c)
# ------------------------------------------------------------------------------------------------
class _State(C.Structure):
    _fields_ = [
        ("nNodes", C.c_int), ("nBonds", C.c_int),
        ("X", _d), ("BONDLIST", _d), ("L0", _d), ("vcf", _d), ("c", _d), ("BONDTYPE", _d), ("s0", _d), ("sc", _d),
        ("rho", _d), ("cons", _d), ("mass", _d),
        ("s0_stride", C.c_int), ("sc_stride", C.c_int),
        ("V", C.c_double), ("DT", C.c_double), ("DAMPING", C.c_double), ("s1", C.c_double), ("k", C.c_double),
        ("alpha", C.c_double), ("sc_steel", C.c_double), ("ff", C.c_double),
        ("law", C.c_int), ("integrator", C.c_int), ("nodal_threads", C.c_int),
        ("nBodies", C.c_int), ("families", C.POINTER(_d)), ("nFams", C.POINTER(C.c_int)), ("centres", _d),
        ("radii", _d), ("moves", C.POINTER(C.c_int)),
        ("x", _d), ("u", _d), ("v", _d), ("a", _d), ("F", _d), ("Fp", _d), ("H", _d), ("flag", _d), ("fail", _d),
        ("l", _d), ("dX", _d), ("dY", _d), ("dZ", _d), ("s", _d), ("bsf", _d), ("fX", _d), ("fY", _d), ("fZ", _d),
        ("sp", _d), ("sXs", _d), ("sYs", _d), ("sZs", _d), ("s0_full", _d), ("sc_full", _d),
        ("schedule", C.c_int), ("nTimeSteps", C.c_int), ("ipi", C.c_int), ("ipi_min", C.c_int),
        ("applied", C.c_double), ("increment", C.c_double), ("eq_tol", C.c_double),
        ("slot_of", C.POINTER(C.c_int)), ("nSlots", C.c_long),
        ("inc", C.c_long), ("inc_iters", C.c_long),
        ("last_disp", C.c_double), ("last_static", C.c_double), ("last_ratio", C.c_double),
        ("last_cn", C.c_double), ("last_inc_end", C.c_int),
    ]


def stencil_slots(case, align: int = 2048):
    """The device's node slots in its stencil layout: the x-fastest index of the node's site in the
    lattice box (coordinates are DX * k), node count padded to a multiple of `align`.  Returns
    (slot_of, nSlots)."""
    k = np.rint(np.asarray(case.X) / case.DX).astype(np.int64)
    k -= k.min(axis=0)
    dims = k.max(axis=0) + 1
    slot = k[:, 0] + dims[0] * (k[:, 1] + dims[1] * k[:, 2])
    n = case.X.shape[0]
    return slot.astype(np.int32), int((n + align - 1) // align * align)


class OracleRun:
    """The reference step sequence on a beam case (any object with the BeamCase fields)."""

    def __init__(self, case, nodal_threads: int = 0, device_order: bool = False):
        """`device_order`: the ADR sums and the static reaction in the order of the device's stencil
        layout (or_slot_sum over stencil_slots(case)) instead of the reference's sequential order."""
        lib()
        if C.sizeof(_State) != lib().or_sizeof_state():
            raise RuntimeError("oracle state layout mismatch")
        self.case = case
        nN, nB = case.X.shape[0], case.BONDLIST.shape[0]
        self._keep = []
        k = self._keep

        def arr(a):
            a = _f(a)
            k.append(a)
            return a

        S = _State()
        S.nNodes, S.nBonds = nN, nB
        self.X = arr(case.X)
        S.X = _p(self.X)
        S.BONDLIST = _p(arr(case.BONDLIST))
        S.L0 = _p(arr(case.L0))
        S.vcf = _p(arr(case.vcf))
        S.c = _p(arr(case.c))
        S.BONDTYPE = _p(arr(case.BONDTYPE))
        s0 = np.atleast_1d(arr(case.s0))
        sc = np.atleast_1d(arr(case.sc))
        k.extend([s0, sc])
        S.s0, S.sc = _p(s0), _p(sc)
        S.s0_stride = 1 if s0.size > 1 else 0
        S.sc_stride = 1 if sc.size > 1 else 0
        self.s0_full = arr(np.broadcast_to(s0, (nB,)) if s0.size == 1 else s0)
        self.sc_full = arr(np.broadcast_to(sc, (nB,)) if sc.size == 1 else sc)
        S.s0_full, S.sc_full = _p(self.s0_full), _p(self.sc_full)
        S.rho = _p(arr(case.DENSITY))
        S.cons = _p(arr(case.CONSTRAINTFLAG))
        S.mass = _p(arr(case.mass)) if case.mass is not None else None
        S.V, S.DT, S.DAMPING = case.cellVolume, case.DT, case.DAMPING
        S.s1, S.k, S.alpha, S.sc_steel, S.ff = case.s1, case.k, case.alpha, case.sc_steel, case.failureFunctionality
        S.law, S.integrator, S.nodal_threads = case.law, case.integrator, nodal_threads
        nb = len(case.bodies)
        S.nBodies = nb
        fams = [arr(b.family) for b in case.bodies]
        self._fam_ptrs = (_d * max(nb, 1))(*[_p(f) for f in fams])
        self._nf = (C.c_int * max(nb, 1))(*[len(f) for f in fams])
        self._cent = arr(np.array([[b.cx, b.cz] for b in case.bodies]).reshape(-1) if nb else np.zeros(2))
        self._rad = arr(np.array([b.R for b in case.bodies]) if nb else np.zeros(1))
        self._mv = (C.c_int * max(nb, 1))(*[1 if b.moves else 0 for b in case.bodies])
        S.families, S.nFams, S.centres, S.radii, S.moves = self._fam_ptrs, self._nf, _p(self._cent), _p(self._rad), self._mv
        z3 = lambda: arr(np.zeros((nN, 3), order="F"))
        zb = lambda: arr(np.zeros(nB))
        self.x = arr(case.X.copy(order="F"))
        self.u, self.v, self.a, self.F, self.Fp = z3(), z3(), z3(), z3(), z3()
        self.H, self.flag = zb(), zb()
        self.fail = arr(np.array(case.fail, dtype=np.float64))
        self.l, self.dX, self.dY, self.dZ, self.s, self.bsf = zb(), zb(), zb(), zb(), zb(), zb()
        self.fX, self.fY, self.fZ, self.sp = zb(), zb(), zb(), zb()
        T = max(nodal_threads, 1)
        self.sX, self.sY, self.sZ = (arr(np.zeros(nN * T)) for _ in range(3))
        for name in ("x", "u", "v", "a", "F", "Fp", "H", "flag", "fail", "l", "dX", "dY", "dZ", "s", "bsf", "fX",
                     "fY", "fZ", "sp"):
            setattr(S, name, _p(getattr(self, name)))
        S.sXs, S.sYs, S.sZs = _p(self.sX), _p(self.sY), _p(self.sZ)
        S.schedule, S.nTimeSteps = case.schedule, case.nTimeSteps
        S.ipi, S.ipi_min = case.itersPerIncrement, getattr(case, "itersPerIncrementMin", 0)
        S.applied, S.increment, S.eq_tol = case.appliedDisplacement, case.increment, getattr(case, "eqTol", 0.0)
        S.inc, S.inc_iters = 1, 0
        if device_order:
            self._slot_of, S.nSlots = stencil_slots(case)
            S.slot_of = self._slot_of.ctypes.data_as(C.POINTER(C.c_int))
        self.S = S
        self.step_count = 0
        self.reactions = []
        self.records = {k: [] for k in ("disp", "reaction_static", "ratio", "cn", "inc_end")}
        f = lib().or_step
        f.restype = C.c_double
        self._step = f

    def step(self, n: int = 1):
        """n steps; the load-point displacement follows the case's schedule inside or_step."""
        fz = (C.c_double * max(len(self.case.bodies), 1))()
        S, rec = self.S, self.records
        for _ in range(n):
            i = self.step_count + 1
            r = self._step(C.byref(S), C.c_int(i), fz)
            self.reactions.append(r)
            rec["disp"].append(S.last_disp)
            rec["reaction_static"].append(S.last_static)
            rec["ratio"].append(S.last_ratio)
            rec["cn"].append(S.last_cn)
            rec["inc_end"].append(S.last_inc_end)
            self.step_count = i
        return self.reactions[-1] if self.reactions else 0.0

    def load_state(self):
        """(current increment, iterations done at it) of the equilibrium staircase."""
        return int(self.S.inc), int(self.S.inc_iters)

    def step_omp(self, n: int, threads: int):
        """Timed baseline: the reference's CPU passes with OpenMP (Euler-Cromer order)."""
        for _ in range(n):
            lib().or_step_omp(C.byref(self.S), C.c_int(threads))


# ------------------------------------------------------------------------------------------------
# numpy restatement of the input builders (independent of the product's C++ generator)
# ------------------------------------------------------------------------------------------------
def lattice_coordinates(nx, ny, nz, DX):
    """buildmaterialpointcoordinates.m:39-59 (x fastest, coordinates DX*k from k = 1)."""
    k1, k2, k3 = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    X = np.stack([DX * k1.reshape(-1, order="F"), DX * k2.reshape(-1, order="F"), DX * k3.reshape(-1, order="F")],
                 axis=1)
    return np.asfortranarray(X)


def brute_force_bonds(X, horizon, chunk=2048):
    """All pairs i<j with Euclidean distance <= horizon, by direct O(N^2) search
    (buildhorizons.m:13-83 semantics, partners in ascending id order)."""
    nN = X.shape[0]
    out_i, out_j = [], []
    for a in range(0, nN, chunk):
        blk = X[a:a + chunk]
        d2 = ((blk[:, None, :] - X[None, :, :]) ** 2).sum(axis=2)
        ii, jj = np.nonzero(np.sqrt(d2) <= horizon)
        ii = ii + a
        keep = jj > ii
        out_i.append(ii[keep])
        out_j.append(jj[keep])
    i = np.concatenate(out_i)
    j = np.concatenate(out_j)
    order = np.lexsort((j, i))
    return np.asfortranarray(np.stack([i[order] + 1, j[order] + 1], axis=1).astype(np.float64))


def undeformed_length(X, BL):
    """buildhorizons.m:89-97."""
    i = BL[:, 0].astype(np.int64) - 1
    j = BL[:, 1].astype(np.int64) - 1
    e = X[i] - X[j]
    return np.sqrt((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2])


def determineintersection_fp(A, B, Cc, D, p0, p1):
    """determineintersection.m:59-150 in floating point (vectorised over bonds p0->p1)."""
    AB, AD = B - A, D - A
    n = np.cross(AB, AD)
    n = n / np.sqrt(np.dot(n, n))
    u = p1 - p0
    w = p0 - A
    Dd = u @ n
    N = -(w @ n)
    par = np.abs(Dd) < 1e-7
    with np.errstate(divide="ignore", invalid="ignore"):
        sI = N / Dd
        I = p0 + sI[:, None] * u
        on = (~par) & (sI >= 0) & (sI <= 1)
        inside = ((A @ AB <= I @ AB) & (I @ AB <= B @ AB) & (A @ AD <= I @ AD) & (I @ AD <= D @ AD))
    return on & inside


def notch_filter(X, BL, DX, xloc, depth):
    """buildnotch.m:13-67: keep the bonds that do not meet the notch rectangle."""
    mn, mx = X.min(axis=0), X.max(axis=0)
    A = np.array([xloc * DX, mn[1], mn[2]])
    B = np.array([xloc * DX, mx[1], mn[2]])
    Cc = np.array([xloc * DX, mx[1], depth * DX])
    D = np.array([xloc * DX, mn[1], depth * DX])
    i = BL[:, 0].astype(np.int64) - 1
    j = BL[:, 1].astype(np.int64) - 1
    cut = determineintersection_fp(A, B, Cc, D, X[i], X[j])
    return np.asfortranarray(BL[~cut])


def family_counts(BL, nN):
    """buildnotch.m:73-85."""
    cnt = np.zeros(nN)
    np.add.at(cnt, BL[:, 0].astype(np.int64) - 1, 1)
    np.add.at(cnt, BL[:, 1].astype(np.int64) - 1, 1)
    return cnt


def volume_correction(L0, horizon, RADIJ):
    """calculatevolumecorrectionfactors.m:16-34."""
    v = np.zeros_like(L0)
    v[L0 <= horizon - RADIJ] = 1
    mid = (L0 > horizon - RADIJ) & (L0 <= horizon)
    v[mid] = (horizon + RADIJ - L0[mid]) / (2 * RADIJ)
    return v


def bond_stiffness(BL, nFAM, c0, V, horizon):
    """buildbonddata.m:48-51 (all concrete)."""
    nbhd = (4.0 / 3.0) * np.pi * (horizon * horizon * horizon)
    i = BL[:, 0].astype(np.int64) - 1
    j = BL[:, 1].astype(np.int64) - 1
    sf = (2 * nbhd) / ((nFAM[i] * V + nFAM[j] * V))
    return sf * c0


def calculatenodalforces_omp_real(BL, F, fX, fY, fZ, sX, sY, sZ, T):
    """Same reduction run by a real OpenMP team (libgomp static schedule)."""
    lib().or_calculatenodalforces_omp_real(C.c_int(BL.shape[0]), C.c_int(F.shape[0]), _p(BL), _p(F), _p(fX),
                                           _p(fY), _p(fZ), _p(sX), _p(sY), _p(sZ), C.c_int(T))
