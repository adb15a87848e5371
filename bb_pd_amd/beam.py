"""Three-point-bend beam inputs in the reference data model.

`build_beam` returns every array the reference's input script saves for the solver
(input/scripts/Gregoire_beam_tests.m:108-251), in the reference layout (FP64, column-major as
numpy arrays of shape (n, 3) in Fortran order, BONDLIST as doubles holding 1-based ids):

* lattice, bond list, notch, L0, vcf, surface-corrected stiffness: the C++ generator
  (bbpd_beam_build), which restates buildmaterialpointcoordinates.m, buildhorizons.m, buildnotch.m,
  calculatevolumecorrectionfactors.m and buildbonddata.m;
* supports and penetrator: buildpenetrator.m:5-24 (families by undeformed x-z distance), supports at
  z = DX - R (Gregoire_beam_tests.m:144-149), penetrator centre corrected to touch the top
  (Gregoire_beam_tests.m:159-167);
* material: rho per node (buildnodaldensity.m), c0 = 12E/(pi delta^4) (Gregoire_beam_tests.m:208);
* per-bond s0 perturbation: s0_b = s0 (1 + a (2U - 1)) with U from SplitMix64 over (seed, bond index);
* exponential-law sc from the fracture energy (Gregoire_beam_tests.m:233-235);
* time step: Euler-Cromer DT (Gregoire_beam_tests.m:244) or ADR with DT = 1 and the fictitious
  mass of adaptivedynamicrelaxation.m:37 (1/4 DT^2 (4/3 pi delta^3) c0 / DX, the same for all nodes).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

SPLITMIX_GAMMA = 0x9E3779B97F4A7C15


def splitmix_uniform(seed: int, idx: np.ndarray) -> np.ndarray:
    """U in [0,1) from SplitMix64 of (seed << 32) ^ idx (idx: uint64 array)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(32)) ^ idx.astype(np.uint64)
        z = z + np.uint64(SPLITMIX_GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Body:
    family: np.ndarray  # 1-based ids (float64), ascending
    cx: float
    cz: float
    R: float
    moves: bool


@dataclass
class BeamCase:
    name: str
    nx: int
    ny: int
    nz: int
    DX: float
    horizon: float
    RADIJ: float
    cellVolume: float
    X: np.ndarray            # (nN, 3) F-order
    BONDLIST: np.ndarray     # (nB, 2) F-order, doubles, 1-based
    L0: np.ndarray
    vcf: np.ndarray
    c: np.ndarray
    nFAM: np.ndarray
    BONDTYPE: np.ndarray
    fail: np.ndarray
    DENSITY: np.ndarray
    CONSTRAINTFLAG: np.ndarray  # (nN, 3) F-order
    c0: float
    E: float
    law: int
    s0: np.ndarray | float
    sc: np.ndarray | float
    s1: float = 0.0
    k: float = 25.0
    alpha: float = 0.25
    sc_steel: float = 1.0
    failureFunctionality: float = 0.0
    integrator: int = _lib.ADR
    DT: float = 1.0
    DAMPING: float = 0.0
    mass: np.ndarray | None = None
    bodies: list = field(default_factory=list)
    schedule: int = 1
    nTimeSteps: int = 1000
    appliedDisplacement: float = -1e-3
    increment: float = -5e-6
    itersPerIncrement: int = 1000
    itersPerIncrementMin: int = 0
    eqTol: float = 0.0
    referenceNode: int = 0
    cmod: tuple = (0, 0)

    @property
    def nNodes(self) -> int:
        return self.X.shape[0]

    @property
    def nBonds(self) -> int:
        return self.BONDLIST.shape[0]


def lattice(nx, ny, nz, DX, horizon_factor=3.015, notch=None, E=37e9):
    """Lattice arrays from the C++ generator: X, BONDLIST, L0, vcf, c, nFAM."""
    L = _lib.lib()
    spec = _lib.BeamSpec(nx, ny, nz, DX, horizon_factor, 0, 0.0, 0.0, E)
    if notch is not None:
        spec.notch, spec.notch_x, spec.notch_depth = 1, float(notch[0]), float(notch[1])
    nN, nB = C.c_int64(), C.c_int64()
    _lib.check(L.bbpd_beam_count(C.byref(spec), C.byref(nN), C.byref(nB)), "bbpd_beam_count")
    nN, nB = nN.value, nB.value
    X = np.zeros((nN, 3), order="F")
    BL = np.zeros((nB, 2), order="F")
    L0 = np.zeros(nB)
    vcf = np.zeros(nB)
    c = np.zeros(nB)
    nFAM = np.zeros(nN)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    _lib.check(L.bbpd_beam_build(C.byref(spec), nB, ptr(X), ptr(BL), ptr(L0), ptr(vcf), ptr(c), ptr(nFAM)),
               "bbpd_beam_build")
    return X, BL, L0, vcf, c, nFAM


def _ec_dt(rho, DX, horizon, c0):
    """Gregoire_beam_tests.m:244: the Euler-Cromer stable time step with a 0.8 safety factor."""
    return (0.8 * math.sqrt(2 * rho * DX / (math.pi * horizon ** 2 * DX * c0))) / 1


def euler_cromer_dt(case: BeamCase) -> float:
    """The Euler-Cromer time step of a beam case (whatever integrator it was built for)."""
    return _ec_dt(float(case.DENSITY[0]), case.DX, case.horizon, case.c0)


def build_penetrator(X: np.ndarray, cx: float, cz: float, R: float, search: float, moves: bool) -> Body:
    """buildpenetrator.m:10-24: family = nodes with undeformed x-z distance <= searchRadius."""
    d = np.sqrt((X[:, 0] - cx) ** 2 + (X[:, 2] - cz) ** 2)
    fam = (np.nonzero(d <= search)[0] + 1).astype(np.float64)
    return Body(fam, cx, cz, R, moves)


def exponential_sc(c: np.ndarray, s0, GF: float, delta: float, k: float = 25.0, alpha: float = 0.25):
    """Gregoire_beam_tests.m:235, per bond (c and s0 may be arrays), left-to-right as written."""
    ek = math.exp(k)
    shape = 2 * k - 2 * ek + alpha * k - alpha * k * ek + 2
    d5 = delta ** 5
    inner = GF - (math.pi * c * d5 * s0 ** 2 * shape) / (10 * k * (ek - 1) * (alpha + 1))
    return (10 * k * (1 - ek) * inner * (alpha + 1)) / (c * d5 * s0 * math.pi * shape)


def build_beam(name: str, nx: int, ny: int, nz: int, DX: float, *, law: int, horizon_factor: float = 3.015,
               E: float = 37e9, rho: float = 2400.0, notch=None, s0: float = 1.0e-4, s0_perturb: float = 0.0,
               seed: int = 0, sc: float | None = None, s1: float = 0.0, GF: float = 125.2, k: float = 25.0,
               alpha: float = 0.25, integrator: int = _lib.ADR, supports_x=None, penetrator_x=None,
               schedule: int = 1, increment: float = -5e-6, iters_per_increment: int = 1000,
               nTimeSteps: int = 20000, appliedDisplacement: float = -1e-3,
               pin_load_line: bool | None = None, eq_tol: float = 0.0,
               iters_per_increment_min: int = 0) -> BeamCase:
    """Three-point-bend beam in the reference's arrays.  `pin_load_line` (default: on under ADR)
    holds u_x of the top node line under the penetrator and u_y of its middle node through
    CONSTRAINTFLAG: frictionless cylinders leave rigid x translation and rotation about z without
    stiffness, and under ADR a drift there inflates the damping quotient's sum of u^2."""
    X, BL, L0, vcf, c, nFAM = lattice(nx, ny, nz, DX, horizon_factor, notch, E)
    nN, nB = X.shape[0], BL.shape[0]
    horizon = horizon_factor * DX
    c0 = (12 * E) / (math.pi * horizon ** 4)
    RADIJ = DX / 2
    V = DX ** 3
    # per-bond s0
    if s0_perturb:
        U = splitmix_uniform(seed, np.arange(nB, dtype=np.uint64))
        s0v = s0 * (1 + s0_perturb * (2 * U - 1))
    else:
        s0v = s0
    if law == _lib.LAW_EXPONENTIAL:
        scv = exponential_sc(c, s0v, GF, horizon, k, alpha) if sc is None else sc
    else:
        scv = sc
    # supports (Gregoire_beam_tests.m:144-149) and penetrator (:159-167)
    length = nx * DX
    if supports_x is None:
        supports_x = (0.05 * length, 0.95 * length)
    if penetrator_x is None:
        penetrator_x = (nx / 2) * DX
    sR, sS = 5 * DX, 10.1 * DX
    pR, pS = 10 * DX, 15.1 * DX
    supports = [build_penetrator(X, sx, -sR + DX, sR, sS, False) for sx in supports_x]
    pz0 = nz * DX + pR
    pen = build_penetrator(X, penetrator_x, pz0, pR, pS, True)
    fam0 = pen.family.astype(np.int64) - 1
    dist = np.sqrt((X[fam0, 0] - penetrator_x) ** 2 + (X[fam0, 2] - pz0) ** 2)
    pen.cz = pz0 - (dist.min() - pR)
    bodies = [pen] + supports
    cons = np.zeros((nN, 3), order="F")
    if pin_load_line is None:
        pin_load_line = integrator == _lib.ADR
    if pin_load_line:
        line = np.isclose(X[:, 0], penetrator_x) & np.isclose(X[:, 2], nz * DX)
        if not line.any():
            raise ValueError("the penetrator is not above a node column: nothing to pin")
        cons[line, 0] = 1
        ids = np.nonzero(line)[0]
        cons[ids[len(ids) // 2], 1] = 1
    # time step
    if integrator == _lib.EULER_CROMER:
        DT = _ec_dt(rho, DX, horizon, c0)
        mass = None
    else:
        DT = 1.0
        mass = np.full(nN, 0.25 * DT * DT * ((4 / 3) * math.pi * horizon ** 3) * c0 / DX)
    # observation nodes: bottom midspan (deflection) and the bottom nodes either side of the notch
    # plane (or of midspan) for the CMOD, all at mid-width
    iy = (ny - 1) // 2
    xm = notch[0] if notch is not None else (nx + 1) / 2
    ixm = int(round(xm)) - 1 if notch is None else int(math.floor(xm)) - 1
    ixl = int(math.floor(xm)) - 1
    ixr = min(ixl + 1, nx - 1)
    nid = lambda ix, iy_, iz: ix + nx * (iy_ + ny * iz) + 1
    ref = nid(min(max(int(round((nx + 1) / 2)) - 1, 0), nx - 1), iy, 0)
    return BeamCase(
        name=name, nx=nx, ny=ny, nz=nz, DX=DX, horizon=horizon, RADIJ=RADIJ, cellVolume=V, X=X, BONDLIST=BL,
        L0=L0, vcf=vcf, c=c, nFAM=nFAM, BONDTYPE=np.zeros(nB), fail=np.ones(nB), DENSITY=np.full(nN, rho),
        CONSTRAINTFLAG=cons, c0=c0, E=E, law=law, s0=s0v,
        sc=scv if scv is not None else 0.0, s1=s1, k=k, alpha=alpha, integrator=integrator, DT=DT, mass=mass,
        bodies=bodies, schedule=schedule, nTimeSteps=nTimeSteps, appliedDisplacement=appliedDisplacement,
        increment=increment, itersPerIncrement=iters_per_increment, itersPerIncrementMin=iters_per_increment_min,
        eqTol=eq_tol, referenceNode=ref,
        cmod=(nid(ixl, iy, 0), nid(ixr, iy, 0)))


# ------------------------------------------------------------------------------------------------
# The workloads: lattice sizes and laws of the five three-point-bend cases
# ------------------------------------------------------------------------------------------------
def case_bilinear_dx5(**kw) -> BeamCase:
    """Unnotched 500x100x50 mm beam, dx = 5 mm (20,000 nodes), bilinear law with a seeded +-1%
    per-bond s0 perturbation (s0 = 1e-4, sc = 2e-3), ADR, 5e-6 m per load increment."""
    args = dict(law=_lib.LAW_BILINEAR, s0=1.0e-4, s0_perturb=0.01, seed=0, sc=2.0e-3,
                supports_x=(25e-3, 475e-3))
    args.update(kw)
    return build_beam("beam500x100x50_dx5_bilinear", 100, 10, 20, 5e-3, **args)


def case_trilinear_dx2p5(**kw) -> BeamCase:
    """Same beam at dx = 2.5 mm (160,000 nodes), centre notch 20 mm deep, trilinear law
    (s0 = 1e-4, knee 5e-4 at 25% force, sc = 3e-3)."""
    args = dict(law=_lib.LAW_TRILINEAR, s0=1.0e-4, s1=5.0e-4, sc=3.0e-3, notch=(100.5, 8.5),
                supports_x=(25e-3, 475e-3))
    args.update(kw)
    return build_beam("beam500x100x50_dx2.5_notched_trilinear", 200, 20, 40, 2.5e-3, **args)


def case_exponential_dx1p25(**kw) -> BeamCase:
    """Mixed-mode beam at dx = 1.25 mm (1.28M nodes), 20 mm notch at 0.3125 half-span off centre,
    exponential softening s0 = 1e-4, k = 25, sc from the fracture energy GF = 125.2 N/m."""
    # span 450 mm -> offset 70.3125 mm = 56.25 dx; notch plane between lattice planes
    nx = 400
    xc = (nx + 1) / 2
    xp = math.floor(xc - 56.25) + 0.5
    args = dict(law=_lib.LAW_EXPONENTIAL, s0=1.0e-4, notch=(xp, 16.5), supports_x=(25e-3, 475e-3))
    args.update(kw)
    return build_beam("beam500x100x50_dx1.25_offset_notch_exponential", nx, 40, 80, 1.25e-3, **args)


def case_size_effect(depth_mm: int, seed: int = 0, **kw) -> BeamCase:
    """Geometrically similar beams of the size-effect family: depth 50 / 100 / 200 / 400 mm, length
    5 x depth, width 50 mm, dx = 5 mm (5k / 20k / 80k / 320k nodes), unnotched, bilinear law with the
    seeded +-1% per-bond s0 perturbation (s0 = 1e-4, sc = 2e-3), ADR; supports at 5% of the length
    from each end as for the other beams."""
    DX = 5e-3
    nz = int(round(depth_mm * 1e-3 / DX))
    nx, ny = 5 * nz, 10
    args = dict(law=_lib.LAW_BILINEAR, s0=1.0e-4, s0_perturb=0.01, seed=seed, sc=2.0e-3)
    args.update(kw)
    return build_beam(f"beam_d{depth_mm}_seed{seed}", nx, ny, nz, DX, **args)


def to_problem(case: BeamCase, keep: list) -> _lib.Problem:
    """Fill the C problem struct; `keep` receives every array the struct points into."""
    def p(a):
        if a is None:
            return None
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 2:
            a = np.asfortranarray(a)
        else:
            a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))

    P = _lib.Problem()
    P.nNodes, P.nBonds = case.nNodes, case.nBonds
    P.undeformedCoordinates = p(case.X)
    P.BONDLIST = p(case.BONDLIST)
    P.UNDEFORMEDLENGTH = p(case.L0)
    P.VOLUMECORRECTIONFACTORS = p(case.vcf)
    P.BONDSTIFFNESS = p(case.c)
    P.BONDTYPE = p(case.BONDTYPE)
    P.fail = p(case.fail)
    P.DENSITY = p(case.DENSITY)
    P.CONSTRAINTFLAG = p(case.CONSTRAINTFLAG)
    P.massVector = p(case.mass)
    P.cellVolume, P.horizon, P.RADIJ = case.cellVolume, case.horizon, case.RADIJ
    P.law = case.law
    if isinstance(case.s0, np.ndarray):
        P.s0 = p(case.s0)
    else:
        P.s0_scalar = float(case.s0)
    if isinstance(case.sc, np.ndarray):
        P.sc = p(case.sc)
    else:
        P.sc_scalar = float(case.sc)
    P.s1, P.k, P.alpha = case.s1, case.k, case.alpha
    P.sc_steel, P.failureFunctionality = case.sc_steel, case.failureFunctionality
    P.integrator, P.DT, P.DAMPING = case.integrator, case.DT, case.DAMPING
    P.nBodies = len(case.bodies)
    for b, body in enumerate(case.bodies):
        P.family[b] = p(body.family)
        P.nFamily[b] = len(body.family)
        P.centreX[b], P.centreZ[b], P.radius[b] = body.cx, body.cz, body.R
        P.moves[b] = 1 if body.moves else 0
    P.schedule, P.nTimeSteps = case.schedule, case.nTimeSteps
    P.appliedDisplacement, P.increment, P.itersPerIncrement = (case.appliedDisplacement, case.increment,
                                                               case.itersPerIncrement)
    P.referenceNode = case.referenceNode
    P.cmodNode1, P.cmodNode2 = case.cmod
    P.eqTol, P.itersPerIncrementMin = case.eqTol, case.itersPerIncrementMin
    return P
