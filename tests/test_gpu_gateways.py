"""The MEX gateways (bb_pd_amd/mex/*Mex.c), called as MATLAB calls them: host arrays in the reference
layout, results written in place into the caller's arrays, through the project's mex.h stand-in.
Each against the oracle on the same inputs: bit for bit, except the exponential law's exp() (1e-14
relative, flags identical).  bbpdStepMex runs a fused solver from a problem struct and is checked
against the oracle's reference step sequence bit for bit."""
import numpy as np
import pytest

from bb_pd_amd import _lib, beam

import mexcall

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float64).view(np.int64),
                          np.ascontiguousarray(b, np.float64).view(np.int64))


def cp(a):
    return np.array(a, dtype=np.float64, order="F", copy=True)


@pytest.fixture(scope="module")
def case():
    print("case: 24x6x10 lattice, dx 5 mm, notched", flush=True)
    return beam.build_beam("t", 24, 6, 10, 5e-3, law=1, sc=2e-3, notch=(12.5, 3.5))


def test_deformedlength_and_forces(gpu, oracle, case):
    rng = np.random.default_rng(1)
    x = np.asfortranarray(case.X + rng.standard_normal((case.nNodes, 3)) * 2e-6)
    nB = case.nBonds
    ref = [np.zeros(nB) for _ in range(5)]
    oracle.calculatedeformedlength(x, case.BONDLIST, case.L0, *ref)
    got = [np.zeros(nB) for _ in range(5)]
    mexcall.call("calculatedeformedlengthMex", cp(x), cp(case.BONDLIST), cp(case.L0), *got)
    for a, b in zip(got, ref):
        assert same(a, b)
    l, dX, dY, dZ, s = got
    H = rng.uniform(0, 1, nB) * (rng.uniform(size=nB) < 0.3)
    sp = np.zeros(nB)
    fo = [np.zeros(nB) for _ in range(3)]
    oracle.calculatebondforces(*fo, case.fail, dX, dY, dZ, l, s, sp, case.c, case.cellVolume, case.vcf, H)
    fg = [np.zeros(nB) for _ in range(3)]
    mexcall.call("calculatebondforcesMex", *fg, cp(case.fail), dX, dY, dZ, l, s, sp, cp(case.c), case.cellVolume,
                 cp(case.vcf), H)
    for a, b in zip(fg, fo):
        assert same(a, b)
    # nodal reduction: the 8-thread OpenMP order with its scratch columns, and the serial order
    nN = case.nNodes
    Fo, Fg = np.zeros((nN, 3), order="F"), np.zeros((nN, 3), order="F")
    so = [np.zeros(nN * 8) for _ in range(3)]
    oracle.calculatenodalforces_openmp(case.BONDLIST, Fo, *fo, *so, 8)
    sg = [np.zeros((nN, 8), order="F") for _ in range(3)]
    mexcall.call("calculatenodalforcesopenmpMex", cp(case.BONDLIST), Fg, *fg, *sg)
    assert same(Fg, Fo)
    for a, b in zip(sg, so):
        assert same(a.reshape(-1, order="F"), b)
    Fs, Fsg = np.zeros((nN, 3), order="F"), np.zeros((nN, 3), order="F")
    oracle.calculatenodalforces_serial(case.BONDLIST, Fs, *fo)
    mexcall.call("calculatenodalforcesMex", cp(case.BONDLIST), Fsg, *fg)
    assert same(Fsg, Fs)
    with pytest.raises(mexcall.MexError, match="nThreads >= 8"):
        mexcall.call("calculatenodalforcesopenmpMex", cp(case.BONDLIST), Fg, *fg,
                     *[np.zeros((nN, 4), order="F") for _ in range(3)])


def test_laws(gpu, oracle):
    n = 20000
    rng = np.random.default_rng(2)
    s = rng.uniform(-1e-3, 5e-3, n)
    BT = rng.choice([0.0, 1.0, 2.0], n, p=[0.8, 0.1, 0.1])
    s0a = 1e-4 * (1 + 0.01 * rng.uniform(-1, 1, n))
    sca = np.full(n, 3e-3)
    H0 = rng.uniform(0, 0.5, n) * (rng.uniform(size=n) < 0.3)
    fl0 = (rng.uniform(size=n) < 0.2).astype(float)
    # exponential (calculatebsfexponentialMex.c)
    bo, Ho, fo = np.zeros(n), H0.copy(), fl0.copy()
    oracle.calculatebsfexponential(s, s0a, sca, 25.0, 0.25, bo, Ho, BT, fo)
    bg, Hg, fg = np.zeros(n), H0.copy(), fl0.copy()
    mexcall.call("calculatebsfexponentialMex", s, s0a, sca, 25.0, 0.25, bg, Hg, BT, fg)
    assert same(fg, fo)
    np.testing.assert_allclose(Hg, Ho, rtol=1e-14, atol=0)
    np.testing.assert_allclose(bg, bo, rtol=1e-14, atol=0)
    # bilinear (calculatebondsofteningfactor.m), per-bond and scalar s0
    for s0 in (s0a, 1e-4):
        Ho, fo = H0.copy(), fl0.copy()
        oracle.calculatebondsofteningfactor(s, s0, 2e-3, fo, Ho, BT)
        Hg, fg = H0.copy(), fl0.copy()
        mexcall.call("calculatebondsofteningfactorMex", s, s0 if np.ndim(s0) else np.array([s0]), 2e-3, fg, Hg, BT)
        assert same(Hg, Ho) and same(fg, fo)
    # trilinear (calculatebsftrilinear.m)
    Ho, fo = H0.copy(), fl0.copy()
    oracle.calculatebsftrilinear(s, 1e-4, 5e-4, 3e-3, Ho, BT, fo)
    Hg, fg = H0.copy(), fl0.copy()
    mexcall.call("calculatebsftrilinearMex", s, 1e-4, 5e-4, 3e-3, Hg, BT, fg)
    assert same(Hg, Ho) and same(fg, fo)
    # brittle latch (calculatebondfailureMex.c)
    fail0 = (rng.uniform(size=n) < 0.9).astype(float)
    fo, fg = fail0.copy(), fail0.copy()
    oracle.calculatebondfailure(fo, 0.0, BT, s, 2e-3, 4e-3)
    mexcall.call("calculatebondfailureMex", fg, 0.0, BT, s, 2e-3, 4e-3)
    assert same(fg, fo)


def test_time_integration_and_damping(gpu, oracle, case):
    rng = np.random.default_rng(3)
    nN = case.nNodes
    F = np.asfortranarray(rng.standard_normal((nN, 3)))
    v0 = np.asfortranarray(rng.standard_normal((nN, 3)) * (rng.uniform(size=(nN, 3)) < 0.8))
    u0 = np.asfortranarray(rng.standard_normal((nN, 3)) * 1e-4)
    cons = np.asfortranarray((rng.uniform(size=(nN, 3)) < 0.1).astype(float))
    rho = rng.uniform(2000, 2500, nN)
    # Euler-Cromer (timeintegrationeulercromerMex.c)
    ao, vo, uo, xo = np.zeros((nN, 3), order="F"), cp(v0), cp(u0), np.zeros((nN, 3), order="F")
    oracle.timeintegrationeulercromer(ao, F, 0.1, vo, rho, cons, 1e-6, uo, xo, case.X)
    ag, vg, ug, xg = np.zeros((nN, 3), order="F"), cp(v0), cp(u0), np.zeros((nN, 3), order="F")
    mexcall.call("timeintegrationeulercromerMex", ag, cp(F), 0.1, vg, rho, cons, 1e-6, ug, xg, cp(case.X))
    for a, b in zip((ag, vg, ug, xg), (ao, vo, uo, xo)):
        assert same(a, b)
    # damping coefficient: the device's order (blocks of 256 nodes by the tree), bit for bit
    m = rng.uniform(1, 2, nN)
    Fp = np.asfortranarray(F + 0.1 * v0 * rng.uniform(0.5, 1.5, (nN, 3)))
    cn = mexcall.call("calculatedampingcoefficientMex", cp(F), m, Fp, 1.0, cp(v0), cp(u0), nlhs=1)
    c1, c2, _ = oracle.dampingsums_device(F, m, Fp, 1.0, v0, u0, np.zeros(nN, np.uint8), None,
                                          np.arange(nN, dtype=np.int32), (nN + 255) // 256 * 256)
    assert cn[0] == oracle.damping_from_sums(c1, c2) and 0 < cn[0] < 1.9
    # ADR update (timeintegrationADR.m), first and later step, forward and previous velocity as one
    for it in (1, 2):
        vo, uo = cp(v0), cp(u0)
        oracle.timeintegrationADR(it, 1.0, vo, F, m, 0.3, vo, uo, cons)
        vg, ug = cp(v0), cp(u0)
        mexcall.call("timeintegrationADRMex", float(it), 1.0, vg, cp(F), m, 0.3, vg, ug, cons)
        assert same(vg, vo) and same(ug, uo)


def test_contact(gpu, oracle, case):
    rng = np.random.default_rng(4)
    u0 = np.asfortranarray(rng.standard_normal((case.nNodes, 3)) * 1e-5)
    v0 = np.asfortranarray(rng.standard_normal((case.nNodes, 3)) * 1e-3)
    for b in case.bodies:
        disp = -3e-5 if b.moves else 0.0
        xo, uo, vo = np.asfortranarray(case.X + u0), cp(u0), cp(v0)
        fz_o = oracle.calculatecontactforce(b.family, b.cx, b.cz, b.R, disp, case.X, xo, uo, vo, 1.2e-6,
                                            case.cellVolume, case.DENSITY)
        xg, ug, vg = np.asfortranarray(case.X + u0), cp(u0), cp(v0)
        pen = {"family": cp(b.family), "centre": np.array([[b.cx, b.cz]], order="F"), "radius": b.R}
        fz = mexcall.call("calculatecontactforceMex", pen, disp, cp(case.X), xg, ug, vg, 1.2e-6, case.cellVolume,
                          cp(case.DENSITY), nlhs=1)
        assert same(xg, xo) and same(ug, uo) and same(vg, vo)
        assert same(fz, [fz_o])


def problem_struct(c):
    p = dict(undeformedCoordinates=cp(c.X), BONDLIST=cp(c.BONDLIST), UNDEFORMEDLENGTH=cp(c.L0),
             VOLUMECORRECTIONFACTORS=cp(c.vcf), BONDSTIFFNESS=cp(c.c), BONDTYPE=cp(c.BONDTYPE), fail=cp(c.fail),
             DENSITY=cp(c.DENSITY), CONSTRAINTFLAG=cp(c.CONSTRAINTFLAG), cellVolume=c.cellVolume,
             horizon=c.horizon, RADIJ=c.RADIJ, law=float(c.law), s0=np.atleast_1d(cp(c.s0)),
             sc=np.atleast_1d(cp(c.sc)), s1=c.s1, k=c.k, alpha=c.alpha, integrator=float(c.integrator), DT=c.DT,
             DAMPING=c.DAMPING, schedule=float(c.schedule), nTimeSteps=float(c.nTimeSteps),
             appliedDisplacement=c.appliedDisplacement, increment=c.increment,
             itersPerIncrement=float(c.itersPerIncrement), referenceNode=float(c.referenceNode),
             cmodNodes=np.array([[float(c.cmod[0]), float(c.cmod[1])]], order="F"),
             bodies=[{"family": cp(b.family), "centre": np.array([[b.cx, b.cz]], order="F"), "radius": b.R,
                      "moves": 1.0 if b.moves else 0.0} for b in c.bodies])
    if c.mass is not None:
        p["massVector"] = cp(c.mass)
    return p


def test_step_gateway(gpu, oracle):
    c = beam.build_beam("t", 24, 6, 10, 5e-3, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01, sc=2e-3,
                        integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=200, appliedDisplacement=-1.5e-4,
                        notch=(12.5, 3.5))
    run = oracle.OracleRun(c, device_order=True)
    h = mexcall.call("bbpdStepMex", "create", problem_struct(c), 0.0, nlhs=1)
    try:
        mexcall.call("bbpdStepMex", "step", h, 120.0)
        run.step(120)
        u, v, F, x = mexcall.call("bbpdStepMex", "nodal", h, nlhs=4)
        H, flag, fail = mexcall.call("bbpdStepMex", "bonds", h, nlhs=3)
        for a, b in ((u, run.u), (v, run.v), (F, run.F), (x, run.x), (H, run.H), (flag, run.flag)):
            assert same(a, b)
        assert flag.sum() > 0
        rows = mexcall.call("bbpdStepMex", "history", h, nlhs=1)
        assert rows.shape == (120, _lib.HISTORY_COLUMNS)
        assert same(rows[:, 0], run.reactions) and same(rows[:, 1], run.records["reaction_static"])
        obs = mexcall.call("bbpdStepMex", "observables", h, nlhs=1)
        assert obs[0, 0] == 120 and obs[0, 6] == flag.sum()
        ck = mexcall.call("bbpdStepMex", "checkpoint", h, nlhs=9)
        mexcall.call("bbpdStepMex", "step", h, 40.0)
        run.step(40)
        mexcall.call("bbpdStepMex", "restore", h, *ck)
        mexcall.call("bbpdStepMex", "graph", h, 20.0, 2.0)
        u2 = mexcall.call("bbpdStepMex", "nodal", h, nlhs=4)[0]
        assert same(u2, run.u)
    finally:
        mexcall.call("bbpdStepMex", "destroy", h)
