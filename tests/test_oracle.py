"""Pins of the oracle (oracle/bbpd_oracle.c) by closed forms, invariants, brute force and real OpenMP.

The reference C needs MATLAB's mex.h and cannot be built here, so each oracle routine is checked
against what the mathematics fixes, chosen so that a dropped term, a wrong sign or index, or a wrong
summation order fails at least one test.
"""
import math

import numpy as np
import pytest

from bb_pd_amd import beam


def z(n):
    return np.zeros(n)


def z3(n):
    return np.zeros((n, 3), order="F")


# ---------------------------------------------------------------- deformed length / stretch
def test_deformedlength_hand_values(oracle):
    X = np.asfortranarray([[0.0, 0, 0], [1, 0, 0], [0, 4, 0], [3, 0, 0]])
    x = np.asfortranarray([[0.0, 0, 0], [1.25, 0, 0], [0, 4, 0], [3, 4, 0]])
    BL = np.asfortranarray([[1.0, 2], [1, 3], [3, 4]])
    L0 = np.array([1.0, 4.0, 5.0])  # |X2-X1|, |X3-X1|, |X4-X3|
    l, dX, dY, dZ, s = z(3), z(3), z(3), z(3), z(3)
    oracle.calculatedeformedlength(x, BL, L0, l, dX, dY, dZ, s)
    assert list(dX) == [1.25, 0.0, 3.0]
    assert list(dY) == [0.0, 4.0, 0.0]
    assert list(l) == [1.25, 4.0, 3.0]          # bond 3: (3,4,0)-(0,4,0) -> length 3
    assert list(s) == [0.25, 0.0, (3.0 - 5.0) / 5.0]
    # column 2 of BONDLIST is the j end: dX = x_j - x_i (sign pin)
    BLr = np.asfortranarray(BL[:, ::-1])
    oracle.calculatedeformedlength(x, BLr, L0, l, dX, dY, dZ, s)
    assert list(dX) == [-1.25, -0.0, -3.0]


def test_stretch_rigid_translation_and_dilation(oracle):
    case = beam.build_beam("t", 6, 5, 4, 1.0, law=1, sc=1.0)
    X = case.X
    x = np.asfortranarray(X + np.array([3.0, -2.0, 7.0]))       # integer translation: exact
    nB = case.nBonds
    l, dX, dY, dZ, s = z(nB), z(nB), z(nB), z(nB), z(nB)
    oracle.calculatedeformedlength(x, case.BONDLIST, case.L0, l, dX, dY, dZ, s)
    assert np.all(s == 0.0)
    e = 1.0 / 16
    x = np.asfortranarray(X * (1 + e))
    oracle.calculatedeformedlength(x, case.BONDLIST, case.L0, l, dX, dY, dZ, s)
    assert np.max(np.abs(s - e)) <= 4 * np.finfo(float).eps


# ---------------------------------------------------------------- bond force
def test_bondforce_hand_value_and_nan(oracle):
    # p = fail*c*(1-H)*(s-sp)*V*vcf = 1*2*0.5*(0.375-0.125)*4*1 = 1 ; f = p*(dX/l)
    fail, c, H, s, sp, vcf = [np.array([v, v]) for v in (1.0, 2.0, 0.5, 0.375, 0.125, 1.0)]
    dX, dY, dZ, l = np.array([3.0, 0.0]), np.array([4.0, 0.0]), np.array([0.0, 0.0]), np.array([5.0, 0.0])
    fX, fY, fZ = z(2), z(2), z(2)
    oracle.calculatebondforces(fX, fY, fZ, fail, dX, dY, dZ, l, s, sp, c, 4.0, vcf, H)
    assert fX[0] == 0.6 and fY[0] == 0.8 and fZ[0] == 0.0
    assert fX[1] == 0.0 and fY[1] == 0.0 and fZ[1] == 0.0   # l = 0 -> 0/0 -> NaN -> 0
    # a failed bond carries nothing; vcf scales linearly
    fail2 = np.array([0.0, 1.0])
    vcf2 = np.array([0.5, 0.5])
    oracle.calculatebondforces(fX, fY, fZ, fail2, dX, dY, dZ, l, s, sp, c, 4.0, vcf2, H)
    assert fX[0] == 0.0


def _random_bonds(rng, nN, nB):
    i = rng.integers(1, nN + 1, nB)
    j = rng.integers(1, nN + 1, nB)
    j = np.where(j == i, (i % nN) + 1, j)
    return np.asfortranarray(np.stack([i, j], 1).astype(float))


# ---------------------------------------------------------------- nodal reduction
def test_nodal_serial_matches_exact_sums(oracle):
    rng = np.random.default_rng(1)
    nN, nB = 50, 400
    BL = _random_bonds(rng, nN, nB)
    f = [rng.standard_normal(nB) for _ in range(3)]
    F = z3(nN)
    oracle.calculatenodalforces_serial(BL, F, *f)
    for d in range(3):
        exact = np.zeros(nN)
        for n in range(nN):
            terms = [f[d][k] for k in range(nB) if BL[k, 0] == n + 1] + \
                    [-f[d][k] for k in range(nB) if BL[k, 1] == n + 1]
            exact[n] = math.fsum(terms)
        np.testing.assert_allclose(F[:, d], exact, rtol=0, atol=1e-12)
        assert abs(math.fsum(F[:, d])) < 1e-12   # Newton's third law


def test_nodal_thread_order_equals_real_openmp(oracle):
    """The sequential T-thread emulation reproduces a real OpenMP team's result bit for bit, and
    differs from the serial order on data with a wide dynamic range (so the check has teeth)."""
    rng = np.random.default_rng(7)
    nN, nB = 40, 997
    BL = _random_bonds(rng, nN, nB)
    f = [rng.standard_normal(nB) * 10.0 ** rng.integers(-8, 9, nB) for _ in range(3)]
    Fs = z3(nN)
    oracle.calculatenodalforces_serial(BL, Fs, *f)
    differs = False
    for T in (1, 2, 3, 8):
        Fe, Fr = z3(nN), z3(nN)
        sc_e = [z(nN * T) for _ in range(3)]
        sc_r = [z(nN * T) for _ in range(3)]
        oracle.calculatenodalforces_openmp(BL, Fe, *f, *sc_e, T)
        oracle.calculatenodalforces_omp_real(BL, Fr, *f, *sc_r, T)
        assert np.array_equal(Fe.view(np.int64), Fr.view(np.int64)), T
        for a, b in zip(sc_e, sc_r):
            assert np.array_equal(a, b)
        if T == 1:
            assert np.array_equal(Fe, Fs)
        elif not np.array_equal(Fe, Fs):
            differs = True
    assert differs


# ---------------------------------------------------------------- Euler-Cromer
def test_eulercromer_single_dof(oracle):
    F = np.asfortranarray([[2.0, 2.0, 2.0]])
    v = np.asfortranarray([[1.0, 1.0, 1.0]])
    u = z3(1)
    X = np.asfortranarray([[1.0, 2.0, 3.0]])
    x = X.copy(order="F")
    a = z3(1)
    cons = np.asfortranarray([[0.0, 1.0, 0.0]])
    oracle.timeintegrationeulercromer(a, F, 2.0, v, np.array([4.0]), cons, 0.5, u, x, X)
    # a = (2 - 2*1)/4 = 0 with damping 2
    assert list(a[0]) == [0.0, 0.0, 0.0]
    F = np.asfortranarray([[2.0, 2.0, 6.0]])
    v = np.asfortranarray([[1.0, 1.0, 1.0]])
    u = z3(1)
    oracle.timeintegrationeulercromer(a, F, 0.0, v, np.array([4.0]), cons, 0.5, u, x, X)
    assert list(a[0]) == [0.5, 0.0, 1.5]                 # DOF 2 constrained
    assert list(v[0]) == [1.25, 1.0, 1.75]               # v += a DT
    assert list(u[0]) == [0.625, 0.5, 0.875]             # u += v(new) DT (Euler-Cromer, not Euler)
    assert list(x[0]) == [1.625, 2.5, 3.875]


# ---------------------------------------------------------------- softening laws
def test_bilinear_closed_forms_and_quirk(oracle):
    s0, sc = 1e-4, 2e-3
    s = np.array([0.5e-4, s0, 3e-4, 1e-3, 1.5e-3, sc, 3e-3])
    n = s.size
    H, flag, BT = z(n), z(n), z(n)
    oracle.calculatebondsofteningfactor(s, s0, sc, flag, H, BT)
    assert list(flag) == [0, 0, 1, 1, 1, 1, 1]
    assert H[0] == 0 and H[1] == 0
    force = (1 - H) * s
    lin = s0 * (sc - s) / (sc - s0)                   # linear descent from s0 at s0 to 0 at sc
    np.testing.assert_allclose(force[2:6], lin[2:6], rtol=1e-12, atol=1e-18)
    assert abs(H[5] - 1) < 1e-15 and H[6] == 1.0
    # irreversibility: unloading keeps H; compression of a flagged bond breaks it (the .m quirk)
    H_before = H.copy()
    s2 = np.array([0.5e-4, s0, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4])
    oracle.calculatebondsofteningfactor(s2, s0, sc, flag, H, BT)
    assert np.array_equal(H, H_before)
    s3 = np.full(n, -1e-5)
    oracle.calculatebondsofteningfactor(s3, s0, sc, flag, H, BT)
    assert list(H) == [0, 0, 1, 1, 1, 1, 1]
    # non-concrete bonds are never flagged; per-bond s0 arrays broadcast like scalars
    H, flag, BT = z(n), z(n), np.full(n, 2.0)
    oracle.calculatebondsofteningfactor(s, np.full(n, s0), sc, flag, H, BT)
    assert not flag.any() and not H.any()


def test_trilinear_closed_forms(oracle):
    s0, s1, sc, beta = 1e-4, 5e-4, 3e-3, 0.25
    eps = 1e-12
    s = np.array([s0 + eps, s1, s1 + eps, 0.5 * (s1 + sc), sc, sc + 1e-6])
    n = s.size
    H, flag, BT = z(n), z(n), z(n)
    oracle.calculatebsftrilinear(s, s0, s1, sc, H, BT, flag)
    force = (1 - H) * s
    assert force[0] == pytest.approx(s0, rel=1e-7)             # peak s0 at s0
    assert force[1] == pytest.approx(beta * s0, rel=1e-12)     # knee beta*s0 at s1 (branch 1)
    assert force[2] == pytest.approx(beta * s0, rel=1e-7)      # continuous into branch 2
    assert force[3] == pytest.approx(0.5 * beta * s0, rel=1e-12)  # linear to 0 at sc
    assert abs(force[4]) < 1e-18 and H[5] == 1.0
    assert list(flag) == [1, 1, 0, 0, 0, 0]                    # flag only in branch 1 (.m:87)


def test_exponential_closed_forms_and_fracture_energy(oracle):
    dx = 1.25e-3
    delta = 3.015 * dx
    c0 = 12 * 37e9 / (math.pi * delta ** 4)
    s0, GF, k, alpha = 1e-4, 125.2, 25.0, 0.25
    sc = float(beam.exponential_sc(np.array([c0]), s0, GF, delta, k, alpha)[0])
    s = np.linspace(s0, sc, 400001)
    s[0] = s0 * (1 + 1e-12)
    n = s.size
    bsf, H, BT, flag = z(n), z(n), z(n), z(n)
    oracle.calculatebsfexponential(s, np.full(n, s0), np.full(n, sc), k, alpha, bsf, H, BT, flag)
    assert H[0] < 1e-9 and abs(H[-1] - 1) < 1e-14
    assert np.all(np.diff(H) >= -1e-15)                        # softening factor is monotone
    # the sc formula (Gregoire_beam_tests.m:235) makes the softening-branch energy equal GF:
    # GF = (pi delta^5 / 5) c * integral_{s0}^{sc} (1 - bsf(s)) s ds
    g = (1 - H) * s
    G = math.pi * delta ** 5 / 5 * c0 * np.trapezoid(g, s)
    assert G == pytest.approx(GF, rel=1e-6)
    # per-step scratch: a bond beyond sc is broken in one step even from below s0 (C semantics)
    s1 = np.array([0.5 * s0, 2 * sc])
    bsf, H, flag = z(2), z(2), z(2)
    oracle.calculatebsfexponential(s1, np.full(2, s0), np.full(2, sc), k, alpha, bsf, H, z(2), flag)
    assert list(H) == [0.0, 1.0] and list(flag) == [0.0, 0.0]


def test_brittle_latch(oracle):
    fail = np.ones(4)
    BT = np.array([0.0, 1.0, 2.0, 0.0])
    s = np.array([2e-3, 2e-3, 2e-3, 0.5e-3])
    oracle.calculatebondfailure(fail, 0.0, BT, s, 1e-3, 5e-3)
    assert list(fail) == [0, 1, 1, 1]          # interface needs 3 sc, steel needs sc_steel
    s = np.array([0.0, 4e-3, 6e-3, 0.0])
    oracle.calculatebondfailure(fail, 0.0, BT, s, 1e-3, 5e-3)
    assert list(fail) == [0, 0, 0, 1]          # irreversible: bond 1 stays failed at s = 0


# ---------------------------------------------------------------- ADR
def test_adr_damping_rayleigh_quotient(oracle):
    # one free DOF of a linear spring: F = -k u.  With v_prev = (u - u_prev)/DT the quotient gives
    # cn = 2 sqrt(k/m); DOFs with zero velocity are skipped.
    kk, m, DT = 0.25, 1.0, 1.0
    u_prev, u = 0.5, 0.75
    F = np.asfortranarray([[-kk * u, 123.0, 0.0]])
    Fp = np.asfortranarray([[-kk * u_prev, -7.0, 0.0]])
    vp = np.asfortranarray([[(u - u_prev) / DT, 0.0, 0.0]])
    uu = np.asfortranarray([[u, 0.0, 0.0]])
    cn = oracle.calculatedampingcoefficient(F, np.array([m]), Fp, DT, vp, uu)
    assert cn == pytest.approx(2 * math.sqrt(kk / m), rel=1e-14)
    F[0, 0], Fp[0, 0] = -4 * u, -4 * u_prev
    assert oracle.calculatedampingcoefficient(F, np.array([m]), Fp, DT, vp, uu) == 1.9   # cn > 2 -> 1.9
    assert oracle.calculatedampingcoefficient(F, np.array([m]), Fp, DT, vp, z3(1)) == 0.0


def test_adr_damping_skips_driven_nodes(oracle):
    # node 0: a free spring DOF (cn = 2 sqrt(k/m)); node 1: a node a contact body moved, with a tiny
    # correction velocity and a large force jump.  Left in, node 1 swamps the quotient; masked out,
    # the quotient is node 0's alone, and with nothing masked it equals the reference routine.
    kk, m, DT = 0.25, np.array([1.0, 1.0]), 1.0
    u_prev, u = 0.5, 0.75
    F = np.asfortranarray([[-kk * u, 0.0, 0.0], [0.0, 0.0, 50.0]])
    Fp = np.asfortranarray([[-kk * u_prev, 0.0, 0.0], [0.0, 0.0, -50.0]])
    vp = np.asfortranarray([[(u - u_prev) / DT, 0.0, 0.0], [0.0, 0.0, 1e-6]])
    uu = np.asfortranarray([[u, 0.0, 0.0], [0.0, 0.0, 2.0]])
    full = oracle.calculatedampingcoefficient(F, m, Fp, DT, vp, uu)
    assert full == 0.0                                  # the driven term is negative and huge
    assert oracle.calculatedampingcoefficient_masked(F, m, Fp, DT, vp, uu, [0, 0]) == full
    cn = oracle.calculatedampingcoefficient_masked(F, m, Fp, DT, vp, uu, [0, 1])
    assert cn == pytest.approx(2 * math.sqrt(kk / m[0]), rel=1e-14)
    assert oracle.calculatedampingcoefficient_masked(F, m, Fp, DT, vp, uu, [1, 1]) == 0.0   # cn2 == 0


def test_adr_update_recurrence(oracle):
    DT, m = 1.0, np.array([4.0])
    F = np.asfortranarray([[8.0, -4.0, 0.0]])
    v = z3(1)
    u = z3(1)
    cons = np.asfortranarray([[0.0, 0.0, 1.0]])
    oracle.timeintegrationADR(1, DT, v, F, m, 0.0, v, u, cons)
    assert list(v[0]) == [1.0, -0.5, 0.0]          # v = DT F / (2m)
    assert list(u[0]) == [1.0, -0.5, 0.0]
    cn = 0.5
    oracle.timeintegrationADR(2, DT, v, F, m, cn, v, u, cons)
    # v2 = ((2 - cn DT) v1 + 2 DT F/m) / (2 + cn DT) = (1.5*1 + 4)/2.5 = 2.2
    assert v[0, 0] == pytest.approx(2.2, rel=1e-15) and u[0, 0] == pytest.approx(3.2, rel=1e-15)
    assert v[0, 2] == 0.0 and u[0, 2] == 0.0


def test_mass_vector(oracle):
    """buildmassmatrix.m:32-40: m = 1/4 DT^2 (nFAM DX^2)(c / DX).  With DX = 0.5, c = 8, DT = 2 and
    nFAM = 10 that is 0.25 * 4 * 2.5 * 16 = 40 (DX in place of DX^2 would give 80, DT in place of
    DT^2 gives 20); the mass is linear in nFAM, scales with DX at fixed c and with DT^2."""
    m = oracle.buildmassmatrix(0.5, 8.0, 2.0, np.array([10.0, 3.0]))
    assert m[0] == 40.0
    assert m[1] == 12.0
    assert oracle.buildmassmatrix(1.0, 8.0, 2.0, np.array([10.0]))[0] == 80.0
    assert oracle.buildmassmatrix(0.5, 8.0, 6.0, np.array([10.0]))[0] == 360.0


# ---------------------------------------------------------------- contact
def test_contact_projection_and_reaction(oracle):
    X = np.asfortranarray([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0], [5.0, 0.0, 0.0]])
    x = X.copy(order="F")
    u, v = z3(3), z3(3)
    rho = np.array([2.0, 2.0, 2.0])
    fam = np.array([1.0, 2.0, 3.0])
    R, cz0, DT, V = 1.0, 0.75, 0.5, 0.125           # node 1 sits 0.25 inside a circle centred above
    fz = oracle.calculatecontactforce(fam, 0.0, cz0, R, 0.0, X, x, u, v, DT, V, rho)
    assert x[0, 0] == 0.0 and x[0, 2] == -0.25       # projected straight down to the circle
    assert u[0, 2] == -0.25 and v[0, 2] == -0.5      # v = (u - u_old)/DT
    assert u[0, 1] == 0.0 and v[0, 1] == 0.0         # no motion across the width
    d = math.hypot(0.3, 0.75)
    assert x[1, 0] == pytest.approx(0.3 / d, rel=1e-15)
    assert x[2, 0] == 5.0                            # outside: untouched
    dvz = [v[0, 2], v[1, 2]]
    expected = sum(-1 * 2.0 * dv / DT * V for dv in dvz)   # reaction = momentum change
    assert fz == pytest.approx(expected, rel=1e-15)
    # the centre move is not cumulative: the same displacement again moves nothing further
    x2, u2, v2 = x.copy(order="F"), u.copy(order="F"), v.copy(order="F")
    oracle.calculatecontactforce(fam, 0.0, cz0, R, -0.1, X, x, u, v, DT, V, rho)
    oracle.calculatecontactforce(fam, 0.0, cz0, R, -0.1, X, x2, u2, v2, DT, V, rho)
    oracle.calculatecontactforce(fam, 0.0, cz0, R, -0.1, X, x2, u2, v2, DT, V, rho)
    assert x2[0, 2] == x[0, 2] == pytest.approx(-0.35, rel=1e-15)


def test_smoothstep(oracle):
    N, a = 1001, -1e-3
    assert oracle.smoothstepdata(1, 1, N, 0, a) == 0.0
    assert oracle.smoothstepdata(N, 1, N, 0, a) == a
    assert oracle.smoothstepdata(501, 1, N, 0, a) == a / 2
    d1 = oracle.smoothstepdata(2, 1, N, 0, a)
    assert abs(d1) < 1e-8 * abs(a)                      # zero slope at the start
    assert abs(a - oracle.smoothstepdata(N - 1, 1, N, 0, a)) < 1e-8 * abs(a)


def test_damage_field_and_energy(oracle):
    BL = np.asfortranarray([[1.0, 2], [1, 3], [2, 3]])
    unbroken = np.array([0.0, 1.0, 0.0])
    d = oracle.calculatedamage(BL, unbroken, np.array([2.0, 2.0, 2.0]))
    assert list(d) == [0.5, 1.0, 0.5]
    # damage energy at s >= sc equals c L0 V^2 times the softening-branch integral of (1-bsf)s
    s0, sc, k, alpha = 1e-4, 2e-3, 25.0, 0.25
    ss = np.linspace(s0, sc, 200001)
    n = ss.size
    bsf, H = z(n), z(n)
    oracle.calculatebsfexponential(ss, np.full(n, s0), np.full(n, sc), k, alpha, bsf, H, z(n), z(n))
    integral = np.trapezoid((1 - H) * ss, ss)
    hist = z(2)
    c, L0, V = np.array([3.0, 3.0]), np.array([2.0, 2.0]), 0.5
    oracle.calculatedamageenergy(c, np.full(2, s0), np.full(2, sc), L0, np.array([3e-3, 0.5e-4]), k, alpha, V, hist)
    assert hist[0] == pytest.approx(3.0 * 2.0 * 0.25 * integral, rel=1e-7)
    assert hist[1] == 0.0


# ---------------------------------------------------------------- the whole step
def test_step_order_single_bond(oracle):
    """Two nodes, one bond, Euler-Cromer: one full oracle step equals the hand sequence."""
    class Case:
        pass
    c = Case()
    c.X = np.asfortranarray([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    c.BONDLIST = np.asfortranarray([[1.0, 2.0]])
    c.L0, c.vcf, c.c, c.BONDTYPE, c.fail = [np.array([v]) for v in (1.0, 1.0, 2.0, 0.0, 1.0)]
    c.s0, c.sc, c.s1, c.k, c.alpha, c.sc_steel, c.failureFunctionality = 0.5, 1.0, 0.0, 25.0, 0.25, 1.0, 0.0
    c.DENSITY = np.array([1.0, 1.0])
    c.CONSTRAINTFLAG = np.asfortranarray([[1.0, 1.0, 1.0], [0.0, 0.0, 0.0]])
    c.mass = None
    c.cellVolume, c.DT, c.DAMPING = 1.0, 0.5, 0.0
    c.law, c.integrator = 1, 0
    c.bodies = []
    c.schedule, c.nTimeSteps, c.itersPerIncrement, c.increment, c.appliedDisplacement = 1, 2, 1, 0.0, 0.0
    run = oracle.OracleRun(c)
    run.x[1, 0] = 1.25                                  # pre-stretched: s = 0.25 (< s0, elastic)
    run.step(1)
    # f = 2 * 0.25 * 1 = 0.5 toward node 1 on node 2; node 1 is fixed
    assert run.F[0, 0] == 0.5 and run.F[1, 0] == -0.5
    assert run.v[1, 0] == -0.25 and run.u[1, 0] == -0.125 and run.x[1, 0] == 0.875
    assert run.u[0, 0] == 0.0


# ---------------------------------------------------------------- the device's summation orders
def test_tree_sum_hand_values(oracle):
    """The halving tree adds v[t] + v[t + n/2] first.  2^53 + 1 rounds back to 2^53 (a tie, to even),
    so left to right [2^53, 1, 1, 0] sums to 2^53, while the tree pairs (2^53 + 1) and (1 + 0) and
    gives 2^53 again; put the ones in the pair v[1], v[3] and the tree adds them first: 2^53 + 2."""
    b = 2.0 ** 53
    assert oracle.tree_sum([b, 1.0, 1.0, 0.0]) == b
    assert oracle.tree_sum([b, 1.0, 0.0, 1.0]) == b + 2
    assert ((b + 1.0) + 0.0) + 1.0 == b                      # the sequential sum of the latter
    assert oracle.tree_sum([3.0]) == 3.0
    v = np.arange(1024, dtype=np.float64)
    assert oracle.tree_sum(v) == 1023 * 1024 / 2             # exact in any order


def test_slot_sum_order(oracle):
    """Blocks of 256 slots by the tree, then lane t of 256 adds the block sums t, t + 256, ... in
    order: with 257 blocks, block 0 and block 256 meet in lane 0 before the lane tree."""
    b = 2.0 ** 53
    v = np.zeros(257 * 256)
    v[0] = b                 # block 0
    v[1] = 1.0               # block 0, pairs with slot 129 at the first tree level
    v[129] = 1.0
    assert oracle.slot_sum(v) == b + 2
    w = np.zeros(257 * 256)
    w[0] = b                 # block 0 -> lane 0
    w[256 * 256] = 1.0       # block 256 -> lane 0, added to block 0's sum: a tie, stays 2^53
    w[256 * 1] = 1.0         # block 1 -> lane 1; the lane tree adds lanes 0 and 1 last... (b + 1)
    assert oracle.slot_sum(w) == b
    w2 = w.copy()
    w2[256 * 256] = 0.0
    w2[256 * 129] = 1.0      # block 129 -> lane 129, paired with lane 1 at the first lane level
    assert oracle.slot_sum(w2) == b + 2
    rng = np.random.default_rng(5)
    x = rng.standard_normal(300 * 256) * 10.0 ** rng.integers(-8, 8, 300 * 256)
    assert abs(oracle.slot_sum(x) - math.fsum(x)) <= 300 * 256 * 2.0 ** -52 * np.abs(x).sum()


def _adr_arrays(n, seed):
    rng = np.random.default_rng(seed)
    F = np.asfortranarray(rng.standard_normal((n, 3)))
    Fp = np.asfortranarray(rng.standard_normal((n, 3)))
    vp = np.asfortranarray(rng.standard_normal((n, 3)))
    vp[rng.random((n, 3)) < 0.1] = 0.0
    u = np.asfortranarray(rng.standard_normal((n, 3)) * 1e-3)
    m = rng.uniform(1.0, 2.0, n)
    return F, m, Fp, vp, u


def test_device_damping_sums(oracle):
    """The device-order ADR sums: for one node (the only nonzero slot) they equal the reference's
    sequential quotient bit for bit; over many nodes they agree with a correctly rounded sum of the
    same per-node terms; skipped nodes and constrained DOFs (the residual) are left out."""
    n = 3000
    F, m, Fp, vp, u = _adr_arrays(n, 1)
    DT = 1.0
    one = np.zeros(n, np.uint8) + 1
    one[17] = 0
    slots = np.arange(n, dtype=np.int32)
    c1, c2, r = oracle.dampingsums_device(F, m, Fp, DT, vp, u, one, None, slots, 4096)
    assert oracle.damping_from_sums(c1, c2) == oracle.calculatedampingcoefficient_masked(F, m, Fp, DT, vp, u, one)
    assert r == F[17, 0] ** 2 + F[17, 1] ** 2 + F[17, 2] ** 2
    none = np.zeros(n, np.uint8)
    cons = np.zeros((n, 3), order="F")
    cons[::7, 2] = 1
    c1, c2, r = oracle.dampingsums_device(F, m, Fp, DT, vp, u, none, cons, slots, 4096)
    t1 = [-(u[i, d] ** 2) * ((F[i, d] / m[i] - Fp[i, d] / m[i]) / (DT * vp[i, d]))
          for i in range(n) for d in range(3) if vp[i, d] != 0]
    t2 = [u[i, d] ** 2 for i in range(n) for d in range(3)]
    t3 = [F[i, d] ** 2 for i in range(n) for d in range(3) if cons[i, d] != 1]
    for got, terms in ((c1, t1), (c2, t2), (r, t3)):
        assert abs(got - math.fsum(terms)) <= 8 * len(terms) * 2.0 ** -53 * math.fsum(abs(t) for t in terms)
    # a permuted slot map changes nothing but the order
    perm = np.random.default_rng(2).permutation(4096)[:n].astype(np.int32)
    c1p, c2p, rp = oracle.dampingsums_device(F, m, Fp, DT, vp, u, none, cons, perm, 4096)
    assert c2p == pytest.approx(c2, rel=1e-13) and rp == pytest.approx(r, rel=1e-13)


def test_static_reaction_order(oracle):
    """1024 lanes, member t + 1024 j added to lane t in order, then the tree: with 2^53 at member 0
    and ones at members 1 and 513 (lanes paired at the first level), 2^53 + 2 per unit volume."""
    b = 2.0 ** 53
    n = 3000
    F = np.zeros((n, 3), order="F")
    F[0, 2], F[1, 2], F[513, 2] = b, 1.0, 1.0
    fam = np.arange(1, n + 1, dtype=np.float64)
    assert oracle.static_reaction(F, fam, 1.0) == b + 2
    F2 = np.zeros((n, 3), order="F")
    F2[0, 2], F2[1024, 2], F2[1, 2] = b, 1.0, 1.0      # member 1024 joins lane 0 first: a tie
    assert oracle.static_reaction(F2, fam, 1.0) == b
    rng = np.random.default_rng(3)
    F3 = np.asfortranarray(rng.standard_normal((n, 3)))
    sub = np.sort(rng.choice(n, 700, replace=False) + 1).astype(np.float64)
    exact = math.fsum(F3[sub.astype(int) - 1, 2] * 0.125)
    assert oracle.static_reaction(F3, sub, 0.125) == pytest.approx(exact, rel=1e-13)


def test_staircase_and_equilibrium_ratio(oracle):
    assert oracle.staircase(1, 1000, -5e-6) == -5e-6
    assert oracle.staircase(1000, 1000, -5e-6) == -5e-6
    assert oracle.staircase(1001, 1000, -5e-6) == -1e-5
    assert oracle.staircase(7, 3, 2.0) == 6.0
    assert oracle.equilibrium_ratio(16.0, 0.5, -4.0) == 0.5        # sqrt(16) * 0.5 / |-4|
    assert oracle.equilibrium_ratio(1.0, 1.0, 0.0) == math.inf


def _adr_beam(schedule, ipi, **kw):
    from bb_pd_amd import _lib, beam
    DX = 5e-3
    case = beam.build_beam("t", 25, 4, 10, DX, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01, sc=2e-3,
                           integrator=_lib.ADR, iters_per_increment=ipi, increment=-4e-6,
                           supports_x=(3 * DX, 23 * DX), penetrator_x=13 * DX, schedule=schedule, **kw)
    return case


def test_equilibrium_staircase_logic(oracle):
    """The equilibrium staircase with tolerance 0 advances every ipi iterations: the same
    displacements, and so the same trajectory bit for bit, as the plain staircase, with the
    increment ends at the same steps.  With a tolerance of +inf every state passes, and it advances
    every ipi_min iterations."""
    a = oracle.OracleRun(_adr_beam(1, 7), device_order=True)
    b = oracle.OracleRun(_adr_beam(2, 7, eq_tol=0.0), device_order=True)
    a.step(40)
    b.step(40)
    assert a.records["disp"] == b.records["disp"]
    assert a.records["inc_end"] == b.records["inc_end"] == [int(i % 7 == 0) for i in range(1, 41)]
    assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v)
    assert b.load_state() == (6, 40 - 35)
    c = oracle.OracleRun(_adr_beam(2, 50, eq_tol=math.inf, iters_per_increment_min=3), device_order=True)
    c.step(12)
    r = c.records
    assert r["reaction_static"][0] == 0.0 and r["ratio"][0] == math.inf   # u = 0 at step 1
    ends = [i + 1 for i, e in enumerate(r["inc_end"]) if e]
    assert ends == [3, 6, 9, 12]
    assert r["disp"] == [-4e-6 * (1 + (i - 1) // 3) for i in range(1, 13)]


def test_device_order_trajectory_close_to_reference_order(oracle):
    """The device summation order changes only the rounding of the ADR sums and of the static
    reaction: over the first 5 ADR steps the two orders' states agree to rounding-level
    differences (later, ADR's damping estimate amplifies rounding differences, which is why a
    bitwise comparison with the device needs the device's order)."""
    case = _adr_beam(1, 20)
    a = oracle.OracleRun(case)
    b = oracle.OracleRun(case, device_order=True)
    a.step(5)
    b.step(5)
    assert np.abs(a.u - b.u).max() <= 1e-12 * np.abs(a.u).max()
    ra, rb = np.array(a.records["reaction_static"]), np.array(b.records["reaction_static"])
    assert np.abs(ra - rb).max() <= 1e-12 * np.abs(ra).max()
    assert np.array(a.records["cn"]) == pytest.approx(np.array(b.records["cn"]), rel=1e-12, abs=1e-300)


# ---------------------------------------------------------------- strain energy
def test_strain_energy_hand_value(oracle):
    """One bond, s = 0.25, c = 2, L0 = 1, V = 1, vcf = 1: micropotential 0.5*2*0.0625 = 0.0625, each
    node's density half of it, total sum(density) * nNodes * V = 0.0625 * 2 = 0.125."""
    BL = np.asfortranarray([[1.0, 2.0]])
    e, d = oracle.calculatestrainenergy(2, np.ones(1), np.array([2.0]), np.array([0.25]), np.ones(1), np.ones(1),
                                        BL, 1.0)
    assert list(d) == [0.03125, 0.03125]
    assert e == 0.125


def test_strain_energy_dilation_bulk_node(oracle):
    """Uniform dilation s = e of every bond with c = c0 and vcf = 1: a bulk node's density is
    1/2 * sum over its bonds of 1/2 c0 e^2 |xi| V, with the sum of |xi| over the integer lattice
    offsets inside the horizon counted here by brute force (DX = 1); doubling e quadruples every
    density exactly."""
    from bb_pd_amd import beam
    case = beam.build_beam("t", 15, 15, 15, 1.0, law=1, sc=1.0, integrator=0, schedule=0)
    e0 = 1e-3
    c = np.full(case.nBonds, case.c0)
    one = np.ones(case.nBonds)
    _, d = oracle.calculatestrainenergy(case.nNodes, case.fail, c, np.full(case.nBonds, e0), case.L0, one,
                                        case.BONDLIST, case.cellVolume)
    r = np.arange(-4, 5)
    o = np.sqrt((r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2).astype(float)).ravel()
    o = o[(o > 0) & (o <= case.horizon)]
    assert o.size == 122
    centre = 7 + 15 * (7 + 15 * 7)
    expect = 0.5 * 0.5 * case.c0 * e0 ** 2 * math.fsum(o) * case.cellVolume
    assert d[centre] == pytest.approx(expect, rel=1e-13)
    _, d2 = oracle.calculatestrainenergy(case.nNodes, case.fail, c, np.full(case.nBonds, 2 * e0), case.L0, one,
                                         case.BONDLIST, case.cellVolume)
    assert np.array_equal(d2, 4 * d)
