"""Fused device-resident solver (layer 2) against the oracle's reference step sequence.

Bars
  * plain FP64, Euler-Cromer, bilinear / trilinear / brittle laws: the whole trajectory (u, v, x,
    nodal force, softening factor, flags, fail, reaction) is bit-identical to the oracle's serial
    reference order, step after step;
  * exponential law: exp() differs from glibc by ulps, so 1e-10 relative (north-star tolerance for
    nodal forces) with identical flags on these short runs;
  * ADR (stencil layout): bit-identical to the oracle run with the device's summation order for the
    damping sums and the static reaction (OracleRun(device_order=True)), per step and free-running,
    including the equilibrium staircase's increment ends and the 20,000-node beam;
  * FP32-mixed: stated error against FP64 on every result, flips of near-threshold bonds counted.
"""
import numpy as np
import pytest

from bb_pd_amd import _lib, beam
from bb_pd_amd.solver import Solver

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.int64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def ec_case(law, nx=24, ny=6, nz=10, steps=300, **kw):
    params = {
        _lib.LAW_BILINEAR: dict(s0=1e-4, s0_perturb=0.01, sc=2e-3),
        _lib.LAW_TRILINEAR: dict(s0=1e-4, s1=5e-4, sc=3e-3),
        _lib.LAW_EXPONENTIAL: dict(s0=1e-4),
        _lib.LAW_LINEAR: dict(sc=1e-3),
    }[law]
    params.update(kw)
    return beam.build_beam("t", nx, ny, nz, 5e-3, law=law, integrator=_lib.EULER_CROMER, schedule=0,
                           nTimeSteps=steps, appliedDisplacement=-1.5e-4, notch=(nx / 2 + 0.5, 3.5), **params)


def compare_state(sol, run, exact=True, rtol=1e-10):
    u, v, F, x = sol.nodal()
    Hs, fl, fa = sol.bonds()
    if exact:
        assert same(u, run.u), "u"
        assert same(v, run.v), "v"
        assert same(x, run.x), "x"
        assert same(F, run.F), "F"
        assert same(Hs, run.H), "H"
    else:
        for name, a, b in (("u", u, run.u), ("v", v, run.v), ("F", F, run.F), ("H", Hs, run.H)):
            scale = max(np.abs(b).max(), 1e-300)
            err = np.abs(a - b).max() / scale
            assert err <= rtol, (name, err)
    assert np.array_equal(fl, run.flag), "flags"
    assert np.array_equal(fa, run.fail), "fail"


@pytest.mark.parametrize("law", [_lib.LAW_BILINEAR, _lib.LAW_TRILINEAR, _lib.LAW_LINEAR])
def test_fp64_euler_cromer_bit_exact(gpu, oracle, law):
    case = ec_case(law)
    print(f"case: EC law {law}, {case.nNodes} nodes {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case)
    with Solver(case, _lib.FP64) as sol:
        for blk in range(3):
            sol.step(100)
            run.step(100)
            compare_state(sol, run, exact=True)
            h = sol.history()
            assert same(h["reaction"], run.reactions[-100:])
            # the last row against the oracle's state: deflection and CMOD are plain reads, the
            # static reaction a tree sum over the body-0 family (calculatereactionforceFast.m:3-5)
            ref = case.referenceNode - 1
            c1, c2 = (i - 1 for i in case.cmod)
            assert same(h["deflection"][-1:], [run.u[ref, 2]])
            assert same(h["cmod"][-1:], [run.u[c2, 0] - run.u[c1, 0]])
            fam = case.bodies[0].family.astype(int) - 1
            Fz = (run.F[fam, 2] * case.cellVolume).sum()
            assert h["reaction_static"][-1] == pytest.approx(Fz, rel=1e-12, abs=1e-12 * np.abs(run.F).max())
        obs = sol.observables()
        assert obs["step"] == 300
        assert obs["n_softening"] == int(run.flag.sum())
        assert obs["n_failed"] == int((run.fail == 0).sum())
        assert obs["n_broken"] == int(((run.H == 1) | (run.fail == 0)).sum())
        if law != _lib.LAW_LINEAR:
            assert obs["n_softening"] > 0 and obs["n_broken"] > 0
        else:
            assert obs["n_failed"] > 0


def oracle_state(run):
    return dict(u=run.u.copy(order="F"), v=run.v.copy(order="F"), x=run.x.copy(order="F"),
                F=run.F.copy(order="F"), Fp=run.Fp.copy(order="F"), H=run.H.copy(), flag=run.flag.copy(),
                fail=run.fail.copy())


def test_fp64_exponential_per_step(gpu, oracle):
    """Exponential law: from the oracle's exact state at steps 100, 200 and 299 the device step
    gives nodal forces within 1e-12 relative (only exp() differs) and identical flags."""
    case = ec_case(_lib.LAW_EXPONENTIAL)
    print(f"case: EC exponential, {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case)
    with Solver(case, _lib.FP64) as sol:
        for k in (100, 200, 299):
            run.step(k - run.step_count)
            sol.restore(k, **oracle_state(run))
            sol.step(1)
            run.step(1)
            compare_state(sol, run, exact=False, rtol=1e-12)
        assert sol.observables()["n_softening"] > 0
    # free-running drift over 300 post-peak steps, printed for the record
    run2 = oracle.OracleRun(case)
    with Solver(case, _lib.FP64) as sol:
        sol.step(300)
        run2.step(300)
        F = sol.nodal()[2]
        drift = np.abs(F - run2.F).max() / np.abs(run2.F).max()
        print(f"exponential law, 300 free-running steps: max |dF|/max|F| = {drift:.2e}")
        assert drift < 1e-6


@pytest.mark.parametrize("integrator", [_lib.EULER_CROMER, _lib.ADR])
def test_checkpoint_resume_bitwise(gpu, integrator):
    """Stop at step 100, checkpoint, resume in a fresh solver: identical to an uninterrupted run."""
    if integrator == _lib.ADR:
        case = beam.build_beam("t", 24, 6, 10, 5e-3, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01, sc=2e-3,
                               integrator=_lib.ADR, iters_per_increment=50, increment=-1e-5)
    else:
        case = ec_case(_lib.LAW_BILINEAR)
    with Solver(case, _lib.FP64) as a:
        a.step(100)
        ck = a.checkpoint()
        a.step(100)
        ref = a.checkpoint()
    with Solver(case, _lib.FP64) as b:
        step = ck.pop("step")
        assert step == 100
        b.restore(step, **ck)
        b.step(100)
        got = b.checkpoint()
    assert got["step"] == ref["step"] == 200
    for k in ("u", "v", "x", "F", "Fp", "H", "flag", "fail"):
        assert same(got[k], ref[k]), k


def test_restore_rewrites_history_words(gpu):
    """The stencil bond kernel takes the history of an offset as +0, without loading it, when the
    slice's history words say that every entry there holds +0; the host rewrites the words with every
    history it uploads.  A damaged state restored into a fresh solver (no damage yet) and an undamaged
    state restored over a damaged one both continue exactly as the uninterrupted runs."""
    case = ec_case(_lib.LAW_BILINEAR)
    with Solver(case, _lib.FP64) as a:
        assert a.layout().startswith("stencil")
        c0 = a.checkpoint()
        a.step(200)
        ck = a.checkpoint()
        a.step(100)
        ref = a.checkpoint()
        assert (ck["H"] > 0).sum() > 0 and ck["flag"].sum() > 0
        # undamaged state over a damaged solver, then the same 300 steps as from set-up
        a.restore(c0.pop("step"), **c0)
        a.step(300)
        again = a.checkpoint()
    with Solver(case, _lib.FP64) as b:
        b.restore(ck.pop("step"), **ck)
        b.step(100)
        got = b.checkpoint()
    for k in ("u", "v", "x", "F", "H", "flag", "fail"):
        assert same(got[k], ref[k]), k
        assert same(again[k], ref[k]), k


def adr_case(ipi=400):
    """Symmetric 25 x 6 x 10 beam on node-column supports under ADR, loaded in increments of 2 um per
    `ipi` iterations.  Frictionless cylinders leave rigid x translation and rotation about z free,
    so u_x is held on the midspan plane (and u_y on one of its lines), as CONSTRAINTFLAG does."""
    DX = 5e-3
    case = beam.build_beam("t", 25, 6, 10, DX, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01, sc=2e-3,
                           integrator=_lib.ADR, iters_per_increment=ipi, increment=-2e-6,
                           supports_x=(3 * DX, 23 * DX), penetrator_x=13 * DX)
    mid = np.isclose(case.X[:, 0], 13 * DX)
    cons = np.zeros((case.nNodes, 3), order="F")
    cons[mid, 0] = 1
    cons[mid & np.isclose(case.X[:, 1], DX), 1] = 1
    case.CONSTRAINTFLAG = cons
    return case


HIST_EXACT = ("reaction", "reaction_static", "deflection", "cmod", "disp", "ratio", "cn")


def compare_history(h, run, n, case):
    """The last n rows of the device history against the oracle's records, bit for bit."""
    assert len(h["reaction_static"]) == n
    ref = case.referenceNode - 1
    assert same(h["reaction"], run.reactions[-n:]), "reaction"
    for k in ("reaction_static", "disp", "ratio", "cn"):
        assert same(h[k], run.records[k][-n:]), k
    assert np.array_equal(h["inc_end"], np.array(run.records["inc_end"][-n:], dtype=bool)), "inc_end"
    assert same(h["deflection"][-1:], [run.u[ref, 2]])
    c1, c2 = (i - 1 for i in case.cmod)
    assert same(h["cmod"][-1:], [run.u[c2, 0] - run.u[c1, 0]])


def test_fp64_adr_per_step(gpu, oracle):
    """ADR: from the oracle's exact state, one device step (force, damping coefficient over the nodes
    no contact body drives, leapfrog update, constraints, contact, static reaction) equals the
    oracle's step bit for bit.  The oracle takes the damping sums and the static reaction in the
    device's summation order (or_slot_sum, or_static_reaction), which is what makes ADR comparable
    bit for bit."""
    case = adr_case()
    print(f"case: ADR bilinear per step, {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        assert sol.layout() == "stencil"
        for k in (0, 1, 40, 399, 400, 401, 700):
            run.step(k - run.step_count)
            st = oracle_state(run)
            sol.restore(k, **st)
            sol.history()
            sol.step(1)
            run.step(1)
            compare_state(sol, run, exact=True)
            compare_history(sol.history(), run, 1, case)
            assert same([sol.observables()["cn"]], run.records["cn"][-1:]), k


def test_fp64_adr_increments(gpu, oracle):
    """ADR free-running over three load increments of 400 iterations: the whole state after every
    increment and every per-step history row (reactions, deflection, CMOD, displacement, equilibrium
    ratio, damping coefficient, increment ends) bit for bit."""
    ipi = 400
    case = adr_case(ipi)
    print(f"case: ADR bilinear 3 increments, {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        for inc in range(3):
            sol.step(ipi)
            run.step(ipi)
            compare_state(sol, run, exact=True)
            h = sol.history()
            compare_history(h, run, ipi, case)
            assert h["inc_end"][-1] and h["reaction_static"][-1] > 100 * (inc + 1)
        obs = sol.observables()
        assert obs["step"] == 3 * ipi
        assert obs["n_softening"] == int(run.flag.sum())


def test_fp64_adr_equilibrium_staircase(gpu, oracle):
    """The equilibrium staircase: each increment ends once the equilibrium ratio falls below 1e-4
    (after at least 20 iterations, at most 400).  The increment ends, the displacements and the
    whole trajectory equal the oracle's bit for bit; restoring the load state resumes it."""
    case = adr_case(400)
    case.schedule, case.eqTol, case.itersPerIncrementMin = 2, 1e-4, 20
    print(f"case: ADR equilibrium staircase, {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        for blk in range(3):
            sol.step(500)
            run.step(500)
            compare_state(sol, run, exact=True)
            compare_history(sol.history(), run, 500, case)
            assert sol.load_state() == run.load_state()
        ends = np.nonzero(run.records["inc_end"])[0]
        assert len(ends) >= 4
        assert min(np.diff(ends)) < 400           # some increments converged before the cap
        ck = sol.checkpoint()
        sol.step(200)
        run.step(200)
        h_ref = sol.history()
        sol.restore(ck["step"], **{k: v for k, v in ck.items() if k != "step"})
        sol.step(200)
        h = sol.history()
        for k in HIST_EXACT:
            assert same(h[k], h_ref[k]), k
        compare_state(sol, run, exact=True)


def test_config0_adr_bit_exact(gpu, oracle):
    """The 20,000-node beam (100 x 10 x 20 lattice, dx = 5 mm, 1,005,544 bonds, bilinear law with the
    seeded +-1% s0 perturbation, sc = 2e-3) under ADR on the staircase of 5 um per 1000 iterations:
    2,500 iterations (two and a half load increments).  Nodal forces, displacements, velocities,
    softening factors, flags and the load-deflection record equal the oracle's bit for bit; the
    asserted bar is the 1e-10 relative tolerance the comparison must meet at least, checked first."""
    case = beam.case_bilinear_dx5()
    assert case.integrator == _lib.ADR and case.itersPerIncrement == 1000
    print(f"case: config-0 beam ADR, {case.nNodes} nodes {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        for blk in range(5):
            sol.step(500)
            run.step(500)
            compare_state(sol, run, exact=False, rtol=1e-10)
            compare_state(sol, run, exact=True)
            compare_history(sol.history(), run, 500, case)
            print(f"  {run.step_count} steps: load {run.records['reaction_static'][-1]:.3f} N, "
                  f"{int(run.flag.sum())} softening bonds", flush=True)
        assert run.records["reaction_static"][-1] > 0


def test_full_beam_dx5_bit_exact(gpu, oracle):
    """The 20,000-node beam: neighbour list, flags and state bit-exact after 200 EC steps."""
    case = beam.case_bilinear_dx5(integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=400,
                                  appliedDisplacement=-2e-4)
    print("case: 100x10x20 EC bilinear", flush=True)
    run = oracle.OracleRun(case)
    with Solver(case, _lib.FP64) as sol:
        sol.step(200)
        run.step(200)
        compare_state(sol, run, exact=True)


def test_graph_replay_equals_eager(gpu):
    case = ec_case(_lib.LAW_BILINEAR)
    with Solver(case, _lib.FP64) as a, Solver(case, _lib.FP64) as b:
        a.step(120)
        b.run_graph(40, 3)
        b.sync()
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        assert same(a.bonds()[0], b.bonds()[0])
        # eager work between replays does not disturb them
        a.step(40)
        b.run_graph(40, 1)
        assert same(a.nodal()[0], b.nodal()[0])


def test_node_order_does_not_matter(gpu, oracle):
    """Renumbered nodes and a shuffled, re-oriented bond list: still bit-exact to the oracle on the
    same (renumbered) input, i.e. the Morton slots and rows follow the given bond order."""
    case = ec_case(_lib.LAW_BILINEAR, steps=200)
    rng = np.random.default_rng(3)
    nN = case.nNodes
    perm = rng.permutation(nN)          # new id of old node n is inv[n]
    inv = np.empty(nN, dtype=np.int64)
    inv[perm] = np.arange(nN)
    case.X = np.asfortranarray(case.X[perm])
    case.DENSITY = case.DENSITY[perm]
    BL = inv[case.BONDLIST.astype(np.int64) - 1] + 1
    order = rng.permutation(case.nBonds)
    BL = BL[order]
    flip = rng.uniform(size=case.nBonds) < 0.5
    BL[flip] = BL[flip][:, ::-1]
    case.BONDLIST = np.asfortranarray(BL.astype(np.float64))
    for name in ("L0", "vcf", "c", "BONDTYPE", "fail", "s0"):
        setattr(case, name, getattr(case, name)[order])
    for b in case.bodies:
        b.family = np.sort(inv[b.family.astype(np.int64) - 1] + 1).astype(np.float64)
    case.referenceNode = int(inv[case.referenceNode - 1] + 1)
    case.cmod = tuple(int(inv[i - 1] + 1) for i in case.cmod)
    run = oracle.OracleRun(case)
    with Solver(case, _lib.FP64) as sol:
        sol.step(150)
        run.step(150)
        compare_state(sol, run, exact=True)


def test_setup_inconsistent_lengths(gpu):
    """An UNDEFORMEDLENGTH that differs from its formula: the FP64 form takes it as given (record
    layout, streamed; its trajectory is checked against the oracle below), the mixed form refuses the
    problem and names the bond."""
    case = ec_case(_lib.LAW_BILINEAR)
    bad = case.L0.copy()
    bad[5] *= 1 + 1e-15
    case.L0 = bad
    with Solver(case, _lib.FP64) as s:
        assert s.layout() == "records (UNDEFORMEDLENGTH or VOLUMECORRECTIONFACTORS differ from their formulas)"
    with pytest.raises(_lib.BBPDError, match=r"UNDEFORMEDLENGTH\(6\)"):
        Solver(case, _lib.MIXED)


def test_mixed_precision_error_report(gpu):
    """FP32-mixed against plain FP64 on the same run: every result within its stated bound."""
    case = ec_case(_lib.LAW_BILINEAR, steps=400)
    with Solver(case, _lib.FP64) as a, Solver(case, _lib.MIXED) as b:
        a.step(400)
        b.step(400)
        ua, va, Fa, xa = a.nodal()
        ub, vb, Fb, xb = b.nodal()
        Ha, fla, _ = a.bonds()
        Hb, flb, _ = b.bonds()
        err_u = np.abs(ua - ub).max() / np.abs(ua).max()
        err_F = np.abs(Fa - Fb).max() / np.abs(Fa).max()
        flips = int((fla != flb).sum())
        broken_a, broken_b = int((Ha == 1).sum()), int((Hb == 1).sum())
        print(f"mixed vs fp64: u {err_u:.2e} F {err_F:.2e} flag flips {flips} broken {broken_a} vs {broken_b}")
        assert err_u < 2e-3
        assert flips <= max(5, 0.02 * fla.sum())
        assert abs(broken_a - broken_b) <= max(5, 0.02 * broken_a)


def _solver_with_layout(case, precision, layout):
    import os
    old = os.environ.get("BBPD_LAYOUT")
    try:
        os.environ["BBPD_LAYOUT"] = layout
        return Solver(case, precision)
    finally:
        if old is None:
            os.environ.pop("BBPD_LAYOUT", None)
        else:
            os.environ["BBPD_LAYOUT"] = old


@pytest.mark.parametrize("law", [_lib.LAW_BILINEAR, _lib.LAW_TRILINEAR, _lib.LAW_LINEAR, _lib.LAW_EXPONENTIAL])
def test_stencil_layout_equals_records(gpu, law):
    """The stencil layout (implicit neighbours, history only per entry) and the record layout give
    bit-identical FP64 trajectories: u, v, x, F, H, flags, fail, reaction history."""
    case = ec_case(law)
    a = _solver_with_layout(case, _lib.FP64, "stencil")
    b = _solver_with_layout(case, _lib.FP64, "records")
    with a, b:
        assert a.layout() == "stencil", a.layout()
        assert b.layout().startswith("records"), b.layout()
        a.step(250)
        b.step(250)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
        assert same(a.history()["reaction"], b.history()["reaction"])
        oa, ob = a.observables(), b.observables()
        for k in ("n_softening", "n_broken", "n_failed"):
            assert oa[k] == ob[k], k
        if law == _lib.LAW_LINEAR:
            assert oa["n_failed"] > 0
        else:
            assert oa["n_softening"] > 0


def test_stencil_bonds_failed_at_setup_equal_records(gpu):
    """Bonds given as failed at set-up (brittle law, fail = 0) hold -0 in the stencil layout's history,
    so set-up must mark their offsets in the slices' history words: the bond kernel takes +0 (intact)
    without a load for an offset whose bit is clear.  Both layouts give the same FP64 trajectory."""
    case = ec_case(_lib.LAW_LINEAR)
    case.fail = np.where(np.arange(case.nBonds) % 7 == 3, 0.0, 1.0)
    a = _solver_with_layout(case, _lib.FP64, "stencil")
    b = _solver_with_layout(case, _lib.FP64, "records")
    with a, b:
        assert a.layout() == "stencil", a.layout()
        assert a.observables()["n_failed"] == int((case.fail == 0).sum())
        a.step(150)
        b.step(150)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)


@pytest.mark.parametrize("horizon_factor", [4.6, 5.9])
def test_stencil_layout_wide_rows(gpu, horizon_factor):
    """A one-node-thick plate with a horizon of 4.6 or 5.9 spacings: bond offsets reach 4 or 5
    spacings along x, so each staged stencil row carries 8 or 10 halo sites beyond its 64 (the
    usual horizon, pi * DX, reaches 3).  The stencil layout still equals the record layout bit for
    bit in FP64."""
    case = beam.build_beam("t", 30, 1, 12, 5e-3, law=_lib.LAW_BILINEAR, s0=1e-4, sc=2e-3,
                           horizon_factor=horizon_factor, integrator=_lib.EULER_CROMER, schedule=0,
                           nTimeSteps=150, appliedDisplacement=-1.5e-4)
    print(f"case: plate, horizon {horizon_factor} dx, {case.nBonds} bonds", flush=True)
    a = _solver_with_layout(case, _lib.FP64, "stencil")
    b = _solver_with_layout(case, _lib.FP64, "records")
    with a, b:
        assert a.layout() == "stencil", a.layout()
        a.step(150)
        b.step(150)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
        assert np.abs(a.nodal()[0]).max() > 0


def _solver_with_env(case, precision, **env):
    import os
    old = {k: os.environ.get(k) for k in env}
    try:
        os.environ.update(env)
        return Solver(case, precision)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("dx", [5e-3, 1.25e-3, 7e-3, 6.25e-4])
def test_fp64_nominal_lengths_equal_per_entry(gpu, dx):
    """FP64 stencil kernel: the undeformed lengths, their reciprocals and the volume correction
    factors taken from per-offset nominal values (one correction step each, checked at set-up on
    every value they take on the lattice) give the ADR trajectory of computing them per entry with
    square root and division, bit for bit, on an exponential-law beam at four lattice spacings."""
    case = beam.build_beam("t", 24, 6, 10, dx, law=_lib.LAW_EXPONENTIAL, s0=1e-4, integrator=_lib.ADR,
                           iters_per_increment=100, increment=-4e-3 * dx,
                           notch=(12.5, 3.5))
    print(f"case: 24 x 6 x 10 beam, dx {dx}, {case.nBonds} bonds", flush=True)
    a = _solver_with_env(case, _lib.FP64)
    b = _solver_with_env(case, _lib.FP64, BBPD_FAST64="0")
    with a, b:
        assert a.layout() == "stencil", a.layout()
        assert b.layout() == "stencil (FP64 lengths computed per entry)", b.layout()
        a.step(600)
        b.step(600)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
        assert same(a.history()["reaction"], b.history()["reaction"])
        assert a.observables()["n_softening"] > 0


def test_stencil_zero_length_bond_equals_records(gpu):
    """Two neighbouring nodes moved onto one another: their bond has length 0, whose force the
    reference turns from NaN into 0 (calculatebondforcesMex.c:38-40).  The FP64 stencil kernel takes
    the deformed length and its reciprocal from the unscaled square-root and reciprocal sequences only
    when every lane of the wavefront is in their range, so this wavefront takes the library's IEEE
    sequences; the two layouts stay bit-identical and every force stays finite."""
    case = ec_case(_lib.LAW_BILINEAR)
    X = case.X
    DX = X[:, 0].min()
    n = int(np.flatnonzero(np.isclose(X[:, 0], 10 * DX) & np.isclose(X[:, 1], 3 * DX) & np.isclose(X[:, 2], 5 * DX))[0])
    m = int(np.flatnonzero(np.isclose(X[:, 0], 11 * DX) & np.isclose(X[:, 1], 3 * DX) & np.isclose(X[:, 2], 5 * DX))[0])
    a = _solver_with_layout(case, _lib.FP64, "stencil")
    b = _solver_with_layout(case, _lib.FP64, "records")
    with a, b:
        assert a.layout() == "stencil", a.layout()
        a.step(20)
        st = a.checkpoint()
        step = st.pop("step")
        st["x"][n] = st["x"][m]
        st["u"][n] = st["x"][n] - X[n]
        a.restore(step, **st)
        b.restore(step, **st)
        a.step(3)
        b.step(3)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
            assert np.isfinite(x).all()
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)


def test_stencil_layout_restore_and_adr(gpu):
    """A checkpoint of the stencil layout restores into the record layout (the history's sign-bit
    flag round-trips) and the two continue bit for bit under Euler-Cromer.  Under ADR the damping
    sum is a block tree over slots, whose order differs between the layouts, so one step from a
    shared state agrees to 1e-12."""
    case = ec_case(_lib.LAW_BILINEAR)
    a = _solver_with_layout(case, _lib.FP64, "stencil")
    b = _solver_with_layout(case, _lib.FP64, "records")
    with a, b:
        assert a.layout() == "stencil"
        a.step(150)
        ca = a.checkpoint()
        assert ca["flag"].sum() > 0
        step = ca.pop("step")
        b.restore(step, **ca)
        a.step(100)
        b.step(100)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
    case = adr_case()
    a = _solver_with_layout(case, _lib.FP64, "stencil")
    b = _solver_with_layout(case, _lib.FP64, "records")
    with a, b:
        assert a.layout() == "stencil"
        a.step(450)
        ca = a.checkpoint()
        step = ca.pop("step")
        b.restore(step, **ca)
        a.step(1)
        b.step(1)
        for x, y in zip(a.nodal(), b.nodal()):
            assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()
        assert a.observables()["cn"] == pytest.approx(b.observables()["cn"], rel=1e-12, abs=1e-14)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)


@pytest.mark.parametrize("law", [_lib.LAW_BILINEAR, _lib.LAW_TRILINEAR, _lib.LAW_EXPONENTIAL, _lib.LAW_LINEAR])
def test_stencil_layout_mixed_close_to_records(gpu, law):
    """Mixed form: the two layouts round differently (c*V*vcf and xi; the stencil kernel takes the
    stretch and the force factor from series in (|y|^2 - |xi|^2) / (2 |xi|^2), tests the law's threshold
    in that variable, and adds groups of entries in FP32 before the FP64 row sum), nothing else: after
    300 Euler-Cromer steps, for each law, both stay as close to the FP64 run as to each other."""
    case = ec_case(law, steps=300)
    a = _solver_with_layout(case, _lib.MIXED, "stencil")
    b = _solver_with_layout(case, _lib.MIXED, "records")
    c = _solver_with_layout(case, _lib.FP64, "records")
    with a, b, c:
        assert a.layout() == "stencil"
        for s in (a, b, c):
            s.step(300)
        ua, ub, uc = a.nodal()[0], b.nodal()[0], c.nodal()[0]
        scale = np.abs(uc).max()
        d_ab, d_ac, d_bc = (np.abs(x - y).max() / scale for x, y in ((ua, ub), (ua, uc), (ub, uc)))
        print(f"mixed stencil vs records {d_ab:.2e}, stencil vs FP64 {d_ac:.2e}, records vs FP64 {d_bc:.2e}")
        assert d_ab <= 1e-4 and d_ac <= 1e-4 and d_bc <= 1e-4
        fa, fb = a.bonds()[1], b.bonds()[1]
        assert abs(int(fa.sum()) - int(fb.sum())) <= max(3, 0.01 * fb.sum())
        na, nb = int((a.bonds()[2] == 0).sum()), int((b.bonds()[2] == 0).sum())
        assert abs(na - nb) <= max(3, 0.01 * nb)
        if law == _lib.LAW_LINEAR:
            assert nb > 0
        else:
            assert fb.sum() > 0


def test_layout_defaults_and_fallback(gpu):
    """Default: the stencil layout for both forms; BBPD_LAYOUT=records asks for the record layout.
    A bond list not sorted by node id takes the record layout even when the stencil one is asked
    for."""
    case = ec_case(_lib.LAW_BILINEAR)
    import os
    assert "BBPD_LAYOUT" not in os.environ
    with Solver(case, _lib.MIXED) as a, Solver(case, _lib.FP64) as b:
        assert a.layout() == "stencil"
        assert b.layout() == "stencil"
    with _solver_with_layout(case, _lib.FP64, "records") as r:
        assert r.layout() == "records (record layout requested)", r.layout()
    order = np.random.default_rng(1).permutation(case.nBonds)
    case.BONDLIST = np.asfortranarray(case.BONDLIST[order])
    for name in ("L0", "vcf", "c", "BONDTYPE", "fail", "s0"):
        setattr(case, name, getattr(case, name)[order])
    with _solver_with_layout(case, _lib.MIXED, "stencil") as s:
        assert s.layout().startswith("records (rows not in ascending"), s.layout()


def test_autotune_leaves_the_trajectory_unchanged(gpu):
    """The autotuner reruns the bond kernel on the live ADR state; every variant computes the same
    arithmetic, so the tuned run equals an untuned one bit for bit, in either form."""
    case = adr_case()
    for prec in (_lib.MIXED, _lib.FP64):
        a = _solver_with_layout(case, prec, "stencil")
        b = _solver_with_layout(case, prec, "stencil")
        with a, b:
            a.step(120)
            choice, ms = a.autotune(3)
            nv = _lib.AUTOTUNE_VARIANTS
            assert 0 <= choice < nv and len(ms) == nv and min(ms) > 0
            a.step(130)
            b.step(250)
            for x, y in zip(a.nodal(), b.nodal()):
                assert same(x, y)
            for x, y in zip(a.bonds(), b.bonds()):
                assert same(x, y)
    with Solver(ec_case(_lib.LAW_BILINEAR), _lib.MIXED) as e:
        with pytest.raises(_lib.BBPDError, match="ADR"):
            e.autotune(3)


def test_launch_variants_agree(gpu):
    """Every launch variant of the stencil bond kernel (minimum waves per SIMD; block orders: XCD
    remap, launch order, XCD stripes) gives the same ADR trajectory bit for bit, in both forms."""
    case = adr_case()
    for prec in (_lib.FP64, _lib.MIXED):
        ref = None
        for v in range(_lib.AUTOTUNE_VARIANTS):
            with _solver_with_env(case, prec, BBPD_LAUNCH_VARIANT=str(v)) as a:
                assert a.layout().startswith("stencil")
                a.step(300)
                out = list(a.nodal()) + list(a.bonds())
            if ref is None:
                ref = out
                continue
            for x, y in zip(ref, out):
                assert same(x, y), (prec, v)


@pytest.mark.parametrize("law", [_lib.LAW_EXPONENTIAL, _lib.LAW_TRILINEAR])
def test_damage_field_and_energy_match_oracle(gpu, oracle, law):
    """Device damage field and damage energy against the oracle, which keeps the reference's per-bond
    running max of calculatedamageenergy over every step (dynamicsolverdisplacementcontrolled.m:121)."""
    case = ec_case(law)
    run = oracle.OracleRun(case)
    hist = np.zeros(case.nBonds)
    s0 = np.ascontiguousarray(run.s0_full)
    sc = np.ascontiguousarray(run.sc_full)
    for _ in range(300):
        run.step(1)
        oracle.calculatedamageenergy(case.c, s0, sc, case.L0, run.s, case.k, case.alpha, case.cellVolume, hist)
    e_ref = hist.sum()
    assert e_ref > 0
    unbroken = (~((run.H == 1) | (run.fail == 0))).astype(np.float64)
    d_ref = oracle.calculatedamage(case.BONDLIST, unbroken, case.nFAM)
    for layout, prec, tol in (("records", _lib.FP64, 1e-9), ("stencil", _lib.FP64, 1e-9),
                              ("stencil", _lib.MIXED, 2e-3)):
        with _solver_with_layout(case, prec, layout) as sol:
            sol.step(300)
            d, e = sol.damage()
            print(f"{layout} {prec}: energy {e:.6e} vs {e_ref:.6e}, damage max {d.max():.3f}")
            assert e == pytest.approx(e_ref, rel=tol)
            if prec == _lib.FP64:
                assert np.array_equal(d, d_ref)
            else:
                assert np.abs(d - d_ref).max() <= 0.05


@pytest.mark.parametrize("law", [_lib.LAW_BILINEAR, _lib.LAW_LINEAR])
@pytest.mark.parametrize("layout", ["stencil", "records"])
def test_strain_energy_matches_oracle(gpu, oracle, law, layout):
    """calculatestrainenergy.m on the device: per-node densities and the total equal the oracle's bit
    for bit (the stretch of the current positions, bonds added in BONDLIST order, nodes summed in
    order); the mixed form's state gives the same energy to FP32 accuracy."""
    case = ec_case(law, steps=200)
    print(f"case: strain energy, law {law}, {layout}", flush=True)
    run = oracle.OracleRun(case)
    with _solver_with_layout(case, _lib.FP64, layout) as sol:
        sol.step(150)
        run.step(150)
        nB = case.nBonds
        l, dX, dY, dZ, s = (np.zeros(nB) for _ in range(5))
        oracle.calculatedeformedlength(run.x, case.BONDLIST, case.L0, l, dX, dY, dZ, s)
        e_ref, d_ref = oracle.calculatestrainenergy(case.nNodes, run.fail, case.c, s, case.L0, case.vcf,
                                                    case.BONDLIST, case.cellVolume)
        d, e = sol.strain_energy()
        assert same(d, d_ref)
        assert same([e], [e_ref]) and e > 0
    with _solver_with_layout(case, _lib.MIXED, layout) as sol:
        sol.step(150)
        d, e = sol.strain_energy()
        assert e == pytest.approx(e_ref, rel=2e-3)


def overlapping_case(integrator):
    """A beam whose contact families share members: a second penetrator with the first one's family and
    a 5% larger radius.  A member inside both circles is projected by the first, then pushed out by the
    second, as the reference's contact calls follow one another (dynamicsolverdisplacementcontrolled.m:
    108-112); in the other order the result differs."""
    import copy
    case = ec_case(_lib.LAW_BILINEAR, steps=200) if integrator == _lib.EULER_CROMER else adr_case(200)
    b2 = copy.deepcopy(case.bodies[0])
    b2.R = b2.R * 1.05
    case.bodies = case.bodies + [b2]
    return case


@pytest.mark.parametrize("integrator", [_lib.EULER_CROMER, _lib.ADR])
def test_overlapping_contact_families_in_order(gpu, oracle, integrator):
    case = overlapping_case(integrator)
    print(f"case: overlapping contact families, integrator {integrator}", flush=True)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        sol.step(150)
        run.step(150)
        compare_state(sol, run, exact=True)
        h = sol.history()
        assert same(h["reaction"], run.reactions)
        assert same(h["reaction_static"], run.records["reaction_static"])
        # the second penetrator moved members the first had placed
        fam = case.bodies[0].family.astype(int) - 1
        d = np.hypot(run.x[fam, 0] - case.bodies[3].cx,
                     run.x[fam, 2] - (case.bodies[3].cz + run.records["disp"][-1]))
        assert np.isclose(d, case.bodies[3].R, rtol=1e-12).any()


def test_overlapping_contact_families_in_a_batch(gpu):
    from bb_pd_amd.batch import Batch
    case = overlapping_case(_lib.ADR)
    other = adr_case(200)
    with Batch(case, _lib.FP64, max_regions=2, max_slots=8192) as b:
        r0 = b.admit(other)
        r1 = b.admit(case)
        b.step(120)
        got = b.nodal(r1)
        h = b.history(r1)
    with Solver(case, _lib.FP64) as s:
        assert s.layout() == "stencil"
        s.step(120)
        for x, y in zip(s.nodal(), got):
            assert same(x, y)
        hs = s.history()
        for k in ("reaction", "reaction_static", "deflection", "cmod"):
            assert same(hs[k], h[k]), k


@pytest.mark.parametrize("law", [_lib.LAW_BILINEAR, _lib.LAW_EXPONENTIAL])
def test_stencil_tabled_and_streamed_parameters(gpu, law):
    """s0 a function of the bond's key (the pair of family sizes, here through BONDSTIFFNESS) and sc
    per bond: the stencil layout keeps s0 in its key table and streams sc, and an entry's law
    threshold comes from its tabled s0 and the region's smallest sc.  The stream count shows in the
    kernel's bytes: none when both are functions of the key, one here, two when neither is.  The FP64
    trajectory equals the record layout's bit for bit; the mixed form's stays within rounding of it."""
    import copy
    base = ec_case(law)
    nB = base.nBonds
    c = np.asarray(base.c, dtype=np.float64)
    rng = np.random.default_rng(7)
    s0_key = 1e-4 * (1 + 0.02 * c / c.max())
    sc0 = np.asarray(base.sc, dtype=np.float64) if np.ndim(base.sc) else np.full(nB, float(base.sc))
    sc_key = sc0.max() * (1 + 0.03 * c / c.max())
    sc_bond = sc0 * (1 + 0.05 * rng.random(nB))
    s0_bond = s0_key * (1 + 0.01 * (2 * rng.random(nB) - 1))

    def variant(s0, sc):
        v = copy.copy(base)
        v.s0, v.sc = s0, sc
        return v

    mixed_case = variant(s0_key, sc_bond)
    by = {}
    for name, case in (("table", variant(s0_key, sc_key)), ("one stream", mixed_case),
                       ("two streams", variant(s0_bond, sc_bond))):
        with _solver_with_layout(case, _lib.FP64, "stencil") as s:
            assert s.layout() == "stencil", s.layout()
            inf = s.info()
            by[name] = (inf["bond_kernel_bytes"], inf["n_entries"])
    ne = by["table"][1]
    assert by["one stream"][0] - by["table"][0] == 8 * ne
    assert by["two streams"][0] - by["one stream"][0] == 8 * ne
    print(f"case: law {law}, s0 tabled, sc streamed, {nB} bonds", flush=True)
    a = _solver_with_layout(mixed_case, _lib.FP64, "stencil")
    b = _solver_with_layout(mixed_case, _lib.FP64, "records")
    m = _solver_with_layout(mixed_case, _lib.MIXED, "stencil")
    with a, b, m:
        for s in (a, b, m):
            s.step(300)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
        assert same(a.history()["reaction"], b.history()["reaction"])
        assert a.observables()["n_softening"] > 0
        ua, um = a.nodal()[0], m.nodal()[0]
        err = np.abs(ua - um).max() / np.abs(ua).max()
        print(f"  mixed vs FP64: u {err:.2e}", flush=True)
        assert err < 2e-3


def test_fp64_given_lengths_and_factors_streamed(gpu, oracle):
    """UNDEFORMEDLENGTH and VOLUMECORRECTIONFACTORS that differ from their formulas (one ulp on every
    third bond's length, a relative 2^-30 on every fifth partial factor) are taken as given: the FP64
    form runs the problem on the record layout with both streamed per entry, bit for bit with the
    oracle, which uses the given arrays (trajectory, strain energy); the mixed form, whose stretch
    comes from the coordinates, refuses them."""
    case = ec_case(_lib.LAW_BILINEAR, steps=300)
    nB = case.nBonds
    k = np.arange(nB)
    case.L0 = np.asarray(case.L0, dtype=np.float64).copy()
    case.L0[k % 3 == 1] = np.nextafter(case.L0[k % 3 == 1], np.inf)
    case.vcf = np.asarray(case.vcf, dtype=np.float64).copy()
    part = (k % 5 == 2) & (case.vcf < 1) & (case.vcf > 0)
    assert part.sum() > 0
    case.vcf[part] = case.vcf[part] * (1 + 2.0 ** -30)
    print(f"case: given lengths and factors, {nB} bonds", flush=True)
    run = oracle.OracleRun(case)
    with Solver(case, _lib.FP64) as sol:
        assert sol.layout() == "records (UNDEFORMEDLENGTH or VOLUMECORRECTIONFACTORS differ from their formulas)", \
            sol.layout()
        for blk in range(3):
            sol.step(100)
            run.step(100)
            compare_state(sol, run, exact=True)
        assert same(sol.history()["reaction"], run.reactions[-300:])
        l, dX, dY, dZ, s = (np.zeros(nB) for _ in range(5))
        oracle.calculatedeformedlength(run.x, case.BONDLIST, case.L0, l, dX, dY, dZ, s)
        e_ref, d_ref = oracle.calculatestrainenergy(case.nNodes, run.fail, case.c, s, case.L0, case.vcf,
                                                    case.BONDLIST, case.cellVolume)
        d, e = sol.strain_energy()
        assert same(d, d_ref) and same([e], [e_ref])
        assert sol.observables()["n_softening"] > 0
    with pytest.raises(_lib.BBPDError, match="UNDEFORMEDLENGTH"):
        Solver(case, _lib.MIXED)
