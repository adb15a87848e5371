"""The five beam workloads at their full sizes on the GPU.

  * dx = 2.5 mm notched trilinear beam (200 x 20 x 40 lattice, 8.9 M bonds): 150 Euler-Cromer steps
    and 30 ADR steps bit for bit against the oracle;
  * dx = 1.25 mm offset-notch exponential beam (400 x 40 x 80 lattice, 74.5 M bonds): one FP64 ADR
    step from a loaded, softening state against the oracle (exp() differs from glibc by ulps: 1e-12
    relative, flags identical), and 60 Euler-Cromer steps of the stencil layout against the record
    layout, bit for bit;
  * the size-effect batch (depths 50 / 100 / 200 / 400 mm x 16 seeds, 64 instances, 358 M bonds)
    admitted at once: one region per depth equals its single run bit for bit, and the 50 mm seed-0
    region equals the oracle bit for bit.
The serving workload (all four depths) is in test_gpu_serve.py.
"""
import copy
import os

import numpy as np
import pytest

from bb_pd_amd import _lib, beam
from bb_pd_amd.batch import Batch
from bb_pd_amd.solver import Solver

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.int64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def assert_state(sol, run, exact=True, rtol=0.0):
    u, v, F, x = sol.nodal()
    H, flag, fail = sol.bonds()
    for name, a, b in (("u", u, run.u), ("v", v, run.v), ("F", F, run.F), ("x", x, run.x), ("H", H, run.H)):
        if exact:
            assert same(a, b), name
        else:
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
            assert err <= rtol, (name, err)
    assert np.array_equal(flag, run.flag), "flags"
    assert np.array_equal(fail, run.fail), "fail"


def load_oracle_state(run, st):
    """Put a device checkpoint into the oracle's arrays (which its C state points into)."""
    for k in ("u", "v", "x", "F", "Fp"):
        np.copyto(getattr(run, k), st[k])
    for k in ("H", "flag", "fail"):
        np.copyto(getattr(run, k), st[k])
    run.step_count = st["step"]


def test_trilinear_dx2p5_euler_cromer_bit_exact(gpu, oracle):
    case = beam.case_trilinear_dx2p5(integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=150,
                                     appliedDisplacement=-1e-4)
    print(f"case: dx 2.5 mm trilinear EC, {case.nNodes} nodes {case.nBonds} bonds", flush=True)
    assert case.nNodes == 160000 and case.nBonds > 8_800_000
    # the device's order for the static reaction (under Euler-Cromer it changes nothing else)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        assert sol.layout().startswith("stencil")
        for blk in range(3):
            sol.step(50)
            run.step(50)
            print(f"  {run.step_count} steps", flush=True)
        assert_state(sol, run)
        h = sol.history()
        assert same(h["reaction"], run.reactions)
        assert same(h["reaction_static"], run.records["reaction_static"])
        obs = sol.observables()
        assert obs["n_softening"] == int(run.flag.sum()) > 0


def test_trilinear_dx2p5_adr_bit_exact(gpu, oracle):
    case = beam.case_trilinear_dx2p5()
    print(f"case: dx 2.5 mm trilinear ADR, {case.nBonds} bonds", flush=True)
    run = oracle.OracleRun(case, device_order=True)
    with Solver(case, _lib.FP64) as sol:
        sol.step(30)
        run.step(30)
        assert_state(sol, run)
        h = sol.history()
        for k in ("reaction_static", "cn", "ratio", "disp"):
            assert same(h[k], run.records[k]), k


def test_exponential_dx1p25_one_step_from_softening_state(gpu, oracle):
    """From a device FP64 state in the load history once softening has started (checked every 1000
    ADR iterations, at most 12000), one step of the device against the oracle's step from the same
    state."""
    case = beam.case_exponential_dx1p25()
    print(f"case: dx 1.25 mm exponential, {case.nNodes} nodes {case.nBonds} bonds", flush=True)
    assert case.nNodes == 1_280_000 and case.nBonds > 74_000_000
    with Solver(case, _lib.FP64) as sol:
        for k in range(12):
            sol.step(1000)
            if sol.observables()["n_softening"] > 0:
                break
        st = sol.checkpoint()
        print(f"  state at step {st['step']}: {int(st['flag'].sum())} softening bonds", flush=True)
        assert int(st["flag"].sum()) > 0, "no softening yet"
        run = oracle.OracleRun(case, device_order=True)
        load_oracle_state(run, st)
        sol.step(1)
        run.step(1)
        print(f"  {int(run.flag.sum())} softening bonds", flush=True)
        assert_state(sol, run, exact=False, rtol=1e-12)
        R = sol.history()["reaction_static"][-1]
        assert R == pytest.approx(run.records["reaction_static"][-1], rel=1e-12)


def _solver_layout(case, precision, layout):
    old = os.environ.get("BBPD_LAYOUT")
    os.environ["BBPD_LAYOUT"] = layout
    try:
        return Solver(case, precision)
    finally:
        if old is None:
            os.environ.pop("BBPD_LAYOUT", None)
        else:
            os.environ["BBPD_LAYOUT"] = old


def test_exponential_dx1p25_stencil_equals_records(gpu):
    case = beam.case_exponential_dx1p25(integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=60,
                                        appliedDisplacement=-6e-5)
    print(f"case: dx 1.25 mm exponential EC, stencil vs records, {case.nBonds} bonds", flush=True)
    a = _solver_layout(case, _lib.FP64, "stencil")
    b = _solver_layout(case, _lib.FP64, "records")
    try:
        assert a.layout().startswith("stencil") and b.layout().startswith("records")
        a.step(60)
        b.step(60)
        for x, y in zip(a.nodal(), b.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
        assert a.observables()["n_softening"] > 0
        ha, hb = a.history(), b.history()
        for k in ("reaction", "reaction_static", "deflection", "cmod"):
            assert same(ha[k], hb[k]), k
    finally:
        a.close()
        b.close()


def size_effect_family():
    cases = []
    for depth in (50, 100, 200, 400):
        base = beam.case_size_effect(depth, 0)
        for seed in range(16):
            c = copy.copy(base)
            if seed:
                U = beam.splitmix_uniform(seed, np.arange(base.nBonds, dtype=np.uint64))
                c.s0 = 1.0e-4 * (1 + 0.01 * (2 * U - 1))
                c.name = f"beam_d{depth}_seed{seed}"
            cases.append(c)
    return cases


def test_size_effect_batch_regions(gpu, oracle):
    cases = size_effect_family()
    steps = 300
    print(f"case: size-effect batch, {len(cases)} instances, {sum(c.nBonds for c in cases)} bonds", flush=True)
    slots = sum((c.nNodes + 255) // 256 * 256 for c in cases)
    with Batch(cases[0], _lib.FP64, max_regions=len(cases), max_slots=slots, history_rows=4096) as b:
        regs = [b.admit(c) for c in cases]
        b.run_graph(50, steps // 50)
        b.sync()
        hist = {i: b.history(regs[i]) for i in (0, 16, 32, 48, 5)}
        nodal = {i: b.nodal(regs[i]) for i in (0, 16, 32, 48, 5)}
        bonds = {i: b.bonds(regs[i]) for i in (0, 16, 32, 48, 5)}
    for i in (0, 16, 32, 48, 5):
        with Solver(cases[i], _lib.FP64) as s:
            assert s.layout() == "stencil"
            s.step(steps)
            for x, y in zip(s.nodal(), nodal[i]):
                assert same(x, y), cases[i].name
            for x, y in zip(s.bonds(), bonds[i]):
                assert same(x, y), cases[i].name
            h = s.history()
            for k in ("reaction", "reaction_static", "deflection", "cmod", "cn", "ratio", "disp"):
                assert same(h[k], hist[i][k]), (cases[i].name, k)
        print(f"  region {cases[i].name} equals its single run", flush=True)
    run = oracle.OracleRun(cases[0], device_order=True)
    run.step(steps)
    u, v, F, x = nodal[0]
    for name, a, ref in (("u", u, run.u), ("v", v, run.v), ("F", F, run.F), ("x", x, run.x)):
        assert same(a, ref), name
    assert same(bonds[0][0], run.H) and np.array_equal(bonds[0][1], run.flag)
    for k in ("reaction_static", "cn", "disp"):
        assert same(hist[0][k], run.records[k]), k


def _free_run_loads(case, precision, chunk=2000, max_steps=120_000, stop=None):
    """Converged loads (static reaction at the load-increment ends) of a free run in graph chunks,
    until `stop(loads)` or max_steps."""
    loads = []
    with Solver(case, precision) as sol:
        step = 0
        while step < max_steps and not (stop and stop(loads)):
            sol.run_graph(chunk, 1)
            sol.sync()
            step += chunk
            h = sol.history()
            loads += list(h["reaction_static"][h["inc_end"]])
        return np.array(loads), step


def test_mixed_peak_load_within_half_percent(gpu):
    """The FP32-mixed form free-running against FP64 on the dx = 2.5 mm notched trilinear beam
    (8.9 M bonds) under the equilibrium staircase (5 um per increment, equilibrium ratio 1e-4, 50 to
    6000 iterations): the peak of the converged load curve agrees within 0.5%, and the curve up to the
    peak within 1% of the peak (the mixed form's stated gate)."""
    case = beam.case_trilinear_dx2p5(schedule=2, iters_per_increment=6000, increment=-5e-6, eq_tol=1e-4,
                                     iters_per_increment_min=50)
    print(f"case: dx2.5 trilinear, equilibrium staircase, {case.nBonds} bonds", flush=True)

    def past_peak(loads):
        if len(loads) < 3:
            return False
        m = int(np.argmax(loads))
        return len(loads) - 1 - m >= 2 and loads[-1] < 0.9 * loads[m]

    l64, n64 = _free_run_loads(case, _lib.FP64, stop=past_peak)
    assert past_peak(l64), f"FP64 run did not pass its peak in {n64} steps: {l64}"
    lm, nm = _free_run_loads(case, _lib.MIXED, stop=lambda l: len(l) >= len(l64))
    n = min(len(l64), len(lm))
    p64, pm = int(np.argmax(l64[:n])), int(np.argmax(lm[:n]))
    err = (lm[pm] - l64[p64]) / l64[p64]
    curve = np.abs(lm[:p64 + 1] - l64[:p64 + 1]).max() / l64[p64]
    print(f"  FP64 peak {l64[p64]:.2f} N at increment {p64 + 1} ({n64} steps); mixed {lm[pm]:.2f} N at "
          f"{pm + 1}: peak error {err:+.2e}, curve to the peak {curve:.2e}", flush=True)
    assert abs(err) <= 5e-3
    assert curve <= 1e-2
