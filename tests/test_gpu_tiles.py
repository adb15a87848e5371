"""The FP32-mixed stencil bond kernel in tiles of 16 x 4 nodes against the same kernel in slices of 64
consecutive slots.  The tiles change which lanes a wavefront takes and how it stages neighbour rows,
not the arithmetic of an entry nor the order of a node's sums: under Euler-Cromer the two give the
same trajectory bit for bit (u, v, x, F, history, flags, reactions).  Under ADR the damping
coefficient is a block sum whose blocks hold other nodes in tiles, so one step from a shared state
agrees to rounding (F itself bit for bit)."""
import os

import numpy as np
import pytest

from bb_pd_amd import _lib, beam
from bb_pd_amd.solver import Solver

pytestmark = pytest.mark.gpu

TILED = "stencil, 16x4 tiles"


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.int64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def solver(case, precision=_lib.MIXED, tiles=True, variant=None):
    keep = {k: os.environ.get(k) for k in ("BBPD_TILES", "BBPD_LAUNCH_VARIANT")}
    try:
        os.environ["BBPD_TILES"] = "1" if tiles else "0"
        if variant is None:
            os.environ.pop("BBPD_LAUNCH_VARIANT", None)
        else:
            os.environ["BBPD_LAUNCH_VARIANT"] = str(variant)
        return Solver(case, precision)
    finally:
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def ec_case(law, nx=32, ny=8, nz=10, steps=300):
    """A notched beam whose box suits the tiles: 32 sites along x (two tiles), 80 lines."""
    params = {
        _lib.LAW_BILINEAR: dict(s0=1e-4, s0_perturb=0.01, sc=2e-3),
        _lib.LAW_TRILINEAR: dict(s0=1e-4, s1=5e-4, sc=3e-3),
        _lib.LAW_EXPONENTIAL: dict(s0=1e-4),
        _lib.LAW_LINEAR: dict(sc=1e-3),
    }[law]
    return beam.build_beam("tiles", nx, ny, nz, 5e-3, law=law, integrator=_lib.EULER_CROMER, schedule=0,
                           nTimeSteps=steps, appliedDisplacement=-1.5e-4, notch=(nx / 2 + 0.5, 3.5), **params)


def assert_same_state(a, b):
    for name, x, y in zip(("u", "v", "F", "x"), a.nodal(), b.nodal()):
        assert same(x, y), name
    for name, x, y in zip(("H", "flag", "fail"), a.bonds(), b.bonds()):
        assert same(x, y), name
    ha, hb = a.history(), b.history()
    for k in ("reaction", "reaction_static", "deflection", "cmod"):
        assert same(ha[k], hb[k]), k


def test_tiles_apply_where_the_box_suits_them(gpu):
    """Tiles are asked for with BBPD_TILES=1 and apply to the mixed form on a box that suits them;
    slices are the default."""
    print("case: layouts of the tile-eligible and other boxes", flush=True)
    c = ec_case(_lib.LAW_BILINEAR, steps=10)
    with solver(c) as s, solver(c, tiles=False) as t, solver(c, _lib.FP64) as f:
        assert s.layout() == TILED
        assert t.layout() == "stencil"
        assert f.layout() == "stencil"          # the FP64 form keeps slices
    keep = os.environ.pop("BBPD_TILES", None)
    try:
        with Solver(c, _lib.MIXED) as d:
            assert d.layout() == "stencil"
    finally:
        if keep is not None:
            os.environ["BBPD_TILES"] = keep
    odd = ec_case(_lib.LAW_BILINEAR, nx=24, steps=10)  # 24 sites along x: not a multiple of 16
    with solver(odd) as s:
        assert s.layout() == "stencil"


@pytest.mark.parametrize("law", [_lib.LAW_BILINEAR, _lib.LAW_TRILINEAR, _lib.LAW_LINEAR, _lib.LAW_EXPONENTIAL])
def test_tiles_equal_slices_euler_cromer(gpu, law):
    case = ec_case(law)
    print(f"case: tiles vs slices, law {law}, {case.nBonds} bonds", flush=True)
    with solver(case) as a, solver(case, tiles=False) as b:
        assert a.layout() == TILED
        a.step(300)
        b.step(300)
        assert_same_state(a, b)
        H, flag, fail = a.bonds()
        print(f"  {int(flag.sum())} flagged, {int((H == 1).sum() + (fail == 0).sum())} broken", flush=True)
        assert flag.sum() > 0 or (fail == 0).sum() > 0


def test_tiles_launch_variants_agree(gpu):
    case = ec_case(_lib.LAW_BILINEAR, steps=200)
    print("case: tiles, six launch variants", flush=True)
    ref = None
    for v in range(_lib.AUTOTUNE_VARIANTS):
        with solver(case, variant=v) as s:
            s.step(200)
            state = list(s.nodal()) + list(s.bonds())
        if ref is None:
            ref = state
        else:
            for x, y in zip(ref, state):
                assert same(x, y), v


def test_tiles_checkpoint_restore(gpu):
    """Restoring a damaged state rebuilds the tiles' history words: the restored run continues as the
    uninterrupted one, and a state taken from slices continues in tiles as it does in slices."""
    case = ec_case(_lib.LAW_BILINEAR, steps=300)
    print("case: tiles checkpoint / restore", flush=True)
    with solver(case) as a, solver(case) as b, solver(case) as c, solver(case, tiles=False) as d:
        a.step(300)
        b.step(150)
        st = b.checkpoint()
        assert st["flag"].sum() > 0
        c.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
        c.step(150)
        for x, y in zip(a.nodal(), c.nodal()):
            assert same(x, y)
        for x, y in zip(a.bonds(), c.bonds()):
            assert same(x, y)
        d.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
        d.step(150)
        for x, y in zip(c.nodal(), d.nodal()):
            assert same(x, y)
        for x, y in zip(c.bonds(), d.bonds()):
            assert same(x, y)


def test_tiles_adr_one_step(gpu):
    """ADR: from one loaded state, a step in tiles and a step in slices give F, H and flags bit for bit
    and u, v within rounding of the damping coefficient's block sums."""
    case = beam.build_beam("tiles_adr", 32, 8, 10, 5e-3, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01,
                           sc=2e-3, integrator=_lib.ADR, iters_per_increment=200, increment=-5e-6)
    print(f"case: tiles ADR, {case.nBonds} bonds", flush=True)
    with solver(case) as a, solver(case, tiles=False) as b:
        a.step(900)
        st = a.checkpoint()
        for s in (a, b):
            s.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
        a.history()
        b.history()
        a.step(1)
        b.step(1)
        ua, va, Fa, _ = a.nodal()
        ub, vb, Fb, _ = b.nodal()
        assert same(Fa, Fb)
        for x, y in zip(a.bonds(), b.bonds()):
            assert same(x, y)
        for x, y in ((ua, ub), (va, vb)):
            assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()
        ca, cb = a.history()["cn"][-1], b.history()["cn"][-1]
        assert ca == pytest.approx(cb, rel=1e-12)


def test_tiles_equal_slices_full_beam(gpu):
    """The 1.28M-node offset-notch exponential beam (400 x 40 x 80 sites, 74.5 M bonds), Euler-Cromer into
    softening: tiles and slices bit for bit."""
    case = beam.case_exponential_dx1p25(integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=60,
                                        appliedDisplacement=-6e-5)
    print(f"case: dx 1.25 mm exponential EC, tiles vs slices, {case.nBonds} bonds", flush=True)
    with solver(case) as a, solver(case, tiles=False) as b:
        assert a.layout() == TILED
        a.step(60)
        b.step(60)
        assert_same_state(a, b)
        assert a.observables()["n_softening"] > 0
