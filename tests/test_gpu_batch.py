"""Batches: problems of different sizes packed into one launch of each kernel, admitted and retired
while others run.  Each region's trajectory equals a single-problem run of the same case bit for bit
(the single runs use the stencil layout, as batches do: same slot order, same block partition of the
ADR damping sums)."""
import os

import numpy as np
import pytest

from bb_pd_amd import _lib, beam
from bb_pd_amd.batch import Batch, BatchFull
from bb_pd_amd.solver import Solver

pytestmark = pytest.mark.gpu

DX = 5e-3


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.int64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def ec(nx, ny, nz, steps=300):
    return beam.build_beam("t", nx, ny, nz, DX, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01, sc=2e-3,
                           integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=steps, appliedDisplacement=-1.5e-4,
                           notch=(nx / 2 + 0.5, 3.5))


def adr(nx, ny, nz, ipi=200, **kw):
    return beam.build_beam("t", nx, ny, nz, DX, law=_lib.LAW_BILINEAR, s0=1e-4, s0_perturb=0.01, sc=2e-3,
                           integrator=_lib.ADR, iters_per_increment=ipi, increment=-2e-6, **kw)


def single(case, precision, steps):
    old = os.environ.get("BBPD_LAYOUT")
    os.environ["BBPD_LAYOUT"] = "stencil"
    try:
        s = Solver(case, precision)
    finally:
        if old is None:
            os.environ.pop("BBPD_LAYOUT")
        else:
            os.environ["BBPD_LAYOUT"] = old
    assert s.layout() == "stencil"
    s.step(steps)
    out = dict(nodal=s.nodal(), bonds=s.bonds(), hist=s.history(), obs=s.observables())
    s.close()
    return out


def assert_region_equals(b, r, ref):
    for x, y in zip(b.nodal(r), ref["nodal"]):
        assert same(x, y)
    for x, y in zip(b.bonds(r), ref["bonds"]):
        assert same(x, y)
    h = b.history(r)
    for k in ("reaction", "reaction_static", "deflection", "cmod"):
        assert same(h[k], ref["hist"][k]), k
    o = b.observables(r)
    for k in ("step", "n_softening", "n_broken", "n_failed"):
        assert o[k] == ref["obs"][k], k


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
def test_batch_regions_equal_single_runs(gpu, precision):
    """Three beams of different sizes; one admitted 60 steps before the others."""
    A, B, C = ec(24, 6, 10), ec(30, 6, 12), ec(20, 5, 10)
    print("case: EC batch of 3", flush=True)
    with Batch(A, precision, max_regions=4, max_slots=40000) as b:
        ra = b.admit(A)
        b.step(60)
        rb = b.admit(B)
        rc = b.admit(C)
        b.step(140)
        assert_region_equals(b, ra, single(A, precision, 200))
        assert_region_equals(b, rb, single(B, precision, 140))
        assert_region_equals(b, rc, single(C, precision, 140))
        assert b.observables(ra)["n_softening"] > 0


def test_batch_adr_equals_single_runs(gpu):
    """ADR: each region's damping coefficient comes from its own nodes only."""
    A = adr(25, 6, 10, supports_x=(3 * DX, 23 * DX), penetrator_x=13 * DX)
    B = adr(33, 6, 12, supports_x=(3 * DX, 31 * DX), penetrator_x=17 * DX)
    print("case: ADR batch of 2", flush=True)
    with Batch(A, _lib.FP64, max_regions=2, max_slots=20000) as b:
        ra, rb = b.admit(A), b.admit(B)
        b.step(450)
        assert_region_equals(b, ra, single(A, _lib.FP64, 450))
        assert_region_equals(b, rb, single(B, _lib.FP64, 450))
        assert b.observables(ra)["cn"] != b.observables(rb)["cn"]


def test_batch_launch_variants_agree(gpu):
    """Every launch variant of the stencil bond kernel (block orders: XCD remap, launch order, XCD
    stripes over the batch's block list) gives the same trajectories bit for bit, in both forms."""
    A = adr(25, 6, 10, supports_x=(3 * DX, 23 * DX), penetrator_x=13 * DX)
    B = adr(33, 6, 12, supports_x=(3 * DX, 31 * DX), penetrator_x=17 * DX)
    nv = _lib.AUTOTUNE_VARIANTS
    for prec in (_lib.FP64, _lib.MIXED):
        ref = None
        for v in range(nv):
            old = os.environ.get("BBPD_LAUNCH_VARIANT")
            os.environ["BBPD_LAUNCH_VARIANT"] = str(v)
            try:
                b = Batch(A, prec, max_regions=2, max_slots=20000)
            finally:
                if old is None:
                    os.environ.pop("BBPD_LAUNCH_VARIANT")
                else:
                    os.environ["BBPD_LAUNCH_VARIANT"] = old
            with b:
                ra, rb = b.admit(A), b.admit(B)
                b.step(300)
                out = [b.nodal(ra), b.nodal(rb), b.bonds(ra), b.bonds(rb)]
            if ref is None:
                ref = out
                continue
            for x, y in zip(ref, out):
                for p, q in zip(x, y):
                    assert same(p, q), (prec, v)


def test_batch_retire_readmit_and_graph(gpu):
    """Retiring frees the slots for the next problem; survivors are undisturbed; a captured graph
    is rebuilt when the membership changes and equals eager stepping."""
    A, B, C = ec(24, 6, 10), ec(20, 5, 10), ec(22, 6, 10)
    with Batch(A, _lib.FP64, max_regions=2, max_slots=5000) as b:
        ra = b.admit(A)
        rb = b.admit(B)
        with pytest.raises(BatchFull):
            b.admit(C)
        b.run_graph(25, 2)
        b.retire(rb)
        rc = b.admit(C)
        assert rc == rb
        b.run_graph(25, 4)
        b.sync()
        assert_region_equals(b, ra, single(A, _lib.FP64, 150))
        assert_region_equals(b, rc, single(C, _lib.FP64, 100))


def test_batch_stop_criterion(gpu):
    """A region is marked finished at the first increment end where the load point has travelled
    the stop displacement; the others keep running."""
    A = adr(25, 6, 10, ipi=100, supports_x=(3 * DX, 23 * DX), penetrator_x=13 * DX)
    with Batch(A, _lib.MIXED, max_regions=2, max_slots=10000) as b:
        ra = b.admit(A, stop_displacement=3 * 2e-6)
        rb = b.admit(A)
        b.step(250)
        st = b.status()
        assert st[ra]["active"] and not st[ra]["finished"]
        b.step(50)
        st = b.status()
        assert st[ra]["finished"] and not st[rb]["finished"]
        assert st[ra]["step"] == st[rb]["step"] == 300
        assert st[ra]["peak"] > 0


def test_batch_rejects_mismatched_problems(gpu):
    A = ec(24, 6, 10)
    with Batch(A, _lib.FP64, max_regions=2, max_slots=5000) as b:
        other = beam.build_beam("t", 24, 6, 10, DX, law=_lib.LAW_TRILINEAR, s0=1e-4, s1=5e-4, sc=3e-3,
                                integrator=_lib.EULER_CROMER, schedule=0, nTimeSteps=300)
        with pytest.raises(_lib.BBPDError, match="law"):
            b.admit(other)
        shuffled = ec(24, 6, 10)
        order = np.random.default_rng(2).permutation(shuffled.nBonds)
        shuffled.BONDLIST = np.asfortranarray(shuffled.BONDLIST[order])
        for name in ("L0", "vcf", "c", "BONDTYPE", "fail", "s0"):
            setattr(shuffled, name, getattr(shuffled, name)[order])
        with pytest.raises(_lib.BBPDError, match="stencil layout"):
            b.admit(shuffled)


def test_prepared_admission(gpu):
    """bbpd_batch_prepare on another host thread while the batch steps, then bbpd_batch_admit_prepared:
    the region runs as a directly admitted one (bit for bit against the single run); a prepared
    admission refused for lack of room stays valid and is admitted once a region retires."""
    import threading
    first, second = ec(24, 6, 10), ec(48, 6, 10)   # 6 and 12 blocks of 256 slots: not both in 4096
    print("case: prepared admission, 2 EC beams", flush=True)
    ref = single(second, _lib.MIXED, 150)
    with Batch(first, _lib.MIXED, max_regions=4, max_slots=4096) as b:
        r0 = b.admit(first)
        out = {}
        t = threading.Thread(target=lambda: out.setdefault("adm", b.prepare(second)))
        t.start()
        b.step(100)                      # the batch steps while the other thread prepares
        t.join()
        b.sync()
        adm = out["adm"]
        with pytest.raises(BatchFull):
            b.admit(adm)
        b.retire(r0)
        r1 = b.admit(adm)
        b.step(150)
        assert_region_equals(b, r1, ref)
