"""Continuous-batching service for fracture solves on one GPU.

Requests arrive over time (Poisson, or all at once for a closed-loop capacity run).  Each is a
three-point-bend beam solve that runs until its load falls below a fraction of its peak after the
peak, or its load point has travelled a set distance.  The service keeps one resident batch
(`Batch`): an arriving request is admitted as soon as a region and enough node slots are free
(admission by device-memory footprint: the batch's slot capacity is sized from an HBM budget);
all admitted requests step together in one launch per kernel, each at its own step count; finished
ones are retired and their slots reused.  Nothing of a request's state leaves the device between
its admission and its retirement.

The reference has no service; its only multi-specimen workflow runs specimens one after another
(solvers/scripts/randomfield/dynamicsolverdisplacementcontrolledRF.m:20-114), stopping each on a
CMOD criterion, the analogue of the stop criterion here.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, beam
from .batch import Batch, BatchFull, BatchTooSmall


@dataclass
class Request:
    rid: int
    params: dict | None = None   # the drawn request; its inputs are built once it has arrived
    arrival: float = 0.0         # s after the start of the run
    case: beam.BeamCase | None = None
    built: float = math.nan      # inputs generated (host), s after the start
    admitted: float = math.nan
    finished: float = math.nan
    region: int = -1
    steps: int = 0
    peak: float = 0.0
    rejected: str = ""
    name: str = ""
    nBonds: int = 0
    history: dict = field(default_factory=dict)

    @property
    def time_to_solution(self) -> float:
        return self.finished - self.arrival


def request_params(rng: np.random.Generator) -> dict:
    """One request of the serving workload: depth in {50, 100, 200, 400} mm, notch ratio in
    {0, 0.2, 0.4} (a midspan notch of that fraction of the depth), E ~ U[30, 45] GPa,
    s0 ~ U[0.8, 1.2]e-4 (scalar)."""
    return dict(depth=int(rng.choice([50, 100, 200, 400])), ratio=float(rng.choice([0.0, 0.2, 0.4])),
                E=float(rng.uniform(30e9, 45e9)), s0=float(rng.uniform(0.8e-4, 1.2e-4)))


def build_request(params: dict, ipi: int, increment_um: float) -> beam.BeamCase:
    """The request's beam: bilinear law with sc = 2e-3, dx = 5 mm, length 5 x depth, width 50 mm, ADR
    on the penetrator staircase of `increment_um` per `ipi` iterations."""
    depth, ratio, E, s0 = params["depth"], params["ratio"], params["E"], params["s0"]
    nz = depth // 5
    nx = 5 * nz
    notch = (nx / 2 + 0.5, ratio * nz + 0.5) if ratio > 0 else None
    return beam.build_beam(f"req_d{depth}_n{ratio}_E{E / 1e9:.1f}", nx, 10, nz, 5e-3, law=_lib.LAW_BILINEAR,
                           s0=s0, sc=2.0e-3, E=E, notch=notch, integrator=_lib.ADR, iters_per_increment=ipi,
                           increment=-increment_um * 1e-6)


def request_case(rng: np.random.Generator, ipi: int, increment_um: float) -> beam.BeamCase:
    return build_request(request_params(rng), ipi, increment_um)


def slot_bytes(precision: int) -> int:
    """Device bytes a batch holds per node slot (node state, bond mask, 128 history entries)."""
    w = 4 if precision == _lib.MIXED else 8
    node = 2 * 32 + 8 * (3 * (3 if precision == _lib.MIXED else 4) + 2 + 3) + 1 + 16
    return node + 128 * w


def serve(requests: list, precision: int, hbm_budget: float, chunk: int, ipi: int = 200,
          increment_um: float = 25.0, stop_fraction: float = 0.2, stop_displacement: float = 5e-4,
          max_regions: int = 64, keep_history: bool = False, keep_cases: bool = False,
          closed_loop: bool = False, build_workers: int = 8, lookahead: int = 16, log=None) -> dict:
    """Run the service over `requests` (their `arrival` times are honoured unless `closed_loop`,
    in which case all are queued at t = 0).  Once a request has arrived, a pool of `build_workers`
    host threads generates its inputs (if it carries only its parameters) and checks and lays them
    out for the batch (`Batch.prepare`), for at most `lookahead` queued requests at a time: that work
    overlaps the device's steps and the other requests', and counts in the request's time to
    solution.  Requests are admitted first come, first served, once prepared and once the batch has
    a region and the node slots; admission then only uploads.  A request's host arrays are dropped
    once it is prepared (unless `keep_cases`): its state lives on the device until it is retired.  A
    request that needs more node slots than the whole batch holds is rejected.  Returns per-request
    times and aggregate figures."""
    import concurrent.futures as cf
    import torch
    slots = int(hbm_budget // slot_bytes(precision))
    pending = sorted(requests, key=lambda r: r.arrival)
    queue: list = []
    active: dict = {}
    done: list = []
    rejected: list = []
    futures: dict = {}
    ready: dict = {}  # rid -> prepared admission
    admit_s: list = []
    build_s: list = []
    prep_s: list = []
    bond_steps = 0
    # wall time launching device chunks (a graph is captured again when the membership changed), waiting
    # for them, admitting, and waiting for a request's inputs
    t_launch = t_sync = t_wait = 0.0
    # the batch's shared scalars (law, integrator, time step, spacing, horizon) from a request of
    # the same workload, built before the clock starts
    shape = pending[0].case if pending[0].case is not None else build_request(pending[0].params, ipi, increment_um)
    pool = cf.ThreadPoolExecutor(max_workers=max(1, build_workers))
    with Batch(shape, precision, max_regions=max_regions, max_slots=slots, history_rows=1 << 17) as b:
        del shape
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        def prepare(r):
            """On a worker thread: the request's inputs, then its admission."""
            t = time.perf_counter()
            case = r.case if r.case is not None else build_request(r.params, ipi, increment_um)
            tb = time.perf_counter() - t
            try:
                adm, err = b.prepare(case), ""
            except BatchTooSmall as e:
                adm, err = None, str(e)
            return case, adm, err, tb, time.perf_counter() - t - tb

        def arrivals():
            t = now()
            while pending and (closed_loop or pending[0].arrival <= t):
                queue.append(pending.pop(0))
            # the queued requests' preparation, oldest first, at most `lookahead` at a time
            for r in queue[:lookahead]:
                if r.rid not in futures and r.rid not in ready:
                    futures[r.rid] = pool.submit(prepare, r)

        def collect(r, wait):
            """Whether the request is prepared (waiting for it if `wait`)."""
            if r.rid in ready:
                return True
            f = futures.get(r.rid)
            if f is None or (not wait and not f.done()):
                return False
            case, adm, err, tb, tp = f.result()
            del futures[r.rid]
            r.built = now()
            r.name, r.nBonds = case.name, case.nBonds
            r.case = case if keep_cases else None
            build_s.append(tb)
            prep_s.append(tp)
            ready[r.rid] = (adm, err)
            return True

        try:
            while pending or queue or active:
                arrivals()
                # admission, first come first served: the head waits for its preparation only when the
                # device has nothing to run; a request that does not fit waits for slots
                while queue:
                    r = queue[0]
                    tw = time.perf_counter()
                    ok = collect(r, wait=not active)
                    t_wait += time.perf_counter() - tw
                    if not ok:
                        break
                    adm, err = ready[r.rid]
                    if adm is None:
                        r.rejected = err
                        del ready[r.rid]
                        rejected.append(queue.pop(0))
                        continue
                    ta = time.perf_counter()
                    try:
                        r.region = b.admit(adm, stop_fraction, stop_displacement)
                    except BatchFull:
                        break
                    admit_s.append(time.perf_counter() - ta)
                    del ready[r.rid]
                    r.admitted = now()
                    active[r.region] = r
                    queue.pop(0)
                    arrivals()
                if not active:
                    if pending and not queue:
                        time.sleep(max(0.0, pending[0].arrival - now()))
                    continue
                td = time.perf_counter()
                b.run_graph(chunk, 1)
                t_launch += time.perf_counter() - td
                arrivals()
                td = time.perf_counter()
                b.sync()
                t_sync += time.perf_counter() - td
                bond_steps += chunk * sum(r.nBonds for r in active.values())
                st = b.status()
                for reg, r in list(active.items()):
                    if st[reg]["finished"]:
                        r.finished = now()
                        r.steps = st[reg]["step"]
                        r.peak = st[reg]["peak"]
                        if keep_history:
                            r.history = b.history(reg)
                        b.retire(reg)
                        del active[reg]
                        done.append(r)
                        if log:
                            log(f"    request {r.rid} ({r.name}) done: {r.steps} steps, "
                                f"{r.time_to_solution:.2f} s")
        finally:
            pool.shutdown(wait=True, cancel_futures=True)
            for adm, _ in ready.values():
                if adm is not None:
                    adm.free()
        wall = now()
        torch.cuda.synchronize()
    tts = np.array([r.time_to_solution for r in done])
    wait_s = np.array([r.admitted - r.arrival for r in done])
    return dict(requests=len(done), rejected=len(rejected), wall_s=wall, requests_per_s=len(done) / wall,
                bond_updates_per_s=bond_steps / wall, p50_s=float(np.percentile(tts, 50)),
                p90_s=float(np.percentile(tts, 90)), p99_s=float(np.percentile(tts, 99)), max_s=float(tts.max()),
                build_s_mean=float(np.mean(build_s)) if build_s else None,
                prepare_s_mean=float(np.mean(prep_s)) if prep_s else None,
                admit_s_mean=float(np.mean(admit_s)) if admit_s else None,
                admission_wait_s_mean=float(wait_s.mean()), slots=slots,
                wall_split=dict(launch_s=t_launch, device_wait_s=t_sync, admission_s=float(np.sum(admit_s)),
                                input_wait_s=t_wait),
                done=done)
