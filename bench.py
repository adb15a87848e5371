"""Benchmark: explicit bond-based peridynamics steps on one MI355X.

Workload: the mixed-mode three-point-bend beam 500x100x50 mm at dx = 1.25 mm (400x40x80 =
1,280,000 nodes, 74.5M bonds, offset notch, exponential softening), ADR with the penetrator
staircase of 5e-6 m per 1000 iterations; one step = stretch -> law -> force -> nodal gather ->
ADR update -> contact.  Both forms run: FP32-mixed (headline `value`) and plain FP64.

Reported (one JSON line on stdout):
  value            bond updates/s of the FP32-mixed form = nBonds * K / seconds (inputs resident)
  fp64             the same for the plain FP64 form
  roofline         dominant kernel (bond kernel, mixed form): algorithmic bytes per launch / average
                   launch time from HIP events on the solver's stream, against 8 TB/s
  cpu_baseline     the oracle's OpenMP restatement of the reference CPU passes timed on the host on
                   a bounded number of steps of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(case, precision, steps, warmup):
    import torch
    from bb_pd_amd.solver import Solver
    t0 = time.time()
    sol = Solver(case, precision)
    log(f"  setup {time.time() - t0:.1f}s, info {sol.info()}")
    sol.step(warmup)
    sol.sync()
    sol.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.step(steps)
    sol.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    avg_ms, launches = sol.kernel_time()
    sol.set_timing(False)
    info = sol.info()
    obs = sol.observables()
    sol.close()
    bytes_per_launch = info["bond_kernel_bytes"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return dict(seconds=dt, ms_per_step=dt * 1e3 / steps, bond_updates_per_s=case.nBonds * steps / dt,
                kernel_ms=avg_ms, launches=launches, bytes_per_launch=bytes_per_launch, achieved_gbs=achieved,
                frac=achieved / HBM_PEAK_GBS, obs=obs, info=info)


def cpu_baseline(case, target_s=20.0):
    """The oracle's OpenMP passes (reference layout, nBonds intermediates) on the same beam."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    run = O.OracleRun(case, nodal_threads=threads)
    run.step_omp(1, threads)                     # first touch
    n, t = 0, 0.0
    t0 = time.perf_counter()
    while t < target_s and n < 50:
        run.step_omp(1, threads)
        n += 1
        t = time.perf_counter() - t0
    return dict(value=case.nBonds * n / t, unit="bond updates/s", cores=threads, kind="port",
                sample=f"{n} steps of the same 74.5M-bond beam (Euler-Cromer order, exponential law, elastic "
                       f"state), oracle/bbpd_oracle.c or_step_omp, {threads} OpenMP threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--form", choices=["both", "fp64", "mixed"], default="both",
                    help="which precision forms to run (profiling runs one)")
    ap.add_argument("--small", action="store_true", help="dx = 2.5 mm notched trilinear beam instead")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="only copy a device buffer of this many GiB 5 times (scales HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy_gib > 0:
        import torch
        n = int(args.calibrate_copy_gib * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one GPU per process: --gpus must be 1 "
                                   "(the multi-GPU slab decomposition is a design, not built)"}))
        sys.exit(2)
    from bb_pd_amd import _lib, beam
    t0 = time.time()
    case = beam.case_trilinear_dx2p5() if args.small else beam.case_exponential_dx1p25()
    log(f"built {case.name}: {case.nNodes} nodes {case.nBonds} bonds in {time.time() - t0:.1f}s")
    fp64 = mixed = None
    if args.form in ("both", "fp64"):
        fp64 = run_form(case, _lib.FP64, args.steps, args.warmup)
        log(f"fp64: {fp64['ms_per_step']:.3f} ms/step, kernel {fp64['kernel_ms']:.3f} ms, "
            f"{fp64['achieved_gbs']:.0f} GB/s")
    if args.form in ("both", "mixed"):
        mixed = run_form(case, _lib.MIXED, args.steps, args.warmup)
        log(f"mixed: {mixed['ms_per_step']:.3f} ms/step, kernel {mixed['kernel_ms']:.3f} ms, "
            f"{mixed['achieved_gbs']:.0f} GB/s")
    if args.form != "both":
        print(json.dumps({"form": args.form, "result": {k: v for k, v in (fp64 or mixed).items()
                                                         if k not in ("obs", "info")}}), flush=True)
        return
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case)
        log(f"cpu baseline: {cpu['value']:.3e} bond updates/s on {cpu['cores']} threads")
    head = mixed
    out = {
        "metric": "bond updates/s (FP32-mixed; FP64 beside) and HBM roofline fraction of the bond kernel",
        "value": head["bond_updates_per_s"],
        "unit": "bond updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bond arithmetic, f64 state",
        "data": "synthetic beam lattice generated in-process",
        "config": {
            "workload": case.name,
            "nodes": case.nNodes,
            "bonds": case.nBonds,
            "law": "exponential (k=25, alpha=0.25)" if case.law == _lib.LAW_EXPONENTIAL else "trilinear",
            "integrator": "ADR",
            "state": f"steps {args.warmup + 1}..{args.warmup + args.steps} of the load history (elastic)",
        },
        "fp64": {"value": fp64["bond_updates_per_s"], "ms_per_step": fp64["ms_per_step"],
                 "kernel_ms": fp64["kernel_ms"], "bytes_per_launch": fp64["bytes_per_launch"],
                 "achieved_gbs": fp64["achieved_gbs"], "frac": fp64["frac"]},
        "mixed_over_fp64": head["bond_updates_per_s"] / fp64["bond_updates_per_s"],
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["frac"], "traffic": None,
                     "kernel": "k_bond_mixed",
                     "kernel_ms": head["kernel_ms"], "bytes_per_launch": head["bytes_per_launch"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
